/*
 * fr_batch.cpp — batched and served planning on one GPU.
 *
 * Many FastRRT plans of mixed grid size advance together: one loop
 * iteration of every running plan is the fixed kernel sequence of
 * fr_launch_iteration() over the engine's flattened (plan, tile) table,
 * captured once as a hipGraph and replayed.  Requests are admitted as slots
 * and the cell budget (HBM) free up, so plans that arrive while others run
 * join at the next iteration (continuous batching).  A plan that reaches its
 * goal is finalised (path_optimize rounds, path extraction) by a second
 * captured graph on a second stream, over a small list of plans, while the
 * others keep searching; it leaves when that graph's event has completed.
 * Each plan draws its random numbers from Philox keyed by its own (seed,
 * stream), so a plan's result does not depend on what else is in the batch.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/fastrrt.h"
#include "fr_engine.h"
#include "fr_math.h"

namespace {

/* batch result status beyond FrStatus: goal reached, but no feasible node
 * near the goal to extract a path from (get_planned_path returns nothing) */
const int FR_RESULT_NO_PATH = 4;

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Pending {
    fastrrt_plan_request req;
    int ticket;
    double t_submit;
};

struct Finished {
    fastrrt_plan_result res;
    std::vector<float> path; /* (x, z, heading) triples */
};

struct Slot {
    bool busy = false;
    bool finalizing = false;
    int ticket = -1;
    long admit_seq = 0; /* first control-block snapshot that shows this plan */
    Pending p;
    std::unique_ptr<FrSlab> slab;
    long cells = 0;
};

/* Device memory as the batch measured it (hipMemGetInfo) and what it holds. */
struct Hbm {
    size_t total = 0, free_at_create = 0, free_last = 0, reserve = 0, engine = 0;
    long requested = 0, budget = 0, stalls = 0;
};

/* Memory left to the rest of the process by default: 2 GiB, at most a 16th
 * of the device. */
const size_t kHbmReserve = 2ull << 30;

/* The cell budget from the measured free memory: the engine's fixed part
 * first, then cells at the bytes a cell costs in a plan's planes (a 256 x
 * 256 grid with a 22 x 40 footprint: its distance planes and path buffers
 * included) plus the engine's per-cell queues; at most `max_cells`. */
Hbm measure_hbm(int max_plans, long max_cells)
{
    if (max_plans <= 0 || max_cells <= 0)
        throw std::runtime_error("invalid batch capacity");
    Hbm h;
    FR_CHECK(hipMemGetInfo(&h.free_at_create, &h.total));
    h.free_last = h.free_at_create;
    h.reserve = std::min<size_t>(kHbmReserve, h.total / 16);
    h.requested = max_cells;
    const double slab_pc = (double)FrSlab::bytes_for(256, 256, 11, 20, 256 * 256 + 1, false) / 65536.0;
    const size_t e1 = FrEngine::bytes_for(max_plans + 1, 2);
    const double eng_pc = (double)(FrEngine::bytes_for(max_plans + 1, (1L << 20) + 2) - e1) / (double)(1L << 20);
    const double avail = (double)h.free_at_create - (double)h.reserve - (double)e1;
    const long fit = avail > 0 ? (long)(avail / (slab_pc + eng_pc)) : 0;
    h.budget = std::min(max_cells, fit);
    if (h.budget < 1)
        throw std::runtime_error("device memory: " + std::to_string(h.free_at_create) + " bytes free, " +
                                 std::to_string(h.reserve) + " reserved: no room for a batch");
    h.engine = FrEngine::bytes_for(max_plans + 1, h.budget + 1);
    return h;
}

} // namespace

static_assert(sizeof(FrCtl) <= FR_PACK_PATH_OFF, "control block must fit before the path of a pinned record");

class FrBatch {
  public:
    /* The engine has one slot more than the batch: a plan that never runs,
     * whose control block stays IDLE, pads the finalisation list. */
    /* The cell budget is max_cells or what the measured free device memory
     * holds, whichever is smaller (measure_hbm). */
    FrBatch(int max_plans, long max_cells)
        : hbm_(measure_hbm(max_plans, max_cells)), eng_(max_plans + 1, hbm_.budget + 1), max_cells_(hbm_.budget),
          slots_(max_plans), dummy_(max_plans), fin_max_(std::min(max_plans, 64))
    {
        try {
            allocate(max_plans);
        } catch (...) {
            /* a half-built batch runs no destructor */
            (void)hipStreamSynchronize(eng_.stream());
            release();
            throw;
        }
        t0_ = now_ms();
    }

    ~FrBatch()
    {
        (void)hipStreamSynchronize(eng_.stream());
        if (st2_)
            (void)hipStreamSynchronize(st2_);
        release();
    }

  private:
    void allocate(int max_plans)
    {
        for (int b = 0; b < 2; ++b) {
            /* written by a copy kernel on the stream (fr_launch_copy) */
            FR_CHECK(hipHostMalloc((void **)&h_snap_[b], sizeof(FrCtl) * max_plans,
                                   hipHostMallocMapped | hipHostMallocCoherent));
            FR_CHECK(hipHostGetDevicePointer((void **)&d_snap_[b], h_snap_[b], 0));
            std::memset(h_snap_[b], 0, sizeof(FrCtl) * max_plans);
            FR_CHECK(hipEventCreateWithFlags(&snap_ev_[b], hipEventDisableTiming));
        }
        FR_CHECK(hipHostMalloc((void **)&h_fin_list_, sizeof(int) * 2 * fin_max_,
                               hipHostMallocMapped | hipHostMallocCoherent));
        FR_CHECK(hipHostGetDevicePointer((void **)&d_fin_list_, h_fin_list_, 0));
        /* finished plans' records, written by the device (k_fin_pack) */
        FR_CHECK(hipHostMalloc((void **)&h_pack_, (size_t)fin_max_ * kPackStride, hipHostMallocMapped));
        FR_CHECK(hipHostGetDevicePointer((void **)&d_pack_, h_pack_, 0));
        FR_CHECK(hipMalloc(&d_list_, sizeof(int) * max_plans));
        FR_CHECK(hipMalloc(&fin_list_, sizeof(int) * 2 * fin_max_)); /* slots, then round limits */
        FR_CHECK(hipStreamCreateWithFlags(&st2_, hipStreamNonBlocking));
        FR_CHECK(hipEventCreateWithFlags(&fin_ev_, hipEventDisableTiming));
        eng_.upload_plans(); /* all-zero descriptors, including the padding plan's */
        FR_CHECK(hipStreamSynchronize(eng_.stream()));
    }

    void release()
    {
        drop_graphs();
        if (fin_exec_)
            (void)hipGraphExecDestroy(fin_exec_);
        if (fin_graph_)
            (void)hipGraphDestroy(fin_graph_);
        if (fin_ev_)
            (void)hipEventDestroy(fin_ev_);
        if (st2_)
            (void)hipStreamDestroy(st2_);
        for (int b = 0; b < 2; ++b) {
            if (h_snap_[b])
                (void)hipHostFree(h_snap_[b]);
            if (snap_ev_[b])
                (void)hipEventDestroy(snap_ev_[b]);
        }
        if (h_fin_list_)
            (void)hipHostFree(h_fin_list_);
        if (h_pack_)
            (void)hipHostFree(h_pack_);
        if (d_list_)
            (void)hipFree(d_list_);
        if (fin_list_)
            (void)hipFree(fin_list_);
        fin_exec_ = nullptr;
        fin_graph_ = nullptr;
        fin_ev_ = nullptr;
        st2_ = nullptr;
        h_snap_[0] = h_snap_[1] = nullptr;
        d_snap_[0] = d_snap_[1] = nullptr;
        snap_ev_[0] = snap_ev_[1] = nullptr;
        h_fin_list_ = nullptr;
        d_fin_list_ = nullptr;
        h_pack_ = d_pack_ = nullptr;
        d_list_ = fin_list_ = nullptr;
    }

  public:
    int submit(const fastrrt_plan_request &r)
    {
        const FrFrame *f = static_cast<const FrFrame *>(r.frame);
        if (!f)
            throw std::runtime_error("request without a frame");
        if ((long)f->W * f->H > max_cells_)
            throw std::runtime_error("request grid exceeds the batch cell budget (" + std::to_string(max_cells_) +
                                     " cells)");
        if (r.region_side != 0 && r.region_side < 16)
            throw std::runtime_error("density region side must be 0 (128 px) or at least 16 px");
        Pending p;
        p.req = r;
        p.ticket = next_ticket_++;
        p.t_submit = now_ms() - t0_;
        queue_.push_back(p);
        return p.ticket;
    }

    int in_flight() const { return (int)queue_.size() + running_; }

    /* New launch shapes for the iteration graph: the queued work drains, the
     * graphs are captured again at the next step. */
    void set_launch(int waves, int tiles, int items)
    {
        FR_CHECK(hipStreamSynchronize(eng_.stream()));
        eng_.set_grids(waves, tiles, items);
        drop_graphs();
    }

    void set_hbm_reserve(size_t bytes) { hbm_.reserve = bytes; }

    void set_block(int threads)
    {
        FR_CHECK(hipStreamSynchronize(eng_.stream()));
        eng_.set_block(threads);
        drop_graphs();
    }

    int block() const { return eng_.launch().block; }

    void memory(fastrrt_batch_hbm *o)
    {
        size_t fr = 0, tot = 0;
        FR_CHECK(hipMemGetInfo(&fr, &tot));
        hbm_.free_last = fr;
        std::memset(o, 0, sizeof(*o));
        o->total_bytes = hbm_.total;
        o->free_bytes_at_create = hbm_.free_at_create;
        o->free_bytes_last = hbm_.free_last;
        o->reserve_bytes = hbm_.reserve;
        o->engine_bytes = hbm_.engine;
        o->slab_bytes = slab_bytes_;
        o->requested_cells = hbm_.requested;
        o->cell_budget = hbm_.budget;
        o->cells_in_use = cells_in_use_;
        o->memory_stalls = hbm_.stalls;
        o->running = running_;
        o->queued = (int)queue_.size();
    }

    void get_launch(int *out) const
    {
        const FrLaunch &L = eng_.launch();
        out[0] = L.grid_waves;
        out[1] = L.grid_tiles;
        out[2] = L.grid_items;
    }

    /* admit -> `iterations` loop iterations -> retire.  The host works one
     * step behind the device: after queueing this step's iterations and a
     * copy of the control blocks, it acts on the previous step's copy while
     * the device runs, so the device does not wait for the host between
     * steps.  A decision taken on that copy lands after the iterations
     * already queued (a plan retired by its deadline may have run one more
     * iteration; the deadline margin covers it). */
    int step(int iterations)
    {
        admit();
        if (running_ > 0) {
            if (running_ > finalizing_) {
                for (int k = 0; k < iterations; ++k)
                    launch_iteration();
                const int b = (int)(snap_seq_ & 1);
                fr_launch_copy(d_snap_[b], eng_.launch().ctl, sizeof(FrCtl) * slots_.size(), eng_.stream());
                FR_CHECK(hipGetLastError());
                FR_CHECK(hipEventRecord(snap_ev_[b], eng_.stream()));
                const long prev = snap_seq_ - 1;
                ++snap_seq_;
                if (prev >= 0 && prev >= snap_done_)
                    process_snapshot(prev);
            } else if (snap_seq_ - 1 >= snap_done_) {
                process_snapshot(snap_seq_ - 1); /* the last copy not acted on */
            } else {
                FR_CHECK(hipEventSynchronize(fin_ev_)); /* only finalisation left */
                poll_and_retire(nullptr, -1);
            }
        }
        return in_flight();
    }

    void process_snapshot(long seq)
    {
        const int b = (int)(seq & 1);
        FR_CHECK(hipEventSynchronize(snap_ev_[b]));
        snap_done_ = seq + 1;
        poll_and_retire(h_snap_[b], seq);
    }

    bool poll(fastrrt_plan_result *res, float *path, int cap)
    {
        if (done_.empty())
            return false;
        Finished &f = done_.front();
        *res = f.res;
        const int n = (int)(f.path.size() / 3);
        if (path && cap >= n) {
            std::memcpy(path, f.path.data(), sizeof(float) * f.path.size());
            res->path_len = n;
        } else if (n > 0) {
            res->path_len = -1;
        }
        done_.pop_front();
        return true;
    }

    void drain()
    {
        while (in_flight() > 0)
            step(4);
    }

    double clock_ms() const { return now_ms() - t0_; }

  private:
    static void fill_params(FrPlanParams &p, const fastrrt_plan_request &r, const FrFrame *f)
    {
        p.W = f->W;
        p.H = f->H;
        p.min_dx = r.min_dist_x;
        p.min_dz = r.min_dist_z;
        p.lbx = r.lower_bound_x;
        p.lbz = r.lower_bound_z;
        p.ubx = r.upper_bound_x;
        p.ubz = r.upper_bound_z;
        p.og_x = (float)(f->W / 2);
        p.og_y = (float)(f->H / 2);
        p.v = r.velocity_m_s;
        p.max_path = r.max_path_size;
        p.dist_tol = r.dist_to_goal_tolerance;
        p.herr_tol = r.heading_error_tolerance_rad;
        p.max_steer_rad = (double)r.max_steering_angle_deg * FR_TO_RAD_D;
        p.lr = (double)(r.vehicle_length / 2.0f);
        const float t = fr_tanf((float)p.max_steer_rad);
        p.max_curv_loop = r.max_curvature < 0
                              ? (float)((double)(2.0f * t) /
                                        (0.5 * (double)r.vehicle_length * (double)sqrtf(4.0f + t)))
                              : r.max_curvature;
        p.max_curv_goal = 0.8f; /* fastrrt.cpp:79 */
        p.sx = r.start_x;
        p.sz = r.start_z;
        p.sh = r.start_heading_rad;
        p.gx = r.goal_x;
        p.gz = r.goal_z;
        p.gh = r.goal_heading_rad;
        p.smart = r.smart ? 1 : 0;
        p.max_iter = r.max_iterations > 0 ? r.max_iterations : 0;
        p.seed = r.seed;
        p.stream = r.stream;
        p.region_side = r.region_side > 0 ? r.region_side : 128;
        p.exp_budget = r.expansion_budget > 0 ? r.expansion_budget : 0;
    }

    /* A plan's planes: a pooled slab of its size class, or a new one when
     * the measured free memory holds it beside the reserve (pooled slabs of
     * other classes are freed first if it does not).  Null when it does not
     * fit: the plan waits until a running plan's slab is given back. */
    std::unique_ptr<FrSlab> take_slab(int W, int H, int mx, int mz)
    {
        auto key = std::make_tuple(W, H, mx, mz);
        auto it = pool_.find(key);
        std::unique_ptr<FrSlab> s;
        if (it != pool_.end() && !it->second.empty()) {
            s = std::move(it->second.back());
            it->second.pop_back(); /* reset on the device at admission (k_slab_reset) */
            return s;
        }
        const size_t need = FrSlab::bytes_for(W, H, mx, mz, path_cap(W, H), false);
        if (!hbm_fits(need)) {
            if (!release_pools())
                return nullptr;
            if (!hbm_fits(need))
                return nullptr;
        }
        s.reset(new FrSlab());
        s->allocate(W, H, mx, mz, path_cap(W, H), false, eng_.stream());
        slab_bytes_ += need;
        return s;
    }

    bool hbm_fits(size_t need)
    {
        size_t fr = 0, tot = 0;
        FR_CHECK(hipMemGetInfo(&fr, &tot));
        hbm_.free_last = fr;
        return fr >= need + hbm_.reserve;
    }

    /* Frees every pooled slab (after the queued device work, which may still
     * read a slab retired since); false when the pool was empty. */
    bool release_pools()
    {
        bool any = false;
        for (auto &kv : pool_)
            any = any || !kv.second.empty();
        if (!any)
            return false;
        FR_CHECK(hipStreamSynchronize(eng_.stream()));
        FR_CHECK(hipStreamSynchronize(st2_));
        for (auto &kv : pool_) {
            for (auto &sl : kv.second)
                slab_bytes_ -= std::min(slab_bytes_, FrSlab::bytes_for(sl->W, sl->H, sl->mx, sl->mz, sl->path_cap,
                                                                        sl->with_intrinsic));
            kv.second.clear();
        }
        return true;
    }

    static int path_cap(int W, int H) { return std::min(W * H + 1, 1 << 16); }

    void give_back(std::unique_ptr<FrSlab> s)
    {
        auto key = std::make_tuple(s->W, s->H, s->mx, s->mz);
        /* keep every slab of the batch's size classes: freeing device memory
         * synchronises the device, and the cell budget bounds the pool */
        auto &v = pool_[key];
        if ((long)v.size() * s->W * s->H < max_cells_) {
            v.push_back(std::move(s));
            return;
        }
        slab_bytes_ -= std::min(slab_bytes_, FrSlab::bytes_for(s->W, s->H, s->mx, s->mz, s->path_cap,
                                                                s->with_intrinsic));
    }

    void admit()
    {
        std::vector<int> admitted;
        while (!queue_.empty()) {
            int free_slot = -1;
            for (int i = 0; i < (int)slots_.size(); ++i)
                if (!slots_[i].busy) {
                    free_slot = i;
                    break;
                }
            if (free_slot < 0)
                break;
            const Pending &p = queue_.front();
            const FrFrame *f = static_cast<const FrFrame *>(p.req.frame);
            const long cells = (long)f->W * f->H;
            if (cells_in_use_ + cells > max_cells_)
                break;
            const int mx = fr_trunc_f((float)p.req.min_dist_x / 2.0f);
            const int mz = fr_trunc_f((float)p.req.min_dist_z / 2.0f);
            std::unique_ptr<FrSlab> slab = take_slab(f->W, f->H, mx, mz);
            if (!slab) {
                hbm_.stalls += 1;
                if (running_ == 0 && admitted.empty()) {
                    /* nothing running will give memory back */
                    const int W = f->W, H = f->H;
                    queue_.pop_front();
                    throw std::runtime_error("device memory: a " + std::to_string(W) + " x " + std::to_string(H) +
                                             " plan does not fit in the measured free memory beside the reserve");
                }
                break;
            }
            Slot &s = slots_[free_slot];
            FrPlanParams prm;
            fill_params(prm, p.req, f);
            prm.gc_lazy = 1;
            prm.opt_rounds = p.req.optimize_rounds;
            s.slab = std::move(slab);
            FrFrameView v;
            v.xyz = f->d_xyz;
            v.costs = f->d_costs;
            v.n_classes = f->n_classes;
            eng_.bind(free_slot, *s.slab, prm, v, false);
            s.busy = true;
            s.ticket = p.ticket;
            s.p = p;
            s.cells = cells;
            cells_in_use_ += cells;
            admitted.push_back(free_slot);
            queue_.pop_front();
        }
        if (admitted.empty())
            return;
        /* descriptors in one copy; slabs reset, planes prepared, control
         * blocks and start nodes written on the device */
        eng_.upload_plans();
        eng_.upload_ints(d_list_, admitted.data(), (int)admitted.size());
        fr_launch_slab_reset(eng_.launch(), d_list_, (int)admitted.size(), eng_.stream());
        eng_.setup(admitted, FR_GOAL_LAZY, true);
        for (int sl : admitted)
            slots_[sl].admit_seq = snap_seq_;
        fr_launch_batch_start(eng_.launch(), d_list_, (int)admitted.size(), eng_.stream());
        FR_CHECK(hipGetLastError());
        running_ += (int)admitted.size();
        rebuild_active();
    }

    void rebuild_active()
    {
        std::vector<int> act;
        for (int i = 0; i < (int)slots_.size(); ++i)
            if (slots_[i].busy)
                act.push_back(i);
        eng_.set_active(act);
    }

    /* One loop iteration of every running plan: one replay of the captured
     * iteration graph, timed whole with kernel timing on.  With the split
     * also on, the iteration is captured as three graphs (before k_expand,
     * k_expand, after) and k_expand is timed with events recorded on the
     * stream between their replays (the host then submits three graphs per
     * iteration, which the whole iteration's time includes). */
    void launch_iteration()
    {
        const bool split = fr_prof_on() && fr_prof_split_on();
        if (!graph_[0] || split != graph_timed_)
            capture(split);
        hipStream_t st = eng_.stream();
        fr_prof_begin(FR_PROF_ITERATION, st);
        if (!graph_timed_) {
            FR_CHECK(hipGraphLaunch(exec_[0], st));
        } else {
            FR_CHECK(hipGraphLaunch(exec_[0], st));
            fr_prof_begin(FR_PROF_EXPAND, st);
            FR_CHECK(hipGraphLaunch(exec_[1], st));
            fr_prof_end(FR_PROF_EXPAND, st);
            FR_CHECK(hipGraphLaunch(exec_[2], st));
        }
        fr_prof_end(FR_PROF_ITERATION, st);
    }

    void capture(bool timed)
    {
        drop_graphs();
        hipStream_t st = eng_.stream();
        const FrLaunch &L = eng_.launch();
        const bool prof = fr_prof_on();
        fr_prof_enable(false); /* no event nodes inside the graphs */
        for (int part = 0; part < (timed ? 3 : 1); ++part) {
            FR_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            /* no k_begin: admission (k_batch_start) and each k_finish leave a
             * plan ready for its next iteration, and the host does not edit a
             * plan between iterations */
            if (!timed) {
                fr_launch_iteration(L, st, false);
            } else if (part == 0) {
                fr_launch_iteration_pre(L, st, false);
            } else if (part == 1) {
                fr_launch_iteration_expand(L, st);
            } else {
                fr_launch_iteration_post(L, st, false);
            }
            FR_CHECK(hipStreamEndCapture(st, &graph_[part]));
            FR_CHECK(hipGraphInstantiate(&exec_[part], graph_[part], nullptr, nullptr, 0));
        }
        fr_prof_enable(prof);
        graph_timed_ = timed;
    }

    void drop_graphs()
    {
        for (int i = 0; i < 3; ++i) {
            if (exec_[i])
                (void)hipGraphExecDestroy(exec_[i]);
            if (graph_[i])
                (void)hipGraphDestroy(graph_[i]);
            exec_[i] = nullptr;
            graph_[i] = nullptr;
        }
    }

    /* snap: control blocks copied after the iterations of step `seq` (null:
     * nothing new to read).  A slot admitted after that copy was taken is
     * only checked against its deadline. */
    void poll_and_retire(const FrCtl *snap, long seq)
    {
        const int ns = (int)slots_.size();
        std::vector<int> found, failed;
        const double t = now_ms() - t0_;
        /* time between polls: a plan whose deadline falls before the next poll
         * is retired now (the timeout of fastrrt.cpp:110-114, per request, as
         * a completion deadline) */
        if (last_poll_ >= 0) {
            const double d = t - last_poll_;
            poll_ema_ = poll_ema_ > 0 ? 0.75 * poll_ema_ + 0.25 * d : d;
        }
        last_poll_ = t;
        bool changed = false;
        if (fin_busy_ && hipEventQuery(fin_ev_) == hipSuccess) {
            collect_finalized();
            changed = true;
        }
        static const FrCtl none{};
        for (int i = 0; i < ns; ++i) {
            if (!slots_[i].busy || slots_[i].finalizing)
                continue;
            const double dl = slots_[i].p.req.deadline_ms;
            const bool seen = snap && seq >= slots_[i].admit_seq;
            const int st = seen ? snap[i].status : FR_SEARCHING;
            if (st == FR_FOUND)
                found.push_back(i);
            else if (st == FR_FAILED)
                failed.push_back(i);
            else if (dl > 0 && t - slots_[i].p.t_submit + 2.0 * poll_ema_ + 0.3 > dl)
                failed.push_back(i); /* with a margin for a slower next iteration */
        }
        for (int sl : failed)
            retire(sl, FR_FAILED, nullptr, 0, snap && seq >= slots_[sl].admit_seq ? snap[sl] : none, false);
        if (!failed.empty()) {
            fr_launch_ctl_idle(eng_.launch().ctl, failed.data(), (int)failed.size(), eng_.stream());
            FR_CHECK(hipGetLastError());
        }
        if (!found.empty() && !fin_busy_)
            launch_finalize(found);
        if (changed || !failed.empty())
            rebuild_active();
    }

    /* path_optimize rounds (fastrrt.cpp:183-196) and getPlannedPath
     * (:204-232) of up to fin_max_ plans that reached their goal, on the
     * second stream.  The control blocks were prepared by k_finish (rounds
     * from the request); the list is padded with the idle plan. */
    /* A plan with a deadline runs as many of its optimisation rounds as fit
     * in the time it has left: a round costs about fin_round_ms_ (measured),
     * and its result reaches the host one poll after the graph ends. */
    int round_limit(int s, double t) const
    {
        const Slot &sl = slots_[s];
        const double dl = sl.p.req.deadline_ms;
        if (dl <= 0)
            return sl.p.req.optimize_rounds;
        const double left = dl - (t - sl.p.t_submit) - 2.0 * poll_ema_ - 0.3;
        const int r = (int)std::floor(left / std::max(fin_round_ms_, 1e-3)) - 1;
        return std::max(0, std::min(r, sl.p.req.optimize_rounds));
    }

    void launch_finalize(const std::vector<int> &found)
    {
        int rounds = 0;
        fin_slots_.clear();
        for (int s : found) {
            if ((int)fin_slots_.size() == fin_max_)
                break;
            fin_slots_.push_back(s);
            rounds = std::max(rounds, slots_[s].p.req.optimize_rounds);
        }
        if (!fin_exec_ || rounds > fin_rounds_)
            capture_finalize(std::max(rounds, fin_rounds_));
        const double t = now_ms() - t0_;
        /* A plan that reaches its goal while this graph runs waits for it:
         * the graph's rounds end in time for the earliest deadline of a plan
         * still searching or waiting (one round and the path extraction
         * after them; a plan past that point finalises with no rounds). */
        int cap = 1 << 30;
        for (int i = 0; i < (int)slots_.size(); ++i) {
            const Slot &sl = slots_[i];
            if (!sl.busy || sl.finalizing || sl.p.req.deadline_ms <= 0 ||
                std::find(fin_slots_.begin(), fin_slots_.end(), i) != fin_slots_.end())
                continue;
            const double left = sl.p.req.deadline_ms - (t - sl.p.t_submit) - 2.0 * poll_ema_ - 0.3;
            cap = std::min(cap, std::max(0, (int)std::floor(left / std::max(fin_round_ms_, 1e-3)) - 2));
        }
        fin_used_rounds_ = 0;
        for (int q = 0; q < fin_max_; ++q) {
            const bool real = q < (int)fin_slots_.size();
            h_fin_list_[q] = real ? fin_slots_[q] : dummy_;
            h_fin_list_[fin_max_ + q] = real ? std::min(cap, round_limit(fin_slots_[q], t)) : 0;
            if (real)
                fin_used_rounds_ = std::max(fin_used_rounds_, h_fin_list_[fin_max_ + q]);
        }
        fin_t_launch_ = t;
        /* the list is read before the graph runs; the next launch waits for
         * this finalisation to end, so the host buffer is free then */
        fr_launch_copy(fin_list_, d_fin_list_, sizeof(int) * 2 * fin_max_, st2_);
        FR_CHECK(hipGetLastError());
        FR_CHECK(hipGraphLaunch(fin_exec_, st2_));
        FR_CHECK(hipEventRecord(fin_ev_, st2_));
        for (int s : fin_slots_)
            slots_[s].finalizing = true;
        finalizing_ += (int)fin_slots_.size();
        fin_busy_ = true;
    }

    /* Grid sizes are small: the kernels stride over cells, path pairs and
     * list entries, and most entries of most launches are gated out. */
    void capture_finalize(int rounds)
    {
        if (fin_exec_)
            (void)hipGraphExecDestroy(fin_exec_);
        if (fin_graph_)
            (void)hipGraphDestroy(fin_graph_);
        fin_exec_ = nullptr;
        fin_graph_ = nullptr;
        const FrLaunch &L = eng_.launch();
        const int n = fin_max_, grid_cells = 16 * FR_BLOCK, grid_path = 16;
        const bool prof = fr_prof_on();
        fr_prof_enable(false);
        FR_CHECK(hipStreamBeginCapture(st2_, hipStreamCaptureModeThreadLocal));
        fr_launch_fin_limits(L, fin_list_, fin_list_ + n, n, st2_);
        for (int r = 0; r < rounds; ++r) {
            fr_launch_goal_scan_multi(L, fin_list_, n, grid_cells, FR_GATE_OPT, st2_);
            fr_launch_best_node(L, fin_list_, n, grid_cells, nullptr, st2_, FR_GATE_OPT);
            fr_launch_walk(L, fin_list_, n, st2_, FR_GATE_OPT);
            fr_launch_opt_round(L, fin_list_, n, grid_path, grid_cells, st2_, FR_GATE_OPT);
        }
        fr_launch_goal_scan_multi(L, fin_list_, n, grid_cells, FR_GATE_FOUND, st2_);
        fr_launch_best_node(L, fin_list_, n, grid_cells, nullptr, st2_, FR_GATE_FOUND);
        fr_launch_walk(L, fin_list_, n, st2_, FR_GATE_FOUND);
        fr_launch_fin_pack(L, fin_list_, n, d_pack_, kPackStride, kPackPath, st2_);
        FR_CHECK(hipStreamEndCapture(st2_, &fin_graph_));
        fr_prof_enable(prof);
        FR_CHECK(hipGraphInstantiate(&fin_exec_, fin_graph_, nullptr, nullptr, 0));
        fin_rounds_ = rounds;
    }

    /* The finalisation graph has completed (its event was queried): every
     * record is in pinned memory and every control block is IDLE again. */
    void collect_finalized()
    {
        /* graph time per round, measured up to this poll (an upper bound) */
        if (fin_used_rounds_ >= 5) {
            const double dt = now_ms() - t0_ - fin_t_launch_;
            fin_round_ms_ = 0.8 * fin_round_ms_ + 0.2 * (dt / (fin_used_rounds_ + 1));
        }
        const int nf = (int)fin_slots_.size();
        for (int q = 0; q < nf; ++q) {
            const uint8_t *rec = h_pack_ + (size_t)q * kPackStride;
            FrCtl c;
            std::memcpy(&c, rec, sizeof(FrCtl));
            const int len = c.goal_reached ? std::max(0, c.path_len) : 0;
            std::vector<float4> p((size_t)len);
            if (len <= kPackPath) {
                if (len)
                    std::memcpy(p.data(), rec + FR_PACK_PATH_OFF, sizeof(float4) * len);
            } else { /* longer than a record holds: read from the plan's slab */
                FR_CHECK(hipMemcpyAsync(p.data(), slots_[fin_slots_[q]].slab->path, sizeof(float4) * len,
                                        hipMemcpyDeviceToHost, st2_));
                FR_CHECK(hipStreamSynchronize(st2_));
            }
            std::vector<float> trip(3 * (size_t)len);
            for (int k = 0; k < len; ++k) {
                trip[3 * k] = (float)(int)p[k].x;
                trip[3 * k + 1] = (float)(int)p[k].y;
                trip[3 * k + 2] = p[k].z;
            }
            slots_[fin_slots_[q]].finalizing = false;
            /* the goal test (graph_best_node.cu:182-222) does not test the
             * footprint, findBestNode does: a reached goal can leave no path */
            retire(fin_slots_[q], len > 0 ? FR_FOUND : FR_RESULT_NO_PATH, trip.data(), len, c, false,
                   len > 0 ? p[len - 1].w : 0.0f);
        }
        finalizing_ -= nf;
        fin_slots_.clear();
        fin_busy_ = false;
    }

    /* The slot's control block goes back to IDLE on the iteration stream
     * (from a constant pinned block), ahead of any later admission. */
    void retire(int s, int status, const float *path, int len, const FrCtl &c, bool reset_ctl = true,
                float path_cost = 0.0f)
    {
        Slot &sl = slots_[s];
        Finished f;
        f.res.user_tag = sl.p.req.user_tag;
        f.res.path_cost = path_cost;
        f.res.status = status;
        f.res.iterations = c.iter;
        f.res.expansions = c.total_expansions;
        f.res.work_rollout_cells = c.w_cells;
        f.res.work_goal_points = c.w_goal_pts;
        f.res.work_goal_foot_points = c.w_foot_pts;
        f.res.path_len = len;
        f.res.submit_ms = sl.p.t_submit;
        f.res.done_ms = now_ms() - t0_;
        if (len > 0)
            f.path.assign(path, path + 3 * (size_t)len);
        done_.push_back(std::move(f));
        if (reset_ctl)
            fr_launch_ctl_idle(eng_.launch().ctl, &s, 1, eng_.stream());
        give_back(std::move(sl.slab));
        cells_in_use_ -= sl.cells;
        sl.busy = false;
        sl.ticket = -1;
        running_ -= 1;
    }

    Hbm hbm_;
    FrEngine eng_;
    long max_cells_;
    long cells_in_use_ = 0;
    size_t slab_bytes_ = 0; /* plans' planes allocated, in use or pooled */
    std::vector<Slot> slots_;
    std::deque<Pending> queue_;
    std::deque<Finished> done_;
    std::map<std::tuple<int, int, int, int>, std::vector<std::unique_ptr<FrSlab>>> pool_;
    FrCtl *h_snap_[2] = {nullptr, nullptr};  /* pinned control-block copies, alternating */
    FrCtl *d_snap_[2] = {nullptr, nullptr};  /* their device addresses */
    hipEvent_t snap_ev_[2] = {nullptr, nullptr};
    long snap_seq_ = 0;                      /* copies queued */
    long snap_done_ = 0;                     /* copies acted on */
    int *d_list_ = nullptr;
    /* finalisation on the second stream */
    int dummy_;
    int fin_max_;
    hipStream_t st2_ = nullptr;
    hipEvent_t fin_ev_ = nullptr;
    hipGraph_t fin_graph_ = nullptr;
    hipGraphExec_t fin_exec_ = nullptr;
    int fin_rounds_ = 0;
    int fin_used_rounds_ = 0;
    double fin_t_launch_ = 0, fin_round_ms_ = 0.06;
    bool fin_busy_ = false;
    int finalizing_ = 0;
    std::vector<int> fin_slots_;
    int *fin_list_ = nullptr;
    int *h_fin_list_ = nullptr, *d_fin_list_ = nullptr;
    static constexpr int kPackPath = 2048;                         /* waypoints per pinned record */
    static constexpr long kPackStride = FR_PACK_PATH_OFF + 16L * kPackPath;
    uint8_t *h_pack_ = nullptr, *d_pack_ = nullptr;
    int running_ = 0;
    int next_ticket_ = 0;
    double t0_ = 0;
    hipGraph_t graph_[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t exec_[3] = {nullptr, nullptr, nullptr};
    bool graph_timed_ = false;
    double last_poll_ = -1, poll_ema_ = 0;
};

#define FR_TRY(ret_on_error, body)                                                                        \
    try {                                                                                                  \
        fr_set_error("");                                                                                  \
        body                                                                                               \
    } catch (const std::exception &e) {                                                                    \
        fr_set_error(e.what());                                                                            \
        return ret_on_error;                                                                               \
    } catch (...) {                                                                                        \
        fr_set_error("unknown error");                                                                     \
        return ret_on_error;                                                                               \
    }

static FrBatch *B(void *p)
{
    if (!p)
        throw std::runtime_error("null batch handle");
    return static_cast<FrBatch *>(p);
}

extern "C" {

void *fastrrt_batch_create(int max_plans, int64_t max_cells)
{
    FR_TRY(nullptr, { return (void *)new FrBatch(max_plans, (long)max_cells); })
}

void fastrrt_batch_destroy(void *batch)
{
    try {
        delete static_cast<FrBatch *>(batch);
    } catch (...) {
    }
}

int fastrrt_batch_submit(void *batch, const fastrrt_plan_request *req)
{
    FR_TRY(-1, {
        if (!req)
            throw std::runtime_error("null request");
        return B(batch)->submit(*req);
    })
}

int fastrrt_batch_step(void *batch, int iterations) { FR_TRY(-1, { return B(batch)->step(iterations); }) }

bool fastrrt_batch_poll(void *batch, fastrrt_plan_result *result, float *path, int cap)
{
    FR_TRY(false, { return B(batch)->poll(result, path, cap); })
}

void fastrrt_batch_drain(void *batch)
{
    try {
        fr_set_error("");
        B(batch)->drain();
    } catch (const std::exception &e) {
        fr_set_error(e.what());
    }
}

double fastrrt_batch_clock_ms(void *batch) { FR_TRY(-1.0, { return B(batch)->clock_ms(); }) }

bool fastrrt_batch_set_launch(void *batch, int grid_waves, int grid_tiles, int grid_items)
{
    FR_TRY(false, {
        B(batch)->set_launch(grid_waves, grid_tiles, grid_items);
        return true;
    })
}

bool fastrrt_batch_memory(void *batch, fastrrt_batch_hbm *out)
{
    FR_TRY(false, {
        if (!out)
            throw std::runtime_error("null output");
        B(batch)->memory(out);
        return true;
    })
}

bool fastrrt_batch_set_hbm_reserve(void *batch, uint64_t bytes)
{
    FR_TRY(false, {
        B(batch)->set_hbm_reserve((size_t)bytes);
        return true;
    })
}

bool fastrrt_batch_set_block_threads(void *batch, int threads)
{
    FR_TRY(false, {
        B(batch)->set_block(threads);
        return true;
    })
}

int fastrrt_batch_get_block_threads(void *batch) { FR_TRY(-1, { return B(batch)->block(); }) }

bool fastrrt_batch_get_launch(void *batch, int *grid_waves_tiles_items)
{
    FR_TRY(false, {
        if (!grid_waves_tiles_items)
            throw std::runtime_error("null output");
        B(batch)->get_launch(grid_waves_tiles_items);
        return true;
    })
}

} /* extern "C" */
