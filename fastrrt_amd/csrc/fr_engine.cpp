/*
 * fr_engine.cpp — slab allocation and the launch engine of the planning loop.
 */
#include "fr_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <stdexcept>

static thread_local std::string g_fr_error;

void fr_set_error(const std::string &msg) { g_fr_error = msg; }
const char *fr_get_error() { return g_fr_error.c_str(); }

/* ------------------------------------------------------------------------ */
/* event profiler                                                           */
/* ------------------------------------------------------------------------ */

namespace {
struct ProfState {
    std::mutex mu;
    bool on = false;
    bool split = false;
    std::deque<std::pair<hipEvent_t, hipEvent_t>> done[FR_PROF_N];
    hipEvent_t open_ev[FR_PROF_N] = {};
    std::vector<hipEvent_t> pool;
    long launches[FR_PROF_N] = {};
    double total_ms[FR_PROF_N] = {};
};
ProfState &prof()
{
    static ProfState p;
    return p;
}
hipEvent_t prof_event(ProfState &p)
{
    if (!p.pool.empty()) {
        hipEvent_t e = p.pool.back();
        p.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    FR_CHECK(hipEventCreate(&e));
    return e;
}
void prof_take(ProfState &p, int k)
{
    const auto pr = p.done[k].front();
    p.done[k].pop_front();
    float ms = 0;
    FR_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
    p.launches[k] += 1;
    p.total_ms[k] += ms;
    p.pool.push_back(pr.first);
    p.pool.push_back(pr.second);
}
/* every recorded pair, waiting for the last ones (reads and resets) */
void prof_drain(ProfState &p, int k)
{
    while (!p.done[k].empty()) {
        FR_CHECK(hipEventSynchronize(p.done[k].front().second));
        prof_take(p, k);
    }
}
/* the pairs whose end event has completed, without waiting: the list stays
 * short and no launch pays for a long drain */
void prof_collect(ProfState &p, int k)
{
    while (!p.done[k].empty()) {
        const hipError_t q = hipEventQuery(p.done[k].front().second);
        if (q == hipErrorNotReady)
            return;
        FR_CHECK(q);
        prof_take(p, k);
    }
}
} // namespace

void fr_prof_enable(bool on)
{
    std::lock_guard<std::mutex> g(prof().mu);
    prof().on = on;
}
bool fr_prof_on() { return prof().on; }
void fr_prof_split(bool on)
{
    std::lock_guard<std::mutex> g(prof().mu);
    prof().split = on;
}
bool fr_prof_split_on() { return prof().split; }
void fr_prof_begin(int k, hipStream_t st)
{
    ProfState &p = prof();
    if (!p.on)
        return;
    std::lock_guard<std::mutex> g(p.mu);
    hipEvent_t e = prof_event(p);
    FR_CHECK(hipEventRecord(e, st));
    p.open_ev[k] = e;
}
void fr_prof_end(int k, hipStream_t st)
{
    ProfState &p = prof();
    if (!p.on)
        return;
    std::lock_guard<std::mutex> g(p.mu);
    if (!p.open_ev[k])
        return;
    hipEvent_t e = prof_event(p);
    FR_CHECK(hipEventRecord(e, st));
    p.done[k].push_back({p.open_ev[k], e});
    p.open_ev[k] = nullptr;
    prof_collect(p, k);
}
void fr_prof_read(int k, long *launches, double *total_ms)
{
    ProfState &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    prof_drain(p, k);
    *launches = p.launches[k];
    *total_ms = p.total_ms[k];
}
void fr_prof_add(int k, double ms)
{
    ProfState &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    p.launches[k] += 1;
    p.total_ms[k] += ms;
}
void fr_prof_reset()
{
    ProfState &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    for (int k = 0; k < FR_PROF_N; ++k) {
        prof_drain(p, k);
        p.launches[k] = 0;
        p.total_ms[k] = 0;
    }
}

static size_t fr_align(size_t v) { return (v + 255) & ~(size_t)255; }

/* ------------------------------------------------------------------------ */
/* FrSlab                                                                   */
/* ------------------------------------------------------------------------ */

FrSlab::~FrSlab() { release(); }

void FrSlab::release()
{
    if (base_)
        (void)hipFree(base_);
    if (dist_base_)
        (void)hipFree(dist_base_);
    base_ = nullptr;
    dist_base_ = nullptr;
    bytes_ = 0;
}

/* The padded distance planes depend on the footprint; they are reallocated
 * alone so that the node grid survives a footprint change, as the reference's
 * grid survives set_plan_data. */
/* Distance planes of a footprint: the plane and its scratch (the row pass),
 * over the grid padded by the footprint's reach. */
static void fr_dist_geometry(int W, int H, int mx, int mz, int *pad, int *PW, int *PH, int *TW, size_t *dbytes,
                             size_t *pwh)
{
    const double hyp = std::sqrt((double)mx * mx + (double)mz * mz);
    *pad = (int)std::ceil(hyp) + 2;
    *PW = W + 2 * *pad;
    *PH = H + 2 * *pad;
    *pwh = (size_t)*PW * *PH;
    *TW = *PW; /* row pitch */
    *dbytes = *pwh;
}

/* The padded distance planes depend on the footprint; they are reallocated
 * alone so that the node grid survives a footprint change, as the reference's
 * grid survives set_plan_data. */
void FrSlab::set_footprint(int mx_, int mz_, hipStream_t st)
{
    if (dist_base_ && mx_ == mx && mz_ == mz)
        return;
    if (dist_base_)
        (void)hipFree(dist_base_);
    dist_base_ = nullptr;
    mx = mx_;
    mz = mz_;
    size_t dbytes = 0, PWH = 0;
    fr_dist_geometry(W, H, mx, mz, &pad, &PW, &PH, &TW, &dbytes, &PWH);
    FR_CHECK(hipMalloc(&dist_base_, fr_align(dbytes) + fr_align(PWH)));
    dist = (uint8_t *)dist_base_;
    dist_tmp = (uint8_t *)dist_base_ + fr_align(dbytes);
    FR_CHECK(hipMemsetAsync(dist_base_, 0, fr_align(dbytes) + fr_align(PWH), st));
}

/* Offsets of the planes in the slab's one allocation (a plane per field,
 * each 256-byte aligned). */
struct FrSlabLayout {
    size_t type, derive, px, pz, heading, cost, intr, direct, gcf, gch, gcc, done, bits, mark, trav, region;
    size_t zero_end, path, pd, claim, escr, bytes;
};

static FrSlabLayout fr_slab_layout(int W, int H, int path_cap, bool with_intrinsic)
{
    const size_t N = (size_t)W * H;
    const int dens = (W / 16 + 1) * (H / 16 + 1); /* fits every region side >= 16 */
    size_t off = 0;
    auto take = [&](size_t b) {
        size_t o = off;
        off += fr_align(b);
        return o;
    };
    FrSlabLayout l;
    l.type = take(N);
    l.derive = take(4 * N);
    l.px = take(4 * N);
    l.pz = take(4 * N);
    l.heading = take(4 * N);
    l.cost = take(4 * N);
    l.intr = with_intrinsic ? take(4 * N) : 0;
    l.direct = with_intrinsic ? take(4 * N) : 0;
    l.gcf = take(N);
    l.gch = take(4 * N);
    l.gcc = take(4 * N);
    l.done = take(4 * ((N + 31) / 32));
    l.bits = take(4 * ((N + 31) / 32));
    l.mark = take(N);
    l.trav = take(N);
    l.region = take(4 * (size_t)2 * dens);
    l.zero_end = off; /* everything above starts at zero */
    l.path = take(16 * (size_t)path_cap); /* written before read */
    l.pd = take(16 * (size_t)path_cap);
    l.claim = take(8 * N);
    l.escr = take(16 * N); /* erase scratch, written before read */
    l.bytes = off;
    return l;
}

size_t FrSlab::bytes_for(int W, int H, int mx, int mz, int path_cap, bool with_intrinsic)
{
    const int cap = path_cap > 0 ? path_cap : W * H + 1;
    int pad, PW, PH, TW;
    size_t dbytes = 0, PWH = 0;
    fr_dist_geometry(W, H, mx, mz, &pad, &PW, &PH, &TW, &dbytes, &PWH);
    return fr_slab_layout(W, H, cap, with_intrinsic).bytes + fr_align(dbytes) + fr_align(PWH);
}

void FrSlab::allocate(int W_, int H_, int mx_, int mz_, int path_cap_, bool with_intrinsic_, hipStream_t st)
{
    release();
    W = W_;
    H = H_;
    mx = mx_;
    mz = mz_;
    with_intrinsic = with_intrinsic_;
    path_cap = path_cap_ > 0 ? path_cap_ : W * H + 1;
    const FrSlabLayout l = fr_slab_layout(W, H, path_cap, with_intrinsic);
    bytes_ = l.bytes;
    FR_CHECK(hipMalloc(&base_, bytes_));
    char *b = (char *)base_;
    type = (uint8_t *)(b + l.type);
    derive = (int32_t *)(b + l.derive);
    px = (int32_t *)(b + l.px);
    pz = (int32_t *)(b + l.pz);
    heading = (float *)(b + l.heading);
    cost = (float *)(b + l.cost);
    intrinsic = with_intrinsic ? (float *)(b + l.intr) : nullptr;
    direct = with_intrinsic ? (float *)(b + l.direct) : nullptr;
    gc_flag = (uint8_t *)(b + l.gcf);
    gc_heading = (float *)(b + l.gch);
    gc_cost = (float *)(b + l.gcc);
    gc_done = (uint32_t *)(b + l.done);
    gc_bits = (uint32_t *)(b + l.bits);
    mark = (uint8_t *)(b + l.mark);
    trav = (uint8_t *)(b + l.trav);
    region = (uint32_t *)(b + l.region);
    path = (float4 *)(b + l.path);
    pd = (float4 *)(b + l.pd);
    claim = (unsigned long long *)(b + l.claim);
    escr = (int32_t *)(b + l.escr);
    zero_end_ = l.zero_end;
    reset(st);
    mx = -1;
    mz = -1;
    set_footprint(mx_, mz_, st);
}

void FrSlab::reset(hipStream_t st)
{
    FR_CHECK(hipMemsetAsync(base_, 0, zero_end_, st));
    FR_CHECK(hipMemsetAsync(claim, 0xFF, 8 * (size_t)W * H, st));
}

FrFrame::~FrFrame()
{
    if (d_xyz)
        (void)hipFree(d_xyz);
    if (d_costs)
        (void)hipFree(d_costs);
}

/* ------------------------------------------------------------------------ */
/* FrEngine                                                                 */
/* ------------------------------------------------------------------------ */

FrEngine::FrEngine(int max_slots, long queue_cap, hipStream_t stream)
    : max_slots_(max_slots), queue_cap_(queue_cap), st_(stream), own_stream_(false)
{
    try {
        allocate();
    } catch (...) {
        /* a half-built engine runs no destructor */
        if (st_)
            (void)hipStreamSynchronize(st_);
        release();
        throw;
    }
}

/* Device buffers of an engine: descriptors, control blocks, the active
 * table and the per-iteration queues (sized by the cell capacity). */
struct FrEngineSizes {
    long max_tiles, pair_cap, demand_cap;
    size_t act_tiles_off, act_bytes, bytes;
};

static FrEngineSizes fr_engine_sizes(int max_slots, long queue_cap)
{
    FrEngineSizes z;
    z.max_tiles = queue_cap / FR_TILE + max_slots + 1;
    z.act_tiles_off = sizeof(int) * (4 + (size_t)(max_slots + 1) / 2 * 2);
    z.act_bytes = z.act_tiles_off + sizeof(int2) * (size_t)z.max_tiles;
    z.pair_cap = std::min<long>(std::max<long>(16 * queue_cap, 1L << 16), 1L << 26);
    z.demand_cap = queue_cap + 1;
    z.bytes = sizeof(FrPlanDev) * (size_t)max_slots + sizeof(FrCtl) * (size_t)max_slots + z.act_bytes +
              sizeof(int) * (size_t)max_slots + sizeof(int) * FR_CNT_N + sizeof(FrExp) * (size_t)(queue_cap + 1) +
              sizeof(FrFresh) * (size_t)(2 * queue_cap + max_slots + 1) + sizeof(int) * 64 +
              sizeof(FrPair) * (size_t)z.pair_cap + 2 * sizeof(int2) * (size_t)z.demand_cap +
              sizeof(FrGcGeo) * (size_t)z.demand_cap;
    return z;
}

size_t FrEngine::bytes_for(int max_slots, long queue_cap) { return fr_engine_sizes(max_slots, queue_cap).bytes; }

void FrEngine::allocate()
{
    const int max_slots = max_slots_;
    const long queue_cap = queue_cap_;
    const FrEngineSizes z = fr_engine_sizes(max_slots, queue_cap);
    if (!st_) {
        FR_CHECK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        own_stream_ = true;
    }
    h_plans_.assign(max_slots, FrPlanDev{});
    /* staging rings in mapped, coherent pinned memory: the iteration's
     * uploads are read by a copy kernel on the stream (fr_launch_copy) */
    const unsigned mapped = hipHostMallocMapped | hipHostMallocCoherent;
    FR_CHECK(hipHostMalloc((void **)&h_plan_stage_, sizeof(FrPlanDev) * max_slots * FR_PLAN_RING, mapped));
    FR_CHECK(hipHostGetDevicePointer((void **)&d_plan_stage_, h_plan_stage_, 0));
    FR_CHECK(hipHostMalloc((void **)&h_int_stage_, sizeof(int) * (max_slots + 1) * FR_STAGE_RING, mapped));
    FR_CHECK(hipHostGetDevicePointer((void **)&d_int_stage_, h_int_stage_, 0));
    for (int k = 0; k < FR_PLAN_RING; ++k)
        FR_CHECK(hipEventCreateWithFlags(&plan_ev_[k], hipEventDisableTiming));
    for (auto &e : int_ev_)
        FR_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const long max_tiles = z.max_tiles;
    FR_CHECK(hipMalloc(&L_.plans, sizeof(FrPlanDev) * max_slots));
    FR_CHECK(hipMalloc(&L_.ctl, sizeof(FrCtl) * max_slots));
    act_max_tiles_ = max_tiles;
    act_tiles_off_ = z.act_tiles_off;
    act_bytes_ = z.act_bytes;
    FR_CHECK(hipMalloc(&d_act_, act_bytes_));
    FR_CHECK(hipHostMalloc((void **)&h_act_stage_, act_bytes_ * FR_ACT_RING, mapped));
    FR_CHECK(hipHostGetDevicePointer((void **)&d_act_stage_, h_act_stage_, 0));
    for (auto &e : act_ev_)
        FR_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    L_.n_tiles = (int *)d_act_;
    L_.n_slots = (int *)d_act_ + 1;
    L_.slots = (int *)d_act_ + 4;
    L_.tiles = (int2 *)((uint8_t *)d_act_ + act_tiles_off_);
    FR_CHECK(hipMalloc(&d_setup_slots_, sizeof(int) * max_slots));
    FR_CHECK(hipMemsetAsync(d_act_, 0, act_bytes_, st_));
    FR_CHECK(hipMalloc(&L_.counters, sizeof(int) * FR_CNT_N));
    FR_CHECK(hipMalloc(&L_.exp_q, sizeof(FrExp) * (queue_cap + 1)));
    FR_CHECK(hipMalloc(&L_.fresh_q, sizeof(FrFresh) * (2 * queue_cap + max_slots + 1)));
    FR_CHECK(hipMalloc(&d_scratch_, sizeof(int) * 64));
    L_.pair_cap = z.pair_cap;
    FR_CHECK(hipMalloc(&L_.pairs, sizeof(FrPair) * L_.pair_cap));
    L_.demand_cap = z.demand_cap;
    FR_CHECK(hipMalloc(&L_.demand, sizeof(int2) * L_.demand_cap));
    FR_CHECK(hipMalloc(&L_.demand2, sizeof(int2) * L_.demand_cap));
    FR_CHECK(hipMalloc(&L_.demand2_geo, sizeof(FrGcGeo) * L_.demand_cap));
    FR_CHECK(hipMemsetAsync(L_.ctl, 0, sizeof(FrCtl) * max_slots, st_));
    FR_CHECK(hipMemsetAsync(L_.counters, 0, sizeof(int) * FR_CNT_N, st_));
    L_.grid_tiles = (int)std::min<long>(max_tiles, 2048);
    L_.grid_items = (int)std::max<long>(1, std::min<long>((queue_cap + FR_BLOCK - 1) / FR_BLOCK, 2048));
    L_.grid_slots = 1;
    L_.grid_waves = 4096; /* blocks of 4 one-item waves; k_expand needs a multiple of 8 blocks */
    L_.block = FR_BLOCK;
    L_.max_slots = max_slots;
}

FrEngine::~FrEngine()
{
    if (st_)
        (void)hipStreamSynchronize(st_);
    release();
}

/* Frees what allocate() made (hipFree / hipHostFree of null are no-ops);
 * the events and the owned stream only when created. */
void FrEngine::release()
{
    (void)hipFree(L_.plans);
    (void)hipFree(L_.ctl);
    (void)hipFree(d_act_);
    (void)hipHostFree(h_act_stage_);
    for (auto e : act_ev_)
        if (e)
            (void)hipEventDestroy(e);
    (void)hipFree(d_setup_slots_);
    (void)hipFree(L_.counters);
    (void)hipFree(L_.exp_q);
    (void)hipFree(L_.fresh_q);
    (void)hipFree(d_scratch_);
    (void)hipHostFree(h_plan_stage_);
    (void)hipHostFree(h_int_stage_);
    for (int k = 0; k < FR_PLAN_RING; ++k)
        if (plan_ev_[k])
            (void)hipEventDestroy(plan_ev_[k]);
    for (auto e : int_ev_)
        if (e)
            (void)hipEventDestroy(e);
    (void)hipFree(L_.pairs);
    (void)hipFree(L_.demand);
    (void)hipFree(L_.demand2);
    (void)hipFree(L_.demand2_geo);
    if (own_stream_ && st_)
        (void)hipStreamDestroy(st_);
}

void FrEngine::set_grids(int waves, int tiles, int items)
{
    /* k_expand, the goal-curve kernels and the fresh-node and pair kernels
     * deal their queues to the eight XCDs in blocks of 8 */
    if (waves < 8 || waves > 65535 || waves % 8 != 0)
        throw std::runtime_error("grid_waves must be a multiple of 8 in [8, 65535]");
    if (tiles < 1 || tiles > 65535 || items < 1 || items > 65535)
        throw std::runtime_error("grid_tiles and grid_items must be in [1, 65535]");
    L_.grid_waves = waves;
    L_.grid_tiles = tiles;
    L_.grid_items = items;
}

void FrEngine::set_block(int threads)
{
    if (threads < 64 || threads > FR_BLOCK || threads % 64 != 0)
        throw std::runtime_error("threads per block must be 64, 128, 192 or 256");
    L_.block = threads;
}

/* The footprint's (2 mz + 1) x (2 mx + 1) samples in FR_GROUP x FR_GROUP
 * groups (the last row and column of groups may be narrower); each group's
 * middle sample, clamped into the group, is its probe. */
static void fr_probe_table(FrPlanDev &d)
{
    const int gr = (2 * d.mz + FR_GROUP) / FR_GROUP, gc = (2 * d.mx + FR_GROUP) / FR_GROUP;
    d.n_groups = (d.mx >= 0 && d.mz >= 0 && gr * gc <= FR_MAX_GROUPS) ? gr * gc : 0;
    for (int g = 0; g < FR_MAX_GROUPS; ++g) { /* entries past the groups: the centre sample */
        d.probe_i[g] = 0.0f;
        d.probe_j[g] = 0.0f;
    }
    for (int g = 0; g < d.n_groups; ++g) {
        const int a = g / gc, b = g % gc;
        const int i0 = -d.mz + a * FR_GROUP, j0 = -d.mx + b * FR_GROUP;
        const int i1 = std::min(i0 + FR_GROUP - 1, d.mz), j1 = std::min(j0 + FR_GROUP - 1, d.mx);
        d.probe_i[g] = (float)std::min(i0 + FR_GROUP / 2, i1);
        d.probe_j[g] = (float)std::min(j0 + FR_GROUP / 2, j1);
    }
}

void FrEngine::bind(int slot, FrSlab &s, const FrPlanParams &p, const FrFrameView &frame, bool upload)
{
    FrPlanDev d{};
    d.W = s.W;
    d.H = s.H;
    d.N = s.W * s.H;
    d.pad = s.pad;
    d.PW = s.PW;
    d.PH = s.PH;
    d.TW = s.TW;
    d.mx = s.mx;
    d.mz = s.mz;
    fr_probe_table(d);
    d.lbx = p.lbx;
    d.lbz = p.lbz;
    d.ubx = p.ubx;
    d.ubz = p.ubz;
    d.og_x = p.og_x;
    d.og_y = p.og_y;
    d.v = p.v;
    d.max_path = p.max_path;
    d.dist_tol = p.dist_tol;
    d.herr_tol = p.herr_tol;
    d.max_curv_loop = p.max_curv_loop;
    d.max_curv_goal = p.max_curv_goal;
    d.max_steer_rad = p.max_steer_rad;
    d.lr = p.lr;
    d.sx = p.sx;
    d.sz = p.sz;
    d.gx = p.gx;
    d.gz = p.gz;
    d.sh = p.sh;
    d.gh = p.gh;
    d.region_side = p.region_side;
    d.dens_w = s.W / p.region_side + 1;
    d.dens_h = s.H / p.region_side + 1;
    d.smart = p.smart;
    d.max_iter = p.max_iter;
    d.exp_budget = p.exp_budget;
    d.seed = p.seed;
    d.stream = p.stream;
    d.type = s.type;
    d.derive = s.derive;
    d.px = s.px;
    d.pz = s.pz;
    d.heading = s.heading;
    d.cost = s.cost;
    d.intrinsic = s.intrinsic;
    d.gc_flag = s.gc_flag;
    d.gc_heading = s.gc_heading;
    d.gc_cost = s.gc_cost;
    d.gc_done = s.gc_done;
    d.gc_bits = s.gc_bits;
    d.gc_lazy = p.gc_lazy;
    d.claim = s.claim;
    d.mark = s.mark;
    d.escr = s.escr;
    d.dist = s.dist;
    d.dist_tmp = s.dist_tmp;
    d.trav = s.trav;
    d.region = s.region;
    d.frame = frame.xyz;
    d.class_costs = frame.costs;
    d.n_classes = frame.n_classes;
    d.path = s.path;
    d.pd = s.pd;
    d.path_cap = s.path_cap;
    d.zero_base = (uint4 *)s.zero_base();
    d.zero_words = (long)(s.zero_bytes() / 16);
    d.opt_limit = p.opt_rounds;
    h_plans_[slot] = d;
    if (upload)
        upload_plan(slot);
}

void FrEngine::upload_plans()
{
    const int k = plan_next_;
    plan_next_ = (plan_next_ + 1) % FR_PLAN_RING;
    FrPlanDev *stage = h_plan_stage_ + (size_t)k * max_slots_;
    FR_CHECK(hipEventSynchronize(plan_ev_[k])); /* that ring entry's copy has read it */
    std::memcpy(stage, h_plans_.data(), sizeof(FrPlanDev) * max_slots_);
    fr_launch_copy(L_.plans, d_plan_stage_ + (size_t)k * max_slots_, sizeof(FrPlanDev) * max_slots_, st_);
    FR_CHECK(hipGetLastError());
    FR_CHECK(hipEventRecord(plan_ev_[k], st_));
}

void FrEngine::upload_ints(int *d_dst, const int *src, int n)
{
    if (n > max_slots_ + 1)
        throw std::runtime_error("upload_ints: too many entries");
    const int k = int_next_;
    int_next_ = (int_next_ + 1) % FR_STAGE_RING;
    int *stage = h_int_stage_ + (size_t)k * (max_slots_ + 1);
    FR_CHECK(hipEventSynchronize(int_ev_[k])); /* that ring entry's copy has read it */
    std::memcpy(stage, src, sizeof(int) * n);
    fr_launch_copy(d_dst, d_int_stage_ + (size_t)k * (max_slots_ + 1), sizeof(int) * n, st_);
    FR_CHECK(hipGetLastError());
    FR_CHECK(hipEventRecord(int_ev_[k], st_));
}

void FrEngine::upload_plan(int slot)
{
    FR_CHECK(hipMemcpyAsync(L_.plans + slot, &h_plans_[slot], sizeof(FrPlanDev), hipMemcpyHostToDevice, st_));
}

void FrEngine::reset_ctl(int slot, int status)
{
    FrCtl c{};
    c.status = status;
    c.pending_start = 1;
    write_ctl(slot, c);
}

void FrEngine::read_ctl(int slot, FrCtl *out)
{
    FR_CHECK(hipMemcpyAsync(out, L_.ctl + slot, sizeof(FrCtl), hipMemcpyDeviceToHost, st_));
    FR_CHECK(hipStreamSynchronize(st_));
}

void FrEngine::write_ctl(int slot, const FrCtl &c)
{
    /* the copy source must outlive the async copy: stage through a member */
    static thread_local FrCtl staged;
    staged = c;
    FR_CHECK(hipMemcpyAsync(L_.ctl + slot, &staged, sizeof(FrCtl), hipMemcpyHostToDevice, st_));
    FR_CHECK(hipStreamSynchronize(st_));
}

void FrEngine::set_active(const std::vector<int> &slots)
{
    if ((int)slots.size() > max_slots_)
        throw std::runtime_error("set_active: too many slots");
    active_ = slots;
    h_tiles_.clear();
    for (int s : slots) {
        const int N = h_plans_[s].N;
        const int nt = (N + FR_TILE - 1) / FR_TILE;
        for (int t = 0; t < nt; ++t)
            h_tiles_.push_back(make_int2(s, t));
    }
    const long nt = (long)h_tiles_.size();
    if (nt > act_max_tiles_)
        throw std::runtime_error("set_active: tile table overflow");
    /* one pinned record, one copy; the ring entry is reused once its copy has read it */
    const int k = act_next_;
    act_next_ = (act_next_ + 1) % FR_ACT_RING;
    uint8_t *stage = h_act_stage_ + (size_t)k * act_bytes_;
    FR_CHECK(hipEventSynchronize(act_ev_[k]));
    int *hdr = (int *)stage;
    hdr[0] = (int)nt;
    hdr[1] = (int)slots.size();
    if (!slots.empty())
        std::memcpy(hdr + 4, slots.data(), sizeof(int) * slots.size());
    if (nt)
        std::memcpy(stage + act_tiles_off_, h_tiles_.data(), sizeof(int2) * nt);
    const size_t bytes = nt ? act_tiles_off_ + sizeof(int2) * nt : act_tiles_off_;
    fr_launch_copy(d_act_, d_act_stage_ + (size_t)k * act_bytes_, bytes, st_);
    FR_CHECK(hipGetLastError());
    FR_CHECK(hipEventRecord(act_ev_[k], st_));
}

void FrEngine::setup(const std::vector<int> &slots, int goal_mode, bool zeroed)
{
    if (slots.empty())
        return;
    int max_N = 0, max_PW = 0, max_PH = 0;
    for (int s : slots) {
        const FrPlanDev &d = h_plans_[s];
        max_N = std::max(max_N, d.N);
        max_PW = std::max(max_PW, d.PW);
        max_PH = std::max(max_PH, d.PH);
    }
    upload_ints(d_setup_slots_, slots.data(), (int)slots.size());
    for (int s : slots) {
        if (goal_mode != FR_GOAL_NONE && h_plans_[s].gc_lazy != (goal_mode == FR_GOAL_LAZY)) {
            h_plans_[s].gc_lazy = goal_mode == FR_GOAL_LAZY;
            upload_plan(s);
        }
    }
    fr_launch_setup(L_, d_setup_slots_, (int)slots.size(), max_N, max_PW, max_PH, goal_mode, zeroed, st_);
    FR_CHECK(hipGetLastError());
}

void FrEngine::goal_complete(int slot)
{
    FR_CHECK(hipMemcpyAsync(d_setup_slots_, &slot, sizeof(int), hipMemcpyHostToDevice, st_));
    fr_launch_goal_complete(L_, d_setup_slots_, 1, h_plans_[slot].N, st_);
    FR_CHECK(hipGetLastError());
    FR_CHECK(hipStreamSynchronize(st_));
}

void FrEngine::iteration()
{
    fr_launch_iteration(L_, st_);
    FR_CHECK(hipGetLastError());
}

void FrEngine::sync() { FR_CHECK(hipStreamSynchronize(st_)); }
