/*
 * fr_engine.h — host-side engine over the planning kernels.
 *
 * FrEngine owns the device tables (plan descriptors, control blocks, the
 * flattened tile table and the work queues) of up to `max_slots` plans and
 * one HIP stream.  FrSlab owns the SoA planes of one plan.
 */
#ifndef FR_ENGINE_H
#define FR_ENGINE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "fr_device.h"

/* FR_CNT_FRESH_SEEN: the fresh queue's length kept for k_finish (which
 * resets the others for the next iteration) */
enum { FR_CNT_EXP = 0, FR_CNT_FRESH = 1, FR_CNT_PAIRS = 2, FR_CNT_DEMAND = 3, FR_CNT_DEMAND2 = 4,
       FR_CNT_FRESH_SEEN = 6, FR_CNT_N = 8 };

/* goal-connection planes at setup: untouched, evaluated when first needed
 * (lazy), or evaluated for every cell (eager) */
enum { FR_GOAL_NONE = 0, FR_GOAL_LAZY = 1, FR_GOAL_EAGER = 2 };

/* A goal curve's set-up (tangents, start heading, point count), handed from
 * the curvature-and-centre pass to the footprint pass. */
struct FrGcGeo {
    float t1x, t1y, t2x, t2y, lh;
    int np;
};

struct FrLaunch {
    FrPlanDev *plans;
    FrCtl *ctl;
    int *slots;
    int2 *tiles;
    int *n_tiles;
    int *counters;
    FrExp *exp_q;
    FrFresh *fresh_q;
    FrPair *pairs;
    long pair_cap;
    int2 *demand;             /* (slot, cell) goal-connection evaluations queued this iteration */
    int2 *demand2;            /* the queued cells whose goal curves passed curvature and centres */
    FrGcGeo *demand2_geo;     /* their curves' set-up, beside them */
    long demand_cap;          /* of both */
    int *n_slots;             /* device: entries of `slots` */
    int grid_tiles, grid_items, grid_slots, grid_waves;
    int block;                /* threads per block of the iteration kernels (k_erase_mark aside) */
    int max_slots;
};

struct FrBestNodeArgs {
    int x, z;
    float heading, radius, max_herr;
};

/* kernel launchers (fr_kernels.hip) */
/* zeroed: the plans' slabs were just reset to zero (k_slab_reset), so the
 * goal planes' reset is skipped */
void fr_launch_setup(const FrLaunch &L, const int *d_slots, int n, int max_N, int max_PW, int max_PH,
                     int goal_mode, bool zeroed, hipStream_t st);
/* evaluate the goal-connection planes of every cell not yet evaluated */
void fr_launch_goal_complete(const FrLaunch &L, const int *d_slots, int n, int max_N, hipStream_t st);
/* entry: run k_begin first (the FastRRT handle, whose host may edit the plan
 * between iterations); without it an iteration starts from the state the
 * previous k_finish or the batch admission left (the batch engine) */
void fr_launch_iteration(const FrLaunch &L, hipStream_t st, bool entry = true);
/* the same iteration in three parts: before, k_expand, after */
void fr_launch_iteration_pre(const FrLaunch &L, hipStream_t st, bool entry = true);
void fr_launch_iteration_expand(const FrLaunch &L, hipStream_t st);
void fr_launch_iteration_post(const FrLaunch &L, hipStream_t st, bool entry = true);
void fr_launch_clear(const FrLaunch &L, int slot, int N, hipStream_t st);
void fr_launch_count(const FrLaunch &L, int slot, int N, int *d_out, hipStream_t st);
void fr_launch_goal_scan(const FrLaunch &L, int slot, int N, int gx, int gz, float gh, float tol, float herr,
                         hipStream_t st);
/* gate: 0 = always (API calls), 1 = plans with an optimisation round pending,
 * 2 = plans whose goal was reached (final extraction) */
enum { FR_GATE_NONE = 0, FR_GATE_OPT = 1, FR_GATE_FOUND = 2 };
void fr_launch_best_node(const FrLaunch &L, const int *d_slots, int n, int max_N, const FrBestNodeArgs *ov,
                         hipStream_t st, int gate = FR_GATE_NONE);
void fr_launch_walk(const FrLaunch &L, const int *d_slots, int n, hipStream_t st, int gate = FR_GATE_NONE);
/* per-entry cap on the optimisation rounds of a batch finalisation */
void fr_launch_fin_limits(const FrLaunch &L, const int *d_list, const int *d_limits, int n, hipStream_t st);
/* pinned-host records of finished plans (control block, then the path from byte FR_PACK_PATH_OFF) */
#define FR_PACK_PATH_OFF 512
void fr_launch_fin_pack(const FrLaunch &L, const int *d_list, int n, uint8_t *pack, long stride, int path_max,
                        hipStream_t st);
void fr_launch_opt_round(const FrLaunch &L, const int *d_slots, int n, int max_path, int max_N, hipStream_t st,
                         int gate = FR_GATE_NONE);
void fr_launch_goal_scan_multi(const FrLaunch &L, const int *d_slots, int n, int max_N, int gate, hipStream_t st);
void fr_launch_batch_start(const FrLaunch &L, const int *d_slots, int n, hipStream_t st);
void fr_launch_slab_reset(const FrLaunch &L, const int *d_slots, int n, hipStream_t st);
/* control blocks of retired plans back to IDLE, one launch for any number of
 * slots (the slot list travels in the kernel arguments) */
void fr_launch_ctl_idle(FrCtl *ctl, const int *slots, int n, hipStream_t st);
/* bytes (a multiple of 4, both ends 4-byte aligned) copied by a kernel on
 * the stream; used between device memory and mapped pinned host memory,
 * where an asynchronous copy call has been seen to hold the calling thread
 * for 6-7 ms now and then */
void fr_launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
void fr_launch_graph_add(const FrLaunch &L, int slot, int x, int z, float h, int px, int pz, float cost, int type,
                         int count_parent, hipStream_t st);
void fr_launch_set_collision(const FrLaunch &L, int slot, int x, int z, int npx, int npz, float h, float cost,
                             hipStream_t st);
void fr_launch_derivate(const FrLaunch &L, int slot, float steer, float path_size, float v, int x, int z,
                        int *d_out, hipStream_t st);
/* CudaGraph::canConnectToGoal: d_out[0] = curve result, d_out[1] = goal pixel blocked */
void fr_launch_hermite_probe(const FrLaunch &L, int slot, int x, int z, int gx, int gz, int goal_heading,
                             float max_curv, float *d_out, hipStream_t st);
/* check_kinematic_connection_start_end (kinematics.cu:233-306) for n queries
 * (start x, start z, end x, end z), one lane each */
void fr_launch_feasible_connections(const FrLaunch &L, int slot, const int4 *d_q, int n, int velocity,
                                    uint8_t *d_ok, double *d_heading, float *d_cost, hipStream_t st);
void fr_launch_accept(const FrLaunch &L, int slot, int N, hipStream_t st);
void fr_launch_solve_collisions(const FrLaunch &L, int slot, int N, int nodes, hipStream_t st);
void fr_launch_region_count(const FrLaunch &L, int slot, int N, uint32_t *d_out, hipStream_t st);

/* Distance-to-goal channel (pixel channel 1) of a frame (fr_frame.hip). */
void fr_launch_frame_distance_to_goal(float *d_xyz, int W, int H, int gx, int gz, hipStream_t st);
/* Safe-zone bit 0x100 of channel 2 for a (2 mz + 1) x (2 mx + 1) footprint;
 * d_near: W * H bytes of scratch (fr_frame.hip). */
void fr_launch_frame_safe_zone(float *d_xyz, const float *d_costs, int n_classes, int W, int H, int mx, int mz,
                               uint8_t *d_near, hipStream_t st);

/* Potential-field intrinsic costs of one grid (fr_apf.hip). */
struct FrApfArgs {
    int W, H;
    const float *xyz;       /* frame, three floats per pixel (class id first) */
    const float *costs;     /* class cost table */
    int n_classes;
    int lbx, lbz, ubx, ubz; /* ego box */
    float *intrinsic;       /* W*H, updated in place */
};
#define FR_APF_MAX_WINDOW 2048
/* half-width of the repulsion window actually read: min(radius, max(W, H) - 1), 0 if radius < 1 */
int fr_apf_window(int radius, int W, int H);
/* scratch: d_bits H*ceil(W/64) words, d_tab (w+1)^2 floats, d_span w+1 ints, w = fr_apf_window */
void fr_launch_apf_repulsion(const FrApfArgs &a, float kr_half, int radius, uint64_t *d_bits, float *d_tab,
                             int *d_span, hipStream_t st);
void fr_launch_apf_attraction(const FrApfArgs &a, float ka_half, int gx, int gz, hipStream_t st);
/* The repulsion with its scratch in a file-scope device array of
 * FR_APF_SCRATCH_BYTES per GPU (no allocation); throws when
 * fr_apf_scratch_bytes(W, H, radius) exceeds it.  Synchronous: returns after
 * the kernels have finished; calls on one device are serialised. */
#define FR_APF_SCRATCH_BYTES (4u << 20)
size_t fr_apf_scratch_bytes(int W, int H, int radius);
void fr_launch_apf_repulsion_noalloc(const FrApfArgs &a, float kr_half, int radius, hipStream_t st);

/* Optional per-kernel timing with HIP events recorded on the launching
 * stream around the kernels named below. */
enum FrProfKernel { FR_PROF_GOAL_CONNECTION = 0, FR_PROF_EXPAND = 1, FR_PROF_CONNECT = 2, FR_PROF_ITERATION = 3,
                    FR_PROF_N = 4 };
void fr_prof_enable(bool on);
bool fr_prof_on();
/* with profiling on, capture the batch iteration in three parts so that
 * k_expand is timed alone (otherwise the one graph is timed whole) */
void fr_prof_split(bool on);
bool fr_prof_split_on();
void fr_prof_begin(int k, hipStream_t st);
void fr_prof_end(int k, hipStream_t st);
/* synchronises the recorded events; returns launches and total ms */
void fr_prof_read(int k, long *launches, double *total_ms);
void fr_prof_reset();
/* one launch of kernel k timed elsewhere (events inside a captured graph) */
void fr_prof_add(int k, double ms);

/* error plumbing: every HIP call of the library goes through FR_CHECK */
void fr_set_error(const std::string &msg);
const char *fr_get_error();
#define FR_CHECK(expr)                                                                                   \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) {                                                                          \
            fr_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                             \
            throw std::runtime_error(std::string("fastrrt: ") + #expr + ": " + hipGetErrorString(_e));   \
        }                                                                                                \
    } while (0)

/* Parameters one plan is created with. */
struct FrPlanParams {
    int W = 0, H = 0;
    int min_dx = 0, min_dz = 0;   /* full min-distance (halved internally) */
    int lbx = -1, lbz = -1, ubx = -1, ubz = -1;
    float og_x = 0, og_y = 0;
    float v = 1.0f, max_path = 40.0f, dist_tol = 20.0f, herr_tol = 0.0f;
    float max_curv_loop = 0.0f, max_curv_goal = 0.8f;
    double max_steer_rad = 0, lr = 0;
    int sx = 0, sz = 0, gx = 0, gz = 0;
    float sh = 0, gh = 0;
    int region_side = 128;
    int smart = 0;
    int max_iter = 0;
    int exp_budget = 0;           /* batch: expected rollouts per iteration (0: every eligible node) */
    uint64_t seed = 0, stream = 0;
    int path_cap = 0;             /* 0: W*H + 1 */
    bool with_intrinsic = false;
    int gc_lazy = 0;              /* goal-connection planes evaluated on demand */
    int opt_rounds = 0;           /* batch: path_optimize rounds after the goal is reached */
};

/* Device planes of one plan (one allocation, sub-allocated). */
class FrSlab {
  public:
    FrSlab() = default;
    ~FrSlab();
    FrSlab(const FrSlab &) = delete;
    FrSlab &operator=(const FrSlab &) = delete;
    void allocate(int W, int H, int mx, int mz, int path_cap, bool with_intrinsic, hipStream_t st);
    void release();
    void set_footprint(int mx, int mz, hipStream_t st);
    /* device bytes allocate() takes for these arguments */
    static size_t bytes_for(int W, int H, int mx, int mz, int path_cap, bool with_intrinsic);
    /* back to the state allocate() leaves: planes zero, claims all-ones */
    void reset(hipStream_t st);
    size_t bytes() const { return bytes_; }
    bool allocated() const { return base_ != nullptr; }
    int W = 0, H = 0, pad = 0, PW = 0, PH = 0, TW = 0, mx = -1, mz = -1, path_cap = 0;
    bool with_intrinsic = false;
    uint8_t *type = nullptr;
    int32_t *derive = nullptr, *px = nullptr, *pz = nullptr;
    float *heading = nullptr, *cost = nullptr, *intrinsic = nullptr, *direct = nullptr;
    uint8_t *gc_flag = nullptr;
    float *gc_heading = nullptr, *gc_cost = nullptr;
    uint32_t *gc_done = nullptr, *gc_bits = nullptr;
    /* planes a new plan needs zero: [zero_base, zero_base + zero_bytes) */
    void *zero_base() const { return base_; }
    size_t zero_bytes() const { return zero_end_; }
    unsigned long long *claim = nullptr;
    int32_t *escr = nullptr; /* 4 * W * H: the erase's node list, index map and jump buffers */
    uint8_t *mark = nullptr, *dist = nullptr, *dist_tmp = nullptr, *trav = nullptr;
    uint32_t *region = nullptr;
    float4 *path = nullptr, *pd = nullptr;

  private:
    void *base_ = nullptr;
    void *dist_base_ = nullptr;
    size_t bytes_ = 0;
    size_t zero_end_ = 0;
};

/* Device pointers a plan reads its obstacles from. */
struct FrFrameView {
    const float *xyz = nullptr;   /* W*H float3 */
    const float *costs = nullptr; /* class-cost table */
    int n_classes = 0;
};

/* A device-resident frame: float3 pixels + class-cost table. */
struct FrFrame {
    int W = 0, H = 0, n_classes = 0;
    float *d_xyz = nullptr;
    float *d_costs = nullptr;
    std::vector<float> h_costs;
    ~FrFrame();
};

class FrEngine {
  public:
    FrEngine(int max_slots, long queue_cap, hipStream_t stream = nullptr);
    ~FrEngine();
    FrEngine(const FrEngine &) = delete;
    FrEngine &operator=(const FrEngine &) = delete;

    hipStream_t stream() const { return st_; }
    const FrLaunch &launch() const { return L_; }
    /* Launch shapes (blocks of FR_BLOCK threads) of the iteration's kernels:
     * the one-wave-per-item kernels, the tile scans, the one-lane-per-item
     * kernels.  Every one of them strides over its work, so the results do
     * not depend on these; a caller re-captures its graphs after a change. */
    void set_grids(int waves, int tiles, int items);
    /* threads per block of the same kernels: 64, 128, 192 or 256 (every one
     * of them works in whole waves and sizes its LDS for 256 threads) */
    void set_block(int threads);
    int max_slots() const { return max_slots_; }
    /* device bytes an engine of these capacities allocates */
    static size_t bytes_for(int max_slots, long queue_cap);

    /* describe plan `slot` (slab + params + frame) and upload its descriptor */
    void bind(int slot, FrSlab &slab, const FrPlanParams &p, const FrFrameView &frame, bool upload = true);
    FrPlanDev &host_plan(int slot) { return h_plans_[slot]; }
    void upload_plan(int slot);
    void upload_plans();                           /* every descriptor, one copy */
    /* copy n ints to device memory through pinned staging (asynchronous) */
    void upload_ints(int *d_dst, const int *src, int n);

    /* control blocks */
    void reset_ctl(int slot, int status);
    void read_ctl(int slot, FrCtl *out);           /* synchronous */
    void write_ctl(int slot, const FrCtl &c);

    /* the set of plans the iteration kernels walk (asynchronous, stream-ordered) */
    void set_active(const std::vector<int> &slots);
    int n_active() const { return (int)active_.size(); }
    const int *d_active() const { return L_.slots; }

    /* per-plan precompute; zeroed: the slabs were just reset to zero */
    void setup(const std::vector<int> &slots, int goal_mode, bool zeroed = false);
    void goal_complete(int slot);                  /* finish a lazy goal-connection pass */
    void iteration();                              /* one loop() for every active plan */
    void sync();

    long queue_cap() const { return queue_cap_; }
    int *d_scratch_int() { return d_scratch_; }

  private:
    void allocate();
    void release();
    int max_slots_;
    long queue_cap_;
    hipStream_t st_;
    bool own_stream_;
    FrLaunch L_{};
    std::vector<FrPlanDev> h_plans_;
    /* pinned ring of FR_PLAN_RING descriptor tables; an upload waits for
     * the copy FR_PLAN_RING uploads back.  One entry: each admission waits
     * for the previous one's copy, which keeps the host from admitting ahead
     * of the device. */
    static const int FR_PLAN_RING = 1;
    FrPlanDev *h_plan_stage_ = nullptr, *d_plan_stage_ = nullptr; /* mapped: host and device addresses */
    hipEvent_t plan_ev_[FR_PLAN_RING] = {};
    int plan_next_ = 0;
    /* pinned ring of FR_STAGE_RING int arrays of max_slots + 1 entries */
    static const int FR_STAGE_RING = 8;
    int *h_int_stage_ = nullptr, *d_int_stage_ = nullptr;
    hipEvent_t int_ev_[FR_STAGE_RING] = {};
    int int_next_ = 0;
    std::vector<int> active_;
    std::vector<int2> h_tiles_;
    /* active lists in one device block [n_tiles, n_slots, 2 pad, slots (max_slots, even), tiles],
     * uploaded in one copy from a pinned ring */
    static const int FR_ACT_RING = 4;
    void *d_act_ = nullptr;
    uint8_t *h_act_stage_ = nullptr, *d_act_stage_ = nullptr;
    size_t act_bytes_ = 0, act_tiles_off_ = 0;
    long act_max_tiles_ = 0;
    hipEvent_t act_ev_[FR_ACT_RING] = {};
    int act_next_ = 0;
    int *d_setup_slots_ = nullptr;
    int *d_scratch_ = nullptr;
};

#endif
