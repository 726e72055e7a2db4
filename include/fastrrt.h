/*
 * fastrrt.h — C ABI of libfastrrt.so (MI355X / gfx950 build).
 *
 * Drop-in for the reference's ctypes boundary: the FastRRT handle ABI of
 * src/python_bind.cpp and the graph ABI of src/python_cuda_graph_bind.cpp
 * keep their symbol names, argument order and meaning.  The one change is
 * the frame: the reference takes a libdriveless SearchFrame*; here the frame
 * is an opaque handle made by fastrrt_frame_create() from host float3 pixels
 * (channel 0 = class id, 1 = distance to goal, 2 = traversability bits) and
 * a class-cost table.  Like the reference, a plan borrows its frame: keep the
 * frame alive until the search is done.
 *
 * Errors: the reference throws C++ exceptions through extern "C"
 * (src/fastrrt.cpp:89, src/cuda/graph_list.cu:100).  Here every entry point
 * catches, keeps the message for fastrrt_last_error(), and returns false /
 * NULL / a no-op.  Thread safety: one handle per host thread, as in the
 * reference; each handle owns its own HIP stream.
 */
#ifndef FASTRRT_C_ABI_H
#define FASTRRT_C_ABI_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Frames (replace libdriveless SearchFrame*)                              */
/* ---------------------------------------------------------------------- */

/* New: the reference passes a SearchFrame* (python_bind.cpp:38, :263). */
void *fastrrt_frame_create(int width, int height, const float *xyz3, const float *class_costs, int n_classes);
void fastrrt_frame_destroy(void *frame);
/* Re-upload the pixels of an existing frame (same size); returns false on error. */
bool fastrrt_frame_update(void *frame, const float *xyz3);
/* The frame's pixels back to the host (W*H*3 floats).  No reference
 * counterpart: the reference reads libdriveless SearchFrame memory directly. */
bool fastrrt_frame_read(void *frame, float *xyz3);
/* libdriveless SearchFrame::processDistanceToGoal (called at
 * unittests/src/test_graph_can_connect_hermite.cpp:51; library absent from
 * the reference tree): channel 1 of every pixel = sqrtf of its integer
 * squared pixel distance to (goal_x, goal_z), the curve length
 * can_connect_to_goal_using_hermite reads. */
bool fastrrt_frame_process_distance_to_goal(void *frame, int goal_x, int goal_z);
/* The safe-zone bit 0x100 of channel 2, which the goal connection reads at a
 * curve's start cell (graph_direct_goal_connection.cu:124-135,
 * isSafeZoneChecked = true at fastrrt.cpp:145) and which libdriveless (absent
 * from the reference tree) sets: set where no obstacle pixel (class cost < 0
 * or a class outside the table) lies within Chebyshev distance
 * ceil(hypot(min_dist_x / 2, min_dist_z / 2)) + 1, so that the footprint is
 * free at every heading; cleared elsewhere.  Pixels outside the frame count
 * as free. */
bool fastrrt_frame_process_safe_zone(void *frame, int min_dist_x, int min_dist_z);
/* Message of the last failed call on this thread ("" if none). */
const char *fastrrt_last_error(void);

/* ---------------------------------------------------------------------- */
/* FastRRT handle (src/python_bind.cpp)                                    */
/* ---------------------------------------------------------------------- */

/* python_bind.cpp:7-30.  class_costs is accepted and unused, as there. */
void *fastrrt_initialize(int width, int height, float perception_width_m, float perception_height_m,
                         float max_steering_angle_deg, float vehicle_length, int lower_bound_x, int lower_bound_z,
                         int upper_bound_x, int upper_bound_z, float *class_costs, float max_curvature);
/* python_bind.cpp:32-36 */
void fastrrt_destroy(void *ptr);
/* python_bind.cpp:38-64; frame = fastrrt_frame_create() handle */
void set_plan_data(void *ptr, void *frame, int start_x, int start_z, float start_heading_rad, int goal_x,
                   int goal_z, float goal_heading_rad, float velocity_m_s, int min_dist_x, int min_dist_z,
                   int timeout_ms, float max_path_size, float dist_to_goal_tolerance,
                   float heading_error_tolerance_rad);
/* python_bind.cpp:66-70 */
bool goal_reached(void *ptr);
/* python_bind.cpp:72-76 */
void search_init(void *ptr, bool copy_intrinsic_costs_from_frame);
/* python_bind.cpp:77-81 */
bool loop(void *ptr, bool smart_expansion);
/* python_bind.cpp:83-87 */
bool path_optimize(void *ptr);
/* python_bind.cpp:89-107: [n, (x, z, type) x n] */
int *export_graph_nodes(void *ptr);
/* python_bind.cpp:109-112 (declared float* there; frees the int[] above) */
void release_export_graph_nodes(float *ptr);
/* python_bind.cpp:138-143: [n, (x, z, heading) x n] */
float *get_planned_path(void *ptr);
/* python_bind.cpp:145-150 */
float *interpolate_planned_path(void *ptr);
/* python_bind.cpp:152-155 */
void release_planned_path_data(float *ptr);
/* python_bind.cpp:157-171 */
float *interpolate_planned_path_p(void *ptr, float *p, int size);
/* python_bind.cpp:173-180 */
float *ideal_curve(void *ptr, int goal_x, int goal_z, float goal_heading_rad);
/* python_bind.cpp:182-186 */
void compute_region_debug_performance(void *ptr);
/* python_bind.cpp:188-192 */
void save_current_graph_state(void *ptr, const char *filename);
/* python_bind.cpp:194-198 */
void load_graph_state(void *ptr, const char *filename);

/* New: Philox key and stream of the next search_init (the reference seeds
 * cuRAND from the wall clock, random_number_gen.cu:23-40). */
void fastrrt_set_seed(void *ptr, uint64_t seed, uint64_t stream);
/* New: side in pixels of the density regions of smart expansion (the
 * reference fixes 128, graph_expansion_search_zones.cu:35-40; a 1024x1024
 * search uses 32).  At least 16; kept across set_plan_data. */
void fastrrt_set_density_region(void *ptr, int side);
/* New: loop() iterations since search_init. */
int fastrrt_iterations(void *ptr);
/* New: work of the last loop(): rollouts, nodes added, nodes goal-checked. */
void fastrrt_iteration_stats(void *ptr, int *n_expansions, int *n_new, int *n_checked);
/* New: binary counterpart of save_current_graph_state (graph.cu:496-520):
 * copies W*H entries of each grid plane to the caller's buffers (any may be
 * NULL); goal_* are the goal-connection planes of set_plan_data. */
void fastrrt_export_state(void *ptr, int32_t *parent_x, int32_t *parent_z, int32_t *type, int32_t *child_count,
                          float *heading, float *cost, uint8_t *goal_flag, float *goal_heading, float *goal_cost);

/* New: HIP-event timing of the hot kernels, recorded on the stream each is
 * launched on.  kernel: 0 = goal-connection precompute, 1 = expansion
 * rollout, 2 = per-loop goal connection, 3 = one whole loop iteration. */
void fastrrt_profile_enable(bool on);
void fastrrt_profile_reset(void);
bool fastrrt_profile_read(int kernel, long *launches, double *total_ms);
/* New: with profiling on, a batch's iteration graph is captured as three
 * graphs (before k_expand, k_expand, after) so that kernel 1 times k_expand
 * alone.  Off (the default), the iteration runs as the one graph it runs
 * unprofiled and kernel 3 times exactly that. */
void fastrrt_profile_split(bool on);

/* ---------------------------------------------------------------------- */
/* Graph handle (src/python_cuda_graph_bind.cpp)                           */
/* ---------------------------------------------------------------------- */

/* python_cuda_graph_bind.cpp:7-34; costs = [count, c0, c1, ...] */
void *cudagraph_initialize(int width, int height, float perception_width_m, float perception_height_m,
                           float max_steering_angle_deg, float vehicle_length, int min_distance_x,
                           int min_distance_z, int lower_bound_x, int lower_bound_z, int upper_bound_x,
                           int upper_bound_z, float *segmentation_class_cost, float max_curvature);
/* :35-39 */
void cudagraph_destroy(void *ptr);
/* :41-46 (apf.cu:23-98,123-138): potential-field repulsion into the intrinsic
 * costs; every obstacle pixel is set to 100*kr/2 before the sums, each pixel
 * sums its contributions in ascending source order; a window wider than
 * 2048 px (radius and grid side both larger) is refused (last error set) */
void compute_apf_repulsion(void *ptr, void *frame, float kr, int radius);
/* No counterpart in the reference: compute_apf_repulsion with the same
 * kernels and the same result, for callers that cannot allocate device
 * memory (code under HIP graph capture on another stream, tight serving
 * loops, harnesses that forbid hipMalloc).  Its scratch (edge bit plane,
 * offset table, disc spans) lives in a file-scope device array of
 * fastrrt_apf_noalloc_scratch_capacity() bytes (4 MiB) per GPU, held until
 * the process ends; the call allocates and frees nothing.  A grid and radius
 * needing more (fastrrt_apf_noalloc_scratch_bytes: any radius up to
 * 512x512, radius <= 1006 at 1024x1024, <= 956 at 2048x2048) are refused:
 * false, with fastrrt_last_error() saying so.
 * Synchronous.  Every call on one GPU shares that GPU's array: calls from
 * several host threads on one GPU are serialised inside the library, and
 * the array keeps the last call's contents (nothing in it is read before
 * the call writes it). */
bool compute_apf_repulsion_noalloc(void *ptr, void *frame, float kr, int radius);
size_t fastrrt_apf_noalloc_scratch_bytes(int width, int height, int radius);
size_t fastrrt_apf_noalloc_scratch_capacity(void);
/* :48-53 (apf.cu:100-121,140-155): potential-field attraction */
void compute_apf_attraction(void *ptr, void *frame, float ka, int goal_x, int goal_z);
/* :55-73: W*H intrinsic costs */
float *get_intrinsic_costs(void *ptr);
/* :75-78 */
void destroy_intrinsic_costs_ptr(float *ptr);
/* :80-84 */
void add(void *ptr, int x, int z, float heading_rad, int parent_x, int parent_z, float cost);
/* :86-90 */
void add_temporary(void *ptr, int x, int z, float heading_rad, int parent_x, int parent_z, float cost);
/* :93-98 */
void derivate_node(void *ptr, void *frame, float steering_angle, int path_size, float velocity, int x, int z);
/* :100-104 */
void accept_derived_nodes(void *ptr, int goal_x, int goal_z, float goal_heading);
/* :105-109 */
bool check_in_graph(void *ptr, int x, int z);
/* :110-114 */
float get_heading(void *ptr, int x, int z);
/* :115-119 */
void clear(void *ptr);
/* :121-140: [n, (x, z, heading, type) x n] */
float *list_all(void *ptr);
/* :142-145 */
void free_list_all(float *p);
/* :147-153 (graph_expansion.cu:246-261, graph_hermite_connection.cu:11-136):
 * the Hermite curve from node (x, z) to the goal, its length read from the
 * frame's distance-to-goal channel (pixel channel 1), goal_heading truncated
 * to an int as at the reference's call; false when the goal pixel is an
 * obstacle, otherwise the curve's float result converted to bool as the
 * reference does (a failed curve, -1 or -2, answers true) */
bool can_connect_to_goal_using_hermite(void *ptr, void *frame, int x, int z, int goal_x, int goal_z,
                                       float goal_heading);
/* New (the reference keeps it on the C++ class, used by
 * unittests/src/test_kinematics.cpp:51-94): CudaGraph::checkFeasibleConnection
 * (kinematics.cu:308-326) -> check_kinematic_connection_start_end (:233-306),
 * an FP64 walk that steers toward the end cell and tests the footprint of
 * every cell it passes; true when it gets within 2 px of the end.  The start
 * heading is the graph's heading at the start cell; velocity is an int there.
 * The graph's class costs, footprint and ego box apply; the frame is read. */
bool cudagraph_check_feasible_connection(void *ptr, void *frame, int start_x, int start_z, int end_x, int end_z,
                                         int velocity_m_s);
/* The same for n queries (start x, start z, end x, end z per query, 4n ints)
 * in one launch, one lane per query: feasible[k] = 0/1 and, if the buffers
 * are not NULL, the walk's final heading and path cost (the reference's out
 * parameters, 0 unless the walk stopped on moving away from the end).
 * Returns n, or -1 on error (fastrrt_last_error). */
int cudagraph_check_feasible_connections(void *ptr, void *frame, const int32_t *queries, int n, int velocity_m_s,
                                         uint8_t *feasible, double *final_heading, float *path_cost);
/* :155-159 */
void solve_collisions(void *ptr);
/* :161-167 */
void process_direct_goal_connection(void *ptr, void *frame, int goal_x, int goal_z, float goal_heading_rad,
                                    float max_curvature);
/* :169-173 */
bool is_directly_connected_to_goal(void *ptr, int x, int z);
/* :175-179 */
float direct_connection_to_goal_cost(void *ptr, int x, int z);
/* :181-186 */
float direct_connection_to_goal_heading(void *ptr, int x, int z);
/* :188-191 */
void dump_graph_to_file(void *ptr, char *filename);
/* :193-197 */
void read_from_dump_file(void *ptr, char *filename);
/* :199-206 */
int *get_parent(void *ptr, int x, int z);
/* :208-210 */
void free_parent_data(int *data_ptr);
/* :212-215 */
float get_cost(void *ptr, int x, int z);
/* :217-220 */
int count_all(void *ptr);

/* New (the reference keeps these on the C++ class only): graph test hooks
 * used by the reference's unit tests (unittests/src/test_graph_*.cpp). */
void cudagraph_set_collision(void *ptr, int x, int z, int new_parent_x, int new_parent_z, float heading,
                             float cost);
int cudagraph_get_type(void *ptr, int x, int z);
int cudagraph_get_child_count(void *ptr, int x, int z);
int cudagraph_count(void *ptr, int type);
bool cudagraph_check_consistent(void *ptr);
/* findBestNode (graph_best_node.cu:134-178); out = (x, z) */
void cudagraph_find_best_node(void *ptr, void *frame, float heading, float radius, int x, int z,
                              float max_heading_error, int *out);
bool cudagraph_check_goal_reached(void *ptr, int goal_x, int goal_z, float heading, float tolerance,
                                  float max_heading_error);
/* canConnectToGoal with the curve's value (graph_hermite_connection.cu:11-136):
 * returns 1 when the goal pixel is an obstacle (no curve), 0 otherwise with
 * *result = the cost, -1 (collision or no points) or -2 (curvature); -1 on error */
int cudagraph_can_connect_hermite(void *ptr, void *frame, int x, int z, int goal_x, int goal_z, float goal_heading,
                                  float *result);
/* derivateNode returning the end cell (graph.h:177); out = (x, z) */
void cudagraph_derivate_node(void *ptr, void *frame, float steering_angle, float path_size, float velocity, int x,
                             int z, int *out);
/* as fastrrt_export_state, for a graph handle */
void cudagraph_export_state(void *ptr, int32_t *parent_x, int32_t *parent_z, int32_t *type, int32_t *child_count,
                            float *heading, float *cost, uint8_t *goal_flag, float *goal_heading, float *goal_cost);

/* ---------------------------------------------------------------------- */
/* Batched and served planning (new; no counterpart in the reference)      */
/* ---------------------------------------------------------------------- */

typedef struct fastrrt_plan_request {
    void *frame;              /* fastrrt_frame_create() handle, borrowed until the plan completes */
    int start_x, start_z;
    float start_heading_rad;
    int goal_x, goal_z;
    float goal_heading_rad;
    float velocity_m_s;
    int min_dist_x, min_dist_z;
    int lower_bound_x, lower_bound_z, upper_bound_x, upper_bound_z; /* ego box, -1 = none */
    float max_steering_angle_deg, vehicle_length, max_curvature;
    float max_path_size, dist_to_goal_tolerance, heading_error_tolerance_rad;
    int max_iterations;       /* loop() iterations before the plan fails (0: none) */
    int optimize_rounds;      /* path_optimize() calls after the goal is reached */
    int smart;                /* smart (density-gated) expansion */
    int region_side;          /* density region side in px for smart expansion (0: 128, as the reference) */
    double deadline_ms;       /* from submission; the plan fails when it passes (<= 0: none) */
    uint64_t seed, stream;    /* Philox key and stream */
    int64_t user_tag;
    int expansion_budget;     /* 0: every eligible node expands each iteration, as the reference;
                                 > 0: eligible nodes kept with probability budget / (eligible nodes of
                                 the previous iteration), a search the reference does not run */
} fastrrt_plan_request;

typedef struct fastrrt_plan_result {
    int64_t user_tag;
    int status;               /* 2 = path found, 3 = failed (iteration cap or deadline),
                                 4 = goal reached but no feasible node to extract a path from */
    int iterations;
    int path_len;             /* waypoints written (-1: the caller's buffer is too small) */
    int expansions;           /* rollouts the search ran */
    double submit_ms, done_ms; /* batch clock */
    float path_cost;          /* cost of the path's last node (cells traversed plus their intrinsic
                                 costs, the grid's cost field; 0 when no path) */
    /* work the search did (no counterpart in the reference): cells its
     * rollouts entered, each one footprint test; points of the goal-connection
     * curves evaluated (a curve's point count each); points of the curves
     * whose footprints were tested */
    int64_t work_rollout_cells, work_goal_points, work_goal_foot_points;
} fastrrt_plan_result;

/* A batch of up to max_plans concurrent plans whose grids total at most
 * max_cells cells.  Each loop iteration of all running plans is one
 * replay of a captured hipGraph. */
void *fastrrt_batch_create(int max_plans, int64_t max_cells);
void fastrrt_batch_destroy(void *batch);
/* Queue a request; returns a ticket >= 0, or -1 (see fastrrt_last_error). */
int fastrrt_batch_submit(void *batch, const fastrrt_plan_request *req);
/* Admit queued requests while slots and cells allow, run `iterations` loop
 * iterations of every running plan, retire finished plans (optimisation and
 * path extraction included).  Returns the plans still queued or running. */
int fastrrt_batch_step(void *batch, int iterations);
/* Pop one finished plan; its path as (x, z, heading) triples into `path`
 * (capacity `cap` waypoints).  Returns false when none is finished. */
bool fastrrt_batch_poll(void *batch, fastrrt_plan_result *result, float *path, int cap);
/* Step until every submitted plan has finished. */
void fastrrt_batch_drain(void *batch);
/* Milliseconds since the batch was created (the clock of submit_ms/done_ms). */
double fastrrt_batch_clock_ms(void *batch);
/* Launch shapes, in blocks of 256 threads, of the batched iteration's
 * kernels: grid_waves for the one-wave-per-item kernels (expansion rollouts,
 * goal-plane curves, goal connection; a multiple of 8, one share per XCD),
 * grid_tiles for the 1024-cell tile scans, grid_items for the
 * one-lane-per-item kernels.  Results do not depend on them.  The defaults
 * are 4096, and for grid_tiles and grid_items the batch's tile and queue
 * capacity up to 2048; fastrrt_amd.autotune measures them per grid size.
 * No counterpart in the reference, which launches one thread per cell.
 * get_launch writes the three values in that order. */
bool fastrrt_batch_set_launch(void *batch, int grid_waves, int grid_tiles, int grid_items);
bool fastrrt_batch_get_launch(void *batch, int *grid_waves_tiles_items);
/* Threads per block of the same kernels (the erase's marking kernel aside):
 * 64, 128, 192 or 256 (the default).  The grids stay counted in blocks.
 * Results do not depend on it; get returns -1 on a null handle. */
bool fastrrt_batch_set_block_threads(void *batch, int threads);
int fastrrt_batch_get_block_threads(void *batch);

/* Device memory of a batch, measured with hipMemGetInfo.  At creation the
 * cell budget is max_cells or what the free memory holds beside the reserve
 * (the engine's queues first, then cells at the bytes a cell of a plan's
 * planes and of the queues takes), whichever is smaller; a request larger
 * than the budget is refused at submission.  Before a new plan's planes are
 * allocated (a plan of a size class with no pooled planes) the free memory
 * is measured again: when it cannot hold them beside the reserve, pooled
 * planes of other sizes are freed, and if it still cannot, the plan waits
 * in the queue (memory_stalls counts such holds) until a running plan gives
 * its planes back; with no plan running, fastrrt_batch_step fails with an
 * error instead.  No counterpart in the reference (one plan per process). */
typedef struct fastrrt_batch_hbm {
    uint64_t total_bytes, free_bytes_at_create, free_bytes_last, reserve_bytes;
    uint64_t engine_bytes;    /* queues and tables of the iteration */
    uint64_t slab_bytes;      /* plans' planes, in use and pooled */
    int64_t requested_cells, cell_budget, cells_in_use;
    int64_t memory_stalls;
    int32_t running, queued;
} fastrrt_batch_hbm;
/* Fills `out`; free_bytes_last is measured by this call. */
bool fastrrt_batch_memory(void *batch, fastrrt_batch_hbm *out);
/* Bytes left free for the rest of the process when planes are allocated
 * (default 2 GiB, at most a 16th of the device). */
bool fastrrt_batch_set_hbm_reserve(void *batch, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
