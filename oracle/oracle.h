/*
 * oracle.h — CPU restatement of FastRRT's parallel planning loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the parity tests,
 * the smoke check and the bench's CPU-baseline leg compare the gfx950 path
 * against.  The product (libfastrrt.so and the fastrrt_amd package) never
 * links, loads or calls it.
 *
 * Each function restates one reference function (cited file:line under the
 * reference tree).  Where the reference is racy on the GPU, the oracle fixes
 * one deterministic outcome, and the HIP path implements the same rule:
 *   - expansion reads a snapshot of the grid taken before the kernel;
 *   - when several expansions end on one cell, the lowest source cell wins;
 *   - subtree erase decides on the pre-erase snapshot, then applies;
 *   - collision re-parenting resets child counts before incrementing parents;
 *   - potential-field repulsion sets every obstacle cell before any sum
 *     reaches it, and each cell sums its contributions in ascending source
 *     cell order;
 *   - both argmin reductions break ties on the lowest cell index (and, inside
 *     one node's goal window, on the last window cell in scan order, which is
 *     what the reference's sequential per-thread scan keeps).
 *
 * Parity status of each part (see the tests under tests/):
 *   kinematic rollout, footprint feasibility, Hermite checks, DAG check,
 *   collision erase, best node, goal test, dump format, Philox: pinned by the
 *   reference's known-answer tests and closed forms.
 *   density gating (smart expansion): pinned by brute-force recount only.
 *   potential-field costs: pinned by the closed form of one isolated obstacle
 *   and by brute-force float64 sums on small grids.
 *   Hermite test hook (canConnectToGoal): pinned by the reference's three
 *   gtests, the straight-connection closed form, the pinned direct-connection
 *   check on the same curves and the curvature measure from finite
 *   differences of the curve.
 *   path optimisation (og_rrt_path_optimize): pinned by hand-worked chains
 *   whose candidate costs are closed forms (straight connections on a free
 *   frame: last improving j per i, first i of maximal gain, strict cost
 *   comparison, cost update of the spliced node) and by an enumeration of
 *   the selection rules over solved scenes (tests/test_oracle_pins.py).
 *   goal-window argmin (og_find_best_goal_direct_connection): pinned by a
 *   brute-force enumeration of the half-open window on small grids and by
 *   known answers for the window edges, the floor of 100 x total and both
 *   tie rules.
 *   FP32 rollout: within 1e-5 relative of a float64 recomputation with libm
 *   (og_kinematic_new_path_pose).
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_graph {
    int W, H;
    /* per cell (reference int4 {parent_x, parent_z, type, derive}) */
    int32_t *px, *pz, *type, *derive;
    /* per cell (reference float4 {heading, cost, intrinsic, direct}) */
    float *heading, *cost, *intrinsic, *direct;
    /* goal direct-connection data (reference float3 {flag, heading, cost}) */
    float *gc_flag, *gc_heading, *gc_cost;
    /* density regions */
    int region_side, dens_w, dens_h;
    uint32_t *region;
    /* physical params (reference double[9]) */
    double rate_w, inv_rate_w, rate_h, inv_rate_h, max_steer_rad, max_steer_deg, lr, max_curv;
    int has_physical;
    /* search params (reference int[...]) */
    int min_dist_x, min_dist_z, lb_x, lb_z, ub_x, ub_z;
    /* class costs */
    float *class_costs;
    int n_classes;
    /* coordinate origin (reference _ogCoordinateStart) */
    float og_x, og_y, og_h;
    /* flags */
    int goal_reached, node_collision;
    /* best direct connection: parent (x, z, heading, cost_to_cell), child (x, z, heading, total) */
    float best_parent[4], best_child[4];
    /* diagnostics */
    uint64_t n_footprint_checks, n_footprint_samples, n_rollouts, n_hermite_walks;
} og_graph;

og_graph *og_graph_new(int W, int H);
void og_graph_free(og_graph *g);
void og_set_region_side(og_graph *g, int side);
void og_set_physical_params(og_graph *g, float perc_w, float perc_h, double max_steer_rad,
                            float vehicle_len, float max_curvature);
void og_set_search_params(og_graph *g, int min_dist_x, int min_dist_z, int lb_x, int lb_z, int ub_x, int ub_z);
void og_set_class_costs(og_graph *g, const float *costs, int n);
void og_set_coordinate_start(og_graph *g, int x, int z, float heading);

void og_add(og_graph *g, int x, int z, float heading, int px, int pz, float cost);
void og_add_temporary(og_graph *g, int x, int z, float heading, int px, int pz, float cost);
void og_remove(og_graph *g, int x, int z);
void og_clear(og_graph *g);
void og_set_collision(og_graph *g, int x, int z, int npx, int npz, float heading, float cost);
void og_get_parent(og_graph *g, int x, int z, int *px, int *pz);
float og_get_cost(og_graph *g, int x, int z);
int og_get_type(og_graph *g, int x, int z);
unsigned og_count(og_graph *g, int type);
unsigned og_count_all(og_graph *g);

/* frame: W*H float3 (class id, distance-to-goal, traversability bits) */
int og_feasible(const og_graph *g, const float *frame, int x, int z, float angle);
void og_kinematic_new_path(og_graph *g, const float *frame, int sx, int sz, float steer, float path_size,
                           float v, float out[4]);
/* as og_kinematic_new_path, also returning the final FP32 map pose (x, y,
 * heading) and the steps executed (test hook of the FP64 recomputation) */
void og_kinematic_new_path_pose(og_graph *g, const float *frame, int sx, int sz, float steer, float path_size,
                                float v, float out[4], double pose[4]);
void og_derivate_node(og_graph *g, const float *frame, float steer, float path_size, float v, int x, int z,
                      int out[2]);
void og_accept_derived_nodes(og_graph *g);
void og_solve_collisions(og_graph *g);
void og_count_regions(og_graph *g);
/* One expansion kernel over all NODE cells with the snapshot/lowest-index
 * rules.  mode 0 = plain (expandTree), 1 = smart (smartExpansion). */
void og_expand(og_graph *g, const float *frame, float max_path, float v, int frontier, int force, int smart,
               uint64_t seed, uint64_t stream, uint32_t iter);

float og_check_direct_connection(og_graph *g, const float *frame, float max_curv, int x, int z,
                                 float local_heading, int gx, int gz, float gh, int safe_zone, int dist_proc);
float og_check_direct_connection_host(og_graph *g, const float *frame, float max_curv, int x, int z,
                                      float local_heading, int gx, int gz, float gh, int safe_zone);
void og_process_direct_goal_connection(og_graph *g, const float *frame, int gx, int gz, float gh,
                                       float max_curv);
int og_find_best_goal_direct_connection(og_graph *g, const float *frame, float radius, int safe_zone);
int og_check_goal_reached(og_graph *g, int gx, int gz, float gh, float tol, float max_herr);
void og_find_best_node(og_graph *g, const float *frame, float heading, float radius, int x, int z,
                       float max_herr, int out[2]);
int og_check_graph_consistent(og_graph *g);
float og_can_connect_hermite(og_graph *g, const float *frame, float max_curv, int x, int z, int gx, int gz,
                             int goal_heading, int *goal_blocked);
void og_apf_repulsion(og_graph *g, const float *frame, float kr, int radius);
void og_apf_attraction(og_graph *g, const float *frame, float ka, int gx, int gz);
int og_check_feasible_connection(og_graph *g, const float *frame, int sx, int sz, int ex, int ez, float v);
int og_check_feasible_connection_ex(og_graph *g, const float *frame, int sx, int sz, int ex, int ez, float v,
                                    double *final_heading, float *path_cost);
int og_interpolate_hermite(int W, int H, int x1, int z1, float h1, int x2, int z2, float h2, float *out,
                           int cap);

void og_dump_graph(og_graph *g, const char *path);
void og_read_dump(og_graph *g, const char *path);

/* Planner (reference FastRRT, src/fastrrt.cpp). */
typedef struct og_rrt og_rrt;
og_rrt *og_rrt_new(int W, int H, float perc_w, float perc_h, float max_steer_deg, float vehicle_len,
                   int lb_x, int lb_z, int ub_x, int ub_z, float max_curvature);
void og_rrt_free(og_rrt *r);
og_graph *og_rrt_graph(og_rrt *r);
void og_rrt_set_seed(og_rrt *r, uint64_t seed, uint64_t stream);
void og_rrt_set_plan_data(og_rrt *r, const float *frame, const float *class_costs, int n_classes, int sx, int sz,
                          float sh, int gx, int gz, float gh, float v, int min_dx, int min_dz, int timeout_ms,
                          float max_path, float dist_tol, float herr_tol);
void og_rrt_search_init(og_rrt *r);
int og_rrt_loop(og_rrt *r, int smart);
int og_rrt_goal_reached(og_rrt *r);
void og_rrt_load_graph_state(og_rrt *r, const char *path);
int og_rrt_get_planned_path(og_rrt *r, float *out, int cap);
int og_rrt_path_optimize(og_rrt *r);
int og_rrt_interpolate_planned_path(og_rrt *r, float *out, int cap);
int og_rrt_iterations(og_rrt *r);

#ifdef __cplusplus
}
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif
float og_sinf(float x);
float og_cosf(float x);
float og_tanf(float x);
float og_atanf(float x);
float og_atan2f(float y, float x);
int og_trunc_d(double v);
void og_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]);
void og_expansion_draws(uint64_t seed, uint64_t stream, uint32_t iter, uint32_t cell, float out[2]);
#ifdef __cplusplus
}
#endif
