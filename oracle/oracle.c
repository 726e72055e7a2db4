/*
 * oracle.c — CPU restatement of FastRRT's parallel planning loop.
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/_build/liboracle.so and used by
 * tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py as the
 * checker.  Nothing in the product links or calls it.
 *
 * Every function below follows the reference function cited beside it
 * (paths relative to the reference tree).  The kernels of the reference run
 * one GPU thread per grid cell; here each kernel body is a loop over cells,
 * and the reference's data races are resolved by the deterministic rules
 * listed in oracle.h.  Arithmetic keeps the reference's FP32/FP64 types and
 * operation order; the translation unit is compiled with -ffp-contract=off.
 * Float transcendental functions come from fastrrt_amd/csrc/fr_math.h,
 * which the CPU suite checks against libm in FP64.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../fastrrt_amd/csrc/fr_math.h"

#define OG_BIG_COST 99999999999LL

static long og_pos(const og_graph *g, int x, int z) { return (long)z * g->W + x; }
static int og_in(const og_graph *g, int x, int z) { return x >= 0 && x < g->W && z >= 0 && z < g->H; }

/* Host-code truncation (x86 cvtt semantics: NaN and overflow -> INT_MIN).
 * Used only where the reference runs the arithmetic on the host:
 * interpolation (src/interpolation.cpp:83-84) and path optimisation
 * (src/cuda/graph_optimization.cu:92-124, a host loop). */
static int og_host_trunc_d(double v)
{
    if (v != v || v >= 2147483648.0 || v <= -2147483649.0)
        return (-2147483647 - 1);
    return (int)v;
}

/* ---------------------------------------------------------------------- */
/* Grid (src/cuda/graph.cu)                                                 */
/* ---------------------------------------------------------------------- */

og_graph *og_graph_new(int W, int H)
{
    og_graph *g = (og_graph *)calloc(1, sizeof(og_graph));
    size_t n = (size_t)W * H;
    g->W = W;
    g->H = H;
    g->px = (int32_t *)calloc(n, 4);
    g->pz = (int32_t *)calloc(n, 4);
    g->type = (int32_t *)calloc(n, 4);
    g->derive = (int32_t *)calloc(n, 4);
    g->heading = (float *)calloc(n, 4);
    g->cost = (float *)calloc(n, 4);
    g->intrinsic = (float *)calloc(n, 4);
    g->direct = (float *)calloc(n, 4);
    g->gc_flag = (float *)calloc(n, 4);
    g->gc_heading = (float *)calloc(n, 4);
    g->gc_cost = (float *)calloc(n, 4);
    g->region = NULL;
    og_set_region_side(g, 128); /* graph_expansion_search_zones.cu:32,370-394 */
    /* coordinate origin defaults to the grid centre (graph.cu:258-261,279) */
    og_set_coordinate_start(g, W / 2, H / 2, 0.0f);
    g->min_dist_x = g->min_dist_z = 0;
    g->lb_x = g->lb_z = g->ub_x = g->ub_z = -1;
    return g;
}

void og_graph_free(og_graph *g)
{
    if (!g)
        return;
    free(g->px); free(g->pz); free(g->type); free(g->derive);
    free(g->heading); free(g->cost); free(g->intrinsic); free(g->direct);
    free(g->gc_flag); free(g->gc_heading); free(g->gc_cost);
    free(g->region); free(g->class_costs);
    free(g);
}

/* graph_expansion_search_zones.cu:370-394 (region side is 128 there) */
void og_set_region_side(og_graph *g, int side)
{
    g->region_side = side;
    g->dens_w = g->W / side + 1;
    g->dens_h = g->H / side + 1;
    free(g->region);
    g->region = (uint32_t *)calloc((size_t)g->dens_w * g->dens_h, 4);
}

/* graph.cu:286-304 */
void og_set_physical_params(og_graph *g, float perc_w, float perc_h, double max_steer_rad, float vehicle_len,
                            float max_curvature)
{
    g->rate_w = (double)((float)g->W / perc_w);
    g->inv_rate_w = (double)(perc_w / (float)g->W);
    g->rate_h = (double)((float)g->H / perc_h);
    g->inv_rate_h = (double)(perc_h / (float)g->H);
    g->max_steer_rad = max_steer_rad;
    g->max_steer_deg = max_steer_rad * FR_TO_DEG_D;
    g->lr = (double)(vehicle_len / 2.0f);
    const float t = fr_tanf((float)max_steer_rad);
    if (max_curvature < 0)
        g->max_curv = (double)(2.0f * t) / (0.5 * (double)vehicle_len * (double)sqrtf(4.0f + t));
    else
        g->max_curv = (double)max_curvature;
    g->has_physical = 1;
}

/* graph.cu:306-314 */
void og_set_search_params(og_graph *g, int min_dist_x, int min_dist_z, int lb_x, int lb_z, int ub_x, int ub_z)
{
    g->min_dist_x = fr_trunc_f((float)min_dist_x / 2.0f);
    g->min_dist_z = fr_trunc_f((float)min_dist_z / 2.0f);
    g->lb_x = lb_x;
    g->lb_z = lb_z;
    g->ub_x = ub_x;
    g->ub_z = ub_z;
}

/* graph.cu:316-338 */
void og_set_class_costs(og_graph *g, const float *costs, int n)
{
    free(g->class_costs);
    g->class_costs = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++)
        g->class_costs[i] = costs[i];
    g->n_classes = n;
}

/* graph.cu:351-362 */
void og_set_coordinate_start(og_graph *g, int x, int z, float heading)
{
    g->og_x = (float)x;
    g->og_y = (float)z;
    g->og_h = heading;
}

/* graph.cu:14-72 (host branch, override = true) */
static void og_set_cell(og_graph *g, long pos, float heading, int px, int pz, float cost, int type)
{
    g->type[pos] = type;
    g->px[pos] = px;
    g->pz[pos] = pz;
    g->heading[pos] = heading;
    g->cost[pos] = cost;
}

/* graph.cu:364-374 */
void og_add(og_graph *g, int x, int z, float heading, int px, int pz, float cost)
{
    if (!og_in(g, x, z))
        return;
    long pos = og_pos(g, x, z);
    if (px != -1 && pz != -1) {
        long pp = og_pos(g, px, pz);
        if (pp >= 0 && pp < (long)g->W * g->H) /* the reference writes out of bounds here */
            g->derive[pp]++;
    }
    og_set_cell(g, pos, heading, px, pz, cost, FR_TYPE_NODE);
}

/* graph.cu:375-381 */
void og_add_temporary(og_graph *g, int x, int z, float heading, int px, int pz, float cost)
{
    if (!og_in(g, x, z))
        return;
    og_set_cell(g, og_pos(g, x, z), heading, px, pz, cost, FR_TYPE_TEMP);
}

/* graph.cu:393-398 */
void og_remove(og_graph *g, int x, int z)
{
    if (!og_in(g, x, z))
        return;
    g->type[og_pos(g, x, z)] = FR_TYPE_NULL;
}

/* graph.cu:400-427: only the type is reset; parents, headings, costs and
 * child counts survive. */
void og_clear(og_graph *g)
{
    long n = (long)g->W * g->H;
    for (long i = 0; i < n; i++)
        g->type[i] = FR_TYPE_NULL;
    g->goal_reached = 0;
    g->node_collision = 0;
}

/* graph.cu:553-563 (host re-parenting decrements the OLD parent) */
void og_set_collision(og_graph *g, int x, int z, int npx, int npz, float heading, float cost)
{
    long pos = og_pos(g, x, z);
    long cpp = og_pos(g, g->px[pos], g->pz[pos]);
    if (g->type[pos] != FR_TYPE_NODE)
        return;
    g->type[pos] = FR_TYPE_COLLISION;
    g->px[pos] = npx;
    g->pz[pos] = npz;
    g->heading[pos] = heading;
    g->cost[pos] = cost;
    if (cpp >= 0 && cpp < (long)g->W * g->H)
        g->derive[cpp]--;
}

/* graph.cu:446-453 */
void og_get_parent(og_graph *g, int x, int z, int *px, int *pz)
{
    if (!og_in(g, x, z) || g->type[og_pos(g, x, z)] == FR_TYPE_NULL) {
        *px = -1;
        *pz = -1;
        return;
    }
    long pos = og_pos(g, x, z);
    *px = g->px[pos];
    *pz = g->pz[pos];
}

/* graph.cu:470-477 */
float og_get_cost(og_graph *g, int x, int z)
{
    if (!og_in(g, x, z))
        return -1;
    return g->cost[og_pos(g, x, z)];
}

/* graph.cu:487-494 */
int og_get_type(og_graph *g, int x, int z)
{
    if (!og_in(g, x, z))
        return -1;
    return g->type[og_pos(g, x, z)];
}

/* graph_list.cu:8-35 */
unsigned og_count(og_graph *g, int type)
{
    unsigned c = 0;
    long n = (long)g->W * g->H;
    for (long i = 0; i < n; i++)
        c += g->type[i] == type;
    return c;
}

/* graph_list.cu:37-66 */
unsigned og_count_all(og_graph *g)
{
    unsigned c = 0;
    long n = (long)g->W * g->H;
    for (long i = 0; i < n; i++)
        c += g->type[i] != FR_TYPE_NULL;
    return c;
}

/* ---------------------------------------------------------------------- */
/* Footprint feasibility (src/cuda/path_feasible.cu)                        */
/* ---------------------------------------------------------------------- */

/* Class id -> obstacle.  A class index outside the cost table reads past the
 * table in the reference (undefined); it is taken as an obstacle here. */
static int og_is_obstacle_px(const og_graph *g, const float *frame, long pos)
{
    int cls = fr_trunc_f(frame[3 * pos]);
    if (cls < 0 || cls >= g->n_classes)
        return 1;
    return g->class_costs[cls] < 0;
}

/* path_feasible.cu:13-78.  isAllAngleFeasible tests bit 0 because of C
 * precedence ("& 256 > 0" is "& 1"), :15. */
int og_feasible(const og_graph *g, const float *frame, int x, int z, float angle)
{
    ((og_graph *)g)->n_footprint_checks++;
    const int W = g->W, H = g->H;
    long pos = og_pos(g, x, z);
    /* a centre outside the grid (reached by the feasible-connection walk):
     * the reference reads its all-heading bit past the frame; taken as clear */
    if (og_in(g, x, z) && (fr_trunc_f(frame[3 * pos + 2]) & 1))
        return 1;
    const int lbx = g->lb_x, lbz = g->lb_z, ubx = g->ub_x, ubz = g->ub_z;
    const int mx = g->min_dist_x, mz = g->min_dist_z;
    float c, s;
    fr_sincosf(angle, &s, &c);
    for (int i = -mz; i <= mz; i++)
        for (int j = -mx; j <= mx; j++) {
            ((og_graph *)g)->n_footprint_samples++;
            const float fx = (float)j * c - (float)i * s + (float)x;
            const float fz = (float)j * s + (float)i * c + (float)z;
            const int xl = fr_trunc_f(fx);
            const int zl = fr_trunc_f(fz);
            if (xl < 0 || xl >= W)
                continue;
            if (zl < 0 || zl >= H)
                continue;
            if (xl >= lbx && xl <= ubx && zl >= ubz && zl <= lbz)
                continue;
            if (og_is_obstacle_px(g, frame, (long)zl * W + xl))
                return 0;
        }
    return 1;
}

/* ---------------------------------------------------------------------- */
/* Kinematic rollout (src/cuda/kinematics.cu:123-231)                       */
/* ---------------------------------------------------------------------- */

/* pose (optional): the final FP32 map pose (x, y, heading) and the number of
 * steps executed, for the FP64 recomputation check of the test suite */
static void og_rollout(og_graph *g, const float *frame, int sx, int sz, float steer, float path_size, float v,
                       float out[4], double pose[4])
{
    g->n_rollouts++;
    out[0] = -1; out[1] = -1; out[2] = 0; out[3] = 0;
    if (!g->has_physical)
        return;
    const double max_steer = g->max_steer_rad;
    const double lr = g->lr;
    const int W = g->W, H = g->H;
    /* convert_waypoint_to_map_pose, kinematics.cu:54-59 */
    const double sp_x = (double)(g->og_y - (float)sz);
    const double sp_y = (double)((float)sx - g->og_x);
    float x = (float)sp_x;
    float y = (float)sp_y;
    if ((double)steer > max_steer)
        steer = (float)max_steer;
    else if ((double)steer < -max_steer)
        steer = (float)(-max_steer);
    const float t = fr_tanf(steer);
    const float dt = (float)0.1;
    const float ds = v * dt;
    const float beta = fr_atanf(t / 2.0f);
    const float hif = (float)((double)(ds * fr_cosf(beta) * t) / (2.0 * lr));
    const int max_size = fr_trunc_f(path_size) + 1;
    int size = 0;
    int last_x = sx, last_z = sz;
    const long start_pos = og_pos(g, sx, sz);
    float heading = g->heading[start_pos];
    float node_cost = g->cost[start_pos];
    int loop_count = 0;
    const int loop_max = fr_trunc_f((float)(100 * max_size) / dt);
    int steps = 0;
    while (size < max_size) {
        if (loop_count++ > loop_max)
            break;
        steps++;
        x += ds * fr_cosf(heading + beta);
        y += ds * fr_sinf(heading + beta);
        heading += hif;
        /* convert_map_pose_to_waypoint, kinematics.cu:25-46 (FP64 sums) */
        const int cx = fr_trunc_d((double)g->og_x + (double)y);
        const int cz = fr_trunc_d((double)g->og_y - (double)x);
        if (cx == last_x && cz == last_z)
            continue;
        if (cx < 0 || cx >= W)
            break;
        if (cz < 0 || cz >= H)
            break;
        size += 1;
        node_cost += g->intrinsic[(long)cz * W + cx] + 1.0f;
        if (!og_feasible(g, frame, cx, cz, heading)) {
            out[0] = -1; out[1] = -1; out[2] = 0; out[3] = 0;
            if (pose) {
                pose[0] = x;
                pose[1] = y;
                pose[2] = heading;
                pose[3] = steps;
            }
            return;
        }
        last_x = cx;
        last_z = cz;
    }
    if (pose) {
        pose[0] = x;
        pose[1] = y;
        pose[2] = heading;
        pose[3] = steps;
    }
    if (size < 1)
        return;
    out[0] = (float)last_x;
    out[1] = (float)last_z;
    out[2] = node_cost;
    out[3] = heading;
}

void og_kinematic_new_path(og_graph *g, const float *frame, int sx, int sz, float steer, float path_size,
                           float v, float out[4])
{
    og_rollout(g, frame, sx, sz, steer, path_size, v, out, NULL);
}

void og_kinematic_new_path_pose(og_graph *g, const float *frame, int sx, int sz, float steer, float path_size,
                                float v, float out[4], double pose[4])
{
    og_rollout(g, frame, sx, sz, steer, path_size, v, out, pose);
}

/* expand_node, graph_expansion.cu:103-145, host sequential semantics
 * (reached through CudaGraph::derivateNode, :224-244). */
void og_derivate_node(og_graph *g, const float *frame, float steer, float path_size, float v, int x, int z,
                      int out[2])
{
    out[0] = -1;
    out[1] = -1;
    if (!og_in(g, x, z) || g->type[og_pos(g, x, z)] != FR_TYPE_NODE)
        return;
    long pos = og_pos(g, x, z);
    float end[4];
    og_kinematic_new_path(g, frame, x, z, steer, path_size, v, end);
    if (end[0] < 0 || end[1] < 0)
        return;
    int ex = fr_trunc_f(end[0]), ez = fr_trunc_f(end[1]);
    if ((float)ex == g->og_x && (float)ez == g->og_y)
        return;
    long epos = og_pos(g, ex, ez);
    if (epos == pos)
        return;
    if (g->type[epos] == FR_TYPE_NULL) {
        g->type[epos] = FR_TYPE_TEMP;
        g->derive[pos]++;
        og_set_cell(g, epos, end[3], x, z, end[2], FR_TYPE_TEMP);
    }
    if (g->type[epos] == FR_TYPE_NODE) {
        og_set_cell(g, epos, end[3], x, z, end[2], FR_TYPE_COLLISION);
        g->node_collision = 1;
        g->derive[pos]--;
    }
    out[0] = ex;
    out[1] = ez;
}

/* graph_expansion.cu:36-82 */
void og_accept_derived_nodes(og_graph *g)
{
    long n = (long)g->W * g->H;
    for (long i = 0; i < n; i++)
        if (g->type[i] == FR_TYPE_TEMP)
            g->type[i] = FR_TYPE_NODE;
}

/* graph_expansion_search_zones.cu:42-66 and :401-449 */
void og_count_regions(og_graph *g)
{
    memset(g->region, 0, sizeof(uint32_t) * (size_t)g->dens_w * g->dens_h);
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            int t = g->type[og_pos(g, x, z)];
            if (t == FR_TYPE_NULL || t == FR_TYPE_PROCESSING)
                continue;
            g->region[(z / g->region_side) * g->dens_w + x / g->region_side]++;
        }
}

/* graph_collision.cu:33-126.  Erase decisions read the pre-erase snapshot;
 * COLLISION -> NODE conversion resets child counts before the parents'
 * counts are incremented. */
void og_solve_collisions(og_graph *g)
{
    const long n = (long)g->W * g->H;
    const int nodes = (int)og_count(g, FR_TYPE_NODE);
    if (nodes <= 2)
        return;
    uint8_t *erase = (uint8_t *)calloc((size_t)n, 1);
    for (long pos = 0; pos < n; pos++) {
        if (g->type[pos] == FR_TYPE_NULL)
            continue;
        long curr = pos;
        int i = nodes + 4;
        while (i-- > 0) {
            int ppx = g->px[curr], ppz = g->pz[curr];
            if (ppx == -1 && ppz == -1)
                break;
            long next = (long)ppz * g->W + ppx;
            int tn = (next >= 0 && next < n) ? g->type[next] : FR_TYPE_NULL;
            if (tn == FR_TYPE_COLLISION || tn == FR_TYPE_NULL) {
                erase[pos] = 1;
                break;
            }
            curr = next;
        }
    }
    for (long pos = 0; pos < n; pos++)
        if (erase[pos])
            g->type[pos] = FR_TYPE_NULL;
    /* reuse the array as the "converted" marker */
    memset(erase, 0, (size_t)n);
    for (long pos = 0; pos < n; pos++)
        if (g->type[pos] == FR_TYPE_COLLISION) {
            g->type[pos] = FR_TYPE_NODE;
            g->derive[pos] = 0;
            erase[pos] = 1;
        }
    for (long pos = 0; pos < n; pos++)
        if (erase[pos]) {
            long pp = (long)g->pz[pos] * g->W + g->px[pos];
            if (pp >= 0 && pp < n)
                g->derive[pp]++;
        }
    free(erase);
    g->node_collision = 0;
}

/* ---------------------------------------------------------------------- */
/* One expansion kernel (graph_expansion.cu:148-222,                        */
/* graph_expansion_search_zones.cu:68-78,161-232)                           */
/* ---------------------------------------------------------------------- */

typedef struct {
    long src, end;
    float heading, cost;
} og_exp;

void og_expand(og_graph *g, const float *frame, float max_path, float v, int frontier, int force, int smart,
               uint64_t seed, uint64_t stream, uint32_t iter)
{
    const long n = (long)g->W * g->H;
    g->node_collision = 0;
    /* phase 1: every eligible NODE cell rolls out against the snapshot */
    og_exp *ex = (og_exp *)malloc(sizeof(og_exp) * (size_t)(n > 0 ? n : 1));
    long nex = 0;
    for (long pos = 0; pos < n; pos++) {
        if (g->type[pos] != FR_TYPE_NODE)
            continue;
        const int z = (int)(pos / g->W), x = (int)(pos - (long)z * g->W);
        if (smart) {
            if (!force) {
                int ok;
                if (frontier)
                    ok = g->derive[pos] == 0;
                else {
                    int dpos = (z / g->region_side) * g->dens_w + x / g->region_side;
                    ok = (double)g->region[dpos] <= 0.5 * (double)g->region_side;
                }
                if (!ok)
                    continue;
            }
        } else {
            if (frontier && g->derive[pos] > 0)
                continue;
        }
        float us, ul;
        fr_expansion_draws(seed, stream, iter, (uint32_t)pos, &us, &ul);
        /* generateRandomNeg / generateRandom, random_number_gen.cu:14-20 */
        const float steer = (2.0f * us - 1.0f) * (float)g->max_steer_rad;
        float psize = max_path * ul;
        if (psize < 5.0f)
            psize = 5.0f;
        float end[4];
        og_kinematic_new_path(g, frame, x, z, steer, psize, v, end);
        if (end[0] < 0 || end[1] < 0)
            continue;
        const int exx = fr_trunc_f(end[0]), ezz = fr_trunc_f(end[1]);
        if ((float)exx == g->og_x && (float)ezz == g->og_y)
            continue;
        const long epos = og_pos(g, exx, ezz);
        if (epos == pos)
            continue;
        ex[nex].src = pos;
        ex[nex].end = epos;
        ex[nex].heading = end[3];
        ex[nex].cost = end[2];
        nex++;
    }
    /* phase 2: commit; the lowest source cell wins each end cell */
    uint8_t *claimed = (uint8_t *)calloc((size_t)n, 1);
    int32_t *snap = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    memcpy(snap, g->type, sizeof(int32_t) * (size_t)n);
    for (long k = 0; k < nex; k++) { /* ex[] is in increasing src order */
        const long e = ex[k].end, s = ex[k].src;
        if (claimed[e])
            continue;
        const int st = snap[e];
        const int sx = (int)(s % g->W), sz = (int)(s / g->W);
        if (st == FR_TYPE_NULL) {
            claimed[e] = 1;
            og_set_cell(g, e, ex[k].heading, sx, sz, ex[k].cost, FR_TYPE_TEMP);
            g->derive[s]++;
        } else if (st == FR_TYPE_NODE) {
            claimed[e] = 1;
            og_set_cell(g, e, ex[k].heading, sx, sz, ex[k].cost, FR_TYPE_COLLISION);
            g->node_collision = 1;
            g->derive[s]--;
        }
    }
    free(snap);
    free(claimed);
    free(ex);
    if (smart)
        og_count_regions(g);
    if (g->node_collision)
        og_solve_collisions(g);
}

/* ---------------------------------------------------------------------- */
/* Hermite goal connection (src/cuda/graph_direct_goal_connection.cu)      */
/* ---------------------------------------------------------------------- */

/* graph.cu:219-234 */
static double og_compute_heading(int x1, int z1, int x2, int z2)
{
    double dz = (double)(z2 - z1);
    double dx = (double)(x2 - x1);
    if (dx == 0 && dz == 0)
        return 0;
    double v1 = 0;
    if (dz != 0)
        v1 = (double)fr_atan2f((float)(-dz), (float)dx);
    else
        v1 = (double)fr_atan2f(0.0f, (float)dx);
    return FR_HALF_PI_D - v1;
}

/* checkDirectConnectionToGoal, graph_direct_goal_connection.cu:16-167.
 * host_nan: 1 when the reference runs this on the host (path optimisation),
 * where float->int truncation of NaN yields INT_MIN instead of 0. */
static float og_check_direct_connection_impl(og_graph *g, const float *frame, float max_curv, int x, int z,
                                             float local_heading, int gx, int gz, float gh, int safe_zone,
                                             int dist_proc, int host_nan)
{
    g->n_hermite_walks++;
    const int W = g->W, H = g->H;
    const long pos = og_pos(g, x, z);
    float distance;
    if (dist_proc)
        distance = frame[3 * pos + 1];
    else {
        const int dx = gx - x, dz = gz - z;
        distance = sqrtf((float)(dx * dx + dz * dz));
    }
    const int num_points = fr_trunc_f(distance);
    const float a1 = (float)((double)local_heading - FR_HALF_PI_D);
    const float a2 = (float)((double)gh - FR_HALF_PI_D);
    float s1, c1, s2, c2;
    fr_sincosf(a1, &s1, &c1);
    fr_sincosf(a2, &s2, &c2);
    const float t1x = distance * c1, t1y = distance * s1;
    const float t2x = distance * c2, t2y = distance * s2;
    int last_x = -1, last_z = -1;
    float node_cost = g->cost[pos];
    for (int i = 0; i < num_points; ++i) {
        const double t = (0.0 + (double)i) / (double)(num_points - 1);
        const double t2 = t * t, t3 = t2 * t;
        const double h00 = 2 * t3 - 3 * t2 + 1;
        const double h10 = t3 - 2 * t2 + t;
        const double h01 = -2 * t3 + 3 * t2;
        const double h11 = t3 - t2;
        const double px = h00 * x + h10 * (double)t1x + h01 * gx + h11 * (double)t2x;
        const double pz = h00 * z + h10 * (double)t1y + h01 * gz + h11 * (double)t2y;
        if (px < 0 || px >= W)
            continue;
        if (pz < 0 || pz >= H)
            continue;
        const int cx = host_nan ? og_host_trunc_d(px) : fr_trunc_d(px);
        const int cz = host_nan ? og_host_trunc_d(pz) : fr_trunc_d(pz);
        if (cx == last_x && cz == last_z)
            continue;
        if (cx < 0 || cx >= W)
            continue;
        if (cz < 0 || cz >= H)
            continue;
        node_cost += g->intrinsic[(long)cz * W + cx] + 1.0f;
        const double t00 = 6 * t2 - 6 * t;
        const double t10 = 3 * t2 - 4 * t + 1;
        const double t01 = -6 * t2 + 6 * t;
        const double t11 = 3 * t2 - 2 * t;
        const double ddx = t00 * x + t10 * (double)t1x + t01 * gx + t11 * (double)t2x;
        const double ddz = t00 * z + t10 * (double)t1y + t01 * gz + t11 * (double)t2y;
        const float heading = (float)((double)fr_atan2f((float)ddz, (float)ddx) + FR_HALF_PI_D);
        const double d00 = 12 * t - 6;
        const double d10 = 6 * t - 4;
        const double d01 = -12 * t + 6;
        const double d11 = 6 * t - 2;
        const double dd2x = d00 * x + d10 * (double)t1x + d01 * gx + d11 * (double)t2x;
        const double dd2z = d00 * z + d10 * (double)t1y + d01 * gz + d11 * (double)t2y;
        if (max_curv > 0) {
            /* pow(v, 1.5) is evaluated as v * sqrt(v) */
            const double v = ddx * ddx + ddz * ddz;
            const float k = (float)(fr_fabs_d(ddx * dd2z - ddz * dd2x) / (v * sqrt(v)));
            if (k > max_curv)
                return -1;
        }
        last_x = cx;
        last_z = cz;
        const int trav = fr_trunc_f(frame[3 * pos + 2]);
        if (safe_zone && (trav & 0x100) > 0)
            continue;
        if (!og_feasible(g, frame, last_x, last_z, heading))
            return -1;
    }
    if (num_points <= 0)
        return -1;
    if (last_x != gx && last_z != gz)
        return -1;
    return node_cost;
}

float og_check_direct_connection(og_graph *g, const float *frame, float max_curv, int x, int z, float local_heading,
                                 int gx, int gz, float gh, int safe_zone, int dist_proc)
{
    return og_check_direct_connection_impl(g, frame, max_curv, x, z, local_heading, gx, gz, gh, safe_zone,
                                           dist_proc, 0);
}

/* the same check with the host's float->int truncation (NaN -> INT_MIN), as
 * path_optimize runs it on the host (graph_optimization.cu:92-110) */
float og_check_direct_connection_host(og_graph *g, const float *frame, float max_curv, int x, int z,
                                      float local_heading, int gx, int gz, float gh, int safe_zone)
{
    return og_check_direct_connection_impl(g, frame, max_curv, x, z, local_heading, gx, gz, gh, safe_zone, 0, 1);
}

/* graph_hermite_connection.cu:11-136 (cudaHermite with fillInterpolation and
 * computeDistance false) behind CudaGraph::canConnectToGoal
 * (graph_expansion.cu:246-261), which runs on the host: host truncation.
 * The curve length is the frame's distance-to-goal channel; the walk starts
 * from the node's own cell (its first point is skipped), the curvature test
 * divides by pow(v, 3/2) with integer 3/2 = 1, i.e. by v, and fails with -2;
 * there is no endpoint test.  goal_heading is an int at the reference's call.
 * Returns the cost, -1 (collision, no points, start outside the grid) or
 * -2 (curvature); *goal_blocked is set when the goal pixel is an obstacle
 * (or outside the grid), where canConnectToGoal answers false before any
 * curve is walked (the curve's float result converts to bool otherwise). */
float og_can_connect_hermite(og_graph *g, const float *frame, float max_curv, int x, int z, int gx, int gz,
                             int goal_heading, int *goal_blocked)
{
    const int W = g->W, H = g->H;
    *goal_blocked = !og_in(g, gx, gz) || og_is_obstacle_px(g, frame, og_pos(g, gx, gz));
    if (*goal_blocked)
        return -1;
    if (!og_in(g, x, z))
        return -1;
    g->n_hermite_walks++;
    const long pos = og_pos(g, x, z);
    const float distance = frame[3 * pos + 1];
    const int num_points = og_host_trunc_d((double)distance);
    const float local_heading = g->heading[pos];
    const float a1 = (float)((double)local_heading - FR_HALF_PI_D);
    const float a2 = (float)((double)goal_heading - FR_HALF_PI_D);
    float s1, c1, s2, c2;
    fr_sincosf(a1, &s1, &c1);
    fr_sincosf(a2, &s2, &c2);
    const float t1x = distance * c1, t1y = distance * s1;
    const float t2x = distance * c2, t2y = distance * s2;
    int last_x = x, last_z = z;
    float node_cost = g->cost[pos];
    for (int i = 0; i < num_points; ++i) {
        const double t = (0.0 + (double)i) / (double)(num_points - 1);
        const double t2 = t * t, t3 = t2 * t;
        const double h00 = 2 * t3 - 3 * t2 + 1;
        const double h10 = t3 - 2 * t2 + t;
        const double h01 = -2 * t3 + 3 * t2;
        const double h11 = t3 - t2;
        const double px = h00 * x + h10 * (double)t1x + h01 * gx + h11 * (double)t2x;
        const double pz = h00 * z + h10 * (double)t1y + h01 * gz + h11 * (double)t2y;
        if (px < 0 || px >= W)
            continue;
        if (pz < 0 || pz >= H)
            continue;
        const int cx = og_host_trunc_d(px);
        const int cz = og_host_trunc_d(pz);
        if (cx == last_x && cz == last_z)
            continue;
        if (cx < 0 || cx >= W)
            continue;
        if (cz < 0 || cz >= H)
            continue;
        node_cost += g->intrinsic[(long)cz * W + cx] + 1.0f;
        const double t00 = 6 * t2 - 6 * t;
        const double t10 = 3 * t2 - 4 * t + 1;
        const double t01 = -6 * t2 + 6 * t;
        const double t11 = 3 * t2 - 2 * t;
        const double ddx = t00 * x + t10 * (double)t1x + t01 * gx + t11 * (double)t2x;
        const double ddz = t00 * z + t10 * (double)t1y + t01 * gz + t11 * (double)t2y;
        const float heading = (float)((double)fr_atan2f((float)ddz, (float)ddx) + FR_HALF_PI_D);
        const double d00 = 12 * t - 6;
        const double d10 = 6 * t - 4;
        const double d01 = -12 * t + 6;
        const double d11 = 6 * t - 2;
        const double dd2x = d00 * x + d10 * (double)t1x + d01 * gx + d11 * (double)t2x;
        const double dd2z = d00 * z + d10 * (double)t1y + d01 * gz + d11 * (double)t2y;
        if (max_curv > 0) {
            const double v = ddx * ddx + ddz * ddz;
            const float k = (float)(fr_fabs_d(ddx * dd2z - ddz * dd2x) / v);
            if (k > max_curv)
                return -2;
        }
        last_x = cx;
        last_z = cz;
        if (!og_feasible(g, frame, last_x, last_z, heading))
            return -1;
    }
    if (num_points <= 0)
        return -1;
    return node_cost;
}

/* ---------------------------------------------------------------------- */
/* Potential-field intrinsic costs (src/cuda/apf.cu)                        */
/* ---------------------------------------------------------------------- */

static int og_in_ego(const og_graph *g, int x, int z)
{
    return x >= g->lb_x && x <= g->ub_x && z >= g->ub_z && z <= g->lb_z;
}

/* apf.cu:17-21: outside the grid counts as an obstacle */
static int og_apf_obstacle(const og_graph *g, const float *frame, int x, int z)
{
    if (!og_in(g, x, z))
        return 1;
    return og_is_obstacle_px(g, frame, og_pos(g, x, z));
}

/* apf.cu:23-98, 123-138.  Every obstacle cell outside the ego box is set to
 * 100*Kr/2; one with a free (or in-grid-free) 4-neighbour then adds
 * Kr/2*(1/p - 1/p0)^2 to every other cell within Euclidean distance p0 =
 * radius that is outside the ego box.  The reference's per-thread set races
 * with other threads' atomic adds; here all sets come first, then the
 * sources add in ascending cell order (one of the reference's outcomes). */
void og_apf_repulsion(og_graph *g, const float *frame, float kr, int radius)
{
    const int W = g->W, H = g->H;
    const float kr_half = (float)(kr * 0.5);
    for (int z = 0; z < H; ++z)
        for (int x = 0; x < W; ++x)
            if (!og_in_ego(g, x, z) && og_is_obstacle_px(g, frame, og_pos(g, x, z)))
                g->intrinsic[og_pos(g, x, z)] = 100 * kr_half;
    for (int z = 0; z < H; ++z)
        for (int x = 0; x < W; ++x) {
            if (og_in_ego(g, x, z) || !og_is_obstacle_px(g, frame, og_pos(g, x, z)))
                continue;
            if (og_apf_obstacle(g, frame, x - 1, z) && og_apf_obstacle(g, frame, x + 1, z) &&
                og_apf_obstacle(g, frame, x, z - 1) && og_apf_obstacle(g, frame, x, z + 1))
                continue;
            const float p0 = (float)radius;
            for (int h = z - radius; h <= z + radius; ++h) {
                if (h < 0)
                    continue;
                if (h >= H)
                    break;
                for (int w = x - radius; w <= x + radius; ++w) {
                    if (w < 0)
                        continue;
                    if (w >= W)
                        break;
                    if (og_in_ego(g, w, h) || (w == x && h == z))
                        continue;
                    const float p = sqrtf((float)((w - x) * (w - x) + (h - z) * (h - z)));
                    if (p > p0)
                        continue;
                    const float f = 1 / p - 1 / p0;
                    const float c = kr_half * f * f;
                    g->intrinsic[og_pos(g, w, h)] += c;
                }
            }
        }
}

/* apf.cu:100-121, 140-155: every free cell loses Ka/2 * squared distance to
 * the goal (the ego box is not skipped here). */
void og_apf_attraction(og_graph *g, const float *frame, float ka, int gx, int gz)
{
    const float ka_half = (float)(ka * 0.5);
    for (int z = 0; z < g->H; ++z)
        for (int x = 0; x < g->W; ++x) {
            const long pos = og_pos(g, x, z);
            if (og_is_obstacle_px(g, frame, pos))
                continue;
            const int dx = gx - x, dz = gz - z;
            const float d = (float)(dx * dx + dz * dz) * ka_half;
            g->intrinsic[pos] = g->intrinsic[pos] - d;
        }
}

/* graph_direct_goal_connection.cu:169-227 */
void og_process_direct_goal_connection(og_graph *g, const float *frame, int gx, int gz, float gh, float max_curv)
{
    if (max_curv < 0)
        max_curv = (float)g->max_curv;
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            const long pos = og_pos(g, x, z);
            g->gc_flag[pos] = 0;
            const float lh = (float)og_compute_heading(x, z, gx, gz);
            const float cost = og_check_direct_connection_impl(g, frame, max_curv, x, z, lh, gx, gz, gh, 0, 0, 0);
            if (cost < 0)
                continue;
            g->gc_flag[pos] = 1.0f;
            g->gc_heading[pos] = lh;
            g->gc_cost[pos] = cost;
        }
}

/* graph_check_goal_reached.cu:17-188 with the tie rules of oracle.h */
int og_find_best_goal_direct_connection(og_graph *g, const float *frame, float radius, int safe_zone)
{
    const float max_curv = (float)g->max_curv;
    g->best_parent[0] = -1;
    g->best_parent[1] = -1;
    long long best = OG_BIG_COST;
    long best_node = -1;
    int best_xp = 0, best_zp = 0;
    float best_lih = 0, best_c = 0, best_total = 0;
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            const long pos = og_pos(g, x, z);
            if (g->type[pos] != FR_TYPE_NODE)
                continue;
            const float heading = g->heading[pos];
            for (int zp = fr_trunc_f((float)z - radius); (float)zp < (float)z + radius; zp++)
                for (int xp = fr_trunc_f((float)x - radius); (float)xp < (float)x + radius; xp++) {
                    if (zp < 0 || zp >= g->H)
                        continue;
                    if (xp < 0 || xp >= g->W)
                        continue;
                    const long q = og_pos(g, xp, zp);
                    if (!(g->gc_flag[q] > 0))
                        continue;
                    const float lih = g->gc_heading[q];
                    const float c = og_check_direct_connection_impl(g, frame, max_curv, x, z, heading, xp, zp, lih,
                                                                    safe_zone, 0, 0);
                    if (c < 0)
                        continue;
                    const float total = c + g->gc_cost[q];
                    const long long lcost = (long long)floor((double)(100.0f * total));
                    /* lower cost wins; equal cost: keep the first node in
                     * cell order, and the last window cell inside it */
                    if (lcost < best || (lcost == best && pos == best_node)) {
                        best = lcost;
                        best_node = pos;
                        best_xp = xp;
                        best_zp = zp;
                        best_lih = lih;
                        best_c = c;
                        best_total = total;
                    }
                }
        }
    if (best >= OG_BIG_COST)
        return 0;
    g->best_parent[0] = (float)(best_node % g->W);
    g->best_parent[1] = (float)(best_node / g->W);
    g->best_parent[2] = g->heading[best_node];
    g->best_parent[3] = best_c;
    g->best_child[0] = (float)best_xp;
    g->best_child[1] = (float)best_zp;
    g->best_child[2] = best_lih;
    g->best_child[3] = best_total;
    return 1;
}

/* graph_best_node.cu:182-251 */
int og_check_goal_reached(og_graph *g, int gx, int gz, float gh, float tol, float max_herr)
{
    if (g->goal_reached)
        return 1;
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            const long pos = og_pos(g, x, z);
            if (g->type[pos] != FR_TYPE_NODE)
                continue;
            const double dx = (double)(gx - x), dz = (double)(gz - z);
            if (sqrt(dx * dx + dz * dz) > (double)tol)
                continue;
            if (fabsf(g->heading[pos] - gh) <= max_herr)
                g->goal_reached = 1;
        }
    return g->goal_reached;
}

/* graph_best_node.cu:16-178.  Cost = floor(1*dist + 3*deg(heading) + 1*cost)
 * with the degree factor taken as an FP64 constant. */
void og_find_best_node(og_graph *g, const float *frame, float heading, float radius, int tx, int tz,
                       float max_herr, int out[2])
{
    long long best = OG_BIG_COST;
    long best_pos = -1;
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            const long pos = og_pos(g, x, z);
            if (g->type[pos] != FR_TYPE_NODE)
                continue;
            const int dx = tx - x, dz = tz - z;
            const float dist = sqrtf((float)(dx * dx + dz * dz));
            if (dist > radius)
                continue;
            const float h = g->heading[pos];
            if (fabsf(h - heading) > max_herr)
                continue;
            if (!og_feasible(g, frame, x, z, h))
                continue;
            const double cd = (double)(1.0f * dist) + (3.0 * FR_TO_DEG_D) * (double)h + (double)(1.0f * g->cost[pos]);
            const long long cost = (long long)floor((double)(float)cd);
            if (cost < best) {
                best = cost;
                best_pos = pos;
            }
        }
    out[0] = -1;
    out[1] = -1;
    if (best >= OG_BIG_COST)
        return;
    out[0] = (int)(best_pos % g->W);
    out[1] = (int)(best_pos / g->W);
}

/* src/graph_check_dag_consistency.cpp:34-76 */
int og_check_graph_consistent(og_graph *g)
{
    const int max_loop = (int)og_count(g, FR_TYPE_NODE) + 4;
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            if (g->type[og_pos(g, x, z)] == FR_TYPE_NULL)
                continue;
            int p_x = x, p_z = z;
            int i = max_loop;
            while (i-- > 0) {
                int ppx, ppz;
                og_get_parent(g, p_x, p_z, &ppx, &ppz);
                if (ppx == p_x && ppz == p_z)
                    return 0;
                if (ppx == -1)
                    break;
                int pt = og_get_type(g, ppx, ppz);
                if (pt != FR_TYPE_NODE && pt != FR_TYPE_COLLISION && pt != FR_TYPE_CONNECT_TO_GOAL)
                    return 0;
                p_x = ppx;
                p_z = ppz;
            }
            if (i <= 0)
                return 0;
        }
    return 1;
}

/* kinematics.cu:75-112 */
static double og_path_heading(double p1x, double p1y, double p2x, double p2y)
{
    double dy = p2y - p1y, dx = p2x - p1x;
    if (dy >= 0 && dx > 0)
        return atan(dy / dx);
    else if (dy >= 0 && dx < 0)
        return FR_PI_D - atan(dy / fabs(dx));
    else if (dy < 0 && dx > 0)
        return -atan(fabs(dy) / dx);
    else if (dy < 0 && dx < 0)
        return atan(dy / dx) - FR_PI_D;
    else if (dx == 0 && dy > 0)
        return FR_HALF_PI_D;
    else if (dx == 0 && dy < 0)
        return -FR_HALF_PI_D;
    return 0.0;
}

/* check_kinematic_connection_start_end, kinematics.cu:233-306 (FP64, host).
 * final_heading / path_cost: the reference's out parameters (set when the
 * walk stops on moving away from the end, 0 otherwise); the intrinsic cost
 * of a cell outside the grid counts as 0 (the reference reads past the
 * grid).  A start outside the grid answers 0. */
int og_check_feasible_connection_ex(og_graph *g, const float *frame, int sx, int sz, int ex, int ez, float v,
                                    double *final_heading, float *path_cost)
{
    *final_heading = 0;
    *path_cost = 0;
    if (!og_in(g, sx, sz))
        return 0;
    const double ddx0 = (double)(ex - sx), ddz0 = (double)(ez - sz);
    const double distance = sqrt(ddx0 * ddx0 + ddz0 * ddz0);
    const double lr = g->lr;
    const double max_steer = g->max_steer_rad;
    const double smx = (double)(g->og_y - (float)sz), smy = (double)((float)sx - g->og_x);
    const double emx = (double)(g->og_y - (float)ez), emy = (double)((float)ex - g->og_x);
    const double dt = 0.1;
    double heading = g->heading[og_pos(g, sx, sz)];
    double path_heading = og_path_heading(smx, smy, emx, emy);
    double steer_a = path_heading - heading;
    if (steer_a < -max_steer) steer_a = -max_steer;
    if (steer_a > max_steer) steer_a = max_steer;
    const double ds = (double)(int)v * dt; /* velocity is an int parameter (:308) */
    const int total_steps = fr_trunc_d(distance / ds);
    double nx = smx, ny = smy;
    int lx = sx, lz = sz;
    double best_end = distance;
    float curr_cost = 0;
    for (int i = 0; i < total_steps; i++) {
        double steer = tan(steer_a);
        double beta = atan(steer / lr);
        nx += ds * cos(heading + beta);
        ny += ds * sin(heading + beta);
        heading += ds * cos(beta) * steer / (2 * lr);
        path_heading = og_path_heading(nx, ny, emx, emy);
        steer_a = path_heading - heading;
        if (steer_a < -max_steer) steer_a = -max_steer;
        if (steer_a > max_steer) steer_a = max_steer;
        const double dx = emx - nx, dy = emy - ny;
        const double dist = sqrt(dx * dx + dy * dy);
        const int cx = fr_trunc_d((double)g->og_x + ny), cz = fr_trunc_d((double)g->og_y - nx);
        if (cx == lx && cz == lz)
            continue;
        curr_cost += (og_in(g, cx, cz) ? g->intrinsic[og_pos(g, cx, cz)] : 0.0f) + 1.0f;
        if (best_end < dist) {
            *final_heading = heading;
            *path_cost = curr_cost;
            return best_end <= 2;
        }
        if (!og_feasible(g, frame, lx, lz, (float)heading))
            return 0;
        lx = cx;
        lz = cz;
        best_end = dist;
    }
    return 1;
}

int og_check_feasible_connection(og_graph *g, const float *frame, int sx, int sz, int ex, int ez, float v)
{
    double fh;
    float pc;
    return og_check_feasible_connection_ex(g, frame, sx, sz, ex, ez, v, &fh, &pc);
}

/* interpolateHermiteCurve, src/interpolation.cpp:36-110 (FP32, host).
 * out: (x, z, heading) triples; returns the number written. */
int og_interpolate_hermite(int W, int H, int x1, int z1, float h1, int x2, int z2, float h2, float *out, int cap)
{
    const double ddx = (double)(x2 - x1), ddz = (double)(z2 - z1);
    const int num_points = fr_trunc_d(sqrt(ddx * ddx + ddz * ddz));
    const float dx = (float)(x2 - x1), dz = (float)(z2 - z1);
    const float d = sqrtf(dx * dx + dz * dz);
    const float a1 = (float)((double)h1 - FR_PI_D / 2);
    const float a2 = (float)((double)h2 - FR_PI_D / 2);
    float s1, c1, s2, c2;
    fr_sincosf(a1, &s1, &c1);
    fr_sincosf(a2, &s2, &c2);
    const float t1x = d * c1, t1y = d * s1, t2x = d * c2, t2y = d * s2;
    int last_x = -1, last_z = -1, n = 0;
    for (int i = 0; i < num_points; ++i) {
        const float t = (float)i / (float)(num_points - 1);
        const float t2 = t * t, t3 = t2 * t;
        const float h00 = 2 * t3 - 3 * t2 + 1;
        const float h10 = t3 - 2 * t2 + t;
        const float h01 = -2 * t3 + 3 * t2;
        const float h11 = t3 - t2;
        const float x = h00 * (float)x1 + h10 * t1x + h01 * (float)x2 + h11 * t2x;
        const float z = h00 * (float)z1 + h10 * t1y + h01 * (float)z2 + h11 * t2y;
        if (x < 0 || x >= W)
            continue;
        if (z < 0 || z >= H)
            continue;
        const int cx = og_host_trunc_d((double)roundf(x));
        const int cz = og_host_trunc_d((double)roundf(z));
        if (cx == last_x && cz == last_z)
            continue;
        if (cx < 0 || cx >= W)
            continue;
        if (cz < 0 || cz >= H)
            continue;
        const float t00 = 6 * t2 - 6 * t;
        const float t10 = 3 * t2 - 4 * t + 1;
        const float t01 = -6 * t2 + 6 * t;
        const float t11 = 3 * t2 - 2 * t;
        const float dxx = t00 * (float)x1 + t10 * t1x + t01 * (float)x2 + t11 * t2x;
        const float dzz = t00 * (float)z1 + t10 * t1y + t01 * (float)z2 + t11 * t2y;
        const float heading = (float)((double)fr_atan2f(dzz, dxx) + FR_HALF_PI_D);
        if (n < cap) {
            out[3 * n] = (float)cx;
            out[3 * n + 1] = (float)cz;
            out[3 * n + 2] = heading;
        }
        n++;
        last_x = cx;
        last_z = cz;
    }
    return n;
}

/* graph.cu:496-546 */
void og_dump_graph(og_graph *g, const char *path)
{
    FILE *fp = fopen(path, "w");
    if (!fp)
        return;
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            long p = og_pos(g, x, z);
            fprintf(fp, "%d %d %d %d %f %f %f %f\n", g->px[p], g->pz[p], g->type[p], g->derive[p], g->heading[p],
                    g->cost[p], g->intrinsic[p], g->direct[p]);
        }
    fclose(fp);
}

void og_read_dump(og_graph *g, const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp)
        return;
    for (int z = 0; z < g->H; z++)
        for (int x = 0; x < g->W; x++) {
            long p = og_pos(g, x, z);
            if (fscanf(fp, "%d %d %d %d %f %f %f %f\n", &g->px[p], &g->pz[p], &g->type[p], &g->derive[p],
                       &g->heading[p], &g->cost[p], &g->intrinsic[p], &g->direct[p]) != 8) {
                fclose(fp);
                return;
            }
        }
    fclose(fp);
}

/* ---------------------------------------------------------------------- */
/* Planner (src/fastrrt.cpp)                                                */
/* ---------------------------------------------------------------------- */

struct og_rrt {
    og_graph *g;
    const float *frame;
    int sx, sz, gx, gz;
    float sh, gh, v, max_path, dist_tol, herr_tol;
    int timeout_ms, has_plan, last_count;
    uint64_t seed, stream;
    uint32_t iter;
    struct timespec t0;
};

og_rrt *og_rrt_new(int W, int H, float perc_w, float perc_h, float max_steer_deg, float vehicle_len, int lb_x,
                   int lb_z, int ub_x, int ub_z, float max_curvature)
{
    og_rrt *r = (og_rrt *)calloc(1, sizeof(og_rrt));
    r->g = og_graph_new(W, H);
    r->herr_tol = (float)(10.0 * FR_TO_RAD_D); /* fastrrt.cpp:39 */
    og_set_physical_params(r->g, perc_w, perc_h, (double)max_steer_deg * FR_TO_RAD_D, vehicle_len, max_curvature);
    r->g->lb_x = lb_x;
    r->g->lb_z = lb_z;
    r->g->ub_x = ub_x;
    r->g->ub_z = ub_z;
    r->seed = 0x5EEDF457ULL;
    return r;
}

void og_rrt_free(og_rrt *r)
{
    if (!r)
        return;
    og_graph_free(r->g);
    free(r);
}

og_graph *og_rrt_graph(og_rrt *r) { return r->g; }
int og_rrt_iterations(og_rrt *r) { return (int)r->iter; }

void og_rrt_set_seed(og_rrt *r, uint64_t seed, uint64_t stream)
{
    r->seed = seed;
    r->stream = stream;
}

static long og_elapsed_ms(og_rrt *r)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (long)((t.tv_sec - r->t0.tv_sec) * 1000 + (t.tv_nsec - r->t0.tv_nsec) / 1000000);
}

/* fastrrt.cpp:64-81 */
void og_rrt_set_plan_data(og_rrt *r, const float *frame, const float *class_costs, int n_classes, int sx, int sz,
                          float sh, int gx, int gz, float gh, float v, int min_dx, int min_dz, int timeout_ms,
                          float max_path, float dist_tol, float herr_tol)
{
    r->has_plan = 1;
    r->sx = sx; r->sz = sz; r->sh = sh;
    r->gx = gx; r->gz = gz; r->gh = gh;
    r->frame = frame;
    r->v = v;
    r->timeout_ms = timeout_ms;
    r->max_path = max_path;
    r->dist_tol = dist_tol;
    r->herr_tol = herr_tol;
    og_set_search_params(r->g, min_dx, min_dz, r->g->lb_x, r->g->lb_z, r->g->ub_x, r->g->ub_z);
    og_set_class_costs(r->g, class_costs, n_classes);
    og_process_direct_goal_connection(r->g, frame, gx, gz, gh, 0.8f);
}

/* fastrrt.cpp:85-98 */
void og_rrt_search_init(og_rrt *r)
{
    if (!r->has_plan)
        return;
    clock_gettime(CLOCK_MONOTONIC, &r->t0);
    og_clear(r->g);
    og_add(r->g, r->sx, r->sz, r->sh, -1, -1, 0);
    r->last_count = 0;
    r->iter = 0;
}

/* fastrrt.cpp:293-297 */
void og_rrt_load_graph_state(og_rrt *r, const char *path)
{
    og_read_dump(r->g, path);
    r->last_count = (int)og_count(r->g, FR_TYPE_TEMP);
}

/* fastrrt.cpp:198-202 */
int og_rrt_goal_reached(og_rrt *r)
{
    return og_check_goal_reached(r->g, r->gx, r->gz, r->gh, r->dist_tol, r->herr_tol);
}

/* fastrrt.cpp:108-181 */
int og_rrt_loop(og_rrt *r, int smart)
{
    og_graph *g = r->g;
    if (r->timeout_ms > 0 && og_elapsed_ms(r) > r->timeout_ms)
        return 0;
    const int frontier = r->last_count >= 100;
    og_expand(g, r->frame, r->max_path, r->v, frontier, r->last_count == 0, smart, r->seed, r->stream, r->iter);
    r->iter++;
    r->last_count = (int)og_count(g, FR_TYPE_TEMP);
    if (r->last_count == 0 && og_count_all(g) == 0) {
        og_add(g, r->sx, r->sz, r->sh, -1, -1, 0);
        return 1;
    }
    og_accept_derived_nodes(g);
    if (og_find_best_goal_direct_connection(g, r->frame, r->dist_tol, 1)) {
        const int ppx = fr_trunc_f(g->best_parent[0]);
        const int ppz = fr_trunc_f(g->best_parent[1]);
        const float base = og_get_cost(g, ppx, ppz);
        og_add(g, fr_trunc_f(g->best_child[0]), fr_trunc_f(g->best_child[1]), g->best_child[2], ppx, ppz,
               g->best_parent[3] + base);
        const int cx = fr_trunc_f(g->best_child[0]);
        const int cz = fr_trunc_f(g->best_child[1]);
        const float child_cost = g->best_child[2] + base;
        og_add(g, r->gx, r->gz, r->gh, cx, cz, child_cost);
    }
    if (og_rrt_goal_reached(r))
        return 0;
    return 1;
}

/* fastrrt.cpp:204-232; out = (x, z, heading) triples */
int og_rrt_get_planned_path(og_rrt *r, float *out, int cap)
{
    if (!r->has_plan)
        return 0;
    if (!og_rrt_goal_reached(r))
        return 0;
    int n[2];
    og_find_best_node(r->g, r->frame, r->gh, r->dist_tol, r->gx, r->gz, r->herr_tol, n);
    int cnt = 0;
    long i = 0;
    /* collect reversed */
    int cap2 = 1 << 20;
    float *tmp = (float *)malloc(sizeof(float) * 3 * 4096);
    int tcap = 4096;
    while (n[0] != -1 && n[1] != -1) {
        if (cnt >= tcap) {
            tcap *= 2;
            tmp = (float *)realloc(tmp, sizeof(float) * 3 * tcap);
        }
        tmp[3 * cnt] = (float)n[0];
        tmp[3 * cnt + 1] = (float)n[1];
        tmp[3 * cnt + 2] = r->g->heading[og_pos(r->g, n[0], n[1])];
        cnt++;
        int ppx, ppz;
        og_get_parent(r->g, n[0], n[1], &ppx, &ppz);
        n[0] = ppx;
        n[1] = ppz;
        if (i++ >= 1000000 || cnt >= cap2) {
            free(tmp);
            return 0;
        }
    }
    for (int k = 0; k < cnt && k < cap; k++) {
        int s = cnt - 1 - k;
        out[3 * k] = tmp[3 * s];
        out[3 * k + 1] = tmp[3 * s + 1];
        out[3 * k + 2] = tmp[3 * s + 2];
    }
    free(tmp);
    return cnt;
}

/* fastrrt.cpp:183-196 -> graph_optimization.cu:25-195 */
int og_rrt_path_optimize(og_rrt *r)
{
    og_graph *g = r->g;
    if (r->timeout_ms > 0 && og_elapsed_ms(r) > r->timeout_ms)
        return 0;
    int cap = 1 << 16;
    float *wp = (float *)malloc(sizeof(float) * 3 * cap);
    int n = og_rrt_get_planned_path(r, wp, cap);
    if (n > cap)
        n = cap;
    /* convertPlannedPath (:76-90): (x, z, heading, cost) */
    float *path = (float *)malloc(sizeof(float) * 4 * (n > 0 ? n : 1));
    float *pd = (float *)calloc((size_t)4 * (n > 0 ? n : 1), sizeof(float));
    for (int k = 0; k < n; k++) {
        path[4 * k] = wp[3 * k];
        path[4 * k + 1] = wp[3 * k + 1];
        path[4 * k + 2] = wp[3 * k + 2];
        path[4 * k + 3] = og_get_cost(g, (int)wp[3 * k], (int)wp[3 * k + 1]);
    }
    free(wp);
    const float max_curv = (float)g->max_curv;
    for (int i = 0; i < n - 2; i++) {
        pd[4 * i] = 0;
        const int x = fr_trunc_f(path[4 * i]), z = fr_trunc_f(path[4 * i + 1]);
        for (int j = i + 2; j < n; j++) {
            const float cur = path[4 * j + 3];
            const float nc = og_check_direct_connection_impl(g, r->frame, max_curv, x, z, path[4 * i + 2],
                                                             fr_trunc_f(path[4 * j]), fr_trunc_f(path[4 * j + 1]),
                                                             path[4 * j + 2], 1, 0, 1);
            if (nc > 0 && nc < cur) {
                pd[4 * i] = 1;
                pd[4 * i + 1] = (float)j;
                pd[4 * i + 2] = nc;
                pd[4 * i + 3] = cur - nc;
            }
        }
    }
    float max_gain = -1;
    int max_pos = -1;
    for (int i = 0; i < n - 2; i++) {
        if (pd[4 * i] == 0.0f)
            continue;
        if (max_gain < pd[4 * i + 3]) {
            max_gain = pd[4 * i + 3];
            max_pos = i;
        }
    }
    if (max_pos < 0) {
        free(path);
        free(pd);
        return 0;
    }
    const int next_pos = fr_trunc_f(pd[4 * max_pos + 1]);
    if (next_pos > max_pos + 1) {
        path[4 * next_pos + 3] = pd[4 * max_pos + 2];
        const int rm = next_pos - max_pos - 1;
        for (int i = max_pos + 1; i + rm < n; ++i)
            for (int c = 0; c < 4; c++)
                path[4 * i + c] = path[4 * (i + rm) + c];
        n -= rm;
    }
    og_clear(g);
    int ppx = fr_trunc_f(path[0]), ppz = fr_trunc_f(path[1]);
    og_add(g, ppx, ppz, path[2], -1, -1, path[3]);
    for (int i = 1; i < n; i++) {
        og_add(g, fr_trunc_f(path[4 * i]), fr_trunc_f(path[4 * i + 1]), path[4 * i + 2], ppx, ppz, path[4 * i + 3]);
        ppx = fr_trunc_f(path[4 * i]);
        ppz = fr_trunc_f(path[4 * i + 1]);
    }
    free(path);
    free(pd);
    return 1;
}

/* fastrrt.cpp:236-240 -> interpolation.cpp:112-135 */
int og_rrt_interpolate_planned_path(og_rrt *r, float *out, int cap)
{
    int pcap = 1 << 16;
    float *wp = (float *)malloc(sizeof(float) * 3 * pcap);
    int n = og_rrt_get_planned_path(r, wp, pcap);
    int total = 0;
    if (n < 2) {
        free(wp);
        return 0;
    }
    for (int i = 1; i < n; i++) {
        int room = cap - total;
        int k = og_interpolate_hermite(r->g->W, r->g->H, (int)wp[3 * (i - 1)], (int)wp[3 * (i - 1) + 1],
                                       wp[3 * (i - 1) + 2], (int)wp[3 * i], (int)wp[3 * i + 1], wp[3 * i + 2],
                                       out + 3 * (total < cap ? total : cap), room > 0 ? room : 0);
        total += k;
    }
    free(wp);
    return total;
}

/* ---------------------------------------------------------------------- */
/* Shared scalar primitives, exported so the CPU suite can check them      */
/* against libm and the Philox known-answer vectors.                       */
/* ---------------------------------------------------------------------- */
float og_sinf(float x) { return fr_sinf(x); }
float og_cosf(float x) { return fr_cosf(x); }
float og_tanf(float x) { return fr_tanf(x); }
float og_atanf(float x) { return fr_atanf(x); }
float og_atan2f(float y, float x) { return fr_atan2f(y, x); }
int og_trunc_d(double v) { return fr_trunc_d(v); }
void og_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    fr_u32x4 c;
    for (int i = 0; i < 4; i++)
        c.v[i] = ctr[i];
    fr_u32x4 r = fr_philox4x32_10(c, k0, k1);
    for (int i = 0; i < 4; i++)
        out[i] = r.v[i];
}
void og_expansion_draws(uint64_t seed, uint64_t stream, uint32_t iter, uint32_t cell, float out[2])
{
    fr_expansion_draws(seed, stream, iter, cell, &out[0], &out[1]);
}
