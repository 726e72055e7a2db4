/*
 * fr_kernels.hip — gfx950 kernels of the FastRRT planning loop.
 *
 * Work decomposition (one kernel per reference phase, all persistent and
 * grid-stride so that one launch covers any number of plans of uneven size):
 *   - cell kernels walk a flattened table of (plan slot, tile) pairs, 1024
 *     cells per tile, 4 cells per lane;
 *   - rollout and goal-connection kernels walk compacted queues built by the
 *     cell kernels (block-level prefix sums + one atomic per block), so the
 *     expensive per-node work scales with the number of nodes, not W*H;
 *   - per-plan decisions (frontier mode, collision solve, reseed, the goal
 *     connection adds) run on the device from each plan's control block, so
 *     a loop iteration needs no host round trip.
 *
 * Results are deterministic and independent of how plans are batched: each
 * expansion draws from Philox keyed by (seed, stream, iteration, cell),
 * every read of the grid in a phase sees the state before that phase, and
 * all races of the reference are resolved by ordered keys (the lowest source
 * cell claims an end cell; argmins break ties on the lowest cell).
 *
 * The footprint test (reference src/cuda/path_feasible.cu:18-78) samples the
 * same (2mz+1)(2mx+1) rotated lattice with the same FP32 arithmetic, but
 * looks each sample up in a padded Chebyshev distance-to-obstacle plane: a
 * sample at distance d certifies the next d-2 samples of its row, which are
 * skipped.  The boolean is identical to sampling every point.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "fr_device.h"
#include "fr_engine.h"

/* ------------------------------------------------------------------------ */
/* device helpers                                                           */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ int fr_host_trunc_d(double v)
{
    /* x86 host conversion: NaN / overflow -> INT_MIN (the reference runs path
     * optimisation on the host, src/cuda/graph_optimization.cu:92-124) */
    if (v != v || v >= 2147483648.0 || v <= -2147483649.0)
        return (-2147483647 - 1);
    return (int)v;
}

__device__ __forceinline__ unsigned long long fr_lcost_key(long long lc)
{
    return (unsigned long long)(lc + (1LL << 62));
}

/* path_feasible.cu:13-78, sample-identical, with distance-plane skipping. */
__device__ bool dev_feasible(const FrPlanDev &P, int x, int z, float h)
{
    const int pos = z * P.W + x;
    if (P.trav[pos] & 1)
        return true;
    float s, c;
    fr_sincosf(h, &s, &c);
    const FR_GLOBAL uint8_t *D = P.dist;
    const int TW = P.TW, pad = P.pad;
    if (c != c || s != s) /* every sample truncates to pixel (0, 0) */
        return D[fr_dist_at(TW, pad, pad)] != 0;
    const float fxc = (float)x, fzc = (float)z;
    const int mx = P.mx, mz = P.mz;
    for (int i = -mz; i <= mz; ++i) {
        const float is = (float)i * s;
        const float ic = (float)i * c;
        int j = -mx;
        while (j <= mx) {
            const float fj = (float)j;
            const float fx = fj * c - is + fxc;
            const float fz = fj * s + ic + fzc;
            const int xl = (int)fx;
            const int zl = (int)fz;
            const int d = D[fr_dist_at(TW, xl + pad, zl + pad)];
            if (d == 0)
                return false;
            j += (d > 2) ? (d - 1) : 1;
        }
    }
    return true;
}

/* The centre pixel's obstacle distance certifies the whole footprint: every
 * sample pixel lies within Chebyshev distance pad - 1 = ceil(hypot(mx, mz)) + 1
 * of the centre (rotation keeps |offset| <= hypot, truncation adds < 1). */
__device__ __forceinline__ bool dev_fp_certified(const FrPlanDev &P, int x, int z, float h)
{
    if (P.trav[z * P.W + x] & 1)
        return true;
    return (h - h) == 0.0f && P.dist[fr_dist_at(P.TW, x + P.pad, z + P.pad)] >= P.pad;
}

/* kinematics.cu:123-231.  Returns 0 when rejected. */
__device__ int dev_rollout(const FrPlanDev &P, int sx, int sz, float steer, float path_size, int *ex, int *ez,
                           float *eh, float *ec)
{
    const int W = P.W, H = P.H;
    float x = P.og_y - (float)sz;
    float y = (float)sx - P.og_x;
    if ((double)steer > P.max_steer_rad)
        steer = (float)P.max_steer_rad;
    else if ((double)steer < -P.max_steer_rad)
        steer = (float)(-P.max_steer_rad);
    const float t = fr_tanf(steer);
    const float dt = (float)0.1;
    const float ds = P.v * dt;
    const float beta = fr_atanf(t / 2.0f);
    const float hif = (float)((double)(ds * fr_cosf(beta) * t) / (2.0 * P.lr));
    const int max_size = fr_trunc_f(path_size) + 1;
    int size = 0;
    int last_x = sx, last_z = sz;
    const int start_pos = sz * W + sx;
    float heading = P.heading[start_pos];
    float node_cost = P.cost[start_pos];
    int loop_count = 0;
    const int loop_max = fr_trunc_f((float)(100 * max_size) / dt);
    const double ogx = (double)P.og_x, ogy = (double)P.og_y;
    while (size < max_size) {
        if (loop_count++ > loop_max)
            break;
        float sn, cs;
        fr_sincosf(heading + beta, &sn, &cs);
        x += ds * cs;
        y += ds * sn;
        heading += hif;
        const int cx = fr_trunc_d(ogx + (double)y);
        const int cz = fr_trunc_d(ogy - (double)x);
        if (cx == last_x && cz == last_z)
            continue;
        if (cx < 0 || cx >= W)
            break;
        if (cz < 0 || cz >= H)
            break;
        size += 1;
        node_cost += (P.intrinsic ? P.intrinsic[cz * W + cx] : 0.0f) + 1.0f;
        if (!dev_fp_certified(P, cx, cz, heading) && !dev_feasible(P, cx, cz, heading))
            return 0;
        last_x = cx;
        last_z = cz;
    }
    if (size < 1)
        return 0;
    *ex = last_x;
    *ez = last_z;
    *eh = heading;
    *ec = node_cost;
    return 1;
}

/* graph_direct_goal_connection.cu:16-167.  The curve is walked twice: the
 * first pass evaluates the curvature bound and the always-sampled centre
 * pixel of every point (both cheap); only curves that survive get footprint
 * tests.  The result is the same as the reference's single pass because any
 * failing point makes the whole connection fail. */
__device__ float dev_check_direct(const FrPlanDev &P, float max_curv, int x, int z, float lh, int gx, int gz,
                                  float gh, bool safe_zone, bool host_nan)
{
    const int W = P.W, H = P.H;
    const int pos = z * W + x;
    const int dxi = gx - x, dzi = gz - z;
    const float distance = sqrtf((float)(dxi * dxi + dzi * dzi));
    const int num_points = fr_trunc_f(distance);
    if (num_points <= 0)
        return -1;
    const float a1 = (float)((double)lh - FR_HALF_PI_D);
    const float a2 = (float)((double)gh - FR_HALF_PI_D);
    float s1, c1, s2, c2;
    fr_sincosf(a1, &s1, &c1);
    fr_sincosf(a2, &s2, &c2);
    const double t1x = (double)(distance * c1), t1y = (double)(distance * s1);
    const double t2x = (double)(distance * c2), t2y = (double)(distance * s2);
    const double dx0 = (double)x, dz0 = (double)z, dgx = (double)gx, dgz = (double)gz;
    const bool skip_fp = safe_zone && ((P.trav[pos] & 2) != 0);
    const double denom = (double)(num_points - 1);
    float node_cost = P.cost[pos];
    int last_x = -1, last_z = -1;
    for (int pass = 0; pass < 2; ++pass) {
        last_x = -1;
        last_z = -1;
        for (int i = 0; i < num_points; ++i) {
            const double t = (0.0 + (double)i) / denom;
            const double t2 = t * t, t3 = t2 * t;
            const double h00 = 2 * t3 - 3 * t2 + 1;
            const double h10 = t3 - 2 * t2 + t;
            const double h01 = -2 * t3 + 3 * t2;
            const double h11 = t3 - t2;
            const double px = h00 * dx0 + h10 * t1x + h01 * dgx + h11 * t2x;
            const double pz = h00 * dz0 + h10 * t1y + h01 * dgz + h11 * t2y;
            if (px < 0 || px >= W)
                continue;
            if (pz < 0 || pz >= H)
                continue;
            const int cx = host_nan ? fr_host_trunc_d(px) : fr_trunc_d(px);
            const int cz = host_nan ? fr_host_trunc_d(pz) : fr_trunc_d(pz);
            if (cx == last_x && cz == last_z)
                continue;
            if (cx < 0 || cx >= W)
                continue;
            if (cz < 0 || cz >= H)
                continue;
            if (pass == 1)
                node_cost += (P.intrinsic ? P.intrinsic[cz * W + cx] : 0.0f) + 1.0f;
            const double t00 = 6 * t2 - 6 * t;
            const double t10 = 3 * t2 - 4 * t + 1;
            const double t01 = -6 * t2 + 6 * t;
            const double t11 = 3 * t2 - 2 * t;
            const double ddx = t00 * dx0 + t10 * t1x + t01 * dgx + t11 * t2x;
            const double ddz = t00 * dz0 + t10 * t1y + t01 * dgz + t11 * t2y;
            last_x = cx;
            last_z = cz;
            if (pass == 0) {
                if (max_curv > 0) {
                    const double d00 = 12 * t - 6;
                    const double d10 = 6 * t - 4;
                    const double d01 = -12 * t + 6;
                    const double d11 = 6 * t - 2;
                    const double dd2x = d00 * dx0 + d10 * t1x + d01 * dgx + d11 * t2x;
                    const double dd2z = d00 * dz0 + d10 * t1y + d01 * dgz + d11 * t2y;
                    const double v = ddx * ddx + ddz * ddz;
                    const double cr = ddx * dd2z - ddz * dd2x;
                    const float k = (float)((cr < 0 ? -cr : cr) / (v * __builtin_sqrt(v)));
                    if (k > max_curv)
                        return -1;
                }
                if (!skip_fp && !(P.trav[cz * W + cx] & 1) && P.dist[fr_dist_at(P.TW, cx + P.pad, cz + P.pad)] == 0)
                    return -1;
            } else if (!skip_fp) {
                const float heading = (float)((double)fr_atan2f((float)ddz, (float)ddx) + FR_HALF_PI_D);
                if (!dev_feasible(P, cx, cz, heading))
                    return -1;
            }
        }
    }
    if (last_x != gx && last_z != gz)
        return -1;
    return node_cost;
}

/* ------------------------------------------------------------------------ */
/* wave-cooperative forms: every lane of a wave calls with the same          */
/* arguments and gets the same result; loads are spread over the 64 lanes    */
/* instead of walked serially by one lane.                                   */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ int dev_lane() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int dev_bcast_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ float dev_bcast_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

/* Lane l gets lane l-1's v; lane 0 gets `first` (DPP wave_shr:1). */
__device__ __forceinline__ float dev_wave_shr1(float v, float first)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(first), __float_as_int(v), 0x138, 0xF, 0xF, false));
}

/* The serial float sum a_l = (((a0 + d_0) + d_1) + ... + d_l) of the lanes'
 * increments, every a_l rounded as the serial loop rounds it: each round
 * adds lane l's increment to lane l-1's partial sum (lane 0 to a0).  Lane l
 * holds its exact sum from round l + 1 on, and it never changes after, since
 * lane l - 1's does not; 64 rounds of two instructions each. */
__device__ __forceinline__ float dev_wave_serial_sum(float a0, float d)
{
    float v = a0;
#pragma unroll
    for (int t = 0; t < 64; ++t)
        v = dev_wave_shr1(v, a0) + d;
    return v;
}

/* inclusive prefix sum over the wave */
__device__ __forceinline__ int dev_wave_incl_scan(int v)
{
    const int lane = dev_lane();
    for (int off = 1; off < 64; off <<= 1) {
        const int n = __shfl_up(v, off, 64);
        if (lane >= off)
            v += n;
    }
    return v;
}

/* path_feasible.cu:18-78 for one lane's own pose (lanes test different
 * poses at once), in two phases.  After the all-angle bit and the centre
 * certificate, the (2 mz + 1) x (2 mx + 1) samples are taken in 7 x 7 groups:
 * the group's middle sample is a probe; every other sample of the group lies
 * within hypot(3, 3) of it before truncation, so within Chebyshev distance 5
 * after, and a probe at distance >= 6 from every obstacle certifies its
 * group.  A probe at distance 0 is itself an obstacle sample.  The probe
 * phase reads every probe (offsets from the plan's table); the group phase
 * samples all 49 slots of each group its probe left open.  The sample
 * coordinates are the reference's expressions, so the result is its boolean;
 * the two phases may run at different times since any collision rejects. */
#define FR_PROBE_BATCH 24

struct FrOpen {
    float s, c;              /* sin, cos of the heading */
    uint32_t open;           /* groups left open by their probes */
};

__device__ __forceinline__ void dev_fp_cell(int j, int i, float s, float c, float fxc, float fzc, int *xl, int *zl)
{
    const float is = (float)i * s, ic = (float)i * c, fj = (float)j;
    *xl = (int)(fj * c - is + fxc);
    *zl = (int)(fj * s + ic + fzc);
}

/* group in row a, column b of the group grid */
__device__ __forceinline__ void dev_group_box(const FrPlanDev &P, int a, int b, int *i0, int *i1, int *j0, int *j1,
                                              int *ic, int *jc)
{
    *i0 = -P.mz + a * FR_GROUP;
    *j0 = -P.mx + b * FR_GROUP;
    *i1 = *i0 + FR_GROUP - 1 < P.mz ? *i0 + FR_GROUP - 1 : P.mz;
    *j1 = *j0 + FR_GROUP - 1 < P.mx ? *j0 + FR_GROUP - 1 : P.mx;
    *ic = *i0 + FR_GROUP / 2 < *i1 ? *i0 + FR_GROUP / 2 : *i1;
    *jc = *j0 + FR_GROUP / 2 < *j1 ? *j0 + FR_GROUP / 2 : *j1;
}

/* Cell of the sample in footprint column fj, row fi (floats of the
 * reference's integer indices): the expressions of path_feasible.cu:48-60. */
__device__ __forceinline__ void dev_fp_cell_f(float fj, float fi, float s, float c, float fxc, float fzc, int *xl,
                                              int *zl)
{
    const float is = fi * s, ic = fi * c;
    *xl = (int)(fj * c - is + fxc);
    *zl = (int)(fj * s + ic + fzc);
}

/* Probe phase: -1 collision, 1 free, 0 open (groups to sample in o).  The
 * probes' offsets come from the plan's table (P.probe_j / P.probe_i). */
__device__ int dev_lane_probe(const FrPlanDev &P, int x, int z, float h, FrOpen *o)
{
    const FR_GLOBAL uint8_t *D = P.dist;
    const int TW = P.TW, pad = P.pad;
    float s, c;
    fr_sincosf(h, &s, &c);
    const bool nan_h = c != c || s != s;
    const int ng = P.n_groups;
    const float fxc = (float)x, fzc = (float)z;
    /* Every byte the test may need first - the cell's bits, the centre and
     * origin distances, the first batch of probes - so that one round trip
     * to memory serves the common cases.  A NaN heading probes the centre
     * (its probes are not used then). */
    const float sp = nan_h ? 0.0f : s, cp = nan_h ? 0.0f : c;
    const int tv = P.trav[z * P.W + x];
    const int dc = D[fr_dist_at(TW, x + pad, z + pad)];
    const int d00 = D[fr_dist_at(TW, pad, pad)];
    int dp[FR_PROBE_BATCH];
#pragma unroll
    for (int k = 0; k < FR_PROBE_BATCH; ++k) {
        /* table entries past the footprint's groups hold the centre sample */
        int px, pz;
        dev_fp_cell_f(P.probe_j[k], P.probe_i[k], sp, cp, fxc, fzc, &px, &pz);
        dp[k] = D[fr_dist_at(TW, px + pad, pz + pad)];
    }
    if (tv & 1)
        return 1;
    if (nan_h)
        return d00 != 0 ? 1 : -1;
    if (dc >= pad)
        return 1;
    if (ng == 0)
        return dev_feasible(P, x, z, h) ? 1 : -1;
    uint32_t open = 0;
#pragma unroll
    for (int k = 0; k < FR_PROBE_BATCH; ++k) {
        if (k >= ng)
            break;
        if (dp[k] == 0)
            return -1;
        if (dp[k] < FR_GROUP_FREE) {
            open |= 1u << k;
        }
    }
    if (ng > FR_PROBE_BATCH) {
        /* the rest of a footprint of more than FR_PROBE_BATCH groups */
#pragma unroll
        for (int k = FR_PROBE_BATCH; k < FR_MAX_GROUPS; ++k) {
            int px, pz;
            dev_fp_cell_f(P.probe_j[k], P.probe_i[k], s, c, fxc, fzc, &px, &pz);
            dp[k - FR_PROBE_BATCH] = D[fr_dist_at(TW, px + pad, pz + pad)];
        }
#pragma unroll
        for (int k = FR_PROBE_BATCH; k < FR_MAX_GROUPS; ++k) {
            if (k >= ng)
                break;
            const int d = dp[k - FR_PROBE_BATCH];
            if (d == 0)
                return -1;
            if (d < FR_GROUP_FREE) {
                open |= 1u << k;
            }
        }
    }
    o->s = s;
    o->c = c;
    o->open = open;
    return open ? 0 : 1;
}

/* A pointer every lane of the wave holds (the callers' plan is uniform), in
 * scalar registers. */
template <typename T>
__device__ __forceinline__ const FR_GLOBAL T *dev_uniform_ptr(const FR_GLOBAL T *p)
{
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return (const FR_GLOBAL T *)(((unsigned long long)hi << 32) | lo);
}

/* Sampling of open group g of the footprint at (x, z): true when one of its
 * samples is an obstacle.  `D`: the plan's distance plane. */
__device__ __forceinline__ bool dev_group_hit(const FrPlanDev &P, const FR_GLOBAL uint8_t *D, int x, int z, float s,
                                              float c, int g)
{
    const int TW = P.TW, pad = P.pad;
    const int gc = (2 * P.mx + FR_GROUP) / FR_GROUP;
    const float fxc = (float)x, fzc = (float)z;
    int i0, i1, j0, j1, ic, jc;
    const int a = g / gc;
    dev_group_box(P, a, g - a * gc, &i0, &i1, &j0, &j1, &ic, &jc);
    /* The row and column products of path_feasible.cu:48-60 once per row and
     * column.  Rows and columns past the footprint (the last row or column of
     * groups may be narrower) repeat the group's last one, so every slot of
     * the 7 x 7 box tests a sample of the group (some twice) and the loads
     * issue back to back, with no per-lane branch (a load under one is
     * waited for before the branch joins). */
    float rs[FR_GROUP], rc[FR_GROUP], cc[FR_GROUP], cs[FR_GROUP];
#pragma unroll
    for (int k = 0; k < FR_GROUP; ++k) {
        const float fi = (float)(i0 + k < i1 ? i0 + k : i1), fj = (float)(j0 + k < j1 ? j0 + k : j1);
        rs[k] = fi * s;
        rc[k] = fi * c;
        cc[k] = fj * c;
        cs[k] = fj * s;
    }
    int lowest = 255;
#pragma unroll
    for (int di = 0; di < FR_GROUP; ++di) {
#pragma unroll
        for (int dj = 0; dj < FR_GROUP; ++dj) {
            const int xl = (int)(cc[dj] - rs[di] + fxc), zl = (int)(cs[dj] + rc[di] + fzc);
            const int v = D[fr_dist_at(TW, xl + pad, zl + pad)];
            lowest = v < lowest ? v : lowest;
        }
    }
    return lowest == 0;
}

/* Per-wave record of up to 64 footprints waiting for their group phase. */
struct FrOpenSet {
    int x[64], z[64];
    float s[64], c[64];
    uint32_t open[64];
    int pref[64];
};

/* Group phase of footprints 0..m-1 of a per-wave set, the open groups of all
 * of them dealt to the lanes in rounds of 64 (a footprint near obstacles
 * with many open groups no longer holds the wave for one round per group).
 * False when any collides. */
__device__ bool dev_groups_dealt(const FrPlanDev &P, FrOpenSet &bg, int m)
{
    /* the set is in LDS: addressed as such (ds_* instead of flat accesses;
     * this function is not inlined, so its reference arrives generic) */
    __attribute__((address_space(3))) FrOpenSet &b = *(__attribute__((address_space(3))) FrOpenSet *)&bg;
    const FR_GLOBAL uint8_t *D = dev_uniform_ptr(P.dist);
    const int lane = dev_lane();
    const int cnt = lane < m ? __popc(b.open[lane]) : 0;
    const int incl = dev_wave_incl_scan(cnt);
    const int total = dev_bcast_i(incl, 63);
    if (!total)
        return true;
    b.pref[lane] = incl - cnt;
    __builtin_amdgcn_wave_barrier();
    for (int t0 = 0; t0 < total; t0 += 64) {
        const int t = t0 + lane;
        bool hit = false;
        if (t < total) {
            /* footprint j: the last with pref[j] <= t (it has open groups) */
            int lo = 0, hi = m - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (b.pref[mid] <= t)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            uint32_t open = b.open[lo];
            for (int r = t - b.pref[lo]; r > 0; --r)
                open &= open - 1;
            hit = dev_group_hit(P, D, b.x[lo], b.z[lo], b.s[lo], b.c[lo], __ffs((int)open) - 1);
        }
        if (__ballot(hit))
            return false;
    }
    return true;
}


/* One Hermite point of graph_direct_goal_connection.cu:16-167: position,
 * in-bounds cell, first derivative. */
struct FrHermite {
    double dx0, dz0, dgx, dgz, t1x, t1y, t2x, t2y, denom, rden;
};

/* t = i / denom, correctly rounded.  For denom <= 32768 (every curve on grids
 * up to 23170 px on a side) it is the product with the rounded reciprocal
 * plus one fused correction, which equals the quotient for every 0 <= i <=
 * denom + 1 (checked exhaustively); denom = 0 gives NaN as 0 / 0 does. */
__device__ __forceinline__ double dev_curve_t(const FrHermite &hm, int i)
{
    const double a = 0.0 + (double)i;
    if (hm.denom <= 32768.0) {
        const double q0 = a * hm.rden;
        return __builtin_fma(__builtin_fma(-q0, hm.denom, a), hm.rden, q0);
    }
    return a / hm.denom;
}

__device__ __forceinline__ bool dev_hermite_point(const FrPlanDev &P, const FrHermite &hm, int i, bool host_nan,
                                                  int *cx, int *cz, double *t_out)
{
    const double t = dev_curve_t(hm, i);
    const double t2 = t * t, t3 = t2 * t;
    const double h00 = 2 * t3 - 3 * t2 + 1;
    const double h10 = t3 - 2 * t2 + t;
    const double h01 = -2 * t3 + 3 * t2;
    const double h11 = t3 - t2;
    const double px = h00 * hm.dx0 + h10 * hm.t1x + h01 * hm.dgx + h11 * hm.t2x;
    const double pz = h00 * hm.dz0 + h10 * hm.t1y + h01 * hm.dgz + h11 * hm.t2y;
    *t_out = t;
    if (px < 0 || px >= P.W)
        return false;
    if (pz < 0 || pz >= P.H)
        return false;
    const int x = host_nan ? fr_host_trunc_d(px) : fr_trunc_d(px);
    const int z = host_nan ? fr_host_trunc_d(pz) : fr_trunc_d(pz);
    *cx = x;
    *cz = z;
    return x >= 0 && x < P.W && z >= 0 && z < P.H;
}

#define FR_OPEN_BUF 128

/* Per-wave buffer of footprints whose probes left groups open. */
struct FrOpenBuf {
    int x[FR_OPEN_BUF], z[FR_OPEN_BUF];
    float s[FR_OPEN_BUF], c[FR_OPEN_BUF];
    uint32_t open[FR_OPEN_BUF];
};

/* Group phase of the first m (<= 64) buffered footprints (their groups dealt
 * to the lanes); the rest move to the front.  False when any collides. */
__device__ bool dev_open_flush(const FrPlanDev &P, FrOpenBuf &b, FrOpenSet &set, int m, int *nb)
{
    const int lane = dev_lane();
    if (lane < m) {
        set.x[lane] = b.x[lane];
        set.z[lane] = b.z[lane];
        set.s[lane] = b.s[lane];
        set.c[lane] = b.c[lane];
        set.open[lane] = b.open[lane];
    }
    __builtin_amdgcn_wave_barrier();
    if (!dev_groups_dealt(P, set, m))
        return false;
    *nb -= m;
    if (lane < *nb) {
        b.x[lane] = b.x[m + lane];
        b.z[lane] = b.z[m + lane];
        b.s[lane] = b.s[m + lane];
        b.c[lane] = b.c[m + lane];
        b.open[lane] = b.open[m + lane];
    }
    return true;
}

/* dev_check_direct with the curve's points spread over the lanes.  A point
 * contributes iff its cell is in the grid and differs from the cell of the
 * previous such point (the reference's last_x / last_z test: duplicates of
 * the last cell never change it, so "last accepted" = "last in-grid").  The
 * cost is summed lane by lane in point order, as the reference adds it.
 * Pass 0 tests curvature and the centre cells; pass 1 the footprints, whose
 * probe phase runs chunk by chunk while the groups left open wait in a
 * per-wave buffer and are sampled 64 at a time (a collision anywhere rejects,
 * so a curve that some probe rejects never has its groups sampled).  `passes`
 * selects the passes run (bit 0, bit 1): a caller may run pass 0 alone and
 * pass 1 later on the curves it kept; the cost is summed in pass 1. */
/* Hermite set-up of graph_direct_goal_connection.cu:16-60 (per lane);
 * false when the curve has no point. */
__device__ __forceinline__ bool dev_curve_setup(int x, int z, float lh, int gx, int gz, float gh, FrHermite *hm,
                                                int *np)
{
    const int dxi = gx - x, dzi = gz - z;
    const float distance = sqrtf((float)(dxi * dxi + dzi * dzi));
    const int num_points = fr_trunc_f(distance);
    *np = num_points;
    if (num_points <= 0)
        return false;
    const float a1 = (float)((double)lh - FR_HALF_PI_D);
    const float a2 = (float)((double)gh - FR_HALF_PI_D);
    float s1, c1, s2, c2;
    fr_sincosf(a1, &s1, &c1);
    fr_sincosf(a2, &s2, &c2);
    hm->t1x = (double)(distance * c1);
    hm->t1y = (double)(distance * s1);
    hm->t2x = (double)(distance * c2);
    hm->t2y = (double)(distance * s2);
    hm->dx0 = (double)x;
    hm->dz0 = (double)z;
    hm->dgx = (double)gx;
    hm->dgz = (double)gz;
    hm->denom = (double)(num_points - 1);
    hm->rden = 1.0 / hm->denom;
    return true;
}

__device__ __forceinline__ double dev_bcast_d(double v, int lane)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ FrHermite dev_bcast_hermite(const FrHermite &h, int lane)
{
    FrHermite o;
    o.dx0 = dev_bcast_d(h.dx0, lane);
    o.dz0 = dev_bcast_d(h.dz0, lane);
    o.dgx = dev_bcast_d(h.dgx, lane);
    o.dgz = dev_bcast_d(h.dgz, lane);
    o.t1x = dev_bcast_d(h.t1x, lane);
    o.t1y = dev_bcast_d(h.t1y, lane);
    o.t2x = dev_bcast_d(h.t2x, lane);
    o.t2y = dev_bcast_d(h.t2y, lane);
    o.denom = dev_bcast_d(h.denom, lane);
    o.rden = dev_bcast_d(h.rden, lane);
    return o;
}

__device__ __forceinline__ float dev_wave_curve(const FrPlanDev &P, const FrHermite &hm, int num_points, float max_curv,
                                                int gx, int gz, bool skip_fp, bool host_nan, float base, int passes);

__device__ __forceinline__ float dev_wave_check_direct(const FrPlanDev &P, float max_curv, int x, int z, float lh, int gx, int gz,
                                       float gh, bool safe_zone, bool host_nan, float base, int passes = 3)
{
    FrHermite hm;
    int num_points;
    if (!dev_curve_setup(x, z, lh, gx, gz, gh, &hm, &num_points))
        return -1;
    const bool skip_fp = safe_zone && ((P.trav[z * P.W + x] & 2) != 0);
    return dev_wave_curve(P, hm, num_points, max_curv, gx, gz, skip_fp, host_nan, base, passes);
}

/* Point b + lane of a curve: its cell, its parameter t, and whether it is
 * accepted (in the grid and in another cell than the previous in-grid point,
 * whose cell last_x / last_z carries from chunk to chunk).  *am: the
 * accepted lanes. */
__device__ __forceinline__ bool dev_curve_point(const FrPlanDev &P, const FrHermite &hm, int b, int num_points,
                                                bool host_nan, int &last_x, int &last_z, int *cx, int *cz, double *t,
                                                unsigned long long *am)
{
    const int lane = dev_lane();
    const int i = b + lane;
    *cx = 0;
    *cz = 0;
    *t = 0;
    const bool cand = i < num_points && dev_hermite_point(P, hm, i, host_nan, cx, cz, t);
    const unsigned long long m = __ballot(cand);
    const unsigned long long lower = m & ((1ull << lane) - 1ull);
    const int src = lower ? 63 - __clzll((long long)lower) : lane;
    int pcx = __shfl(*cx, src, 64), pcz = __shfl(*cz, src, 64);
    if (!lower) {
        pcx = last_x;
        pcz = last_z;
    }
    const bool acc = cand && !(*cx == pcx && *cz == pcz);
    if (m) {
        const int hi = 63 - __clzll((long long)m);
        last_x = dev_bcast_i(*cx, hi);
        last_z = dev_bcast_i(*cz, hi);
    }
    *am = __ballot(acc);
    return acc;
}

__device__ __forceinline__ void dev_curve_d1(const FrHermite &hm, double t, double *ddx, double *ddz)
{
    const double t2 = t * t;
    const double t00 = 6 * t2 - 6 * t;
    const double t10 = 3 * t2 - 4 * t + 1;
    const double t01 = -6 * t2 + 6 * t;
    const double t11 = 3 * t2 - 2 * t;
    *ddx = t00 * hm.dx0 + t10 * hm.t1x + t01 * hm.dgx + t11 * hm.t2x;
    *ddz = t00 * hm.dz0 + t10 * hm.t1y + t01 * hm.dgz + t11 * hm.t2y;
}

/* curvature test of graph_direct_goal_connection.cu (float of the double
 * curvature against max_curv).  The float estimate from the hardware
 * reciprocal square root is within 1e-5 of the exact value, so only values
 * within 0.1% of the bound (and non-finite ones) take the exact division. */
__device__ __forceinline__ bool dev_curve_too_sharp(const FrHermite &hm, double t, float max_curv)
{
    double ddx, ddz;
    dev_curve_d1(hm, t, &ddx, &ddz);
    const double d00 = 12 * t - 6;
    const double d10 = 6 * t - 4;
    const double d01 = -12 * t + 6;
    const double d11 = 6 * t - 2;
    const double dd2x = d00 * hm.dx0 + d10 * hm.t1x + d01 * hm.dgx + d11 * hm.t2x;
    const double dd2z = d00 * hm.dz0 + d10 * hm.t1y + d01 * hm.dgz + d11 * hm.t2y;
    const double v = ddx * ddx + ddz * ddz;
    const double cr = ddx * dd2z - ddz * dd2x;
    const double acr = cr < 0 ? -cr : cr;
    const float vf = (float)v;
    if (vf > 1e-20f && vf < 1e24f) { /* rs^3 stays a normal float */
        const float rs = __builtin_amdgcn_rsqf(vf);
        const float kf = (float)acr * (rs * rs * rs);
        if (kf < max_curv * 0.999f)
            return false;
        if (kf > max_curv * 1.001f && kf <= 3.0e38f)
            return true;
    }
    const float k = (float)(acr / (v * __builtin_sqrt(v)));
    return k > max_curv;
}

/* True when no point of the curve can fail the curvature test: an upper
 * bound of the curvature over all t in [0, 1] is at most 0.75 max_curv.
 * With P'(t) = A t^2 + B t + C and P''(t) = 2 A t + B (A = 6 (p0 - p1) +
 * 3 (m0 + m1), B = 6 (p1 - p0) - 4 m0 - 2 m1, C = m0), the cross product
 * P' x P'' = -(A x B) t^2 + 2 (C x A) t + C x B is bounded on [0, 1] by the
 * sum of its coefficients' magnitudes, and |P'| from below by its value at
 * the middles of four quarters less the largest |P''| (at an end: P'' is
 * linear) times the distance to a middle.  The 25 % margin dwarfs the
 * rounding of the per-point evaluation (relative 1e-12 when |P'| >= 0.5),
 * so a skipped test would have passed.  NaN set-ups give false. */
__device__ __forceinline__ bool dev_curve_gentle(const FrHermite &hm, float max_curv)
{
    const double ax = 6.0 * (hm.dx0 - hm.dgx) + 3.0 * (hm.t1x + hm.t2x);
    const double az = 6.0 * (hm.dz0 - hm.dgz) + 3.0 * (hm.t1y + hm.t2y);
    const double bx = 6.0 * (hm.dgx - hm.dx0) - 4.0 * hm.t1x - 2.0 * hm.t2x;
    const double bz = 6.0 * (hm.dgz - hm.dz0) - 4.0 * hm.t1y - 2.0 * hm.t2y;
    const double cx = hm.t1x, cz = hm.t1y;
    const double axb = ax * bz - az * bx, cxa = cx * az - cz * ax, cxb = cx * bz - cz * bx;
    const double qmax = __builtin_fabs(axb) + 2.0 * __builtin_fabs(cxa) + __builtin_fabs(cxb);
    const double ex = 2.0 * ax + bx, ez = 2.0 * az + bz;
    const double m2 = __builtin_fmax(bx * bx + bz * bz, ex * ex + ez * ez);
    double lo2 = 1e300;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double t = 0.125 + 0.25 * j;
        const double px = (ax * t + bx) * t + cx, pz = (az * t + bz) * t + cz;
        lo2 = __builtin_fmin(lo2, px * px + pz * pz);
    }
    const double L = __builtin_sqrt(lo2) - 0.125 * __builtin_sqrt(m2);
    return L >= 0.5 && qmax <= 0.75 * (double)max_curv * (L * L * L);
}

/* Exact first look of pass 0 at a set-up curve of more than 64 points, from
 * 64 of its points spread over it (lane l takes point l * s, s = ceil(n / 64)).
 * Rejection: pass 0 rejects a curve when an in-grid point's centre pixel is
 * an obstacle pixel not feasible at every heading, and every in-grid point
 * has the cell of a point that pass 0 tests (the first of its run of equal
 * cells), so a sample there is conclusive.  Certification (`cert_ok`: no
 * per-point curvature test, goal inside the grid): every point lies within
 * g = max(s / 2 + 1, n - 1 - last sample) steps of t = 1 / (n - 1) of a
 * sample, so within delta = M g / (n - 1) of it, M >= |P'| on [0, 1] being
 * |A| + |B| + |C| for P'(t) = A t^2 + B t + C; its cell is then within
 * Chebyshev distance ceil(delta) of the sample's, so samples all in the grid
 * with distance planes above ceil(delta) + 1 leave no obstacle pixel on the
 * curve, and the last in-grid point is the goal (t = 1 evaluates to it
 * exactly).  Returns -1 (rejected), 1 (passes, last cell the goal) or 0
 * (walk the curve).  Most rejected curves end here in one chunk instead of
 * the two or three the walk takes to reach and test their first obstacle.
 * (The same look at the footprints of pass 1 cost more than it saved: the
 * curves that reach pass 1 mostly pass it.) */
#ifndef FR_SPARSE
#define FR_SPARSE 1
#endif
__device__ __forceinline__ int dev_curve_sparse_pass0(const FrPlanDev &P, const FrHermite &hm, int num_points,
                                                      bool host_nan, bool cert_ok)
{
    const int s = (num_points + 63) >> 6;
    const int i = dev_lane() * s;
    int cx = 0, cz = 0;
    double t = 0;
    const bool in = i < num_points && dev_hermite_point(P, hm, i, host_nan, &cx, &cz, &t);
    int tv = 1, dv = 255;
    if (in) {
        tv = P.trav[cz * P.W + cx];
        dv = P.dist[fr_dist_at(P.TW, cx + P.pad, cz + P.pad)];
    }
    if (__ballot(in && !(tv & 1) && dv == 0))
        return -1;
    if (!cert_ok)
        return 0;
    const double ax = 6.0 * (hm.dx0 - hm.dgx) + 3.0 * (hm.t1x + hm.t2x);
    const double az = 6.0 * (hm.dz0 - hm.dgz) + 3.0 * (hm.t1y + hm.t2y);
    const double bx = 6.0 * (hm.dgx - hm.dx0) - 4.0 * hm.t1x - 2.0 * hm.t2x;
    const double bz = 6.0 * (hm.dgz - hm.dz0) - 4.0 * hm.t1y - 2.0 * hm.t2y;
    const double M = __builtin_sqrt(ax * ax + az * az) + __builtin_sqrt(bx * bx + bz * bz) +
                     __builtin_sqrt(hm.t1x * hm.t1x + hm.t1y * hm.t1y);
    const int last = s * ((num_points - 1) / s < 63 ? (num_points - 1) / s : 63);
    const int tail = num_points - 1 - last;
    const int g = s / 2 + 1 > tail ? s / 2 + 1 : tail;
    const double delta = M * (double)g / (double)(num_points - 1) * (1.0 + 1e-9) + 1e-6;
    if (!(delta < 60.0)) /* NaN, or too coarse for the capped distance plane */
        return 0;
    const int thr = (int)__builtin_ceil(delta) + 1;
    return __ballot(i < num_points && (!in || dv <= thr)) ? 0 : 1;
}

/* Pass 0, chunk by chunk of 64 points: each chunk's geometry and curvature,
 * then its centre cells' loads; the previous chunk's bytes are tested while
 * these are in flight.  Most rejected curves meet their first obstacle pixel
 * within the first chunk, so the walk usually ends there. */
__device__ __forceinline__ bool dev_wave_curve_pass0(const FrPlanDev &P, const FrHermite &hm, int num_points,
                                                     float max_curv, bool skip_fp, bool host_nan, int &last_x,
                                                     int &last_z)
{
    last_x = -1;
    last_z = -1;
    if (FR_SPARSE && !skip_fp && num_points > 64) {
        const bool cert_ok = !(max_curv > 0) && hm.dgx >= 0 && hm.dgx < P.W && hm.dgz >= 0 && hm.dgz < P.H;
        const int sp = dev_curve_sparse_pass0(P, hm, num_points, host_nan, cert_ok);
        if (sp < 0)
            return false;
        if (sp > 0) {
            last_x = (int)hm.dgx;
            last_z = (int)hm.dgz;
            return true;
        }
    }
    int tv_prev = 1, dv_prev = 1;
    for (int b = 0; b < num_points; b += 64) {
        double t;
        unsigned long long am;
        int cx, cz;
        const bool acc = dev_curve_point(P, hm, b, num_points, host_nan, last_x, last_z, &cx, &cz, &t, &am);
        const bool bad = acc && max_curv > 0 && dev_curve_too_sharp(hm, t, max_curv);
        if (__ballot(bad))
            return false;
        int tv = 1, dv = 1;
        if (acc && !skip_fp) {
            tv = P.trav[cz * P.W + cx];
            dv = P.dist[fr_dist_at(P.TW, cx + P.pad, cz + P.pad)];
        }
        if (__ballot(!(tv_prev & 1) && dv_prev == 0))
            return false;
        tv_prev = tv;
        dv_prev = dv;
    }
    return !__ballot(!(tv_prev & 1) && dv_prev == 0);
}

/* The passes of dev_wave_check_direct over a set-up curve. */
__device__ __forceinline__ float dev_wave_curve(const FrPlanDev &P, const FrHermite &hm, int num_points, float max_curv,
                                                int gx, int gz, bool skip_fp, bool host_nan, float base, int passes)
{
    const int W = P.W;
    const int lane = dev_lane();
    int last_x = -1, last_z = -1;
    if ((passes & 1) && !dev_wave_curve_pass0(P, hm, num_points, max_curv, skip_fp, host_nan, last_x, last_z))
        return -1;
    float node_cost = base;
    if (passes & 2) {
        __shared__ FrOpenBuf s_open[FR_BLOCK / 64];
        __shared__ FrOpenSet s_oset[FR_BLOCK / 64];
        const int wslot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        FrOpenBuf &ob = s_open[wslot];
        FrOpenSet &os = s_oset[wslot];
        int nb = 0;
        const unsigned long long below = (1ull << lane) - 1ull;
        last_x = -1;
        last_z = -1;
        for (int b = 0; b < num_points; b += 64) {
            int cx, cz;
            double t;
            unsigned long long am;
            const bool acc = dev_curve_point(P, hm, b, num_points, host_nan, last_x, last_z, &cx, &cz, &t, &am);
            if (!am)
                continue;
            float add = 0;
            int pr = 1;
            FrOpen o;
            if (acc) {
                add = (P.intrinsic ? P.intrinsic[cz * W + cx] : 0.0f) + 1.0f;
                /* a point whose centre distance certifies every heading (or
                 * whose pixel is feasible at every heading) needs no heading:
                 * the derivative, atan2 and sincos run only for the others.
                 * With one point (t = 0/0) the heading is NaN, which the
                 * certificate does not cover. */
                const bool cert = !skip_fp && num_points > 1 &&
                                  ((P.trav[cz * W + cx] & 1) ||
                                   P.dist[fr_dist_at(P.TW, cx + P.pad, cz + P.pad)] >= P.pad);
                if (!skip_fp && !cert) {
                    double ddx, ddz;
                    dev_curve_d1(hm, t, &ddx, &ddz);
                    const float hd = (float)((double)fr_atan2f((float)ddz, (float)ddx) + FR_HALF_PI_D);
                    pr = dev_lane_probe(P, cx, cz, hd, &o);
                }
            }
            if (__ballot(pr < 0))
                return -1;
            unsigned long long cm = am;
            while (cm) {
                const int l = __ffsll((long long)cm) - 1;
                cm &= cm - 1;
                node_cost += dev_bcast_f(add, l);
            }
            const unsigned long long om = __ballot(pr == 0);
            if (pr == 0) {
                const int r = nb + __popcll(om & below);
                ob.x[r] = cx;
                ob.z[r] = cz;
                ob.s[r] = o.s;
                ob.c[r] = o.c;
                ob.open[r] = o.open;
            }
            nb += __popcll(om);
            if (nb >= 64 && !dev_open_flush(P, ob, os, 64, &nb))
                return -1;
        }
        while (nb > 0)
            if (!dev_open_flush(P, ob, os, nb < 64 ? nb : 64, &nb))
                return -1;
    }
    if (last_x != gx && last_z != gz)
        return -1;
    return node_cost;
}

/* graph.cu:219-234 */
__device__ __forceinline__ double dev_compute_heading(int x1, int z1, int x2, int z2)
{
    const double dz = (double)(z2 - z1), dx = (double)(x2 - x1);
    if (dx == 0 && dz == 0)
        return 0;
    double v1;
    if (dz != 0)
        v1 = (double)fr_atan2f((float)(-dz), (float)dx);
    else
        v1 = (double)fr_atan2f(0.0f, (float)dx);
    return FR_HALF_PI_D - v1;
}

/* graph.cu:364-374 executed on the device */
__device__ void dev_add(const FrPlanDev &P, int x, int z, float h, int ppx, int ppz, float cost)
{
    if (x < 0 || x >= P.W || z < 0 || z >= P.H)
        return;
    const int pos = z * P.W + x;
    if (ppx != -1 && ppz != -1) {
        const long pp = (long)ppz * P.W + ppx;
        if (pp >= 0 && pp < P.N)
            P.derive[pp]++;
    }
    P.type[pos] = FR_TYPE_NODE;
    P.px[pos] = ppx;
    P.pz[pos] = ppz;
    P.heading[pos] = h;
    P.cost[pos] = cost;
}

__device__ __forceinline__ bool dev_goal_hit(const FrPlanDev &P, int x, int z, float h)
{
    const double dx = (double)(P.gx - x), dz = (double)(P.gz - z);
    if (__builtin_sqrt(dx * dx + dz * dz) > (double)P.dist_tol)
        return false;
    return fabsf(h - P.gh) <= P.herr_tol;
}

/* Block-wide exclusive scan of one int per thread; returns the block total. */
__device__ int block_exclusive_scan(int v, int *out_prefix)
{
    __shared__ int s_wave[FR_BLOCK / 64];
    __shared__ int s_total;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        int n = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += n;
    }
    if (lane == 63)
        s_wave[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < FR_BLOCK / 64; ++w) {
            int t = s_wave[w];
            s_wave[w] = acc;
            acc += t;
        }
        s_total = acc;
    }
    __syncthreads();
    *out_prefix = s_wave[wid] + incl - v;
    int total = s_total;
    __syncthreads();
    return total;
}

/* ------------------------------------------------------------------------ */
/* per-plan setup                                                           */
/* ------------------------------------------------------------------------ */

/* Frame -> traversability bits and the raw obstacle mask.  An obstacle is a
 * pixel whose class cost is negative (path_feasible.cu:472-475); pixels in
 * the ego box are skipped by the reference (:469-470) and count as free; a
 * class id outside the table counts as an obstacle. */
__device__ __forceinline__ void dev_prep_px(const FrPlanDev &P, int pos, float c0, float c2, uint8_t *tv, uint8_t *mk)
{
    const int z = pos / P.W, x = pos - z * P.W;
    const int trav = fr_trunc_f(c2);
    *tv = (uint8_t)((trav & 1) | ((trav & 0x100) ? 2 : 0));
    const int cls = fr_trunc_f(c0);
    int obst = (cls < 0 || cls >= P.n_classes) ? 1 : (P.class_costs[cls] < 0);
    if (x >= P.lbx && x <= P.ubx && z >= P.ubz && z <= P.lbz)
        obst = 0;
    *mk = (uint8_t)obst;
}

/* Four pixels per thread: their 48 bytes of frame in three 16-byte loads,
 * both byte planes written a 32-bit word at a time (the planes and the
 * frame start 256-byte aligned). */
__global__ void __launch_bounds__(FR_BLOCK) k_setup_prep(const FrPlanDev *plans, const int *slots)
{
    const FrPlanDev &P = plans[slots[blockIdx.y]];
    const int quads = P.N >> 2;
    const float4 *F = (const float4 *)P.frame;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += gridDim.x * blockDim.x) {
        const float4 a = F[3 * q], b = F[3 * q + 1], c = F[3 * q + 2];
        /* pixels: (a.x a.y a.z) (a.w b.x b.y) (b.z b.w c.x) (c.y c.z c.w) */
        uint8_t tv[4], mk[4];
        dev_prep_px(P, 4 * q, a.x, a.z, &tv[0], &mk[0]);
        dev_prep_px(P, 4 * q + 1, a.w, b.y, &tv[1], &mk[1]);
        dev_prep_px(P, 4 * q + 2, b.z, c.x, &tv[2], &mk[2]);
        dev_prep_px(P, 4 * q + 3, c.y, c.w, &tv[3], &mk[3]);
        ((uint32_t *)P.trav)[q] = (uint32_t)tv[0] | ((uint32_t)tv[1] << 8) | ((uint32_t)tv[2] << 16) |
                                  ((uint32_t)tv[3] << 24);
        ((uint32_t *)P.mark)[q] = (uint32_t)mk[0] | ((uint32_t)mk[1] << 8) | ((uint32_t)mk[2] << 16) |
                                  ((uint32_t)mk[3] << 24);
    }
    /* the last N mod 4 pixels */
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int pos = 4 * quads + t;
    if (t < (P.N & 3)) {
        const float *f = P.frame + 3 * (long)pos;
        dev_prep_px(P, pos, f[0], f[2], &P.trav[pos], &P.mark[pos]);
    }
}

/* Row pass of the padded Chebyshev distance plane: horizontal distance to
 * the nearest obstacle pixel of the same row, capped at dev_dist_cap.  One
 * wave per padded row: a forward sweep carries the last obstacle to the
 * left, a backward sweep the next one to the right (ballots over 64-pixel
 * chunks). */
/* Every reader of the distance plane compares it with values up to pad
 * (the centre certificate: D >= pad), so distances are kept up to
 * min(pad, FR_DIST_CAP) only; the column pass then scans at most that many
 * rows either side. */
__device__ __forceinline__ int dev_dist_cap(const FrPlanDev &P) { return P.pad < FR_DIST_CAP ? P.pad : FR_DIST_CAP; }

__global__ void __launch_bounds__(FR_BLOCK) k_setup_rowdist(const FrPlanDev *plans, const int *slots)
{
    const FrPlanDev &P = plans[slots[blockIdx.y]];
    const int cap = dev_dist_cap(P);
    const int wpb = blockDim.x >> 6;
    const int lane = dev_lane();
    for (int zp = blockIdx.x * wpb + (int)(threadIdx.x >> 6); zp < P.PH; zp += gridDim.x * wpb) {
        uint8_t *out = P.dist_tmp + (long)zp * P.PW;
        const int z = zp - P.pad;
        if (z < 0 || z >= P.H) {
            for (int xp = lane; xp < P.PW; xp += 64)
                out[xp] = (uint8_t)cap;
            continue;
        }
        const uint8_t *m = P.mark + (long)z * P.W;
        int last = -100000; /* grid x of the last obstacle to the left */
        for (int b = 0; b < P.PW; b += 64) {
            const int xp = b + lane, x = xp - P.pad;
            const bool ob = xp < P.PW && x >= 0 && x < P.W && m[x] != 0;
            const unsigned long long msk = __ballot(ob);
            const unsigned long long upto = msk & (lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull));
            const int prev = upto ? b + 63 - __clzll((long long)upto) - P.pad : last;
            if (xp < P.PW) {
                const int d = x - prev;
                out[xp] = (uint8_t)(d < cap ? d : cap);
            }
            if (msk)
                last = b + 63 - __clzll((long long)msk) - P.pad;
        }
        int next = 100000; /* grid x of the next obstacle to the right */
        const int nb = (P.PW + 63) / 64;
        for (int c = nb - 1; c >= 0; --c) {
            const int b = c * 64;
            const int xp = b + lane, x = xp - P.pad;
            const bool ob = xp < P.PW && x >= 0 && x < P.W && m[x] != 0;
            const unsigned long long msk = __ballot(ob);
            const unsigned long long from = msk & ~((1ull << lane) - 1ull);
            const int nx = from ? b + __ffsll((long long)from) - 1 - P.pad : next;
            if (xp < P.PW) {
                const int d = nx - x;
                const int cur = out[xp];
                if (d < cur)
                    out[xp] = (uint8_t)d;
            }
            if (msk)
                next = b + __ffsll((long long)msk) - 1 - P.pad;
        }
    }
}

/* Column pass: D(x, z) = min_dz max(|dz|, rowdist(x, z + dz)), capped.  A
 * dz at or beyond the running minimum cannot lower it.  One block per 64 x
 * 64 tile of the padded plane: the tile's columns, cap rows either side
 * included, are staged in LDS (rows beyond the plane read 255), and each
 * lane scans its column there for 16 rows of the tile. */
#define FR_COL_TILE 64
__global__ void __launch_bounds__(256) k_setup_coldist(const FrPlanDev *plans, const int *slots)
{
    __shared__ uint8_t s_t[(FR_COL_TILE + 2 * FR_DIST_CAP) * FR_COL_TILE];
    const FrPlanDev &P = plans[slots[blockIdx.y]];
    const int cap = dev_dist_cap(P);
    const int ntx = (P.PW + FR_COL_TILE - 1) / FR_COL_TILE;
    const int nty = (P.PH + FR_COL_TILE - 1) / FR_COL_TILE;
    const uint8_t *T = P.dist_tmp;
    const int c = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
        const int ty = tile / ntx, tx = tile - ty * ntx;
        const int x0 = tx * FR_COL_TILE, z0 = ty * FR_COL_TILE;
        const int rows = FR_COL_TILE + 2 * cap; /* staged rows z0 - cap .. z0 + 63 + cap */
        const int xp = x0 + c;
        __syncthreads(); /* the previous tile's reads are done */
        for (int r = wv; r < rows; r += 4) {
            const int zp = z0 - cap + r;
            uint8_t v = 255;
            if (zp >= 0 && zp < P.PH && xp < P.PW)
                v = T[(long)zp * P.PW + xp];
            s_t[r * FR_COL_TILE + c] = v;
        }
        __syncthreads();
        if (xp < P.PW) {
            for (int k = 0; k < FR_COL_TILE / 4; ++k) {
                const int r = wv * (FR_COL_TILE / 4) + k; /* row within the tile */
                const int zp = z0 + r;
                if (zp >= P.PH)
                    break;
                const uint8_t *col = s_t + (r + cap) * FR_COL_TILE + c;
                int best = col[0];
                for (int dz = 1; dz < best; ++dz) {
                    const int up = col[-dz * FR_COL_TILE], dn = col[dz * FR_COL_TILE];
                    int v = up < dn ? up : dn;
                    v = v > dz ? v : dz;
                    best = v < best ? v : best;
                }
                P.dist[fr_dist_at(P.TW, xp, zp)] = (uint8_t)best;
            }
        }
    }
}

/* graph_direct_goal_connection.cu:169-198 for one cell: Hermite to the goal
 * (one wave).  The start cost is the cost plane as it was at setup, kept in
 * gc_cost by k_setup_gc_reset, so an evaluation made later gives the same
 * planes as the reference's pass at set-up time. */
/* Result of one cell's evaluation in every plane.  (The lazy kernels store
 * less: k_gc_demand has set the evaluated bit of every cell it queues, and a
 * queued cell's flag bit is still clear, each cell being evaluated once
 * after the reset, so there only a flagged cell touches the bit planes.) */
__device__ void dev_gc_store(const FrPlanDev &P, int pos, float lh, float c)
{
    if (dev_lane() == 0) {
        P.gc_flag[pos] = c < 0 ? 0 : 1;
        if (c >= 0) {
            P.gc_heading[pos] = lh;
            P.gc_cost[pos] = c;
            atomicOr(&P.gc_bits[pos >> 5], 1u << (pos & 31));
        } else {
            atomicAnd(&P.gc_bits[pos >> 5], ~(1u << (pos & 31)));
        }
        atomicOr(&P.gc_done[pos >> 5], 1u << (pos & 31));
    }
}

__device__ __forceinline__ void dev_gc_cell(const FrPlanDev &P, int pos)
{
    const int z = pos / P.W, x = pos - z * P.W;
    const float lh = (float)dev_compute_heading(x, z, P.gx, P.gz);
    const float c = dev_wave_check_direct(P, P.max_curv_goal, x, z, lh, P.gx, P.gz, P.gh, false, false,
                                          P.gc_cost[pos]);
    dev_gc_store(P, pos, lh, c);
}

/* snapshot of the cost plane, flags cleared, nothing evaluated */
__global__ void __launch_bounds__(FR_BLOCK) k_setup_gc_reset(const FrPlanDev *plans, const int *slots)
{
    const FrPlanDev &P = plans[slots[blockIdx.y]];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x) {
        P.gc_cost[pos] = P.cost[pos];
        P.gc_flag[pos] = 0;
        if ((pos & 31) == 0) {
            P.gc_done[pos >> 5] = 0;
            P.gc_bits[pos >> 5] = 0;
        }
    }
}

/* every cell (skip_done: every cell not yet evaluated), one wave per cell */
__global__ void __launch_bounds__(FR_BLOCK) k_setup_goal_connection(const FrPlanDev *plans, const int *slots, int skip_done)
{
    const FrPlanDev &P = plans[slots[blockIdx.y]];
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    for (int pos = blockIdx.x * wpb + (int)(threadIdx.x >> 6); pos < P.N; pos += nw) {
        if (skip_done && (P.gc_done[pos >> 5] & (1u << (pos & 31))))
            continue;
        dev_gc_cell(P, pos);
    }
}

/* ------------------------------------------------------------------------ */
/* loop iteration                                                           */
/* ------------------------------------------------------------------------ */

/* The state a loop iteration starts from, for one plan (one wave: the lanes
 * clear the density counts the iteration will write, lane 0 the control
 * fields; `iter`: the iteration about to run).  Run by k_finish for the
 * plan's next iteration; k_begin runs it again at the FastRRT handle's
 * entries (a repeat only advances the claim epoch once more). */
__device__ void dev_ctl_ready(const FrPlanDev &P, FrCtl &C, int iter)
{
    const int lane = dev_lane();
    if (P.smart) {
        const int nr = P.dens_w * P.dens_h;
        FR_GLOBAL uint32_t *wr = P.region + ((iter + 1) & 1) * nr;
        for (int r = lane; r < nr; r += 64)
            wr[r] = 0;
    }
    if (lane != 0)
        return;
    C.frontier = C.last_temp >= 100;
    C.force = C.last_temp == 0;
    C.collision = 0;
    C.n_node_pre = 0;
    C.n_conv = 0;
    C.n_temp = 0;
    C.n_all = 0;
    C.reseeded = 0;
    C.found = 0;
    C.n_expansions = 0;
    C.n_fresh = 0;
    C.best_lcost = (long long)~0ull; /* holds fr_lcost_key() values */
    C.best_pick = ~0ull;
    C.pick_slow = 0;
    C.epoch += 1;
    C.chain_only = 0;
}

__device__ __forceinline__ void dev_counters_reset(int *counters)
{
    counters[FR_CNT_EXP] = 0;
    counters[FR_CNT_FRESH] = 0;
    counters[FR_CNT_PAIRS] = 0;
    counters[FR_CNT_DEMAND] = 0;
    counters[FR_CNT_DEMAND2] = 0;
}

/* Entry reset (one wave per plan) before a loop iteration of the FastRRT
 * handle, whose host may have edited the plan since the last iteration. */
__global__ void __launch_bounds__(FR_BLOCK) k_begin(const FrPlanDev *plans, FrCtl *ctl, const int *slots, const int *n_slots_p, int *counters)
{
    const int i = blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    const int n_slots = *n_slots_p;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        dev_counters_reset(counters);
    if (i >= n_slots)
        return;
    const int s = slots[i];
    FrCtl &C = ctl[s];
    if (C.status != FR_SEARCHING)
        return;
    dev_ctl_ready(plans[s], C, C.iter);
}

/* Tile scans: one wave per tile of FR_TILE = 1024 cells, lane l holding
 * the type bytes of cells base + 16 l .. base + 16 l + 15 (one 16-byte
 * load: the type plane is 256-byte aligned and padded to a multiple of 256
 * bytes; cells at or past N read as NULL).  A tile without the cells a scan
 * looks for costs one load and one ballot. */
__device__ __forceinline__ uint4 dev_tile_types(const FrPlanDev &P, int base)
{
    const int p0 = base + 16 * dev_lane();
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (p0 < P.N) {
        v = *reinterpret_cast<const uint4 *>(P.type + p0);
        const int left = P.N - p0;
        if (left < 16) {
            uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if (i >= left)
                    w[i >> 2] &= ~(0xFFu << (8 * (i & 3)));
            v = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
    return v;
}

__device__ __forceinline__ int dev_byte(uint4 v, int i)
{
    const uint32_t w = i < 4 ? v.x : (i < 8 ? v.y : (i < 12 ? v.z : v.w));
    return (int)((w >> (8 * (i & 3))) & 0xFFu);
}

/* 16-bit mask of the bytes of v equal to t */
__device__ __forceinline__ uint32_t dev_bytes_eq(uint4 v, int t)
{
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        m |= (uint32_t)(dev_byte(v, i) == t) << i;
    return m;
}

/* 16-bit mask of the non-zero bytes of v */
__device__ __forceinline__ uint32_t dev_bytes_nz(uint4 v)
{
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        m |= (uint32_t)(dev_byte(v, i) != 0) << i;
    return m;
}

/* Wave-wide reservation of `cnt` consecutive queue entries per lane in lane
 * order; returns this lane's first index (one atomic per wave). */
__device__ __forceinline__ int dev_wave_reserve(int *counter, int cnt, int *total_out)
{
    const int incl = dev_wave_incl_scan(cnt);
    const int total = dev_bcast_i(incl, 63);
    *total_out = total;
    int at = 0;
    if (total && dev_lane() == 0)
        at = atomicAdd(counter, total);
    return dev_bcast_i(at, 0) + incl - cnt;
}

/* First index and stride of a one-wave-per-item loop over [0, n) whose
 * waves take the items of their XCD's contiguous eighth (block b runs on
 * XCD b % 8) when the grid is a multiple of 8 blocks, and every
 * gridDim-th wave's item otherwise. */
#define FR_XCD_ITEMS(T, n, first, stride, end)                                                             \
    const bool first##_x8 = (gridDim.x & 7u) == 0;                                                          \
    const int first##_x = first##_x8 ? (int)(blockIdx.x & 7u) : 0;                                          \
    const int first##_d = first##_x8 ? 8 : 1;                                                               \
    const T end = (T)((long)(n) * (first##_x + 1) / first##_d);                                             \
    const T first = (T)((long)(n) * first##_x / first##_d) + (T)(blockIdx.x / first##_d) * wpb +            \
                    (T)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));                             \
    const T stride = (T)(gridDim.x / first##_d) * wpb;

/* Gate of __CUDA_random_node_expansion (graph_expansion.cu:157-164) and of
 * __CUDA_smart_node_expansion (graph_expansion_search_zones.cu:461-473):
 * queue one rollout per eligible NODE cell. */
__global__ void __launch_bounds__(FR_BLOCK) k_select(const FrPlanDev *plans, FrCtl *ctl, const int2 *tiles, const int *n_tiles_p,
                         FrExp *q, int *counters)
{
    const int n_tiles = *n_tiles_p;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    const int lane = dev_lane();
    for (int t = blockIdx.x * wpb + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); t < n_tiles; t += nw) {
        const int2 tl = tiles[t];
        const int slot = __builtin_amdgcn_readfirstlane(tl.x);
        const FrPlanDev &P = plans[slot];
        FrCtl &C = ctl[slot];
        if (C.status != FR_SEARCHING)
            continue;
        const int base = __builtin_amdgcn_readfirstlane(tl.y) * FR_TILE;
        const uint4 v = dev_tile_types(P, base);
        const uint32_t nodes_m = dev_bytes_eq(v, FR_TYPE_NODE);
        if (!__ballot(nodes_m != 0))
            continue;
        const int nr = P.dens_w * P.dens_h;
        const uint32_t *rd = P.region + (C.iter & 1) * nr;
        uint32_t ok_m = 0;
        for (uint32_t m = nodes_m; m; m &= m - 1) {
            const int i = __ffs((int)m) - 1;
            const int pos = base + 16 * lane + i;
            bool ok;
            if (P.smart) {
                if (C.force)
                    ok = true;
                else if (C.frontier)
                    ok = P.derive[pos] == 0;
                else {
                    const int z = pos / P.W, x = pos - z * P.W;
                    const int dp = (z / P.region_side) * P.dens_w + x / P.region_side;
                    ok = (double)rd[dp] <= 0.5 * (double)P.region_side;
                }
            } else {
                ok = !(C.frontier && P.derive[pos] > 0);
            }
            if (ok)
                ok_m |= 1u << i;
        }
        if (P.exp_budget > 0) {
            /* Optional thinning (no counterpart in the reference, which
             * expands every eligible node): each eligible node is kept with
             * probability budget / (eligible nodes of the previous
             * iteration), drawn from the third word of its expansion's
             * Philox block, so the plan stays independent of the batch. */
            const int elig = dev_bcast_i(dev_wave_incl_scan(__popc(ok_m)), 63);
            if (lane == 0 && elig)
                atomicAdd(&C.n_eligible, elig);
            const int prev = C.prev_eligible;
            if (prev > P.exp_budget) {
                const uint32_t thr = (uint32_t)((double)P.exp_budget / (double)prev * 4294967296.0);
                uint32_t keep = 0;
                for (uint32_t m = ok_m; m; m &= m - 1) {
                    const int i = __ffs((int)m) - 1;
                    if (fr_expansion_gate(P.seed, P.stream, (uint32_t)C.iter, (uint32_t)(base + 16 * lane + i)) < thr)
                        keep |= 1u << i;
                }
                ok_m = keep;
            }
        }
        int total;
        int w = dev_wave_reserve(&counters[FR_CNT_EXP], __popc(ok_m), &total);
        const int node_total = dev_bcast_i(dev_wave_incl_scan(__popc(nodes_m)), 63);
        for (uint32_t m = ok_m; m; m &= m - 1) {
            FrExp e;
            e.slot = slot;
            e.src = base + 16 * lane + __ffs((int)m) - 1;
            e.end = -1;
            e.heading = 0;
            e.cost = 0;
            q[w++] = e;
        }
        if (lane == 0) {
            if (node_total)
                atomicAdd(&C.n_node_pre, node_total);
            if (total)
                atomicAdd(&C.n_expansions, total);
        }
    }
}

/* expand_node (graph_expansion.cu:103-145) with check_kinematic_new_path
 * (kinematics.cu:123-231), one wave per rollout.  The step chain is walked
 * 128 steps at a time, two consecutive steps per lane: the heading and
 * position recurrences are float sums whose rounding order matters, so they
 * are accumulated step by step, as systolic sums along the wave (each round
 * adds a lane's two increments to the previous lane's last value); the
 * per-step sincos and the cell of each step are spread over the lanes, and the new
 * cells' footprints are tested 64 at a time once the walk's geometry is
 * known (dev_roll_flush).  A step's cell is new iff it differs from the
 * previous step's cell
 * (the reference's last_x / last_z: a repeated cell never changes it, a new
 * in-grid cell becomes it, a new out-of-grid cell ends the walk).  The walk
 * ends before step k when k > loop_max, or after the step that is out of the
 * grid or that completes max_size cells; the end heading is the chain value
 * after the last step executed.  Every rollout reads the pre-expansion grid;
 * the end cell is claimed here and written by k_commit. */
#define FR_ROLL_BUF 192 /* < 64 cells left after a flush + at most one new cell per step of a 128-step chunk */

/* 128 positions of a walk chunk in walk order (two steps per lane: position
 * 2 l and 2 l + 1 on lane l), as two 64-bit words. */
struct FrMask128 {
    unsigned long long lo, hi;
};

/* bits 0..31 of x to the even positions */
__device__ __forceinline__ unsigned long long dev_spread32(unsigned long long x)
{
    x &= 0xFFFFFFFFull;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

/* a: ballot of the lanes' first steps, b: of their second steps */
__device__ __forceinline__ FrMask128 dev_interleave(unsigned long long a, unsigned long long b)
{
    FrMask128 m;
    m.lo = dev_spread32(a) | (dev_spread32(b) << 1);
    m.hi = dev_spread32(a >> 32) | (dev_spread32(b >> 32) << 1);
    return m;
}

__device__ __forceinline__ int dev_m128_first(const FrMask128 &m)
{
    return m.lo ? __ffsll((long long)m.lo) - 1 : (m.hi ? 64 + __ffsll((long long)m.hi) - 1 : 128);
}

/* position of the n-th (n >= 1) set position; n is at most the count */
__device__ __forceinline__ int dev_m128_nth(const FrMask128 &m, int n)
{
    const int c = __popcll(m.lo);
    unsigned long long w = n <= c ? m.lo : m.hi;
    const int base = n <= c ? 0 : 64;
    for (int r = 1; r < (n <= c ? n : n - c); ++r)
        w &= w - 1;
    return base + __ffsll((long long)w) - 1;
}

/* the positions below lim */
__device__ __forceinline__ FrMask128 dev_m128_below(const FrMask128 &m, int lim)
{
    FrMask128 r;
    r.lo = lim >= 64 ? m.lo : (m.lo & ((1ull << lim) - 1ull));
    r.hi = lim >= 128 ? m.hi : (lim <= 64 ? 0ull : (m.hi & ((1ull << (lim - 64)) - 1ull)));
    return r;
}


/* Footprint tests and cost of the first m (<= 64) buffered cells of a
 * rollout, one cell per lane, all lanes' loads in flight together; the cost
 * is added cell by cell in walk order.  False when any footprint collides. */
__device__ bool dev_roll_flush(const FrPlanDev &P, const int *bx, const int *bz, const float *bh, int m,
                               float &node_cost, FrOpenSet &set)
{
    const int lane = dev_lane();
    float add = 0;
    int pr = 1;
    FrOpen o = {};
    if (lane < m) {
        const int cx = bx[lane], cz = bz[lane];
        add = (P.intrinsic ? P.intrinsic[cz * P.W + cx] : 0.0f) + 1.0f;
        /* certified cells skip the sincos and probe loads */
        if (!dev_fp_certified(P, cx, cz, bh[lane]))
            pr = dev_lane_probe(P, cx, cz, bh[lane], &o);
        set.x[lane] = cx;
        set.z[lane] = cz;
        set.s[lane] = o.s;
        set.c[lane] = o.c;
        set.open[lane] = pr == 0 ? o.open : 0u;
    }
    if (__ballot(pr < 0))
        return false;
    __builtin_amdgcn_wave_barrier();
    if (__ballot(pr == 0) && !dev_groups_dealt(P, set, m))
        return false;
    for (int r = 0; r < m; ++r)
        node_cost += dev_bcast_f(add, r);
    return true;
}

__global__ void __launch_bounds__(FR_BLOCK) k_expand(const FrPlanDev *__restrict__ plans, FrCtl *__restrict__ ctl,
                                                     FrExp *__restrict__ q, const int *__restrict__ counters)
{
    __shared__ int s_bx[FR_BLOCK / 64][FR_ROLL_BUF];
    __shared__ int s_bz[FR_BLOCK / 64][FR_ROLL_BUF];
    __shared__ float s_bh[FR_BLOCK / 64][FR_ROLL_BUF];
    __shared__ FrOpenSet s_set[FR_BLOCK / 64];
    const int n = counters[FR_CNT_EXP];
    const int wpb = blockDim.x >> 6;
    /* wave index made uniform, so a rollout's descriptor and its plan's
     * fields are read through scalar loads */
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = dev_lane();
    int *bx = s_bx[wv], *bz = s_bz[wv];
    float *bh = s_bh[wv];
    /* Workgroups are dealt to the eight XCDs round robin (block b runs on
     * XCD b % 8); each XCD takes one contiguous eighth of the queue, which
     * holds a plan's rollouts together, so a plan's planes are read through
     * one XCD's L2.  The grid is a multiple of 8 blocks. */
    const int xcd = (int)(blockIdx.x & 7u), jb = (int)(blockIdx.x >> 3);
    const int nbx = (int)(gridDim.x >> 3);
    if (jb >= nbx)
        return;
    const int lo = (int)((long long)n * xcd / 8), hi = (int)((long long)n * (xcd + 1) / 8);
    for (int i = lo + jb * wpb + wv; i < hi; i += nbx * wpb) {
        const FrExp e = q[i];
        const int slot = __builtin_amdgcn_readfirstlane(e.slot);
        const FrPlanDev &P = plans[slot];
        FrCtl &C = ctl[slot];
        const int src = __builtin_amdgcn_readfirstlane(e.src);
        const int W = P.W, H = P.H;
        const int sz = src / W, sx = src - sz * W;
        /* the node's heading and cost are fetched before the draws and the
         * trigonometry, so their latency overlaps that arithmetic */
        float h = P.heading[src];
        float node_cost = P.cost[src];
        float us, ul;
        fr_expansion_draws(P.seed, P.stream, (uint32_t)C.iter, (uint32_t)src, &us, &ul);
        float steer = (2.0f * us - 1.0f) * (float)P.max_steer_rad;
        float psize = P.max_path * ul;
        if (psize < 5.0f)
            psize = 5.0f;
        /* kinematics.cu:150-180 */
        float xf = P.og_y - (float)sz;
        float yf = (float)sx - P.og_x;
        if ((double)steer > P.max_steer_rad)
            steer = (float)P.max_steer_rad;
        else if ((double)steer < -P.max_steer_rad)
            steer = (float)(-P.max_steer_rad);
        const float t = fr_tanf(steer);
        const float dt = (float)0.1;
        const float ds = P.v * dt;
        const float beta = fr_atanf(t / 2.0f);
        const float hif = (float)((double)(ds * fr_cosf(beta) * t) / (2.0 * P.lr));
        const int max_size = fr_trunc_f(psize) + 1;
        const int loop_max = fr_trunc_f((float)(100 * max_size) / dt);
        const double ogx = (double)P.og_x, ogy = (double)P.og_y;
        int size = 0, last_x = sx, last_z = sz, prev_x = sx, prev_z = sz;
        float end_h = h;
        bool ok = true;
        int nb = 0; /* cells buffered, footprints not yet tested */
        for (int k0 = 0;; k0 += 128) {
            /* two steps per lane, a = k0 + 2 lane and b = a + 1: each round
             * of the serial sums adds lane l's two increments after lane
             * l - 1's second sum (the chunk's start on lane 0) */
            const int ka = k0 + 2 * lane, kb = ka + 1;
            float vA = h, vB = h;
#pragma unroll
            for (int t = 0; t < 64; ++t) {
                vA = dev_wave_shr1(vB, h) + hif;
                vB = vA + hif;
            }
            const float hA1 = vA, hB1 = vB, hA0 = dev_wave_shr1(vB, h), hB0 = vA;
            const float hh = dev_bcast_f(vB, 63);
            float snA, csA, snB, csB;
            fr_sincosf(hA0 + beta, &snA, &csA);
            fr_sincosf(hB0 + beta, &snB, &csB);
            const float dxA = ds * csA, dyA = ds * snA, dxB = ds * csB, dyB = ds * snB;
            float uA = xf, uB = xf, wA = yf, wB = yf;
#pragma unroll
            for (int t = 0; t < 64; ++t) {
                uA = dev_wave_shr1(uB, xf) + dxA;
                wA = dev_wave_shr1(wB, yf) + dyA;
                uB = uA + dxB;
                wB = wA + dyB;
            }
            const float xx = dev_bcast_f(uB, 63), yy = dev_bcast_f(wB, 63);
            const int cxA = fr_trunc_d(ogx + (double)wA), czA = fr_trunc_d(ogy - (double)uA);
            const int cxB = fr_trunc_d(ogx + (double)wB), czB = fr_trunc_d(ogy - (double)uB);
            int px = __shfl_up(cxB, 1, 64), pz = __shfl_up(czB, 1, 64);
            if (lane == 0) {
                px = prev_x;
                pz = prev_z;
            }
            const bool freshA = !(cxA == px && czA == pz);
            const bool freshB = !(cxB == cxA && czB == czA);
            const bool outA = cxA < 0 || cxA >= W || czA < 0 || czA >= H;
            const bool outB = cxB < 0 || cxB >= W || czB < 0 || czB >= H;
            const FrMask128 m_stop = dev_interleave(__ballot(ka > loop_max), __ballot(kb > loop_max));
            const FrMask128 m_oob = dev_interleave(__ballot(freshA && outA), __ballot(freshB && outB));
            const FrMask128 m_in = dev_interleave(__ballot(freshA && !outA), __ballot(freshB && !outB));
            /* first position at which the walk ends: before its step (loop
             * cap), or after it (out of the grid, or max_size reached) */
            const int t_before = dev_m128_first(m_stop);
            int t_at = dev_m128_first(m_oob);
            const int need = max_size - size;
            if (__popcll(m_in.lo) + __popcll(m_in.hi) >= need) {
                const int t_reach = dev_m128_nth(m_in, need);
                t_at = t_reach < t_at ? t_reach : t_at;
            }
            const bool ends_before = t_before <= t_at && t_before < 128;
            const bool ends_at = !ends_before && t_at < 128;
            const int lim = ends_before ? t_before : (ends_at ? t_at + 1 : 128);
            const FrMask128 acc = dev_m128_below(m_in, lim);
            /* the new cells join the buffer in walk order */
            const unsigned long long aw = lane < 32 ? acc.lo : acc.hi;
            const int sh = 2 * (lane & 31);
            const bool accA = (aw >> sh) & 1ull, accB = (aw >> (sh + 1)) & 1ull;
            const int rA = nb + (lane < 32 ? 0 : __popcll(acc.lo)) + __popcll(aw & ((1ull << sh) - 1ull));
            if (accA) {
                bx[rA] = cxA;
                bz[rA] = czA;
                bh[rA] = hA1;
            }
            if (accB) {
                const int rB = rA + (accA ? 1 : 0);
                bx[rB] = cxB;
                bz[rB] = czB;
                bh[rB] = hB1;
            }
            const int n_acc = __popcll(acc.lo) + __popcll(acc.hi);
            nb += n_acc;
            size += n_acc;
            if (n_acc) {
                const int pl = acc.hi ? 127 - __clzll((long long)acc.hi) : 63 - __clzll((long long)acc.lo);
                const int L = pl >> 1;
                const int lxA = dev_bcast_i(cxA, L), lzA = dev_bcast_i(czA, L);
                const int lxB = dev_bcast_i(cxB, L), lzB = dev_bcast_i(czB, L);
                last_x = (pl & 1) ? lxB : lxA;
                last_z = (pl & 1) ? lzB : lzA;
            }
            bool done = true;
            if (ends_before) {
                const int L = t_before >> 1;
                const float a0 = dev_bcast_f(hA0, L), b0 = dev_bcast_f(hB0, L);
                end_h = (t_before & 1) ? b0 : a0;
            } else if (ends_at) {
                const int L = t_at >> 1;
                const float a1 = dev_bcast_f(hA1, L), b1 = dev_bcast_f(hB1, L);
                end_h = (t_at & 1) ? b1 : a1;
            } else {
                done = false;
                h = hh;
                xf = xx;
                yf = yy;
                prev_x = dev_bcast_i(cxB, 63);
                prev_z = dev_bcast_i(czB, 63);
            }
            while (nb >= 64 || (done && nb > 0)) {
                const int m = nb < 64 ? nb : 64;
                if (!dev_roll_flush(P, bx, bz, bh, m, node_cost, s_set[wv])) {
                    ok = false;
                    break;
                }
                nb -= m;
                /* the rest to the front, 64 entries at a time (an entry is
                 * read before any write reaches its place) */
                for (int j = lane; j < nb; j += 64) {
                    bx[j] = bx[64 + j];
                    bz[j] = bz[64 + j];
                    bh[j] = bh[64 + j];
                }
            }
            if (done || !ok)
                break;
        }
        if (lane != 0)
            continue;
        atomicAdd(&C.w_cells, size);
        FrExp &o = q[i];
        o.end = -1;
        if (!ok || size < 1)
            continue;
        if ((float)last_x == P.og_x && (float)last_z == P.og_y)
            continue;
        const int epos = last_z * W + last_x;
        if (epos == src)
            continue;
        const int st = P.type[epos];
        if (st != FR_TYPE_NULL && st != FR_TYPE_NODE)
            continue;
        o.end = epos;
        o.heading = end_h;
        o.cost = node_cost;
        const unsigned long long key = ((unsigned long long)(0xFFFFFFFFu - C.epoch) << 32) | (unsigned)src;
        atomicMin(&P.claim[epos], key);
    }
}

__global__ void __launch_bounds__(FR_BLOCK) k_commit(const FrPlanDev *plans, FrCtl *ctl, const FrExp *q, const int *counters)
{
    const int n = counters[FR_CNT_EXP];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const FrExp e = q[i];
        if (e.end < 0)
            continue;
        const FrPlanDev &P = plans[e.slot];
        FrCtl &C = ctl[e.slot];
        const unsigned long long key = ((unsigned long long)(0xFFFFFFFFu - C.epoch) << 32) | (unsigned)e.src;
        if (P.claim[e.end] != key)
            continue;
        const int st = P.type[e.end];
        const int sz = e.src / P.W, sx = e.src - sz * P.W;
        P.px[e.end] = sx;
        P.pz[e.end] = sz;
        P.heading[e.end] = e.heading;
        P.cost[e.end] = e.cost;
        if (st == FR_TYPE_NULL) {
            P.type[e.end] = FR_TYPE_TEMP;
            P.derive[e.src] += 1;
        } else {
            P.type[e.end] = FR_TYPE_COLLISION;
            P.derive[e.src] -= 1;
            C.collision = 1;
            atomicAdd(&C.n_conv, 1);
        }
    }
}

/* __CUDA_count_nodes_in_density_region (graph_expansion_search_zones.cu:42-66) */
/* Region counts of smart plans after the expansion (wave `wave` of `nw`
 * over the tiles); run by the density blocks of k_erase_mark. */
__device__ void dev_density_tiles(const FrPlanDev *plans, const FrCtl *ctl, const int2 *tiles, int n_tiles, int wave,
                                  int nw)
{
    const int lane = dev_lane();
    for (int t = wave; t < n_tiles; t += nw) {
        const int2 tl = tiles[t];
        const int slot = __builtin_amdgcn_readfirstlane(tl.x);
        const FrPlanDev &P = plans[slot];
        const FrCtl &C = ctl[slot];
        if (C.status != FR_SEARCHING || !P.smart)
            continue;
        const int base = __builtin_amdgcn_readfirstlane(tl.y) * FR_TILE;
        const uint4 v = dev_tile_types(P, base);
        /* counted: every type but NULL and PROCESSING */
        const uint32_t cm = dev_bytes_nz(v) & ~dev_bytes_eq(v, FR_TYPE_PROCESSING);
        if (!__ballot(cm != 0))
            continue;
        const int nr = P.dens_w * P.dens_h;
        uint32_t *wr = P.region + ((C.iter + 1) & 1) * nr;
        int cur = -1, acc = 0;
        for (uint32_t m = cm; m; m &= m - 1) {
            const int pos = base + 16 * lane + __ffs((int)m) - 1;
            const int z = pos / P.W, x = pos - z * P.W;
            const int r = (z / P.region_side) * P.dens_w + x / P.region_side;
            if (r != cur) {
                if (acc)
                    atomicAdd(&wr[cur], (uint32_t)acc);
                cur = r;
                acc = 0;
            }
            acc++;
        }
        if (acc)
            atomicAdd(&wr[cur], (uint32_t)acc);
    }
}

__device__ __forceinline__ bool dev_solve_runs(const FrCtl &C)
{
    return C.collision && (C.n_node_pre - C.n_conv) > 2;
}

/* __CUDA_solveGraphCollision_erase_trees (graph_collision.cu:33-83), deciding
 * on the pre-erase grid: marks only. */
#define FR_ERASE_BLOCK 1024
#define FR_ERASE_CAP 4096

/* Blocks [0, n_mark_blocks): one block per plan that runs the erase.  The
 * blocks after them count the smart plans' density regions
 * (graph_expansion_search_zones.cu:42-66, after the expansion and before the
 * erase, which this kernel only marks), so that pass needs no launch of its
 * own.  The plan's non-NULL cells are
 * gathered into LDS in cell order; each one's verdict (some ancestor is
 * COLLISION or NULL) is then found by pointer jumping over parent indices in
 * LDS, doubling the covered chain length each round, instead of each cell
 * walking its chain through global memory.  The rounds cover at least the
 * walk's bound of nodes + 4 steps; a chain ends at the root (its own parent
 * in the jump) or at its first bad parent, as the walk does.  A plan with
 * more than FR_ERASE_CAP non-NULL cells jumps over buffers in global memory
 * instead of LDS. */
__global__ void __launch_bounds__(FR_ERASE_BLOCK) k_erase_mark(const FrPlanDev *plans, const FrCtl *ctl,
                                                               const int *slots, const int *n_slots_p,
                                                               const int2 *tiles, const int *n_tiles_p,
                                                               int n_mark_blocks)
{
    __shared__ int s_pos[FR_ERASE_CAP];
    __shared__ uint16_t s_anc[2][FR_ERASE_CAP];
    __shared__ uint8_t s_bad[2][FR_ERASE_CAP];
    __shared__ int s_wsum[FR_ERASE_BLOCK / 64];
    if ((int)blockIdx.x >= n_mark_blocks) {
        const int wpb = FR_ERASE_BLOCK / 64;
        dev_density_tiles(plans, ctl, tiles, *n_tiles_p,
                          ((int)blockIdx.x - n_mark_blocks) * wpb +
                              __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)),
                          ((int)gridDim.x - n_mark_blocks) * wpb);
        return;
    }
    if ((int)blockIdx.x >= *n_slots_p)
        return;
    const int slot = slots[blockIdx.x];
    const FrPlanDev &P = plans[slot];
    const FrCtl &C = ctl[slot];
    if (C.status != FR_SEARCHING || !dev_solve_runs(C))
        return;
    const int nodes = C.n_node_pre - C.n_conv;
    const int t = threadIdx.x, lane = dev_lane(), wid = t >> 6;
    const int nchunks = (P.N + 15) / 16;
    const int per = (nchunks + FR_ERASE_BLOCK - 1) / FR_ERASE_BLOCK;
    const int c0 = t * per, c1 = c0 + per < nchunks ? c0 + per : nchunks;
    /* count this thread's non-NULL cells (four chunks' loads in flight at a
     * time), block-wide exclusive scan */
    int cnt = 0;
    for (int cb = c0; cb < c1; cb += 4) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            v[u] = cb + u < c1 ? *reinterpret_cast<const uint4 *>(P.type + 16 * (cb + u)) : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = cb + u;
            uint32_t nz = dev_bytes_nz(v[u]);
            if (16 * c + 16 > P.N)
                nz &= c < c1 ? (1u << (P.N - 16 * c)) - 1u : 0u;
            cnt += __popc(nz);
        }
    }
    const int incl = dev_wave_incl_scan(cnt);
    if (lane == 63)
        s_wsum[wid] = incl;
    __syncthreads();
    int base = 0, total = 0;
    for (int w = 0; w < FR_ERASE_BLOCK / 64; ++w) {
        base += w < wid ? s_wsum[w] : 0;
        total += s_wsum[w];
    }
    const int off = base + incl - cnt;
    if (total > FR_ERASE_CAP) {
        /* The same pointer jumping over buffers in global memory (the plan's
         * erase scratch): the node list in cell order, each listed cell's
         * list index, and two jump buffers holding the ancestor index with
         * the bad bit on top.  One block, so the block barrier orders the
         * rounds. */
        int *gpos = P.escr, *gidx = P.escr + P.N;
        int *gj[2] = {P.escr + 2 * (long)P.N, P.escr + 3 * (long)P.N};
        int w = off;
        for (int c = c0; cnt && c < c1; ++c) { /* a thread without cells skips its chunks */
            const uint4 v = *reinterpret_cast<const uint4 *>(P.type + 16 * c);
            uint32_t nz = dev_bytes_nz(v);
            if (16 * c + 16 > P.N)
                nz &= (1u << (P.N - 16 * c)) - 1u;
            for (uint32_t m = nz; m; m &= m - 1) {
                const int pos = 16 * c + __ffs((int)m) - 1;
                gpos[w] = pos;
                gidx[pos] = w;
                ++w;
            }
        }
        __syncthreads();
        for (int k = t; k < total; k += FR_ERASE_BLOCK) {
            const int pos = gpos[k];
            const int ppx = P.px[pos], ppz = P.pz[pos];
            int v = k;
            if (!(ppx == -1 && ppz == -1)) {
                const long next = (long)ppz * P.W + ppx;
                const int tn = (next >= 0 && next < P.N) ? P.type[next] : FR_TYPE_NULL;
                if (tn == FR_TYPE_COLLISION || tn == FR_TYPE_NULL)
                    v = (int)((unsigned)k | 0x80000000u);
                else /* non-NULL, so in the list */
                    v = gidx[next];
            }
            gj[0][k] = v;
        }
        __syncthreads();
        const int grounds = 32 - __clz(nodes + 4);
        int gcur = 0;
        for (int r = 0; r < grounds; ++r) {
            const int *src = gj[gcur];
            int *dst = gj[gcur ^ 1];
            /* four entries per thread at a time, so their loads overlap */
            for (int k0 = t; k0 < total; k0 += 4 * FR_ERASE_BLOCK) {
                unsigned mine[4], up[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * FR_ERASE_BLOCK;
                    mine[u] = k < total ? (unsigned)src[k] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * FR_ERASE_BLOCK;
                    up[u] = k < total ? (unsigned)src[mine[u] & 0x7FFFFFFFu] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * FR_ERASE_BLOCK;
                    if (k < total)
                        dst[k] = (int)((up[u] & 0x7FFFFFFFu) | ((mine[u] | up[u]) & 0x80000000u));
                }
            }
            __syncthreads();
            gcur ^= 1;
        }
        for (int k = t; k < total; k += FR_ERASE_BLOCK)
            P.mark[gpos[k]] = (uint8_t)((unsigned)gj[gcur][k] >> 31);
        return;
    }
    /* the non-NULL cells in cell order (a thread without cells skips its chunks) */
    int w = off;
    for (int c = c0; cnt && c < c1; ++c) {
        const uint4 v = *reinterpret_cast<const uint4 *>(P.type + 16 * c);
        uint32_t nz = dev_bytes_nz(v);
        if (16 * c + 16 > P.N)
            nz &= (1u << (P.N - 16 * c)) - 1u;
        for (uint32_t m = nz; m; m &= m - 1)
            s_pos[w++] = 16 * c + __ffs((int)m) - 1;
    }
    __syncthreads();
    /* first step of every chain: root (own parent), bad parent, or the parent's index */
    for (int k = t; k < total; k += FR_ERASE_BLOCK) {
        const int pos = s_pos[k];
        const int ppx = P.px[pos], ppz = P.pz[pos];
        int anc = k, bad = 0;
        if (!(ppx == -1 && ppz == -1)) {
            const long next = (long)ppz * P.W + ppx;
            const int tn = (next >= 0 && next < P.N) ? P.type[next] : FR_TYPE_NULL;
            if (tn == FR_TYPE_COLLISION || tn == FR_TYPE_NULL) {
                bad = 1;
            } else { /* non-NULL, so in the list */
                int lo = 0, hi = total - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_pos[mid] < (int)next)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                anc = lo;
            }
        }
        s_anc[0][k] = (uint16_t)anc;
        s_bad[0][k] = (uint8_t)bad;
    }
    __syncthreads();
    const int rounds = 32 - __clz(nodes + 4);
    int cur = 0;
    for (int r = 0; r < rounds; ++r) {
        for (int k = t; k < total; k += FR_ERASE_BLOCK) {
            const int a = s_anc[cur][k];
            s_bad[cur ^ 1][k] = s_bad[cur][k] | s_bad[cur][a];
            s_anc[cur ^ 1][k] = s_anc[cur][a];
        }
        __syncthreads();
        cur ^= 1;
    }
    /* marks are read for non-NULL cells only (k_accept) */
    for (int k = t; k < total; k += FR_ERASE_BLOCK)
        P.mark[s_pos[k]] = s_bad[cur][k];
}

/* One sweep over the tiles: the erase's verdicts and
 * __CUDA_solveGraphCollision_set_nodes (graph_collision.cu:85-104:
 * COLLISION -> NODE with no children, the parent gaining one), then count(TEMP),
 * countAll() and __CUDA_accept_derived_nodes (fastrrt.cpp:130-142,
 * graph_expansion.cu:36-65) on the tile as the erase left it.  Converted
 * nodes, accepted nodes, the re-added start and (after external edits)
 * every node are queued for the goal test and the goal connection search.
 * A converted node's parent is never converted itself (a COLLISION parent
 * marks the child for erasing), so the parents' increments need no order
 * against the converted nodes' resets. */
__global__ void __launch_bounds__(FR_BLOCK) k_accept(const FrPlanDev *plans, FrCtl *ctl, const int2 *tiles, const int *n_tiles_p,
                         FrFresh *fq, int *counters)
{
    const int n_tiles = *n_tiles_p;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    const int lane = dev_lane();
    for (int t = blockIdx.x * wpb + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); t < n_tiles; t += nw) {
        const int2 tl = tiles[t];
        const int slot = __builtin_amdgcn_readfirstlane(tl.x);
        const FrPlanDev &P = plans[slot];
        FrCtl &C = ctl[slot];
        if (C.status != FR_SEARCHING)
            continue;
        const int base = __builtin_amdgcn_readfirstlane(tl.y) * FR_TILE;
        const uint4 v = dev_tile_types(P, base);
        uint32_t nz = dev_bytes_nz(v);
        if (!__ballot(nz != 0))
            continue;
        uint32_t temps = dev_bytes_eq(v, FR_TYPE_TEMP);
        uint32_t nodes = dev_bytes_eq(v, FR_TYPE_NODE);
        uint32_t conv = 0;
        if (dev_solve_runs(C)) {
            /* k_erase_mark marked non-NULL cells only */
            const uint32_t coll = dev_bytes_eq(v, FR_TYPE_COLLISION);
            for (uint32_t m = nz; m; m &= m - 1) {
                const int i = __ffs((int)m) - 1;
                const int pos = base + 16 * lane + i;
                if (P.mark[pos]) {
                    P.type[pos] = FR_TYPE_NULL;
                    nz &= ~(1u << i);
                    temps &= ~(1u << i);
                    nodes &= ~(1u << i);
                } else if ((coll >> i) & 1u) {
                    P.type[pos] = FR_TYPE_NODE;
                    P.derive[pos] = 0;
                    const long pp = (long)P.pz[pos] * P.W + P.px[pos];
                    if (pp >= 0 && pp < P.N)
                        atomicAdd(&P.derive[pp], 1);
                    conv |= 1u << i;
                    nodes |= 1u << i;
                }
            }
        }
        const int start_pos = (P.sx >= 0 && P.sx < P.W && P.sz >= 0 && P.sz < P.H) ? P.sz * P.W + P.sx : -1;
        uint32_t fresh = temps;
        if (C.full_check)
            fresh |= nodes;
        else if (C.pending_start && start_pos >= base + 16 * lane && start_pos < base + 16 * lane + 16)
            fresh |= nodes & (1u << (start_pos - base - 16 * lane));
        for (uint32_t m = temps; m; m &= m - 1)
            P.type[base + 16 * lane + __ffs((int)m) - 1] = FR_TYPE_NODE;
        /* a converted node is queued once, as converted (flagged), and once
         * more when the accept queues every node */
        int total;
        int w = dev_wave_reserve(&counters[FR_CNT_FRESH], __popc(conv) + __popc(fresh), &total);
        const int ttot = dev_bcast_i(dev_wave_incl_scan(__popc(temps)), 63);
        const int atot = dev_bcast_i(dev_wave_incl_scan(__popc(nz)), 63);
        for (int pass = 0; pass < 2; ++pass) {
            for (uint32_t m = pass ? fresh : conv; m; m &= m - 1) {
                FrFresh f;
                f.slot = slot;
                f.pos = base + 16 * lane + __ffs((int)m) - 1;
                f.flags = pass ? 0 : FR_FRESH_CONVERTED;
                f.pair_base = 0;
                f.pair_cnt = 0;
                f.pad0 = 0;
                f.key = ~0ull;
                fq[w++] = f;
            }
        }
        if (lane == 0) {
            if (ttot)
                atomicAdd(&C.n_temp, ttot);
            if (atot)
                atomicAdd(&C.n_all, atot);
        }
    }
}

/* ---- goal connection of the newly accepted nodes --------------------------
 * __CUDA__check_goal_reached_with_direct_connection_cost and
 * __CUDA__check_goal_reached_with_direct_connection
 * (graph_check_goal_reached.cu:17-139): every node is tried against every
 * goal-flagged cell of its (2r)^2 window; the pair with the least
 * floor(100 * total cost) wins.  Here each (node, flagged cell) pair is one
 * lane.  A node's window is enumerated in the reference's scan order so that
 * ties inside one node keep the last cell, as the reference's sequential
 * per-thread scan does; ties across nodes keep the lowest cell index.
 * Only nodes that became NODE this iteration are tried: every older node was
 * tried when it appeared, and had it connected the search would have ended.
 */

#define FR_RANK_BITS 21
#define FR_RANK_MASK ((1ull << FR_RANK_BITS) - 1)

__device__ __forceinline__ unsigned long long dev_node_key(long long lc, int rank)
{
    const long long lim = (1LL << 42) - 1;
    if (lc > lim)
        lc = lim;
    if (lc < -lim)
        lc = -lim;
    unsigned long long r = (unsigned long long)(rank < 0 ? 0 : rank);
    if (r > FR_RANK_MASK)
        r = FR_RANK_MASK;
    return ((unsigned long long)(lc + (1LL << 42)) << FR_RANK_BITS) | (FR_RANK_MASK - r);
}

__device__ __forceinline__ long long dev_key_lcost(unsigned long long key)
{
    return (long long)(key >> FR_RANK_BITS) - (1LL << 42);
}

__device__ __forceinline__ int dev_key_rank(unsigned long long key) { return (int)(FR_RANK_MASK - (key & FR_RANK_MASK)); }

/* window bounds of graph_check_goal_reached.cu:295-296 (float radius) */
__device__ __forceinline__ void dev_window(const FrPlanDev &P, int x, int z, int *z0, int *z1, int *x0, int *x1)
{
    const float r = P.dist_tol;
    int a = fr_trunc_f((float)z - r), b = a;
    while ((float)b < (float)z + r)
        b++;
    int c = fr_trunc_f((float)x - r), d = c;
    while ((float)d < (float)x + r)
        d++;
    *z0 = a < 0 ? 0 : a;
    *z1 = b > P.H ? P.H : b;
    *x0 = c < 0 ? 0 : c;
    *x1 = d > P.W ? P.W : d;
}

/* A window by 32-cell words of a bit plane: row r of the window covers the
 * cells [(z0 + r) W + x0, (z0 + r) W + x0 + wx), i.e. at most wpr words; pair
 * p = r * wpr + k is the k-th of them, with the mask of the window's bits.
 * Pairs in increasing order visit the window's cells in row-major order. */
struct FrWin {
    int z0, x0, wx, wz, wpr, npairs;
};

__device__ __forceinline__ bool dev_win_make(const FrPlanDev &P, int x, int z, FrWin *w)
{
    int z0, z1, x0, x1;
    dev_window(P, x, z, &z0, &z1, &x0, &x1);
    w->z0 = z0;
    w->x0 = x0;
    w->wx = x1 - x0;
    w->wz = z1 - z0;
    if (w->wx <= 0 || w->wz <= 0)
        return false;
    w->wpr = (w->wx + 62) / 32;
    w->npairs = w->wz * w->wpr;
    return true;
}

__device__ __forceinline__ uint32_t dev_win_pair(const FrPlanDev &P, const FrWin &w, int p, long *word)
{
    *word = 0;
    if (p >= w.npairs)
        return 0u;
    const int r = p / w.wpr, k = p - r * w.wpr;
    const long qa = (long)(w.z0 + r) * P.W + w.x0, qb = qa + w.wx - 1;
    const long wa = qa >> 5, wb = qb >> 5, wi = wa + k;
    if (wi > wb)
        return 0u;
    uint32_t m = 0xFFFFFFFFu;
    if (wi == wa)
        m &= 0xFFFFFFFFu << (int)(qa & 31);
    if (wi == wb)
        m &= 0xFFFFFFFFu >> (31 - (int)(qb & 31));
    *word = wi;
    return m;
}


/* node -> flagged cell check; returns the lcost or FR_BIG_COST.  `wave`:
 * every lane of the wave calls with the same node and cell. */
__device__ __forceinline__ long long dev_pair_lcost(const FrPlanDev &P, int node, int cell, float *c_out,
                                                    bool wave = false)
{
    const int z = node / P.W, x = node - z * P.W;
    const int zp = cell / P.W, xp = cell - zp * P.W;
    const float c = wave ? dev_wave_check_direct(P, P.max_curv_loop, x, z, P.heading[node], xp, zp,
                                                 P.gc_heading[cell], true, false, P.cost[node])
                         : dev_check_direct(P, P.max_curv_loop, x, z, P.heading[node], xp, zp, P.gc_heading[cell],
                                            true, false);
    if (c_out)
        *c_out = c;
    if (c < 0)
        return FR_BIG_COST;
    const float total = c + P.gc_cost[cell];
    return (long long)floor((double)(100.0f * total));
}

/* Cells of the new nodes' windows whose goal-connection planes are not yet
 * evaluated (lazy plans): claimed a word at a time with atomicOr on the
 * evaluated-bit plane, queued. */
__global__ void __launch_bounds__(FR_BLOCK) k_gc_demand(const FrPlanDev *plans, const FrFresh *fq, int2 *dq,
                                                         long dq_cap, int *counters)
{
    const int n = counters[FR_CNT_FRESH];
    const int wpb = blockDim.x >> 6;
    const int lane = dev_lane();
    FR_XCD_ITEMS(int, n, i0, nw, iend)
    for (int i = i0; i < iend; i += nw) {
        const FrFresh f = fq[i];
        const FrPlanDev &P = plans[f.slot];
        if (!P.gc_lazy || P.type[f.pos] != FR_TYPE_NODE)
            continue;
        const int z = f.pos / P.W, x = f.pos - z * P.W;
        FrWin w;
        if (!dev_win_make(P, x, z, &w))
            continue;
        for (int b = 0; b < w.npairs; b += 64) {
            long word;
            const uint32_t m = dev_win_pair(P, w, b + lane, &word);
            uint32_t mine = 0;
            if (m) {
                const uint32_t need = m & ~P.gc_done[word];
                if (need)
                    mine = need & ~atomicOr(&P.gc_done[word], need);
            }
            const int c = __popc(mine);
            const int incl = dev_wave_incl_scan(c);
            const int total = dev_bcast_i(incl, 63);
            if (!total)
                continue;
            int base = 0;
            if (lane == 0)
                base = atomicAdd(&counters[FR_CNT_DEMAND], total);
            long o = (long)dev_bcast_i(base, 0) + (incl - c);
            for (uint32_t mm = mine; mm; mm &= mm - 1) {
                const int bit = __ffs((int)mm) - 1;
                if (o < dq_cap) /* each cell is queued once per plan: sum N <= dq_cap */
                    dq[o] = make_int2(f.slot, (int)(word * 32 + bit));
                ++o;
            }
        }
    }
}

/* The lazily evaluated cells of one iteration are split in blocks of up to
 * 64 consecutive queue entries, one block per wave (more cells per wave as
 * the queue grows): the lanes set up the block's goal curves together, one
 * cell per lane, then the wave walks the block's curves one by one. */
#ifndef FR_GC_EVAL_MIN
#define FR_GC_EVAL_MIN 2
#endif
#ifndef FR_GC_FOOT_MIN
#define FR_GC_FOOT_MIN 1
#endif
/* Workgroups are dealt to the eight XCDs round robin (block b runs on XCD
 * b % 8): each XCD takes one contiguous eighth of the queue, whose entries
 * mostly belong to a few plans, so a plan's planes are read through one
 * XCD's L2.  The grid is a multiple of 8 blocks.  Defines lo, hi (this XCD's
 * range), nwaves (waves per XCD), cpw and gw (this wave's index on its XCD). */
#define FR_GC_XCD_RANGE(n, lo_min)                                                                           \
    const int xcd = (int)(blockIdx.x & 7u), jb = (int)(blockIdx.x >> 3);                                     \
    const long nwaves = (long)(gridDim.x >> 3) * wpb;                                                          \
    const long lo = (n) * xcd / 8, hi = (n) * (xcd + 1) / 8;                                                   \
    const int cpw = dev_gc_block(hi - lo, nwaves, lo_min);                                                     \
    const long gw = (long)jb * wpb + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
__device__ __forceinline__ int dev_gc_block(long n, long nwaves, int lo)
{
    const long c = (n + nwaves - 1) / nwaves;
    return c < lo ? lo : (c > 64 ? 64 : (int)c);
}

/* Per lane: set-up of the goal curve of queue entry d (base: the cost the
 * curve starts from, read when `with_base`). */
__device__ __forceinline__ bool dev_gc_lane_setup(const FrPlanDev *plans, int2 d, bool with_base, FrHermite *hm, int *np,
                                                  float *lh, float *base)
{
    const FrPlanDev &P = plans[d.x];
    const int z = d.y / P.W, x = d.y - z * P.W;
    *lh = (float)dev_compute_heading(x, z, P.gx, P.gz);
    *base = with_base ? P.gc_cost[d.y] : 0.0f;
    return dev_curve_setup(x, z, *lh, P.gx, P.gz, P.gh, hm, np);
}

/* Curvature and centre cells of the queued cells' goal curves; a rejection
 * is stored, the other cells are queued for k_gc_foot with their curves'
 * set-up (at most as many as k_gc_eval reads, so dq2 has dq's capacity). */
__global__ void __launch_bounds__(FR_BLOCK) k_gc_eval(const FrPlanDev *__restrict__ plans, FrCtl *__restrict__ ctl,
                                                      const int2 *__restrict__ dq, long dq_cap, int2 *__restrict__ dq2,
                                                      FrGcGeo *__restrict__ geo2, int *__restrict__ counters)
{
    long n = counters[FR_CNT_DEMAND];
    if (n > dq_cap)
        n = dq_cap;
    const int wpb = blockDim.x >> 6;
    FR_GC_XCD_RANGE(n, FR_GC_EVAL_MIN)
    const int lane = dev_lane();
    for (long b0 = lo + gw * cpw; b0 < hi; b0 += nwaves * cpw) {
        const int cnt = (int)(hi - b0 < cpw ? hi - b0 : cpw);
        int2 d = make_int2(0, 0);
        FrHermite hm = {};
        int np = 0;
        float lh = 0, base = 0;
        bool ok = false, gentle = false;
        if (lane < cnt) {
            d = dq[b0 + lane];
            ok = dev_gc_lane_setup(plans, d, false, &hm, &np, &lh, &base);
            /* the curve's first point is the cell itself: an obstacle pixel
             * there rejects it without a walk */
            const FrPlanDev &P = plans[d.x];
            const int z = d.y / P.W, x = d.y - z * P.W;
            if (ok && !(P.trav[d.y] & 1) && P.dist[fr_dist_at(P.TW, x + P.pad, z + P.pad)] == 0)
                ok = false;
            gentle = ok && dev_curve_gentle(hm, P.max_curv_goal);
        }
        const unsigned long long okm = __ballot(ok), gentlem = __ballot(gentle);
        unsigned long long kept = 0; /* cells of the block queued for k_gc_foot */
        for (int k = 0; k < cnt; ++k) {
            const int slot = dev_bcast_i(d.x, k), pos = dev_bcast_i(d.y, k);
            const FrPlanDev &P = plans[slot];
            bool keep = false;
            if ((okm >> k) & 1ull) {
                const FrHermite hk = dev_bcast_hermite(hm, k);
                /* a curve whose curvature bound clears the limit skips the per-point test */
                const float mc = ((gentlem >> k) & 1ull) ? 0.0f : P.max_curv_goal;
                keep = dev_wave_curve(P, hk, dev_bcast_i(np, k), mc, P.gx, P.gz, false, false, 0.0f, 1) >= 0;
            }
            if (keep)
                kept |= 1ull << k;
        }
        /* the block's rejections stored together, one lane per cell (their
         * bits stay clear, see dev_gc_store) */
        if (lane < cnt && !((kept >> lane) & 1ull))
            plans[d.x].gc_flag[d.y] = 0;
        if (lane < cnt && np > 0)
            atomicAdd(&ctl[d.x].w_goal_pts, np);
        /* one queue reservation per block */
        if (kept) {
            int at = 0;
            if (lane == 0)
                at = atomicAdd(&counters[FR_CNT_DEMAND2], __popcll(kept));
            at = dev_bcast_i(at, 0);
            if ((kept >> lane) & 1ull) {
                const int o = at + __popcll(kept & ((1ull << lane) - 1ull));
                dq2[o] = d;
                FrGcGeo g;
                g.t1x = (float)hm.t1x;
                g.t1y = (float)hm.t1y;
                g.t2x = (float)hm.t2x;
                g.t2y = (float)hm.t2y;
                g.lh = lh;
                g.np = np;
                geo2[o] = g;
            }
        }
    }
}

/* Footprints and cost of the cells k_gc_eval kept, from the set-up it
 * handed over (the tangents are floats widened to double, so the curve is
 * the one k_gc_eval walked). */
__global__ void __launch_bounds__(FR_BLOCK) k_gc_foot(const FrPlanDev *__restrict__ plans, FrCtl *__restrict__ ctl,
                                                      const int2 *__restrict__ dq2, const FrGcGeo *__restrict__ geo2,
                                                      const int *__restrict__ counters)
{
    const long n = counters[FR_CNT_DEMAND2];
    const int wpb = blockDim.x >> 6;
    FR_GC_XCD_RANGE(n, FR_GC_FOOT_MIN)
    const int lane = dev_lane();
    for (long b0 = lo + gw * cpw; b0 < hi; b0 += nwaves * cpw) {
        const int cnt = (int)(hi - b0 < cpw ? hi - b0 : cpw);
        int2 d = make_int2(0, 0);
        FrHermite hm = {};
        int np = 0;
        float lh = 0, base = 0;
        if (lane < cnt) {
            d = dq2[b0 + lane];
            const FrGcGeo g = geo2[b0 + lane];
            const FrPlanDev &P = plans[d.x];
            const int z = d.y / P.W, x = d.y - z * P.W;
            hm.t1x = (double)g.t1x;
            hm.t1y = (double)g.t1y;
            hm.t2x = (double)g.t2x;
            hm.t2y = (double)g.t2y;
            hm.dx0 = (double)x;
            hm.dz0 = (double)z;
            hm.dgx = (double)P.gx;
            hm.dgz = (double)P.gz;
            np = g.np;
            hm.denom = (double)(np - 1);
            hm.rden = 1.0 / hm.denom;
            lh = g.lh;
            base = P.gc_cost[d.y];
        }
        float res = -1;
        for (int k = 0; k < cnt; ++k) {
            const int slot = dev_bcast_i(d.x, k);
            const FrPlanDev &P = plans[slot];
            const FrHermite hk = dev_bcast_hermite(hm, k);
            const float c = dev_wave_curve(P, hk, dev_bcast_i(np, k), P.max_curv_goal, P.gx, P.gz, false, false,
                                           dev_bcast_f(base, k), 2);
            res = lane == k ? c : res;
        }
        /* the block's results stored together, one lane per cell */
        if (lane < cnt) {
            const FrPlanDev &P = plans[d.x];
            const int pos = d.y;
            atomicAdd(&ctl[d.x].w_foot_pts, np);
            P.gc_flag[pos] = res < 0 ? 0 : 1;
            if (res >= 0) {
                P.gc_heading[pos] = lh;
                P.gc_cost[pos] = res;
                atomicOr(&P.gc_bits[pos >> 5], 1u << (pos & 31));
            }
        }
    }
}

/* Goal test of the new nodes (graph_best_node.cu:182-222), window counts and
 * the pair queue; one wave per node, the window read by words of the flag
 * bit plane, pairs in the window's row-major order. */
__global__ void __launch_bounds__(FR_BLOCK) k_connect_emit(const FrPlanDev *plans, FrCtl *ctl, FrFresh *fq,
                                                            FrPair *pairs, long pair_cap, int *counters)
{
    const int n = counters[FR_CNT_FRESH];
    const int wpb = blockDim.x >> 6;
    const int lane = dev_lane();
    FR_XCD_ITEMS(int, n, i0, nw, iend)
    for (int i = i0; i < iend; i += nw) {
        FrFresh &f = fq[i];
        const int slot = f.slot, fpos = f.pos;
        const FrPlanDev &P = plans[slot];
        FrCtl &C = ctl[slot];
        const bool node = P.type[fpos] == FR_TYPE_NODE;
        if (lane == 0) {
            f.key = ~0ull;
            f.pair_cnt = 0;
        }
        if (!node)
            continue;
        const int z = fpos / P.W, x = fpos - z * P.W;
        if (lane == 0) {
            atomicAdd(&C.n_fresh, 1);
            if (dev_goal_hit(P, x, z, P.heading[fpos]))
                C.goal_reached = 1;
            /* the node's queue entry, found from its cell by k_finish (the
             * erase's scratch plane is free until the next erase; a node
             * queued twice has identical entries) */
            P.escr[fpos] = i;
        }
        FrWin w;
        if (!dev_win_make(P, x, z, &w))
            continue;
        int cnt = 0;
        for (int b = 0; b < w.npairs; b += 64) {
            long word;
            const uint32_t m = dev_win_pair(P, w, b + lane, &word);
            const int c = m ? __popc(m & P.gc_bits[word]) : 0;
            cnt += dev_bcast_i(dev_wave_incl_scan(c), 63);
        }
        if (cnt == 0)
            continue;
        int base = 0;
        if (lane == 0)
            base = atomicAdd(&counters[FR_CNT_PAIRS], cnt);
        base = dev_bcast_i(base, 0);
        if ((long)base + cnt > pair_cap) {
            if (lane == 0)
                f.flags |= FR_FRESH_INLINE;
            continue;
        }
        long o = base;
        for (int b = 0; b < w.npairs; b += 64) {
            long word;
            const uint32_t m = dev_win_pair(P, w, b + lane, &word);
            const uint32_t on = m ? (m & P.gc_bits[word]) : 0u;
            const int c = __popc(on);
            const int incl = dev_wave_incl_scan(c);
            long q = o + (incl - c);
            for (uint32_t mm = on; mm; mm &= mm - 1) {
                FrPair pr;
                pr.fresh = i;
                pr.cell = (int)(word * 32 + __ffs((int)mm) - 1);
                pairs[q++] = pr;
            }
            o += dev_bcast_i(incl, 63);
        }
        if (lane == 0) {
            f.pair_base = base;
            f.pair_cnt = cnt;
        }
    }
}

/* The plan's pick among its connections (the least lcost, then the lowest
 * node cell: the pick of graph_check_goal_reached.cu:115-139 with ties
 * ordered), as one 64-bit minimum; connections outside its fields (lcost not
 * in [0, 2^40), cell beyond 2^24) leave the pick to k_finish's scan. */
__device__ __forceinline__ void dev_pick(FrCtl &C, long long lc, int pos)
{
    if (lc >= 0 && lc < (1LL << 40) && pos < (1 << 24))
        atomicMin(&C.best_pick, ((unsigned long long)lc << 24) | (unsigned)pos);
    else
        C.pick_slow = 1;
}

/* Nodes whose pairs did not fit the pair queue scan their window in-thread
 * (thread `t` of `nt`); run by k_connect_eval after its pairs. */
__device__ void dev_connect_inline(const FrPlanDev *plans, FrCtl *ctl, FrFresh *fq, int n, int t, int nt)
{
    for (int i = t; i < n; i += nt) {
        FrFresh &f = fq[i];
        if (!(f.flags & FR_FRESH_INLINE))
            continue;
        const FrPlanDev &P = plans[f.slot];
        if (P.type[f.pos] != FR_TYPE_NODE)
            continue;
        const int z = f.pos / P.W, x = f.pos - z * P.W;
        int z0, z1, x0, x1;
        dev_window(P, x, z, &z0, &z1, &x0, &x1);
        int rank = 0;
        unsigned long long best = ~0ull;
        long long bl = FR_BIG_COST;
        for (int zp = z0; zp < z1; ++zp)
            for (int xp = x0; xp < x1; ++xp) {
                const int q = zp * P.W + xp;
                if (!P.gc_flag[q])
                    continue;
                const long long lc = dev_pair_lcost(P, f.pos, q, nullptr);
                if (lc < FR_BIG_COST) {
                    const unsigned long long k = dev_node_key(lc, rank);
                    if (k < best) {
                        best = k;
                        bl = lc;
                    }
                }
                rank++;
            }
        f.key = best;
        if (bl < FR_BIG_COST) {
            atomicMin((unsigned long long *)&ctl[f.slot].best_lcost, fr_lcost_key(bl));
            dev_pick(ctl[f.slot], bl, f.pos);
        }
    }
}

/* one wave per (node, flagged cell) pair, then the nodes whose pairs did
 * not fit (their flag keeps them out of the pair loop) */
__global__ void __launch_bounds__(FR_BLOCK) k_connect_eval(const FrPlanDev *plans, FrCtl *ctl, FrFresh *fq, FrPair *pairs, long pair_cap,
                               int *counters)
{
    long n = counters[FR_CNT_PAIRS];
    if (n > pair_cap)
        n = pair_cap;
    const int n_fresh = counters[FR_CNT_FRESH];
    const int wpb = blockDim.x >> 6;
    FR_XCD_ITEMS(long, n, p0, nw, pend)
    for (long p = p0; p < pend; p += nw) {
        const FrPair pr = pairs[p];
        if (pr.fresh < 0 || pr.fresh >= n_fresh)
            continue; /* slot reserved by a node whose pairs did not fit */
        FrFresh &f = fq[pr.fresh];
        if (f.flags & FR_FRESH_INLINE)
            continue;
        if ((long)f.pair_base + f.pair_cnt <= p || p < f.pair_base)
            continue; /* a pair slot past the capacity check */
        const FrPlanDev &P = plans[f.slot];
        float c;
        const long long lc = dev_pair_lcost(P, f.pos, pr.cell, &c, true);
        if (lc >= FR_BIG_COST || dev_lane() != 0)
            continue;
        pairs[p].cost = c; /* read back by k_finish for the winning pair */
        atomicMin(&f.key, dev_node_key(lc, (int)(p - f.pair_base)));
        atomicMin((unsigned long long *)&ctl[f.slot].best_lcost, fr_lcost_key(lc));
        dev_pick(ctl[f.slot], lc, f.pos);
    }
    dev_connect_inline(plans, ctl, fq, n_fresh, (int)(blockIdx.x * blockDim.x + threadIdx.x),
                       (int)(gridDim.x * blockDim.x));
    if (blockIdx.x == 0 && threadIdx.x == 0)
        counters[FR_CNT_FRESH_SEEN] = n_fresh;
}

/* the rank-th goal-flagged cell of a node's window, in row-major order
 * (one wave, by words of the flag bit plane) */
__device__ int dev_wave_window_cell(const FrPlanDev &P, int node, int rank)
{
    const int z = node / P.W, x = node - z * P.W;
    FrWin w;
    if (!dev_win_make(P, x, z, &w))
        return -1;
    const int lane = dev_lane();
    int seen = 0;
    for (int b = 0; b < w.npairs; b += 64) {
        long word;
        const uint32_t m = dev_win_pair(P, w, b + lane, &word);
        const uint32_t on = m ? (m & P.gc_bits[word]) : 0u;
        const int c = __popc(on);
        const int incl = dev_wave_incl_scan(c);
        const int total = dev_bcast_i(incl, 63);
        if (rank < seen + total) {
            const int r = rank - seen;
            const bool here = r >= incl - c && r < incl;
            const unsigned long long hm = __ballot(here);
            const int l = __ffsll((long long)hm) - 1;
            int cell = -1;
            if (lane == l) {
                uint32_t mm = on;
                for (int t = r - (incl - c); t > 0; --t)
                    mm &= mm - 1;
                cell = (int)(word * 32 + __ffs((int)mm) - 1);
            }
            return dev_bcast_i(cell, l);
        }
        seen += total;
    }
    return -1;
}

/* k_finish's bookkeeping for one plan (lane 0): the reseed or the goal
 * connection's two adds, the goal test, the status. */
__device__ void dev_finish_plan(const FrPlanDev &P, FrCtl &C, const FrFresh *fq, unsigned long long bk, bool reseed,
                                int cell, float c_cost, int npos, int nx, int nz)
{
    C.last_temp = C.n_temp;
    C.iter += 1;
    C.total_expansions += C.n_expansions;
    if (reseed) {
        dev_add(P, P.sx, P.sz, P.sh, -1, -1, 0.0f);
        C.reseeded = 1;
        C.pending_start = 1;
        C.full_check = 0;
    } else {
        C.pending_start = 0;
        C.full_check = 0;
        if (bk != ~0ull) {
            const int cz = cell / P.W, cx = cell - cz * P.W;
            const float base = P.cost[npos];
            const float ch = P.gc_heading[cell];
            dev_add(P, cx, cz, ch, nx, nz, c_cost + base);
            dev_add(P, P.gx, P.gz, P.gh, cx, cz, ch + base);
            C.found = 1;
            if (dev_goal_hit(P, cx, cz, ch))
                C.goal_reached = 1;
            if (P.gx >= 0 && P.gx < P.W && P.gz >= 0 && P.gz < P.H && dev_goal_hit(P, P.gx, P.gz, P.gh))
                C.goal_reached = 1;
            if (!C.goal_reached)
                C.full_check = 1;
        }
    }
    if (C.goal_reached) {
        C.status = FR_FOUND;
        /* batch finalisation: path_optimize rounds of the request */
        C.opt_limit = P.opt_limit;
        C.opt_active = P.opt_limit > 0;
        C.opt_rounds = 0;
    } else if (P.max_iter > 0 && C.iter >= P.max_iter) {
        C.status = FR_FAILED;
    }
}

/* End of FastRRT::loop (fastrrt.cpp:130-180): reseed, goal-connection adds,
 * goal test, iteration bookkeeping.  One wave per plan: the lanes share the
 * winner's Hermite check, lane 0 writes; then the plan's next iteration is
 * made ready (dev_ctl_ready), the finished iteration's figures kept in the
 * stat_* fields. */
__global__ void __launch_bounds__(FR_BLOCK) k_finish(const FrPlanDev *plans, FrCtl *ctl, const int *slots, const int *n_slots_p,
                         const FrFresh *fq, const FrPair *pairs, int *counters, int scan)
{
    const int i = blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    const int n_slots = *n_slots_p;
    /* the queues' counters for the next iteration (every reader has run) */
    if (blockIdx.x == 0 && threadIdx.x == 0)
        dev_counters_reset(counters);
    if (i >= n_slots)
        return;
    const int s = slots[i];
    FrCtl &C = ctl[s];
    const FrPlanDev &P = plans[s];
    if (C.status != FR_SEARCHING)
        return;
    const int next_iter = C.iter + 1;
    const bool lead = dev_lane() == 0;
    const bool reseed = C.n_temp == 0 && C.n_all == 0;
    /* the winning node's queue entry: from the plan's pick, or (connections
     * outside the pick's fields, or `scan`: the FastRRT handle, whose few
     * entries make the scan cheap) the least (node cell, entry) among the
     * entries whose best connection has the plan's least lcost */
    int win = -1;
    if (!C.pick_slow && !scan) {
        if (C.best_pick != ~0ull)
            win = P.escr[(int)(C.best_pick & 0xFFFFFFull)];
    } else {
        const int n = counters[FR_CNT_FRESH_SEEN];
        unsigned long long key = ~0ull;
        for (int k = dev_lane(); k < n; k += 64) {
            const FrFresh f = fq[k];
            if (f.slot == s && f.key != ~0ull &&
                fr_lcost_key(dev_key_lcost(f.key)) == (unsigned long long)C.best_lcost) {
                const unsigned long long kk = ((unsigned long long)(unsigned)f.pos << 32) | (unsigned)k;
                key = kk < key ? kk : key;
            }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const unsigned long long o = __shfl_xor(key, off, 64);
            key = o < key ? o : key;
        }
        if (key != ~0ull)
            win = (int)(key & 0xFFFFFFFFull);
    }
    unsigned long long bk = win >= 0 ? 0ull : ~0ull;
    int nx = 0, nz = 0, cell = -1, npos = 0;
    float c_cost = 0;
    if (!reseed && win >= 0) {
        const FrFresh f = fq[win];
        npos = f.pos;
        nz = f.pos / P.W;
        nx = f.pos - nz * P.W;
        if (!(f.flags & FR_FRESH_INLINE)) {
            /* the winning pair of k_connect_eval, its curve cost kept */
            const FrPair pr = pairs[f.pair_base + dev_key_rank(f.key)];
            cell = pr.cell;
            c_cost = pr.cost;
        } else {
            cell = dev_wave_window_cell(P, f.pos, dev_key_rank(f.key));
            if (cell >= 0)
                (void)dev_pair_lcost(P, f.pos, cell, &c_cost, true);
            else
                bk = ~0ull; /* unreachable: the rank counts flags present since the pair pass */
        }
    }
    if (lead) {
        dev_finish_plan(P, C, fq, bk, reseed, cell, c_cost, npos, nx, nz);
        C.stat_expansions = C.n_expansions;
        C.stat_fresh = C.n_fresh;
        C.stat_reseeded = C.reseeded;
        C.prev_eligible = C.n_eligible;
        C.n_eligible = 0;
    }
    __builtin_amdgcn_wave_barrier();
    dev_ctl_ready(P, C, next_iter);
}

/* ------------------------------------------------------------------------ */
/* whole-grid helpers (graph API, path extraction, optimisation)           */
/* ------------------------------------------------------------------------ */

/* graph.cu:400-414: type <- NULL everywhere */
__global__ void __launch_bounds__(FR_BLOCK) k_clear(const FrPlanDev *plans, int slot)
{
    const FrPlanDev &P = plans[slot];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x)
        P.type[pos] = FR_TYPE_NULL;
}

/* graph_list.cu:8-66: out[0..5] = count by type */
__global__ void __launch_bounds__(FR_BLOCK) k_count_types(const FrPlanDev *plans, int slot, int *out)
{
    const FrPlanDev &P = plans[slot];
    __shared__ int s_c[8];
    if (threadIdx.x < 8)
        s_c[threadIdx.x] = 0;
    __syncthreads();
    int c[6] = {0, 0, 0, 0, 0, 0};
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x) {
        const int t = P.type[pos];
        if (t < 6)
            c[t]++;
    }
    for (int k = 0; k < 6; ++k)
        if (c[k])
            atomicAdd(&s_c[k], c[k]);
    __syncthreads();
    if (threadIdx.x < 6 && s_c[threadIdx.x])
        atomicAdd(&out[threadIdx.x], s_c[threadIdx.x]);
}

/* graph_best_node.cu:182-222 over every NODE */
__global__ void __launch_bounds__(FR_BLOCK) k_goal_scan(const FrPlanDev *plans, FrCtl *ctl, int slot, int gx, int gz, float gh, float tol,
                            float herr)
{
    const FrPlanDev &P = plans[slot];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x) {
        if (P.type[pos] != FR_TYPE_NODE)
            continue;
        const int z = pos / P.W, x = pos - z * P.W;
        const double dx = (double)(gx - x), dz = (double)(gz - z);
        if (__builtin_sqrt(dx * dx + dz * dz) > (double)tol)
            continue;
        if (fabsf(P.heading[pos] - gh) <= herr)
            ctl[slot].goal_reached = 1;
    }
}

/* Cells a whole-grid scan of a plan in batch finalisation visits.  After an
 * optimisation round has cleared the grid and re-added the path chain
 * (k_opt_rebuild), the NODE cells are exactly the chain's cells, so the scan
 * walks the chain (block 0) instead of the grid; a cell listed twice is
 * evaluated twice with the same values.  Returns false when done. */
struct FrScan {
    bool chain;
    int k, step, n;
};

__device__ __forceinline__ FrScan dev_scan_begin(const FrPlanDev &P, const FrCtl &C, int gate)
{
    FrScan sc;
    sc.chain = gate != 0 && C.chain_only;
    if (sc.chain) {
        sc.k = blockIdx.x == 0 ? (int)threadIdx.x : 0x7FFFFFFF;
        sc.step = blockDim.x;
        sc.n = C.path_len;
    } else {
        sc.k = blockIdx.x * blockDim.x + threadIdx.x;
        sc.step = gridDim.x * blockDim.x;
        sc.n = P.N;
    }
    return sc;
}

__device__ __forceinline__ bool dev_scan_next(const FrPlanDev &P, FrScan &sc, int *pos)
{
    while (sc.k < sc.n) {
        const int k = sc.k;
        sc.k += sc.step;
        if (!sc.chain) {
            *pos = k;
            return true;
        }
        const int x = fr_trunc_f(P.path[k].x), z = fr_trunc_f(P.path[k].y);
        if (x < 0 || x >= P.W || z < 0 || z >= P.H)
            continue;
        *pos = z * P.W + x;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool dev_gate(const FrCtl &C, int gate)
{
    if (gate == 1)
        return C.status == FR_FOUND && C.opt_active && C.goal_reached;
    if (gate == 2)
        return C.status == FR_FOUND && C.goal_reached;
    return true;
}

__global__ void __launch_bounds__(FR_BLOCK) k_best_node_init(FrCtl *ctl, const int *slots, int n, int gate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    FrCtl &C = ctl[slots[i]];
    if (!dev_gate(C, gate))
        return;
    C.bn_cost = (long long)~0ull;
    C.bn_pos = 0x7FFFFFFF;
}

/* __CUDA_KERNEL_findBestNodeWithHeading_bestCost (graph_best_node.cu:21-75).
 * Ties: lowest cell index (second kernel). */
__global__ void __launch_bounds__(FR_BLOCK) k_best_node_cost(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int tx_override,
                                 int tz_override, float h_override, float r_override, float herr_override,
                                 int use_override, int gate)
{
    const int s = slots[blockIdx.y];
    const FrPlanDev &P = plans[s];
    FrCtl &C = ctl[s];
    if (!dev_gate(C, gate))
        return;
    const int tx = use_override ? tx_override : P.gx;
    const int tz = use_override ? tz_override : P.gz;
    const float th = use_override ? h_override : P.gh;
    const float radius = use_override ? r_override : P.dist_tol;
    const float herr = use_override ? herr_override : P.herr_tol;
    FrScan sc = dev_scan_begin(P, C, gate);
    int pos;
    while (dev_scan_next(P, sc, &pos)) {
        if (P.type[pos] != FR_TYPE_NODE)
            continue;
        const int z = pos / P.W, x = pos - z * P.W;
        const int dx = tx - x, dz = tz - z;
        const float dist = sqrtf((float)(dx * dx + dz * dz));
        if (dist > radius)
            continue;
        const float h = P.heading[pos];
        if (fabsf(h - th) > herr)
            continue;
        if (!dev_feasible(P, x, z, h))
            continue;
        const double cd = (double)(1.0f * dist) + (3.0 * FR_TO_DEG_D) * (double)h + (double)(1.0f * P.cost[pos]);
        const long long cost = (long long)floor((double)(float)cd);
        atomicMin((unsigned long long *)&C.bn_cost, fr_lcost_key(cost));
    }
}

__global__ void __launch_bounds__(FR_BLOCK) k_best_node_pick(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int tx_override,
                                 int tz_override, float h_override, float r_override, float herr_override,
                                 int use_override, int gate)
{
    const int s = slots[blockIdx.y];
    const FrPlanDev &P = plans[s];
    FrCtl &C = ctl[s];
    if (!dev_gate(C, gate))
        return;
    const unsigned long long best = (unsigned long long)C.bn_cost;
    if (best >= fr_lcost_key(FR_BIG_COST))
        return;
    const int tx = use_override ? tx_override : P.gx;
    const int tz = use_override ? tz_override : P.gz;
    const float th = use_override ? h_override : P.gh;
    const float radius = use_override ? r_override : P.dist_tol;
    const float herr = use_override ? herr_override : P.herr_tol;
    FrScan sc = dev_scan_begin(P, C, gate);
    int pos;
    while (dev_scan_next(P, sc, &pos)) {
        if (P.type[pos] != FR_TYPE_NODE)
            continue;
        const int z = pos / P.W, x = pos - z * P.W;
        const int dx = tx - x, dz = tz - z;
        const float dist = sqrtf((float)(dx * dx + dz * dz));
        if (dist > radius)
            continue;
        const float h = P.heading[pos];
        if (fabsf(h - th) > herr)
            continue;
        const double cd = (double)(1.0f * dist) + (3.0 * FR_TO_DEG_D) * (double)h + (double)(1.0f * P.cost[pos]);
        const long long cost = (long long)floor((double)(float)cd);
        if (fr_lcost_key(cost) != best)
            continue;
        if (!dev_feasible(P, x, z, h))
            continue;
        atomicMin(&C.bn_pos, pos);
    }
}

/* fastrrt.cpp:214-231: walk the parents of the best node, reversed into
 * path[] as (x, z, heading, cost).  getParent (graph.cu:446-453) yields
 * (-1, -1) for an out-of-bounds or NULL cell, so such a cell is still pushed
 * and ends the walk.  One thread per plan; a path longer than the buffer is
 * reported as empty, as the reference does after 10^6 steps. */
__global__ void __launch_bounds__(FR_BLOCK) k_walk(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int n_slots, int gate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_slots)
        return;
    const int s = slots[i];
    const FrPlanDev &P = plans[s];
    FrCtl &C = ctl[s];
    if (!dev_gate(C, gate))
        return;
    C.path_len = 0;
    if ((unsigned long long)C.bn_cost >= fr_lcost_key(FR_BIG_COST) || C.bn_pos == 0x7FFFFFFF)
        return;
    int x = C.bn_pos % P.W, z = C.bn_pos / P.W;
    int n = 0;
    while (x != -1 && z != -1) {
        if (n >= P.path_cap) {
            C.path_len = 0;
            return;
        }
        const bool inb = x >= 0 && x < P.W && z >= 0 && z < P.H;
        const int pos = inb ? z * P.W + x : 0;
        P.path[n++] = make_float4((float)x, (float)z, inb ? P.heading[pos] : 0.0f, inb ? P.cost[pos] : -1.0f);
        if (!inb || P.type[pos] == FR_TYPE_NULL)
            break;
        x = P.px[pos];
        z = P.pz[pos];
    }
    for (int a = 0, b = n - 1; a < b; ++a, --b) {
        float4 t = P.path[a];
        P.path[a] = P.path[b];
        P.path[b] = t;
    }
    C.path_len = n;
}

/* graph_optimization.cu:25-56 for all (i, j >= i + 2), one wave per pair:
 * each i keeps the last improving j with its connection cost, as one 64-bit
 * key (j << 32 | cost bits) in pd[i].xy, reduced with atomicMax (0: none). */
__global__ void __launch_bounds__(FR_BLOCK) k_opt_pairs(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int gate)
{
    const int s = slots[blockIdx.y];
    const FrPlanDev &P = plans[s];
    const FrCtl &C = ctl[s];
    if (!dev_gate(C, gate))
        return;
    const int n = C.path_len;
    if (n < 3)
        return;
    const long npairs = (long)n * n;
    const float max_curv = P.max_curv_loop;
    const int wpb = blockDim.x >> 6;
    const long nw = (long)gridDim.x * wpb;
    for (long k = blockIdx.x * (long)wpb + (threadIdx.x >> 6); k < npairs; k += nw) {
        const int i = (int)(k / n), j = (int)(k - (long)i * n);
        if (i >= n - 2 || j < i + 2)
            continue;
        const float4 a = P.path[i], b = P.path[j];
        const float cur = b.w;
        const int ax = fr_trunc_f(a.x), az = fr_trunc_f(a.y);
        const float nc = dev_wave_check_direct(P, max_curv, ax, az, a.z, fr_trunc_f(b.x), fr_trunc_f(b.y), b.z,
                                               true, true, P.cost[az * P.W + ax]);
        if (dev_lane() == 0 && nc > 0 && nc < cur)
            atomicMax((unsigned long long *)&P.pd[i].x,
                      ((unsigned long long)(unsigned)j << 32) | (unsigned)__float_as_int(nc));
    }
}

__global__ void __launch_bounds__(FR_BLOCK) k_opt_reset(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int gate)
{
    const int s = slots[blockIdx.y];
    const FrPlanDev &P = plans[s];
    if (!dev_gate(ctl[s], gate))
        return;
    const int n = ctl[s].path_len;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        *(unsigned long long *)&P.pd[i].x = 0ull;
}

/* graph_optimization.cu:126-183: the first i with the largest gain wins and
 * the path is spliced (one thread per plan); k_opt_clear then empties the
 * grid and k_opt_rebuild re-adds the chain.  opt_changed = 2 marks the plans
 * this round changed until the rebuild. */
__global__ void __launch_bounds__(FR_BLOCK) k_opt_apply(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int n_slots, int gate)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_slots)
        return;
    const int s = slots[t];
    const FrPlanDev &P = plans[s];
    FrCtl &C = ctl[s];
    if (!dev_gate(C, gate))
        return;
    int n = C.path_len;
    C.opt_changed = 0;
    int max_pos = -1, next_pos = -1;
    float max_gain = -1, max_nc = 0;
    for (int i = 0; i < n - 2; ++i) {
        const unsigned long long key = *(const unsigned long long *)&P.pd[i].x;
        if (!key)
            continue;
        const int j = (int)(key >> 32);
        const float nc = __int_as_float((int)(unsigned)(key & 0xFFFFFFFFull));
        const float gain = P.path[j].w - nc;
        if (max_gain < gain) {
            max_gain = gain;
            max_pos = i;
            max_nc = nc;
            next_pos = j;
        }
    }
    if (max_pos < 0) {
        C.opt_active = 0;
        return;
    }
    if (next_pos > max_pos + 1) {
        P.path[next_pos].w = max_nc;
        const int rm = next_pos - max_pos - 1;
        for (int i = max_pos + 1; i + rm < n; ++i)
            P.path[i] = P.path[i + rm];
        n -= rm;
    }
    C.path_len = n;
    C.opt_changed = 2;
    C.opt_rounds += 1;
    if (gate == 1 && C.opt_rounds >= C.opt_limit)
        C.opt_active = 0;
    C.goal_reached = 0;
}

/* graph.cu:400-414 for the plans the round changed */
__global__ void __launch_bounds__(FR_BLOCK) k_opt_clear(const FrPlanDev *plans, const FrCtl *ctl, const int *slots)
{
    const int s = slots[blockIdx.y];
    if (ctl[s].opt_changed != 2)
        return;
    const FrPlanDev &P = plans[s];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x)
        P.type[pos] = FR_TYPE_NULL;
}

__global__ void __launch_bounds__(FR_BLOCK) k_opt_rebuild(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int n_slots)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_slots)
        return;
    const int s = slots[t];
    FrCtl &C = ctl[s];
    if (C.opt_changed != 2)
        return;
    const FrPlanDev &P = plans[s];
    const int n = C.path_len;
    int ppx = fr_trunc_f(P.path[0].x), ppz = fr_trunc_f(P.path[0].y);
    dev_add(P, ppx, ppz, P.path[0].z, -1, -1, P.path[0].w);
    for (int i = 1; i < n; ++i) {
        const int cx = fr_trunc_f(P.path[i].x), cz = fr_trunc_f(P.path[i].y);
        dev_add(P, cx, cz, P.path[i].z, ppx, ppz, P.path[i].w);
        ppx = cx;
        ppz = cz;
    }
    C.opt_changed = 1;
    C.chain_only = 1;
}

/* ------------------------------------------------------------------------ */
/* launchers                                                                */
/* ------------------------------------------------------------------------ */

static int fr_grid_for(long work, int per_block, int cap)
{
    long b = (work + per_block - 1) / per_block;
    if (b < 1)
        b = 1;
    if (b > cap)
        b = cap;
    return (int)b;
}

void fr_launch_setup(const FrLaunch &L, const int *d_slots, int n, int max_N, int max_PW, int max_PH,
                     int goal_mode, bool zeroed, hipStream_t st)
{
    if (n <= 0)
        return;
    dim3 b(FR_BLOCK);
    hipLaunchKernelGGL(k_setup_prep, dim3(fr_grid_for((max_N + 3) / 4, FR_BLOCK, 1024), n), b, 0, st, L.plans,
                       d_slots);
    hipLaunchKernelGGL(k_setup_rowdist, dim3(fr_grid_for(max_PH, FR_BLOCK / 64, 1024), n), b, 0, st, L.plans,
                       d_slots);
    /* one block per 64 x 64 tile of the largest padded plane */
    const long tiles = ((long)max_PW + FR_COL_TILE - 1) / FR_COL_TILE * (((long)max_PH + FR_COL_TILE - 1) / FR_COL_TILE);
    hipLaunchKernelGGL(k_setup_coldist, dim3((unsigned)std::min(tiles, 4096L), n), dim3(256), 0, st, L.plans,
                       d_slots);
    if (goal_mode == FR_GOAL_NONE)
        return;
    /* a slab reset to zero already holds the planes k_setup_gc_reset writes */
    if (!zeroed)
        hipLaunchKernelGGL(k_setup_gc_reset, dim3(fr_grid_for(max_N, FR_BLOCK, 1024), n), b, 0, st, L.plans,
                           d_slots);
    if (goal_mode == FR_GOAL_EAGER) {
        fr_prof_begin(FR_PROF_GOAL_CONNECTION, st);
        hipLaunchKernelGGL(k_setup_goal_connection, dim3(fr_grid_for(max_N, FR_BLOCK / 64, 16384), n), b, 0, st,
                           L.plans, d_slots, 0);
        fr_prof_end(FR_PROF_GOAL_CONNECTION, st);
    }
}

void fr_launch_goal_complete(const FrLaunch &L, const int *d_slots, int n, int max_N, hipStream_t st)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_setup_goal_connection, dim3(fr_grid_for(max_N, FR_BLOCK / 64, 16384), n), dim3(FR_BLOCK), 0,
                       st, L.plans, d_slots, 1);
}

/* One loop iteration of every active plan, in three parts so that a caller
 * that captures graphs can time k_expand with events between them. */
void fr_launch_iteration_pre(const FrLaunch &L, hipStream_t st, bool entry)
{
    dim3 b(L.block);
    if (entry)
        hipLaunchKernelGGL(k_begin, dim3(fr_grid_for(L.max_slots, L.block / 64, 65535)), b, 0, st, L.plans, L.ctl,
                           L.slots, L.n_slots, L.counters);
    hipLaunchKernelGGL(k_select, dim3(L.grid_tiles), b, 0, st, L.plans, L.ctl, L.tiles, L.n_tiles, L.exp_q,
                       L.counters);
}

void fr_launch_iteration_expand(const FrLaunch &L, hipStream_t st)
{
    hipLaunchKernelGGL(k_expand, dim3(L.grid_waves), dim3(L.block), 0, st, L.plans, L.ctl, L.exp_q, L.counters);
}

void fr_launch_iteration_post(const FrLaunch &L, hipStream_t st, bool entry)
{
    dim3 b(L.block);
    hipLaunchKernelGGL(k_commit, dim3(L.grid_items), b, 0, st, L.plans, L.ctl, L.exp_q, L.counters);
    const int dens_blocks = (L.grid_tiles + FR_ERASE_BLOCK / FR_BLOCK - 1) / (FR_ERASE_BLOCK / FR_BLOCK);
    hipLaunchKernelGGL(k_erase_mark, dim3(L.max_slots + dens_blocks), dim3(FR_ERASE_BLOCK), 0, st, L.plans, L.ctl,
                       L.slots, L.n_slots, L.tiles, L.n_tiles, L.max_slots);
    hipLaunchKernelGGL(k_accept, dim3(L.grid_tiles), b, 0, st, L.plans, L.ctl, L.tiles, L.n_tiles, L.fresh_q,
                       L.counters);
    fr_prof_begin(FR_PROF_CONNECT, st);
    fr_prof_begin(FR_PROF_GOAL_CONNECTION, st);
    hipLaunchKernelGGL(k_gc_demand, dim3(L.grid_waves), b, 0, st, L.plans, L.fresh_q, L.demand, L.demand_cap,
                       L.counters);
    hipLaunchKernelGGL(k_gc_eval, dim3(L.grid_waves), b, 0, st, L.plans, L.ctl, L.demand, L.demand_cap, L.demand2,
                       L.demand2_geo, L.counters);
    hipLaunchKernelGGL(k_gc_foot, dim3(L.grid_waves), b, 0, st, L.plans, L.ctl, L.demand2, L.demand2_geo,
                       L.counters);
    fr_prof_end(FR_PROF_GOAL_CONNECTION, st);
    hipLaunchKernelGGL(k_connect_emit, dim3(L.grid_waves), b, 0, st, L.plans, L.ctl, L.fresh_q, L.pairs,
                       L.pair_cap, L.counters);
    hipLaunchKernelGGL(k_connect_eval, dim3(L.grid_waves), b, 0, st, L.plans, L.ctl, L.fresh_q, L.pairs,
                       L.pair_cap, L.counters);
    fr_prof_end(FR_PROF_CONNECT, st);
    hipLaunchKernelGGL(k_finish, dim3(fr_grid_for(L.max_slots, L.block / 64, 65535)), b, 0, st, L.plans, L.ctl,
                       L.slots, L.n_slots, L.fresh_q, L.pairs, L.counters, entry ? 1 : 0);
}

void fr_launch_iteration(const FrLaunch &L, hipStream_t st, bool entry)
{
    fr_prof_begin(FR_PROF_ITERATION, st);
    fr_launch_iteration_pre(L, st, entry);
    fr_prof_begin(FR_PROF_EXPAND, st);
    fr_launch_iteration_expand(L, st);
    fr_prof_end(FR_PROF_EXPAND, st);
    fr_launch_iteration_post(L, st, entry);
    fr_prof_end(FR_PROF_ITERATION, st);
}

void fr_launch_clear(const FrLaunch &L, int slot, int N, hipStream_t st)
{
    hipLaunchKernelGGL(k_clear, dim3(fr_grid_for(N, FR_BLOCK, 2048)), dim3(FR_BLOCK), 0, st, L.plans, slot);
}

void fr_launch_count(const FrLaunch &L, int slot, int N, int *d_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_count_types, dim3(fr_grid_for(N, FR_BLOCK, 1024)), dim3(FR_BLOCK), 0, st, L.plans,
                       slot, d_out);
}

void fr_launch_goal_scan(const FrLaunch &L, int slot, int N, int gx, int gz, float gh, float tol, float herr,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_goal_scan, dim3(fr_grid_for(N, FR_BLOCK, 2048)), dim3(FR_BLOCK), 0, st, L.plans, L.ctl,
                       slot, gx, gz, gh, tol, herr);
}

void fr_launch_best_node(const FrLaunch &L, const int *d_slots, int n, int max_N, const FrBestNodeArgs *ov,
                         hipStream_t st, int gate)
{
    const int use = ov != nullptr;
    const int tx = use ? ov->x : 0, tz = use ? ov->z : 0;
    const float h = use ? ov->heading : 0, r = use ? ov->radius : 0, e = use ? ov->max_herr : 0;
    dim3 g(fr_grid_for(max_N, FR_BLOCK, 1024), n);
    hipLaunchKernelGGL(k_best_node_init, dim3(fr_grid_for(n, 64, 65535)), dim3(64), 0, st, L.ctl, d_slots, n, gate);
    hipLaunchKernelGGL(k_best_node_cost, g, dim3(FR_BLOCK), 0, st, L.plans, L.ctl, d_slots, tx, tz, h, r, e, use,
                       gate);
    hipLaunchKernelGGL(k_best_node_pick, g, dim3(FR_BLOCK), 0, st, L.plans, L.ctl, d_slots, tx, tz, h, r, e, use,
                       gate);
}

/* Start of a batch finalisation: list entry q may run at most limits[q]
 * optimisation rounds (the time left before its deadline). */
__global__ void __launch_bounds__(FR_BLOCK) k_fin_limits(FrCtl *ctl, const int *list, const int *limits, int n)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        FrCtl &C = ctl[list[q]];
        if (C.status != FR_FOUND)
            continue;
        if (limits[q] < C.opt_limit)
            C.opt_limit = limits[q] > 0 ? limits[q] : 0;
        if (C.opt_rounds >= C.opt_limit)
            C.opt_active = 0;
    }
}

void fr_launch_fin_limits(const FrLaunch &L, const int *d_list, const int *d_limits, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_fin_limits, dim3(1), dim3(FR_BLOCK), 0, st, L.ctl, d_list, d_limits, n);
}

/* End of a batch finalisation: each listed plan's control block and path
 * (when it fits) written straight into pinned host memory, record q at
 * pack + q * stride (the control block, then float4 waypoints from byte
 * FR_PACK_PATH_OFF), and the control block returned to IDLE, so the host
 * collects the results without copies.  One block per list entry. */
__global__ void __launch_bounds__(FR_BLOCK) k_fin_pack(const FrPlanDev *plans, FrCtl *ctl, const int *list, int n,
                                                       uint8_t *pack, long stride, int path_max)
{
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        const int s = list[q];
        uint8_t *rec = pack + (long)q * stride;
        const FrCtl c = ctl[s];
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(&c);
        uint32_t *hw = reinterpret_cast<uint32_t *>(rec);
        for (int k = threadIdx.x; k < (int)(sizeof(FrCtl) / 4); k += blockDim.x)
            hw[k] = cw[k];
        const int len = c.goal_reached ? (c.path_len > 0 ? c.path_len : 0) : 0;
        if (len <= path_max) {
            const float4 *src = reinterpret_cast<const float4 *>(plans[s].path);
            float4 *dst = reinterpret_cast<float4 *>(rec + FR_PACK_PATH_OFF);
            for (int k = threadIdx.x; k < len; k += blockDim.x)
                dst[k] = src[k];
        }
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) {
            FrCtl idle = {};
            idle.status = FR_IDLE;
            ctl[s] = idle;
        }
        __syncthreads();
    }
}

void fr_launch_fin_pack(const FrLaunch &L, const int *d_list, int n, uint8_t *pack, long stride, int path_max,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_fin_pack, dim3(n), dim3(FR_BLOCK), 0, st, L.plans, L.ctl, d_list, n, pack, stride, path_max);
}

void fr_launch_walk(const FrLaunch &L, const int *d_slots, int n, hipStream_t st, int gate)
{
    hipLaunchKernelGGL(k_walk, dim3(fr_grid_for(n, 64, 65535)), dim3(64), 0, st, L.plans, L.ctl, d_slots, n, gate);
}

void fr_launch_opt_round(const FrLaunch &L, const int *d_slots, int n, int max_path, int max_N, hipStream_t st,
                         int gate)
{
    dim3 g1(fr_grid_for(max_path, FR_BLOCK, 64), n);
    hipLaunchKernelGGL(k_opt_reset, g1, dim3(FR_BLOCK), 0, st, L.plans, L.ctl, d_slots, gate);
    dim3 g2(fr_grid_for((long)max_path * max_path, FR_BLOCK / 64, 8192), n);
    hipLaunchKernelGGL(k_opt_pairs, g2, dim3(FR_BLOCK), 0, st, L.plans, L.ctl, d_slots, gate);
    hipLaunchKernelGGL(k_opt_apply, dim3(fr_grid_for(n, 64, 65535)), dim3(64), 0, st, L.plans, L.ctl, d_slots, n,
                       gate);
    hipLaunchKernelGGL(k_opt_clear, dim3(fr_grid_for(max_N, FR_BLOCK, 256), n), dim3(FR_BLOCK), 0, st, L.plans,
                       L.ctl, d_slots);
    hipLaunchKernelGGL(k_opt_rebuild, dim3(fr_grid_for(n, 64, 65535)), dim3(64), 0, st, L.plans, L.ctl, d_slots,
                       n);
}

/* graph_best_node.cu:182-222 for several plans (blockIdx.y), skipping plans
 * whose flag is already set (the flag is sticky, :229-230) */
__global__ void __launch_bounds__(FR_BLOCK) k_goal_scan_multi(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int gate)
{
    const int s = slots[blockIdx.y];
    const FrPlanDev &P = plans[s];
    FrCtl &C = ctl[s];
    if (C.status != FR_FOUND || (gate == 1 && !C.opt_active) || C.goal_reached)
        return;
    FrScan sc = dev_scan_begin(P, C, gate);
    int pos;
    while (dev_scan_next(P, sc, &pos)) {
        if (P.type[pos] != FR_TYPE_NODE)
            continue;
        const int z = pos / P.W, x = pos - z * P.W;
        if (dev_goal_hit(P, x, z, P.heading[pos]))
            C.goal_reached = 1;
    }
}

void fr_launch_goal_scan_multi(const FrLaunch &L, const int *d_slots, int n, int max_N, int gate, hipStream_t st)
{
    hipLaunchKernelGGL(k_goal_scan_multi, dim3(fr_grid_for(max_N, FR_BLOCK, 256), n), dim3(FR_BLOCK), 0, st,
                       L.plans, L.ctl, d_slots, gate);
}

/* A new plan's slab back to the state of a fresh allocation: its planes
 * zero, its claim keys all-ones (one block row per plan). */
__global__ void __launch_bounds__(FR_BLOCK) k_slab_reset(const FrPlanDev *plans, const int *slots)
{
    const FrPlanDev &P = plans[slots[blockIdx.y]];
    const long stride = (long)gridDim.x * blockDim.x;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (long w = t; w < P.zero_words; w += stride)
        P.zero_base[w] = make_uint4(0u, 0u, 0u, 0u);
    for (long c = t; c < P.N; c += stride)
        P.claim[c] = ~0ull;
}

/* search_init for a batch of freshly admitted plans (fastrrt.cpp:85-98): a
 * fresh control block, and the start node on the zeroed slab. */
__global__ void __launch_bounds__(FR_BLOCK) k_batch_start(const FrPlanDev *plans, FrCtl *ctl, const int *slots, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int s = slots[i];
    const FrPlanDev &P = plans[s];
    /* the state k_finish leaves for a next iteration (dev_ctl_ready), as the
     * first iteration of the plan starts from it (the density counts are
     * zero after the slab reset) */
    FrCtl c = {};
    c.status = FR_SEARCHING;
    c.pending_start = 1;
    c.force = 1;
    c.best_lcost = (long long)~0ull;
    c.best_pick = ~0ull;
    c.epoch = 1;
    ctl[s] = c;
    dev_add(P, P.sx, P.sz, P.sh, -1, -1, 0.0f);
}

void fr_launch_slab_reset(const FrLaunch &L, const int *d_slots, int n, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_slab_reset, dim3(64, n), dim3(FR_BLOCK), 0, st, L.plans, d_slots);
}

void fr_launch_batch_start(const FrLaunch &L, const int *d_slots, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_batch_start, dim3(fr_grid_for(n, 64, 65535)), dim3(64), 0, st, L.plans, L.ctl, d_slots, n);
}

/* Retired plans' control blocks back to the IDLE state of an empty slot.
 * The slots come in the kernel arguments, so retiring many plans at once
 * costs one launch rather than one host-to-device copy per plan. */
#define FR_IDLE_ARGS 240
struct FrSlotArgs {
    int n;
    int s[FR_IDLE_ARGS];
};

__global__ void __launch_bounds__(256) k_ctl_idle(FrCtl *ctl, FrSlotArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    FrCtl c = {};
    c.status = FR_IDLE;
    ctl[a.s[i]] = c;
}

void fr_launch_ctl_idle(FrCtl *ctl, const int *slots, int n, hipStream_t st)
{
    for (int base = 0; base < n; base += FR_IDLE_ARGS) {
        FrSlotArgs a;
        a.n = std::min(FR_IDLE_ARGS, n - base);
        for (int k = 0; k < a.n; ++k)
            a.s[k] = slots[base + k];
        hipLaunchKernelGGL(k_ctl_idle, dim3(1), dim3(256), 0, st, ctl, a);
    }
}

/* A copy by 4-byte words (16 at a time where both ends allow). */
__global__ void __launch_bounds__(256) k_copy_words(uint32_t *dst, const uint32_t *src, long nwords)
{
    const long stride = (long)gridDim.x * blockDim.x;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0) {
        const long n4 = nwords >> 2;
        for (long i = t; i < n4; i += stride)
            ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
        for (long i = 4 * n4 + t; i < nwords; i += stride)
            dst[i] = src[i];
    } else {
        for (long i = t; i < nwords; i += stride)
            dst[i] = src[i];
    }
}

void fr_launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    if (bytes == 0)
        return;
    if ((bytes & 3) || (((uintptr_t)dst | (uintptr_t)src) & 3))
        throw std::runtime_error("fr_launch_copy: sizes and addresses must be multiples of 4 bytes");
    const long nw = (long)(bytes / 4);
    const long blocks = std::min((nw / 4 + 255) / 256 + 1, 1024L);
    hipLaunchKernelGGL(k_copy_words, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t *)dst,
                       (const uint32_t *)src, nw);
}

/* ------------------------------------------------------------------------ */
/* single-plan graph operations (the CudaGraph API, python_cuda_graph_bind) */
/* ------------------------------------------------------------------------ */

/* graph.cu:364-381 (add / addTemporary); count_parent=0 for addTemporary */
__global__ void __launch_bounds__(FR_BLOCK) k_op_add(const FrPlanDev *plans, int slot, int x, int z, float h, int ppx, int ppz, float cost,
                         int type, int count_parent)
{
    const FrPlanDev &P = plans[slot];
    if (x < 0 || x >= P.W || z < 0 || z >= P.H)
        return;
    const int pos = z * P.W + x;
    if (count_parent && ppx != -1 && ppz != -1) {
        const long pp = (long)ppz * P.W + ppx;
        if (pp >= 0 && pp < P.N)
            P.derive[pp]++;
    }
    P.type[pos] = (uint8_t)type;
    P.px[pos] = ppx;
    P.pz[pos] = ppz;
    P.heading[pos] = h;
    P.cost[pos] = cost;
}

/* graph.cu:553-563 */
__global__ void __launch_bounds__(FR_BLOCK) k_op_set_collision(const FrPlanDev *plans, int slot, int x, int z, int npx, int npz, float h,
                                   float cost)
{
    const FrPlanDev &P = plans[slot];
    if (x < 0 || x >= P.W || z < 0 || z >= P.H)
        return;
    const int pos = z * P.W + x;
    const long cpp = (long)P.pz[pos] * P.W + P.px[pos];
    if (P.type[pos] != FR_TYPE_NODE)
        return;
    P.type[pos] = FR_TYPE_COLLISION;
    P.px[pos] = npx;
    P.pz[pos] = npz;
    P.heading[pos] = h;
    P.cost[pos] = cost;
    if (cpp >= 0 && cpp < P.N)
        P.derive[cpp]--;
}

/* CudaGraph::canConnectToGoal (graph_expansion.cu:246-261), a host call in
 * the reference, over cudaHermite (graph_hermite_connection.cu:11-136) with
 * fillInterpolation and computeDistance false: the curve length is the
 * frame's distance-to-goal channel, the walk starts from the node's own cell,
 * the curvature is |x'z'' - z'x''| / v (pow(v, 3/2) with integer 3/2 = 1)
 * against the bound, failing with -2; no endpoint test.  out[0] = the
 * curve's result (cost, -1 or -2), out[1] = 1 when the goal pixel is an
 * obstacle or outside the grid (the reference answers false before any
 * curve).  One thread: a test hook, not a loop phase. */
__global__ void __launch_bounds__(64) k_op_hermite(const FrPlanDev *plans, int slot, int x, int z, int gx, int gz,
                                                   int goal_heading, float max_curv, float *out)
{
    if (threadIdx.x != 0)
        return;
    const FrPlanDev &P = plans[slot];
    const int W = P.W, H = P.H;
    out[0] = -1;
    out[1] = 0;
    if (gx < 0 || gx >= W || gz < 0 || gz >= H) {
        out[1] = 1;
        return;
    }
    {
        const int cls = fr_trunc_f(P.frame[3 * ((long)gz * W + gx)]);
        if (cls < 0 || cls >= P.n_classes || P.class_costs[cls] < 0) {
            out[1] = 1;
            return;
        }
    }
    if (x < 0 || x >= W || z < 0 || z >= H)
        return;
    const int pos = z * W + x;
    const float distance = P.frame[3 * (long)pos + 1];
    const int num_points = fr_host_trunc_d((double)distance);
    const float a1 = (float)((double)P.heading[pos] - FR_HALF_PI_D);
    const float a2 = (float)((double)goal_heading - FR_HALF_PI_D);
    float s1, c1, s2, c2;
    fr_sincosf(a1, &s1, &c1);
    fr_sincosf(a2, &s2, &c2);
    const double t1x = (double)(distance * c1), t1y = (double)(distance * s1);
    const double t2x = (double)(distance * c2), t2y = (double)(distance * s2);
    const double dx0 = (double)x, dz0 = (double)z, dgx = (double)gx, dgz = (double)gz;
    int last_x = x, last_z = z;
    float node_cost = P.cost[pos];
    for (int i = 0; i < num_points; ++i) {
        const double t = (0.0 + (double)i) / (double)(num_points - 1);
        const double t2 = t * t, t3 = t2 * t;
        const double h00 = 2 * t3 - 3 * t2 + 1;
        const double h10 = t3 - 2 * t2 + t;
        const double h01 = -2 * t3 + 3 * t2;
        const double h11 = t3 - t2;
        const double px = h00 * dx0 + h10 * t1x + h01 * dgx + h11 * t2x;
        const double pz = h00 * dz0 + h10 * t1y + h01 * dgz + h11 * t2y;
        if (px < 0 || px >= W)
            continue;
        if (pz < 0 || pz >= H)
            continue;
        const int cx = fr_host_trunc_d(px);
        const int cz = fr_host_trunc_d(pz);
        if (cx == last_x && cz == last_z)
            continue;
        if (cx < 0 || cx >= W)
            continue;
        if (cz < 0 || cz >= H)
            continue;
        node_cost += (P.intrinsic ? P.intrinsic[cz * W + cx] : 0.0f) + 1.0f;
        const double t00 = 6 * t2 - 6 * t;
        const double t10 = 3 * t2 - 4 * t + 1;
        const double t01 = -6 * t2 + 6 * t;
        const double t11 = 3 * t2 - 2 * t;
        const double ddx = t00 * dx0 + t10 * t1x + t01 * dgx + t11 * t2x;
        const double ddz = t00 * dz0 + t10 * t1y + t01 * dgz + t11 * t2y;
        const float heading = (float)((double)fr_atan2f((float)ddz, (float)ddx) + FR_HALF_PI_D);
        if (max_curv > 0) {
            const double d00 = 12 * t - 6;
            const double d10 = 6 * t - 4;
            const double d01 = -12 * t + 6;
            const double d11 = 6 * t - 2;
            const double dd2x = d00 * dx0 + d10 * t1x + d01 * dgx + d11 * t2x;
            const double dd2z = d00 * dz0 + d10 * t1y + d01 * dgz + d11 * t2y;
            const double v = ddx * ddx + ddz * ddz;
            const float k = (float)(fr_fabs_d(ddx * dd2z - ddz * dd2x) / v);
            if (k > max_curv) {
                out[0] = -2;
                return;
            }
        }
        last_x = cx;
        last_z = cz;
        if (!dev_feasible(P, cx, cz, heading))
            return;
    }
    if (num_points > 0)
        out[0] = node_cost;
}

/* ------------------------------------------------------------------------ */
/* Path feasibility: check_kinematic_connection_start_end                   */
/* (kinematics.cu:233-306) behind CudaGraph::checkFeasibleConnection        */
/* (:308-326).  FP64 throughout, one lane per (start, end) query.           */
/* ------------------------------------------------------------------------ */

/* compute_path_heading, kinematics.cu:75-92 */
__device__ __forceinline__ double dev_path_heading(double p1x, double p1y, double p2x, double p2y)
{
    const double dy = p2y - p1y, dx = p2x - p1x;
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

/* clip, kinematics.cu:113-120 */
__device__ __forceinline__ double dev_clip(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* The footprint test at any centre.  The walk below can leave the grid;
 * there the reference still tests the footprint of the previous cell
 * (path_feasible.cu:18-78: out-of-grid samples are skipped).  In the grid
 * this is dev_feasible; outside it every sample is tested against the
 * frame (ego box skipped, a class outside the table is an obstacle), and
 * the all-heading bit of a centre outside the grid counts as clear (the
 * reference reads it past the frame). */
__device__ bool dev_feasible_anywhere(const FrPlanDev &P, int x, int z, float h)
{
    if (x >= 0 && x < P.W && z >= 0 && z < P.H)
        return dev_feasible(P, x, z, h);
    float s, c;
    fr_sincosf(h, &s, &c);
    for (int i = -P.mz; i <= P.mz; ++i)
        for (int j = -P.mx; j <= P.mx; ++j) {
            const int xl = fr_trunc_f((float)j * c - (float)i * s + (float)x);
            const int zl = fr_trunc_f((float)j * s + (float)i * c + (float)z);
            if (xl < 0 || xl >= P.W || zl < 0 || zl >= P.H)
                continue;
            if (xl >= P.lbx && xl <= P.ubx && zl >= P.ubz && zl <= P.lbz)
                continue;
            const int cls = fr_trunc_f(P.frame[3 * ((long)zl * P.W + xl)]);
            if (cls < 0 || cls >= P.n_classes || P.class_costs[cls] < 0)
                return false;
        }
    return true;
}

/* kinematics.cu:233-306.  The start heading is the grid's heading at the
 * start cell; velocity is the reference's int argument (:308).  Outputs:
 * the verdict, and the heading and cost reached when the walk stopped on
 * moving away from the end (both 0 otherwise, as in the reference).  The
 * intrinsic cost of a cell outside the grid counts as 0 (the reference
 * reads past the grid there). */
__device__ bool dev_feasible_connection(const FrPlanDev &P, int sx, int sz, int ex, int ez, int velocity,
                                        double *final_heading, float *path_cost)
{
    *final_heading = 0;
    *path_cost = 0;
    const double ddx0 = (double)(ex - sx), ddz0 = (double)(ez - sz);
    const double distance = sqrt(ddx0 * ddx0 + ddz0 * ddz0);
    const double lr = P.lr;
    const double max_steer = P.max_steer_rad;
    const double smx = (double)(P.og_y - (float)sz), smy = (double)((float)sx - P.og_x);
    const double emx = (double)(P.og_y - (float)ez), emy = (double)((float)ex - P.og_x);
    const double dt = 0.1;
    double heading = (double)P.heading[sz * P.W + sx];
    double steer_a = dev_clip(dev_path_heading(smx, smy, emx, emy) - heading, -max_steer, max_steer);
    const double ds = (double)(float)velocity * dt;
    const int total_steps = fr_trunc_d(distance / ds);
    double nx = smx, ny = smy;
    int lx = sx, lz = sz;
    double best_end = distance;
    float curr_cost = 0;
    for (int i = 0; i < total_steps; i++) {
        const double steer = tan(steer_a);
        const double beta = atan(steer / lr);
        nx += ds * cos(heading + beta);
        ny += ds * sin(heading + beta);
        heading += ds * cos(beta) * steer / (2 * lr);
        steer_a = dev_clip(dev_path_heading(nx, ny, emx, emy) - heading, -max_steer, max_steer);
        const double dx = emx - nx, dy = emy - ny;
        const double dist = sqrt(dx * dx + dy * dy);
        const int cx = fr_trunc_d((double)P.og_x + ny), cz = fr_trunc_d((double)P.og_y - nx);
        if (cx == lx && cz == lz)
            continue;
        const bool in = cx >= 0 && cx < P.W && cz >= 0 && cz < P.H;
        curr_cost += ((in && P.intrinsic) ? P.intrinsic[cz * P.W + cx] : 0.0f) + 1.0f;
        if (best_end < dist) {
            *final_heading = heading;
            *path_cost = curr_cost;
            return best_end <= 2;
        }
        if (!dev_feasible_anywhere(P, lx, lz, (float)heading))
            return false;
        lx = cx;
        lz = cz;
        best_end = dist;
    }
    return true;
}

/* One lane per query q = (start x, start z, end x, end z); a start outside
 * the grid answers false. */
__global__ void __launch_bounds__(FR_BLOCK) k_op_feasible_connection(const FrPlanDev *plans, int slot, const int4 *q,
                                                                     int n, int velocity, uint8_t *ok,
                                                                     double *final_heading, float *path_cost)
{
    const FrPlanDev &P = plans[slot];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int4 a = q[k];
        double fh = 0;
        float pc = 0;
        bool r = false;
        if (a.x >= 0 && a.x < P.W && a.y >= 0 && a.y < P.H)
            r = dev_feasible_connection(P, a.x, a.y, a.z, a.w, velocity, &fh, &pc);
        ok[k] = r ? 1 : 0;
        final_heading[k] = fh;
        path_cost[k] = pc;
    }
}

/* CudaGraph::derivateNode -> expand_node, host sequential semantics
 * (graph_expansion.cu:103-145,224-244).  out = end cell or (-1, -1). */
__global__ void __launch_bounds__(FR_BLOCK) k_op_derivate(const FrPlanDev *plans, int slot, float steer, float path_size, float v, int x, int z,
                              int *out)
{
    FrPlanDev P = plans[slot];
    out[0] = -1;
    out[1] = -1;
    if (x < 0 || x >= P.W || z < 0 || z >= P.H)
        return;
    const int pos = z * P.W + x;
    if (P.type[pos] != FR_TYPE_NODE)
        return;
    P.v = v;
    int ex, ez;
    float eh, ec;
    if (!dev_rollout(P, x, z, steer, path_size, &ex, &ez, &eh, &ec))
        return;
    if ((float)ex == P.og_x && (float)ez == P.og_y)
        return;
    const int epos = ez * P.W + ex;
    if (epos == pos)
        return;
    if (P.type[epos] == FR_TYPE_NULL) {
        P.type[epos] = FR_TYPE_TEMP;
        P.derive[pos]++;
        P.px[epos] = x;
        P.pz[epos] = z;
        P.heading[epos] = eh;
        P.cost[epos] = ec;
    }
    if (P.type[epos] == FR_TYPE_NODE) {
        P.type[epos] = FR_TYPE_COLLISION;
        P.px[epos] = x;
        P.pz[epos] = z;
        P.heading[epos] = eh;
        P.cost[epos] = ec;
        P.derive[pos]--;
        out[2] = 1; /* node collision flag */
    }
    out[0] = ex;
    out[1] = ez;
}

__global__ void __launch_bounds__(FR_BLOCK) k_op_accept(const FrPlanDev *plans, int slot)
{
    const FrPlanDev &P = plans[slot];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x)
        if (P.type[pos] == FR_TYPE_TEMP)
            P.type[pos] = FR_TYPE_NODE;
}

__global__ void __launch_bounds__(FR_BLOCK) k_op_solve_mark(const FrPlanDev *plans, int slot, int nodes)
{
    const FrPlanDev &P = plans[slot];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x) {
        uint8_t m = 0;
        if (P.type[pos] != FR_TYPE_NULL) {
            long curr = pos;
            int i = nodes + 4;
            while (i-- > 0) {
                const int ppx = P.px[curr], ppz = P.pz[curr];
                if (ppx == -1 && ppz == -1)
                    break;
                const long next = (long)ppz * P.W + ppx;
                const int tn = (next >= 0 && next < P.N) ? P.type[next] : FR_TYPE_NULL;
                if (tn == FR_TYPE_COLLISION || tn == FR_TYPE_NULL) {
                    m = 1;
                    break;
                }
                curr = next;
            }
        }
        P.mark[pos] = m;
    }
}

__global__ void __launch_bounds__(FR_BLOCK) k_op_solve_apply(const FrPlanDev *plans, int slot)
{
    const FrPlanDev &P = plans[slot];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x) {
        if (P.mark[pos]) {
            P.type[pos] = FR_TYPE_NULL;
            P.mark[pos] = 0;
        } else if (P.type[pos] == FR_TYPE_COLLISION) {
            P.type[pos] = FR_TYPE_NODE;
            P.derive[pos] = 0;
            P.mark[pos] = 2;
        }
    }
}

__global__ void __launch_bounds__(FR_BLOCK) k_op_solve_parents(const FrPlanDev *plans, int slot)
{
    const FrPlanDev &P = plans[slot];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x) {
        if (P.mark[pos] != 2)
            continue;
        P.mark[pos] = 0;
        const long pp = (long)P.pz[pos] * P.W + P.px[pos];
        if (pp >= 0 && pp < P.N)
            atomicAdd(&P.derive[pp], 1);
    }
}

__global__ void __launch_bounds__(FR_BLOCK) k_op_region_count(const FrPlanDev *plans, int slot, uint32_t *out)
{
    const FrPlanDev &P = plans[slot];
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < P.N; pos += gridDim.x * blockDim.x) {
        const int ty = P.type[pos];
        if (ty == FR_TYPE_NULL || ty == FR_TYPE_PROCESSING)
            continue;
        const int z = pos / P.W, x = pos - z * P.W;
        atomicAdd(&out[(z / P.region_side) * P.dens_w + x / P.region_side], 1u);
    }
}

void fr_launch_graph_add(const FrLaunch &L, int slot, int x, int z, float h, int px, int pz, float cost, int type,
                         int count_parent, hipStream_t st)
{
    hipLaunchKernelGGL(k_op_add, dim3(1), dim3(1), 0, st, L.plans, slot, x, z, h, px, pz, cost, type, count_parent);
}

void fr_launch_set_collision(const FrLaunch &L, int slot, int x, int z, int npx, int npz, float h, float cost,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_op_set_collision, dim3(1), dim3(1), 0, st, L.plans, slot, x, z, npx, npz, h, cost);
}

void fr_launch_derivate(const FrLaunch &L, int slot, float steer, float path_size, float v, int x, int z, int *d_out,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_op_derivate, dim3(1), dim3(1), 0, st, L.plans, slot, steer, path_size, v, x, z, d_out);
}

void fr_launch_hermite_probe(const FrLaunch &L, int slot, int x, int z, int gx, int gz, int goal_heading,
                             float max_curv, float *d_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_op_hermite, dim3(1), dim3(64), 0, st, L.plans, slot, x, z, gx, gz, goal_heading, max_curv,
                       d_out);
}

void fr_launch_feasible_connections(const FrLaunch &L, int slot, const int4 *d_q, int n, int velocity,
                                    uint8_t *d_ok, double *d_heading, float *d_cost, hipStream_t st)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_op_feasible_connection, dim3(fr_grid_for(n, FR_BLOCK, 4096)), dim3(FR_BLOCK), 0, st, L.plans,
                       slot, d_q, n, velocity, d_ok, d_heading, d_cost);
}

void fr_launch_accept(const FrLaunch &L, int slot, int N, hipStream_t st)
{
    hipLaunchKernelGGL(k_op_accept, dim3(fr_grid_for(N, FR_BLOCK, 2048)), dim3(FR_BLOCK), 0, st, L.plans, slot);
}

void fr_launch_solve_collisions(const FrLaunch &L, int slot, int N, int nodes, hipStream_t st)
{
    dim3 g(fr_grid_for(N, FR_BLOCK, 2048));
    hipLaunchKernelGGL(k_op_solve_mark, g, dim3(FR_BLOCK), 0, st, L.plans, slot, nodes);
    hipLaunchKernelGGL(k_op_solve_apply, g, dim3(FR_BLOCK), 0, st, L.plans, slot);
    hipLaunchKernelGGL(k_op_solve_parents, g, dim3(FR_BLOCK), 0, st, L.plans, slot);
}

void fr_launch_region_count(const FrLaunch &L, int slot, int N, uint32_t *d_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_op_region_count, dim3(fr_grid_for(N, FR_BLOCK, 2048)), dim3(FR_BLOCK), 0, st, L.plans,
                       slot, d_out);
}
