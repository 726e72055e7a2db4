/*
 * fr_device.h — device-side plan descriptors for the gfx950 planning loop.
 *
 * One "plan" is one FastRRT search: a node grid of W x H cells kept in HBM
 * as structure-of-arrays planes (type, parent, heading, cost, child count),
 * the goal-connection planes computed once per plan, the per-plan obstacle
 * distance plane the footprint test reads, and a small control block.  The
 * same kernels advance one plan (the FastRRT handle: one plan per launch
 * sequence, host-synchronised per loop() call) or hundreds of plans of
 * uneven size at once (the batch engine: persistent grid-stride kernels over
 * a flattened (plan, tile) table, captured as a hipGraph).
 */
#ifndef FR_DEVICE_H
#define FR_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fr_math.h"

#define FR_TILE 1024          /* cells per tile of the flattened tile table */
#define FR_BLOCK 256          /* threads per block of the tile kernels */
#define FR_DIST_CAP 63        /* largest cap of the Chebyshev obstacle-distance plane (min(pad, 63) is kept) */
#define FR_BIG_COST 99999999999LL
#define FR_GROUP 7            /* footprint samples are certified in FR_GROUP x FR_GROUP groups */
#define FR_GROUP_FREE 6       /* a group's middle sample at this distance certifies the group */
#define FR_MAX_GROUPS 32      /* footprints of more groups are sampled row by row */

/* The planes live in device global memory.  In the device pass their
 * pointers carry the global address space, so every access compiles to a
 * global_* instruction: a flat access also counts against the LDS/scalar
 * wait counter, so each LDS or scalar wait would wait for it as well.  The
 * layout is the same in both passes (64-bit pointers). */
#if defined(__HIP_DEVICE_COMPILE__)
#define FR_GLOBAL __attribute__((address_space(1)))
#else
#define FR_GLOBAL
#endif

enum FrStatus {
    FR_IDLE = 0,
    FR_SEARCHING = 1,
    FR_FOUND = 2,     /* goal reached; path extraction pending or done */
    FR_FAILED = 3,    /* iteration cap / deadline */
};

/* Parameters and plane pointers of one plan slot (read-only to kernels). */
struct FrPlanDev {
    int W, H, N;
    int pad, PW, PH;          /* padded distance plane (W + 2 pad) x (H + 2 pad) */
    int TW;                   /* row pitch (bytes) of the padded distance plane */
    int mx, mz;               /* half footprint (min_dist / 2, truncated) */
    int lbx, lbz, ubx, ubz;   /* ego box, pixels inside are ignored */
    float og_x, og_y;         /* map-frame origin in grid pixels */
    float v, max_path, dist_tol, herr_tol;
    float max_curv_loop;      /* curvature bound of the per-loop goal connection */
    float max_curv_goal;      /* curvature bound of the per-plan precompute */
    double max_steer_rad, lr;
    int sx, sz, gx, gz;
    float sh, gh;
    int region_side, dens_w, dens_h;
    int smart;
    int max_iter;
    int exp_budget;           /* > 0: eligible nodes thinned to about this many rollouts per iteration */
    int n_groups;             /* 7 x 7 sample groups of the footprint (0: more than FR_MAX_GROUPS) */
    float probe_j[FR_MAX_GROUPS], probe_i[FR_MAX_GROUPS]; /* each group's middle sample (column, row) */
    uint64_t seed, stream;
    /* node grid */
    FR_GLOBAL uint8_t *type;
    FR_GLOBAL int32_t *derive, *px, *pz;
    FR_GLOBAL float *heading, *cost;
    FR_GLOBAL float *intrinsic;         /* may be null: all-zero intrinsic costs */
    /* goal connection */
    FR_GLOBAL uint8_t *gc_flag;
    FR_GLOBAL float *gc_heading, *gc_cost;
    FR_GLOBAL uint32_t *gc_done;        /* bit per cell: goal-connection planes evaluated */
    FR_GLOBAL uint32_t *gc_bits;        /* bit per cell: gc_flag, for window scans by word */
    int gc_lazy;              /* evaluate a cell when a node's window first covers it */
    /* scratch planes */
    FR_GLOBAL unsigned long long *claim;
    FR_GLOBAL int32_t *escr;            /* 4 * N ints: the erase's node list, cell -> list index, two jump buffers */
    FR_GLOBAL uint8_t *mark;
    FR_GLOBAL uint8_t *dist;            /* padded Chebyshev distance to the nearest obstacle pixel, row-major
                                 (fr_dist_at) */
    FR_GLOBAL uint8_t *dist_tmp;
    FR_GLOBAL uint8_t *trav;            /* bit0: feasible for all headings; bit1: safe-zone bit 0x100 */
    FR_GLOBAL uint32_t *region;
    /* input frame (float3 per pixel) and class-cost table */
    const FR_GLOBAL float *frame;
    const FR_GLOBAL float *class_costs;
    int n_classes;
    /* path buffers */
    FR_GLOBAL float4 *path;
    FR_GLOBAL float4 *pd;
    int path_cap;
    int opt_limit;            /* path_optimize rounds of batch finalisation */
    /* planes a new plan starts from zero (batch admission resets them) */
    FR_GLOBAL uint4 *zero_base;
    long zero_words;
};

/* Byte offset of padded cell (xp, zp) in the distance plane: row-major with
 * a pitch of TW bytes, one multiply-add per lookup (an 8 x 8-tiled layout
 * measured 1 % slower per batched iteration: its address takes seven
 * instructions, and the footprint's bytes hit L2 either way). */
#if defined(__HIPCC__)
__host__ __device__ __forceinline__
#else
inline
#endif
unsigned fr_dist_at(int TW, int xp, int zp)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24((unsigned)zp, (unsigned)TW) + (unsigned)xp; /* full-rate 24-bit product */
#else
    return (unsigned)(zp * TW + xp);
#endif
}

/* Mutable per-plan control block. */
struct FrCtl {
    int status;
    int iter;
    int last_temp;
    int goal_reached;
    int collision;
    int n_node_pre;           /* NODE cells at expansion start */
    int n_conv;               /* NODE -> COLLISION conversions this iteration */
    int n_temp;               /* TEMP cells after erase */
    int n_all;                /* non-NULL cells after erase */
    int reseeded;             /* the start was re-added this iteration */
    int pending_start;        /* the start cell must be goal-checked next iteration */
    int frontier, force;
    unsigned epoch;
    int n_expansions;         /* rollouts queued this iteration */
    int n_fresh;              /* nodes goal-checked this iteration */
    long long best_lcost;     /* goal direct connection: min floor(100 * total) */
    unsigned long long best_pick; /* min (lcost << 24 | node cell) over the connections (~0: none) */
    int pick_slow;            /* a connection outside best_pick's fields: k_finish picks by scanning */
    int found;                /* direct connection added this iteration */
    /* path extraction / optimisation */
    long long bn_cost;
    int bn_pos;
    int path_len;
    int opt_rounds;
    int opt_changed;
    float opt_gain_f;
    int opt_limit;            /* batch finalisation: max successful rounds */
    unsigned long long opt_key;
    int full_check;           /* next connect pass must scan all NODE cells */
    int total_expansions;     /* rollouts since the search started */
    int chain_only;           /* the grid holds only the rebuilt path chain (after an optimisation round) */
    int opt_active;           /* batch finalisation: optimisation rounds pending */
    int n_eligible;           /* nodes the expansion gate passed this iteration (before any thinning) */
    int prev_eligible;        /* the same, previous iteration */
    /* the last finished iteration's figures (the per-iteration fields above
     * are reset for the next iteration when one finishes) */
    int stat_expansions, stat_fresh, stat_reseeded;
    /* work done for the plan since admission: cells its rollouts entered
     * (each one footprint test), Hermite points of the goal-connection
     * curves evaluated by the centre pass, and of those the footprint pass
     * tested (each a curve's point count) */
    int w_cells, w_goal_pts, w_foot_pts;
};

/* One queued expansion (one rollout). */
struct FrExp {
    int slot;
    int src;
    int end;                  /* -1: rejected */
    float heading;
    float cost;
};

/* One node to goal-check / connect this iteration. */
struct FrFresh {
    int slot;
    int pos;
    int flags;                /* bit0: converted from COLLISION; bit1: window evaluated inline */
    int pair_base;            /* first (node, window cell) pair of this node */
    int pair_cnt;
    int pad0;
    unsigned long long key;   /* node-local best connection (cost, then last in scan order) */
};

/* One (node, goal-flagged window cell) Hermite check. */
struct FrPair {
    int fresh;
    int cell;
    float cost;               /* node -> cell curve cost (set when the curve is valid) */
};

#define FR_FRESH_CONVERTED 1
#define FR_FRESH_INLINE 2

#endif
