/*
 * fr_math.h — scalar primitives shared by the gfx950 kernels, the host
 * planner and the CPU oracle.
 *
 * The planning loop turns FP32 poses into integer grid cells with C
 * truncation after hundreds of 0.1-px integration steps
 * (reference: src/cuda/kinematics.cu:184-221) and samples a rotated
 * footprint through cosf/sinf (src/cuda/path_feasible.cu:450-461).
 * A one-ULP difference between two implementations of cosf flips cells, so
 * the device and the host must evaluate exactly the same arithmetic.  The
 * float transcendental functions below therefore widen the float argument
 * to FP64, reduce it with a two-part Cody-Waite split of pi/2, evaluate a
 * truncated Taylor series in FP64 with plain (non-fused) operations, and
 * round the FP64 result once to FP32.  Every translation unit that includes
 * this header is compiled with FP contraction off, so gcc on x86 and clang
 * on gfx950 produce identical bits.  Accuracy (|error| well below one FP32
 * ULP except at rounding ties) is checked against libm in FP64 by the CPU
 * test suite.
 *
 * Also here: C truncation with the device conversion rules (NaN -> 0,
 * saturation), and the Philox4x32-10 counter-based generator of
 * Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
 * (SC'11), which replaces the reference's per-cell XORWOW state
 * (src/cuda/random_number_gen.cu:5-45) with a stateless draw keyed by
 * (seed, stream, iteration, cell).
 *
 * The header is valid C99 (oracle, gcc) and HIP C++ (hipcc).
 */
#ifndef FR_MATH_H
#define FR_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define FR_HD __host__ __device__ static inline
#else
#define FR_HD static inline
#endif

#ifdef __cplusplus
#define FR_NAN_D (__builtin_nan(""))
#else
#define FR_NAN_D (__builtin_nan(""))
#endif

/* Node types (reference: include/graph.h:31-36). */
#define FR_TYPE_NULL 0
#define FR_TYPE_NODE 1
#define FR_TYPE_TEMP 2
#define FR_TYPE_PROCESSING 3
#define FR_TYPE_COLLISION 4
#define FR_TYPE_CONNECT_TO_GOAL 5

#define FR_PI_D 3.141592653589793238462643383279502884
#define FR_HALF_PI_D 1.570796326794896619231321691639751442
#define FR_TO_DEG_D (180.0 / FR_PI_D)
#define FR_TO_RAD_D (FR_PI_D / 180.0)

/* pi/2 split: the high part carries 33 significant bits so that k * P1 is
 * exact for |k| < 2^20; the low part is the FP64 rounding of the rest. */
#define FR_PIO2_HI 1.57079632673412561416625976562500000000e+00
#define FR_PIO2_LO 6.07710050650619224932663035837264e-11
#define FR_TWO_OVER_PI 6.36619772367581382433e-01

FR_HD int fr_isnan_d(double v) { return v != v; }

/* C (int) truncation as the device performs it: NaN -> 0, saturating. */
FR_HD int fr_trunc_d(double v)
{
    if (v != v)
        return 0;
    if (v >= 2147483647.0)
        return 2147483647;
    if (v <= -2147483648.0)
        return (-2147483647 - 1);
    return (int)v;
}

FR_HD int fr_trunc_f(float v) { return fr_trunc_d((double)v); }

FR_HD double fr_floor_d(double v)
{
    /* floor without libm: valid for |v| < 2^52 (all uses here are). */
    double t = (double)(int64_t)v;
    return (t > v) ? t - 1.0 : t;
}

FR_HD double fr_fabs_d(double v) { return v < 0 ? -v : v; }

/* Taylor kernels on |r| <= pi/4; truncation error < 5e-17. */
FR_HD double fr_sin_kernel(double r)
{
    double r2 = r * r;
    double p = -7.6471637318198164759e-13;           /* -1/15! */
    p = p * r2 + 1.6059043836821614599e-10;           /* 1/13! */
    p = p * r2 - 2.5052108385441718775e-08;           /* -1/11! */
    p = p * r2 + 2.7557319223985890653e-06;           /* 1/9! */
    p = p * r2 - 1.9841269841269841253e-04;           /* -1/7! */
    p = p * r2 + 8.3333333333333332177e-03;           /* 1/5! */
    p = p * r2 - 1.6666666666666665741e-01;           /* -1/3! */
    return r + r * (r2 * p);
}

FR_HD double fr_cos_kernel(double r)
{
    double r2 = r * r;
    double p = 4.7794773323873852974e-14;             /* 1/16! */
    p = p * r2 - 1.1470745597729724714e-11;           /* -1/14! */
    p = p * r2 + 2.0876756987868098979e-09;           /* 1/12! */
    p = p * r2 - 2.7557319223985888276e-07;           /* -1/10! */
    p = p * r2 + 2.4801587301587301566e-05;           /* 1/8! */
    p = p * r2 - 1.3888888888888889419e-03;           /* -1/6! */
    p = p * r2 + 4.1666666666666664354e-02;           /* 1/4! */
    p = p * r2 - 0.5;                                 /* -1/2! */
    return 1.0 + r2 * p;
}

/* sin and cos of a float argument, evaluated in FP64, not yet rounded. */
FR_HD void fr_sincos_d(double x, double *s, double *c)
{
    if (x != x || x - x != 0.0) { /* NaN or infinity */
        *s = FR_NAN_D;
        *c = FR_NAN_D;
        return;
    }
    double k = fr_floor_d(x * FR_TWO_OVER_PI + 0.5);
    double r = (x - k * FR_PIO2_HI) - k * FR_PIO2_LO;
    double sr = fr_sin_kernel(r);
    double cr = fr_cos_kernel(r);
    int64_t ki = (int64_t)k;
    int q = (int)(ki & 3);
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

FR_HD float fr_sinf(float x)
{
    double s, c;
    fr_sincos_d((double)x, &s, &c);
    return (float)s;
}

FR_HD float fr_cosf(float x)
{
    double s, c;
    fr_sincos_d((double)x, &s, &c);
    return (float)c;
}

FR_HD void fr_sincosf(float x, float *s, float *c)
{
    double sd, cd;
    fr_sincos_d((double)x, &sd, &cd);
    *s = (float)sd;
    *c = (float)cd;
}

FR_HD float fr_tanf(float x)
{
    double s, c;
    fr_sincos_d((double)x, &s, &c);
    return (float)(s / c);
}

/* atan of a in [0, 1] (pi/2 - atan a when inv), in FP64.  With j*pi/16 the
 * nearest of five angles and c = tan(j*pi/16) (rounded to FP64),
 * atan(a) = j*pi/16 + atan(t), t = (a - c) / (1 + a c), |t| <= tan(pi/32),
 * and the odd Taylor series of atan(t) to t^17 finishes (the first term left
 * out is below 4e-21).  The tangents and thresholds tan((2j+1)pi/32) come from
 * a 60-digit evaluation. */
FR_HD double fr_atan_unit(double a, int inv)
{
    const int j = (a > 0.09849140335716425) + (a > 0.3033466836073424) + (a > 0.5345111359507917) +
                  (a > 0.8206787908286604);
    double c = 0.0;
    c = j == 1 ? 0.198912367379658 : c;
    c = j == 2 ? 0.41421356237309503 : c;
    c = j == 3 ? 0.6681786379192989 : c;
    c = j == 4 ? 1.0 : c;
    const double t = (a - c) / (1.0 + a * c);
    const double t2 = t * t;
    double p = 1.0 / 17.0;
    p = 1.0 / 15.0 - t2 * p;
    p = 1.0 / 13.0 - t2 * p;
    p = 1.0 / 11.0 - t2 * p;
    p = 1.0 / 9.0 - t2 * p;
    p = 1.0 / 7.0 - t2 * p;
    p = 1.0 / 5.0 - t2 * p;
    p = 1.0 / 3.0 - t2 * p;
    double r = (double)j * 0.19634954084936207 + (t - t * (t2 * p));
    if (inv)
        r = FR_HALF_PI_D - r;
    return r;
}

FR_HD double fr_atan_d(double x)
{
    if (x != x)
        return x;
    int neg = x < 0;
    double a = neg ? -x : x;
    double r;
    if (a > 1.0) {
        if (a - a != 0.0) /* infinity */
            return neg ? -FR_HALF_PI_D : FR_HALF_PI_D;
        r = fr_atan_unit(1.0 / a, 1);
    } else {
        r = fr_atan_unit(a, 0);
    }
    return neg ? -r : r;
}

FR_HD float fr_atanf(float x) { return (float)fr_atan_d((double)x); }

FR_HD int fr_signbit_d(double v)
{
    union { double d; uint64_t u; } b;
    b.d = v;
    return (int)(b.u >> 63);
}

/* atan2 with the IEEE special cases, evaluated in FP64. */
FR_HD double fr_atan2_d(double y, double x)
{
    if (x != x || y != y)
        return FR_NAN_D;
    double ay = fr_fabs_d(y), ax = fr_fabs_d(x);
    double r;
    if (y == 0.0) {
        r = fr_signbit_d(x) ? FR_PI_D : 0.0;
    } else if (x == 0.0) {
        r = FR_HALF_PI_D;
    } else if (ax - ax != 0.0 && ay - ay != 0.0) {
        r = fr_signbit_d(x) ? 3.0 * FR_PI_D / 4.0 : FR_PI_D / 4.0;
    } else if (ax - ax != 0.0) {
        r = fr_signbit_d(x) ? FR_PI_D : 0.0;
    } else if (ay - ay != 0.0) {
        r = FR_HALF_PI_D;
    } else {
        /* one division: the smaller magnitude over the larger */
        r = ay > ax ? fr_atan_unit(ax / ay, 1) : fr_atan_unit(ay / ax, 0);
        if (fr_signbit_d(x))
            r = FR_PI_D - r;
    }
    return fr_signbit_d(y) ? -r : r;
}

FR_HD float fr_atan2f(float y, float x) { return (float)fr_atan2_d((double)y, (double)x); }

/* ---------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon et al., SC'11, section 4.2).                      */
/* ---------------------------------------------------------------------- */
#define FR_PHILOX_M0 0xD2511F53u
#define FR_PHILOX_M1 0xCD9E8D57u
#define FR_PHILOX_W0 0x9E3779B9u
#define FR_PHILOX_W1 0xBB67AE85u

typedef struct { uint32_t v[4]; } fr_u32x4;

FR_HD fr_u32x4 fr_philox4x32_10(fr_u32x4 ctr, uint32_t k0, uint32_t k1)
{
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)FR_PHILOX_M0 * (uint64_t)ctr.v[0];
        uint64_t p1 = (uint64_t)FR_PHILOX_M1 * (uint64_t)ctr.v[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        fr_u32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += FR_PHILOX_W0;
        k1 += FR_PHILOX_W1;
    }
    return ctr;
}

/* Uniform float in (0, 1] from the top 24 bits: exact in FP32. */
FR_HD float fr_u01(uint32_t r) { return (float)((r >> 8) + 1u) * (1.0f / 16777216.0f); }

/* The two draws one expansion consumes: counter = (cell, iteration,
 * stream_lo, stream_hi), key = seed.  u[0] -> steering, u[1] -> length. */
FR_HD void fr_expansion_draws(uint64_t seed, uint64_t stream, uint32_t iter, uint32_t cell,
                              float *u_steer, float *u_len)
{
    fr_u32x4 c;
    c.v[0] = cell;
    c.v[1] = iter;
    c.v[2] = (uint32_t)stream;
    c.v[3] = (uint32_t)(stream >> 32);
    fr_u32x4 r = fr_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    *u_steer = fr_u01(r.v[0]);
    *u_len = fr_u01(r.v[1]);
}

/* Third word of the same Philox block: the draw of the optional expansion
 * thinning (unused when every eligible node expands). */
FR_HD uint32_t fr_expansion_gate(uint64_t seed, uint64_t stream, uint32_t iter, uint32_t cell)
{
    fr_u32x4 c;
    c.v[0] = cell;
    c.v[1] = iter;
    c.v[2] = (uint32_t)stream;
    c.v[3] = (uint32_t)(stream >> 32);
    return fr_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32)).v[2];
}

#endif /* FR_MATH_H */
