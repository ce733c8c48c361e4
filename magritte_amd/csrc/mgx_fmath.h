/*
 * Deterministic single-precision transcendental functions.
 *
 * Magritte's formal solver calls the float functions expf, expm1f, logf, powf
 * and erff of the C library (src/solver/solver.tpp:1109, 1120, 2450;
 * src/solver/solver_interp_helper.tpp:37-39, 67-68, 98;
 * src/model/thermodynamics/thermodynamics.tpp:13).  Their results feed the
 * traced point counts (through logf) and every opacity, source function and
 * optical depth.  To let the CPU restatement and the gfx950 kernels produce the
 * same bits, both evaluate the functions below: each widens its float
 * argument to double, evaluates a fixed polynomial in a fixed order with
 * correctly rounded IEEE operations -- explicit fused multiply-adds (fma() of
 * the C library on the host, v_fma_f64 on the device) and plain operations
 * elsewhere (the file must be compiled with -ffp-contract=off, so the compiler
 * forms no other FMA) -- and rounds once to float.
 *
 * Accuracy: the double results are within a few units in the last place of
 * double, so the float result is the correctly rounded one except when the
 * exact value lies within ~1e-15 (relative) of a float rounding boundary.
 *
 * The header compiles as C (gcc, for the CPU restatement) and as HIP
 * (hipcc, host and device).
 */
#pragma once

#include <math.h>
#include <stdint.h>

#define MGX_FMA(a, b, c) fma((a), (b), (c))

#if defined(__HIPCC__)
#define MGX_FM_HD __host__ __device__ static inline
#if defined(__HIP_DEVICE_COMPILE__)
#define MGX_FM_CONST static __constant__ const
#else
#define MGX_FM_CONST static const
#endif
#else
#define MGX_FM_HD static inline
#define MGX_FM_CONST static const
#endif

#include "mgx_fmath_coeffs.h"

MGX_FM_HD uint64_t mgx_fm_bits(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u;
}

MGX_FM_HD double mgx_fm_from_bits(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, sizeof x);
    return x;
}

/* 2^k for -1022 <= k <= 1023 (exact). */
MGX_FM_HD double mgx_fm_pow2(int k)
{
    return mgx_fm_from_bits((uint64_t)(k + 1023) << 52);
}

/* Round to nearest integer (ties to even) for |x| < 2^51. */
MGX_FM_HD double mgx_fm_rint(double x)
{
    const double shift = 0x1.8p52;
    return (x + shift) - shift;
}

/* exp(x) in double for x in [-708, 709]; 0 below, +inf above. */
MGX_FM_HD double mgx_exp_d(double x)
{
    if (x != x) return x;
    if (x > 709.0) return 1.0 / 0.0;
    if (x < -708.0) return 0.0;
    const double kd = mgx_fm_rint(x * MGX_LOG2E);
    const double r = MGX_FMA(-kd, MGX_LN2_LO, MGX_FMA(-kd, MGX_LN2_HI, x));
    double p = MGX_EXP_C[13];
    for (int n = 12; n >= 0; n--) p = MGX_FMA(p, r, MGX_EXP_C[n]);
    /* p in [0.70, 1.42]; the product stays normal for the clamped range */
    const int k = (int)kd;
    if (k < -1000) return (p * mgx_fm_pow2(k + 200)) * mgx_fm_pow2(-200);
    if (k > 1000) return (p * mgx_fm_pow2(k - 200)) * mgx_fm_pow2(200);
    return p * mgx_fm_pow2(k);
}

/* Natural logarithm in double for finite x > 0 (also handles 0, <0, inf, nan). */
MGX_FM_HD double mgx_log_d(double x)
{
    if (x != x) return x;
    if (x < 0.0) return 0.0 / 0.0;
    if (x == 0.0) return -1.0 / 0.0;
    if (x == 1.0 / 0.0) return x;
    int e = 0;
    if (x < 0x1p-1022) { /* subnormal: scale into the normal range */
        x = x * 0x1p54;
        e = -54;
    }
    const uint64_t u = mgx_fm_bits(x);
    e += (int)((u >> 52) & 0x7ff) - 1023;
    double m = mgx_fm_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > MGX_SQRT2) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = MGX_LOG_C[11];
    for (int n = 10; n >= 0; n--) p = MGX_FMA(p, s2, MGX_LOG_C[n]);
    const double ed = (double)e;
    return MGX_FMA(ed, MGX_LN2_HI, MGX_FMA(ed, MGX_LN2_LO, s * p));
}

MGX_FM_HD float mgx_expf(float x) { return (float)mgx_exp_d((double)x); }

MGX_FM_HD float mgx_logf(float x) { return (float)mgx_log_d((double)x); }

MGX_FM_HD float mgx_expm1f(float x)
{
    const double xd = (double)x;
    if (xd != xd) return x;
    if (xd > -0.5 && xd < 0.5) {
        double p = MGX_EXPM1_C[17];
        for (int n = 16; n >= 0; n--) p = MGX_FMA(p, xd, MGX_EXPM1_C[n]);
        return (float)(p * xd);
    }
    return (float)(mgx_exp_d(xd) - 1.0);
}

/* x^y for x > 0 as used by the log interpolation (other x handled loosely). */
MGX_FM_HD float mgx_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    const double xd = (double)x;
    if (xd == 0.0) return (y > 0.0f) ? 0.0f : 1.0f / 0.0f;
    return (float)mgx_exp_d((double)y * mgx_log_d(xd));
}

/* erf with the interval coefficient table passed in (the kernels keep a copy in LDS) */
MGX_FM_HD float mgx_erff_tab(float x, const double *tab)
{
    double xd = (double)x;
    if (xd != xd) return x;
    const double sgn = (xd < 0.0) ? -1.0 : 1.0;
    const double a = xd * sgn;
    double r;
    if (a < 0.25) {
        const double a2 = a * a;
        double p = MGX_ERF0_C[11];
        for (int n = 10; n >= 0; n--) p = MGX_FMA(p, a2, MGX_ERF0_C[n]);
        r = a * p;
    } else if (a < 4.0) {
        const int j = (int)(a * 8.0); /* 2..31 */
        const double t = a - ((double)j * 0.125 + 0.0625);
        const double *c = &tab[(j - 2) * (MGX_ERF_DEG + 1)];
        double p = c[MGX_ERF_DEG];
        for (int n = MGX_ERF_DEG - 1; n >= 0; n--) p = MGX_FMA(p, t, c[n]);
        r = p;
    } else {
        r = 1.0;
    }
    return (float)(sgn * r);
}

MGX_FM_HD float mgx_erff(float x) { return mgx_erff_tab(x, MGX_ERF_TAB); }
