/*
 * The float transcendentals of the reference's C library, restated.
 *
 * Magritte's formal solver calls glibc 2.35's expf, expm1f, erff, logf and
 * powf (src/solver/solver.tpp:1109, 1120, 2450;
 * src/solver/solver_interp_helper.tpp:37-39, 67-68, 98;
 * src/model/thermodynamics/thermodynamics.tpp:13).  These are not correctly
 * rounded, and the solver takes the difference of two erff values in the line
 * wings, so a one-ulp difference in either moves an optical depth by up to
 * 1e-3 relative.  To reproduce the reference's numbers, the functions here
 * follow the published algorithms that glibc 2.35 compiles for x86-64:
 *
 *  - erff, expm1f: Sun's fdlibm single-precision routines (s_erff.c,
 *    s_expm1f.c, as kept in glibc's sysdeps/ieee754/flt-32): rational and
 *    polynomial approximations evaluated in float, with the float
 *    coefficients of fdlibm's double routines rounded to float.  erff's tail
 *    calls the library's internal expf, which on x86-64 is the baseline
 *    (non-FMA) build of the expf below.
 *  - expf: the table-driven method of Szabolcs Nagy (ARM optimized-routines,
 *    glibc 2.27 and later): x*32/ln2 = k + r, 2^(k/32) from a 32-entry table,
 *    2^(r/32) by a cubic in double, one rounding to float.  The exported
 *    expf of x86-64 glibc is the build with fused multiply-adds (chosen at run
 *    time on every CPU with FMA), the one erff calls is the build without.
 *    The table holds 2^(i/32) rounded to double, less i<<47 in its bits.
 *  - logf, powf: the same family.  z = x / 2^k in [0.699, 1.398] falls in one
 *    of 16 subintervals; log(x) = k ln2 + log(c) + log1p(z/c - 1) with 1/c
 *    from a table and log1p by a polynomial in double (degree 4 for logf; for
 *    powf, log2 by a degree-5 polynomial, then y*log2(x) into the expf
 *    method above in base 2).  The table's 1/c values are the published ones
 *    (1/c near 1/centre of the subinterval; subinterval 3 takes 1/centre
 *    rounded to double, and the one containing 1 takes c = 1); log(c) and
 *    log2(c) are those values' logarithms rounded to double (mpmath).
 *
 * Every function here is plain IEEE arithmetic in a fixed order: the file must
 * be compiled with -ffp-contract=off (fused operations are written as fma()).
 * It compiles as C (gcc, for the CPU restatement) and as HIP (host and
 * device); float operations are single precision on both (SSE on x86-64,
 * v_*_f32 on gfx950, with IEEE division and denormals).
 *
 * Validated against the C library of the machine it runs on
 * (oracle/fmath_exhaustive.c, tests/test_fmath.py): expf, expm1f and erff on
 * all 2^32 float arguments, logf on all 2^31 positive ones; powf on sampled
 * arguments of the log interpolation (bases over 10^+-6, exponents 1 - i/m).
 */
#pragma once

#include <math.h>
#include <stdint.h>

#include "mgx_fmath.h"

MGX_FM_HD uint32_t mgx_rf_asuint(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u;
}

MGX_FM_HD float mgx_rf_asfloat(uint32_t u)
{
    float x;
    __builtin_memcpy(&x, &u, sizeof x);
    return x;
}

/* 2^(i/32) rounded to double, minus i << 47 in its bit pattern (so that adding
 * k << 47 for k = 32q + i gives 2^(k/32) with the exponent carried). */
MGX_FM_CONST uint64_t MGX_REF_EXP2_TAB[32] = {
    0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL,
    0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL,
    0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL,
    0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL,
    0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL,
    0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL,
    0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL,
    0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL};

/*
 * Where the tables are read.  On the GPU a per-lane index into constant memory
 * is a global load in the middle of a dependent chain, so device code reads
 * copies in LDS: every kernel that evaluates expf, erff, logf or powf calls
 * mgx_ref_tables_init() (below) as its first statement, in all threads of the
 * block (it ends with a barrier).  Host code reads the constant tables.
 */
#if defined(__HIP_DEVICE_COMPILE__)
static __shared__ uint64_t mgx_ref_lds_exp2[32];
static __shared__ double mgx_ref_lds_log[16][3];
#define MGX_REF_EXP2_AT(i) mgx_ref_lds_exp2[(i)]
#define MGX_REF_LOG_AT(i, j) mgx_ref_lds_log[(i)][(j)]
#else
#define MGX_REF_EXP2_AT(i) MGX_REF_EXP2_TAB[(i)]
#define MGX_REF_LOG_AT(i, j) MGX_REF_LOG_TAB[(i)][(j)]
#endif
/* LDS the tables take in each workgroup of a kernel that uses them */
#define MGX_REF_LDS_TABLE_BYTES (32 * 8 + 16 * 3 * 8)

/* Cubic for 2^(r/32), r in [-1/2, 1/2]: coefficients of r^3, r^2, r (the
 * minimax fit of optimized-routines' exp2f, scaled by 1/32^3, 1/32^2, 1/32). */
#define MGX_REF_EXPF_C0 (0x1.c6af84b912394p-5 / 32 / 32 / 32)
#define MGX_REF_EXPF_C1 (0x1.ebfce50fac4f3p-3 / 32 / 32)
#define MGX_REF_EXPF_C2 (0x1.62e42ff0c52d6p-1 / 32)
#define MGX_REF_EXPF_INVLN2N (0x1.71547652b82fep+0 * 32)
#define MGX_REF_EXPF_SHIFT 0x1.8p+52

/* glibc expf; FUSED selects the build with fused multiply-adds (the exported
 * symbol on x86-64 CPUs with FMA) or the baseline build (erff's internal call).
 * The main path runs for every argument and the results of |x| >= 88, inf and
 * nan are selected afterwards (the same values as glibc's early returns; on
 * the GPU this avoids divergent branches around a path no lane takes). */
MGX_FM_HD float mgx_ref_expf_impl(float x, int fused)
{
    const double xd = (double)x;
    const uint32_t ux = mgx_rf_asuint(x);
    const uint32_t abstop = (ux >> 20) & 0x7ff;
    double kd, r, zz = 0.0;
    uint64_t ki;
    if (fused) {
        /* both uses of x*InvLn2N are sums, so the FMA build fuses them */
        kd = MGX_FMA(MGX_REF_EXPF_INVLN2N, xd, MGX_REF_EXPF_SHIFT);
        ki = mgx_fm_bits(kd);
        kd -= MGX_REF_EXPF_SHIFT;
        r = MGX_FMA(MGX_REF_EXPF_INVLN2N, xd, -kd);
    } else {
        zz = MGX_REF_EXPF_INVLN2N * xd;
        kd = zz + MGX_REF_EXPF_SHIFT;
        ki = mgx_fm_bits(kd);
        kd -= MGX_REF_EXPF_SHIFT;
        r = zz - kd;
    }
    uint64_t t = MGX_REF_EXP2_AT(ki % 32);
    t += ki << 47;
    const double s = mgx_fm_from_bits(t);
    double z, y;
    const double r2 = r * r;
    if (fused) {
        z = MGX_FMA(MGX_REF_EXPF_C0, r, MGX_REF_EXPF_C1);
        y = MGX_FMA(MGX_REF_EXPF_C2, r, 1.0);
        y = MGX_FMA(z, r2, y);
    } else {
        z = MGX_REF_EXPF_C0 * r + MGX_REF_EXPF_C1;
        y = MGX_REF_EXPF_C2 * r + 1.0;
        y = z * r2 + y;
    }
    y = y * s;
    float res = (float)y;
    /* |x| >= 88 or nan: -inf -> 0, inf/nan -> x + x, overflow -> inf, underflow -> 0 */
    if (abstop >= 0x42b) {
        res = (ux == 0xff800000u) ? 0.0f
            : (abstop >= 0x7f8)   ? x + x
            : (x > 0x1.62e42ep6f) ? mgx_rf_asfloat(0x7f800000u)
            : (x < -0x1.9fe368p6f) ? 0.0f
                                   : res;
    }
    return res;
}

MGX_FM_HD float mgx_ref_expf(float x) { return mgx_ref_expf_impl(x, 1); }

/* fdlibm s_expm1f.c */
MGX_FM_HD float mgx_ref_expm1f(float x)
{
    const float one = 1.0f, huge = 1.0e+30f, tiny = 1.0e-30f;
    const float o_threshold = 8.8721679688e+01f; /* 0x42b17180 */
    const float ln2_hi = 6.9313812256e-01f;      /* 0x3f317180 */
    const float ln2_lo = 9.0580006145e-06f;      /* 0x3717f7d1 */
    const float invln2 = 1.4426950216e+00f;      /* 0x3fb8aa3b */
    const float Q1 = -3.3333335072e-02f;         /* 0xbd088889 */
    const float Q2 = 1.5873016091e-03f;          /* 0x3ad00d01 */
    const float Q3 = -7.9365076090e-05f;         /* 0xb8a670cd */
    const float Q4 = 4.0082177293e-06f;          /* 0x36867e54 */
    const float Q5 = -2.0109921195e-07f;         /* 0xb457edbb */
    float y, hi, lo, c = 0.0f, t, e, hxs, hfx, r1;
    int32_t k;
    uint32_t hx = mgx_rf_asuint(x);
    const uint32_t xsb = hx & 0x80000000u;
    hx &= 0x7fffffffu;

    if (hx >= 0x4195b844u) {         /* |x| >= 27 ln2 */
        if (hx >= 0x42b17218u) {     /* |x| >= 88.72... */
            if (hx > 0x7f800000u) return x + x;
            if (hx == 0x7f800000u) return (xsb == 0) ? x : -1.0f;
            if (x > o_threshold) return huge * huge;
        }
        if (xsb != 0) return tiny - one; /* x < -27 ln2 */
    }

    if (hx > 0x3eb17218u) {          /* |x| > 0.5 ln2 */
        if (hx < 0x3f851592u) {      /* and |x| < 1.5 ln2 */
            if (xsb == 0) { hi = x - ln2_hi; lo = ln2_lo; k = 1; }
            else { hi = x + ln2_hi; lo = -ln2_lo; k = -1; }
        } else {
            k = (int32_t)(invln2 * x + ((xsb == 0) ? 0.5f : -0.5f));
            t = (float)k;
            hi = x - t * ln2_hi;
            lo = t * ln2_lo;
        }
        x = hi - lo;
        c = (hi - x) - lo;
    } else if (hx < 0x33000000u) {   /* |x| < 2^-25 */
        t = huge + x;
        return x - (t - (huge + x));
    } else {
        k = 0;
    }

    hfx = 0.5f * x;
    hxs = x * hfx;
    r1 = one + hxs * (Q1 + hxs * (Q2 + hxs * (Q3 + hxs * (Q4 + hxs * Q5))));
    t = 3.0f - r1 * hfx;
    e = hxs * ((r1 - t) / (6.0f - x * t));
    if (k == 0) return x - (x * e - hxs);
    e = (x * (e - c) - c);
    e -= hxs;
    if (k == -1) return 0.5f * (x - e) - 0.5f;
    if (k == 1) {
        if (x < -0.25f) return -2.0f * (e - (x + 0.5f));
        return one + 2.0f * (x - e);
    }
    if (k <= -2 || k > 56) {         /* exp(x) - 1 ~ exp(x) */
        y = one - (e - x);
        y = mgx_rf_asfloat(mgx_rf_asuint(y) + ((uint32_t)k << 23));
        return y - one;
    }
    if (k < 23) {
        t = mgx_rf_asfloat(0x3f800000u - (0x1000000u >> k)); /* 1 - 2^-k */
        y = t - (e - x);
        y = mgx_rf_asfloat(mgx_rf_asuint(y) + ((uint32_t)k << 23));
    } else {
        t = mgx_rf_asfloat((uint32_t)(0x7f - k) << 23);       /* 2^-k */
        y = x - (e + t);
        y += one;
        y = mgx_rf_asfloat(mgx_rf_asuint(y) + ((uint32_t)k << 23));
    }
    return y;
}

/* a / b rounded to float.  Host: the IEEE division.  Device: v_rcp_f32, one
 * Newton step and a residual correction (Markstein's sequence: q0 = a y,
 * r = a - b q0 exactly by FMA, q = q0 + r y) instead of the compiler's scaled
 * division, valid where neither the quotient nor the operands come near the
 * float overflow or underflow thresholds, as in erff's divisions.  That it
 * rounds every one of erff's quotients as the IEEE division does is checked on
 * all 2^32 arguments of erff (mgx_fmath_digest, tests/test_gpu_fmath.py). */
MGX_FM_HD float mgx_ref_fdiv(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    const float q = a * y;
    const float r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
#else
    return a / b;
#endif
}

/* fdlibm s_erff.c */
MGX_FM_HD float mgx_ref_erff(float x)
{
    const float tiny = 1e-30f, one = 1.0f;
    const float erx = 8.4506291151e-01f;
    const float efx = 1.2837916613e-01f;
    const float pp0 = 1.2837916613e-01f, pp1 = -3.2504209876e-01f, pp2 = -2.8481749818e-02f,
                pp3 = -5.7702702470e-03f, pp4 = -2.3763017452e-05f;
    const float qq1 = 3.9791721106e-01f, qq2 = 6.5022252500e-02f, qq3 = 5.0813062117e-03f,
                qq4 = 1.3249473704e-04f, qq5 = -3.9602282413e-06f;
    const float pa0 = -2.3621185683e-03f, pa1 = 4.1485610604e-01f, pa2 = -3.7220788002e-01f,
                pa3 = 3.1834661961e-01f, pa4 = -1.1089469492e-01f, pa5 = 3.5478305072e-02f,
                pa6 = -2.1663755178e-03f;
    const float qa1 = 1.0642088205e-01f, qa2 = 5.4039794207e-01f, qa3 = 7.1828655899e-02f,
                qa4 = 1.2617121637e-01f, qa5 = 1.3637083583e-02f, qa6 = 1.1984500103e-02f;
    const float ra0 = -9.8649440333e-03f, ra1 = -6.9385856390e-01f, ra2 = -1.0558626175e+01f,
                ra3 = -6.2375331879e+01f, ra4 = -1.6239666748e+02f, ra5 = -1.8460508728e+02f,
                ra6 = -8.1287437439e+01f, ra7 = -9.8143291473e+00f;
    const float sa1 = 1.9651271820e+01f, sa2 = 1.3765776062e+02f, sa3 = 4.3456588745e+02f,
                sa4 = 6.4538726807e+02f, sa5 = 4.2900814819e+02f, sa6 = 1.0863500214e+02f,
                sa7 = 6.5702495575e+00f, sa8 = -6.0424413532e-02f;
    const float rb0 = -9.8649431020e-03f, rb1 = -7.9928326607e-01f, rb2 = -1.7757955551e+01f,
                rb3 = -1.6063638306e+02f, rb4 = -6.3756646729e+02f, rb5 = -1.0250950928e+03f,
                rb6 = -4.8351919556e+02f;
    const float sb1 = 3.0338060379e+01f, sb2 = 3.2579251099e+02f, sb3 = 1.5367296143e+03f,
                sb4 = 3.1998581543e+03f, sb5 = 2.5530502930e+03f, sb6 = 4.7452853394e+02f,
                sb7 = -2.2440952301e+01f;
    float R, S, P, Q, s, y, z, r;
    const uint32_t hxu = mgx_rf_asuint(x);
    const int neg = (hxu >> 31) != 0;
    const uint32_t ix = hxu & 0x7fffffffu;
    if (ix >= 0x7f800000u) { /* erf(nan) = nan, erf(+-inf) = +-1 */
        const int i = neg ? 2 : 0;
        return (float)(1 - i) + one / x;
    }
    if (ix < 0x3f580000u) {          /* |x| < 0.84375 */
        if (ix < 0x31800000u) {      /* |x| < 2^-28 */
            if (ix < 0x04000000u) return 0.0625f * (16.0f * x + (16.0f * efx) * x);
            return x + efx * x;
        }
        z = x * x;
        r = pp0 + z * (pp1 + z * (pp2 + z * (pp3 + z * pp4)));
        s = one + z * (qq1 + z * (qq2 + z * (qq3 + z * (qq4 + z * qq5))));
        y = mgx_ref_fdiv(r, s);
        return x + x * y;
    }
    if (ix < 0x3fa00000u) {          /* 0.84375 <= |x| < 1.25 */
        s = fabsf(x) - one;
        P = pa0 + s * (pa1 + s * (pa2 + s * (pa3 + s * (pa4 + s * (pa5 + s * pa6)))));
        Q = one + s * (qa1 + s * (qa2 + s * (qa3 + s * (qa4 + s * (qa5 + s * qa6)))));
        const float pq = mgx_ref_fdiv(P, Q);
        if (!neg) return erx + pq;
        return -erx - pq;
    }
    if (ix >= 0x40c00000u) {         /* |x| >= 6 */
        if (!neg) return one - tiny;
        return tiny - one;
    }
    x = fabsf(x);
    s = mgx_ref_fdiv(one, x * x);
    if (ix < 0x4036db6eu) {          /* |x| < 1/0.35 */
        R = ra0 + s * (ra1 + s * (ra2 + s * (ra3 + s * (ra4 + s * (ra5 + s * (ra6 + s * ra7))))));
        S = one + s * (sa1 + s * (sa2 + s * (sa3 + s * (sa4 + s * (sa5 + s * (sa6 + s * (sa7 + s * sa8)))))));
    } else {
        R = rb0 + s * (rb1 + s * (rb2 + s * (rb3 + s * (rb4 + s * (rb5 + s * rb6)))));
        S = one + s * (sb1 + s * (sb2 + s * (sb3 + s * (sb4 + s * (sb5 + s * (sb6 + s * sb7))))));
    }
    z = mgx_rf_asfloat(mgx_rf_asuint(x) & 0xfffff000u);
    r = mgx_ref_expf_impl(-z * z - 0.5625f, 0) * mgx_ref_expf_impl((z - x) * (z + x) + mgx_ref_fdiv(R, S), 0);
    const float rx = mgx_ref_fdiv(r, x);
    if (!neg) return one - rx;
    return rx - one;
}

/* {1/c, log(c), log2(c)} for the 16 subintervals of [0x1.66p-1, 0x1.66p0). */
MGX_FM_CONST double MGX_REF_LOG_TAB[16][3] = {
    {0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2, -0x1.efec65b963019p-2},
    {0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2, -0x1.b0b6832d4fca4p-2},
    {0x1.49539f0f010b0p+0, -0x1.01eae7f513a67p-2, -0x1.7418b0a1fb77bp-2},
    {0x1.3c995a47babe7p+0, -0x1.b31d8575bce3bp-3, -0x1.39de8e1559f6ep-2},
    {0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3, -0x1.01d9bf3f2b631p-2},
    {0x1.25e227b0b8ea0p+0, -0x1.1aa2bc79c8100p-3, -0x1.97c1d1b3b7af0p-3},
    {0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4, -0x1.2f9e393af3c9fp-3},
    {0x1.12358f08ae5bap+0, -0x1.1973c5a611cccp-4, -0x1.960cbbf788d5cp-4},
    {0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5, -0x1.a6f9db6475fcep-5},
    {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
    {0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5, 0x1.338ca9f24f53dp-4},
    {0x1.ca4b31f026aa0p-1, 0x1.c5e53aa362eb4p-4, 0x1.476a9543891bap-3},
    {0x1.b2036576afce6p-1, 0x1.526e57720db08p-3, 0x1.e840b4ac4e4d2p-3},
    {0x1.9c2d163a1aa2dp-1, 0x1.bc2860d224770p-3, 0x1.40645f0c6651cp-2},
    {0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2, 0x1.88e9c2c1b9ff8p-2},
    {0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2, 0x1.ce0a44eb17bccp-2}};

#define MGX_REF_LN2 0x1.62e42fefa39efp-1

#if defined(__HIPCC__)
/* copies the tables into LDS (device code; see MGX_REF_EXP2_AT) */
__device__ static inline void mgx_ref_tables_init()
{
#if defined(__HIP_DEVICE_COMPILE__)
    for (uint32_t k = threadIdx.x; k < 32; k += blockDim.x) mgx_ref_lds_exp2[k] = MGX_REF_EXP2_TAB[k];
    for (uint32_t k = threadIdx.x; k < 48; k += blockDim.x)
        mgx_ref_lds_log[k / 3][k % 3] = MGX_REF_LOG_TAB[k / 3][k % 3];
    __syncthreads();
#endif
}
#endif

/* glibc logf (the FMA build) */
MGX_FM_HD float mgx_ref_logf(float x)
{
    uint32_t ix = mgx_rf_asuint(x);
    if (ix == 0x3f800000u) return 0.0f;
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
        /* x < 0x1p-126, inf or nan */
        if (ix * 2 == 0) return -mgx_rf_asfloat(0x7f800000u);
        if (ix == 0x7f800000u) return x;
        if ((ix & 0x80000000u) || ix * 2 >= 0xff000000u) return mgx_rf_asfloat(0x7fc00000u);
        ix = mgx_rf_asuint(x * 0x1p23f); /* subnormal: normalise */
        ix -= 23u << 23;
    }
    const uint32_t tmp = ix - 0x3f330000u;
    const int i = (int)((tmp >> 19) % 16);
    const int k = (int32_t)tmp >> 23;
    const uint32_t iz = ix - (tmp & 0xff800000u);
    const double invc = MGX_REF_LOG_AT(i, 0), logc = MGX_REF_LOG_AT(i, 1);
    const double z = (double)mgx_rf_asfloat(iz);
    const double r = MGX_FMA(z, invc, -1.0);
    const double y0 = MGX_FMA((double)k, MGX_REF_LN2, logc);
    const double r2 = r * r;
    double y = MGX_FMA(0x1.5575b0be00b6ap-2, r, -0x1.ffffef20a4123p-2);
    y = MGX_FMA(-0x1.00ea348b88334p-2, r2, y);
    y = MGX_FMA(y, r2, y0 + r);
    return (float)y;
}

/* log2 of the float with bits ix (normal, positive, or with a sign-cleared and
 * renormalised exponent), in double: glibc powf's log2_inline. */
MGX_FM_HD double mgx_ref_powf_log2(uint32_t ix)
{
    const uint32_t tmp = ix - 0x3f330000u;
    const int i = (int)((tmp >> 19) % 16);
    const uint32_t top = tmp & 0xff800000u;
    const uint32_t iz = ix - top;
    const int k = (int32_t)top >> 23;
    const double invc = MGX_REF_LOG_AT(i, 0), logc = MGX_REF_LOG_AT(i, 2);
    const double z = (double)mgx_rf_asfloat(iz);
    const double r = MGX_FMA(z, invc, -1.0);
    const double y0 = logc + (double)k;
    const double r2 = r * r;
    double y = MGX_FMA(0x1.27616c9496e0bp-2, r, -0x1.71969a075c67ap-2);
    const double p = MGX_FMA(0x1.ec70a6ca7baddp-2, r, -0x1.7154748bef6c8p-1);
    const double r4 = r2 * r2;
    double q = MGX_FMA(0x1.71547652ab82bp0, r, y0);
    q = MGX_FMA(p, r2, q);
    y = MGX_FMA(y, r4, q);
    return y;
}

/* 2^xd rounded to float, with the sign bias of a negative result: glibc
 * powf's exp2_inline (the expf method in base 2). */
MGX_FM_HD float mgx_ref_powf_exp2(double xd, uint64_t sign_bias)
{
    const double shift = 0x1.8p+52 / 32;
    double kd = xd + shift;
    const uint64_t ki = mgx_fm_bits(kd);
    kd -= shift;
    const double r = xd - kd;
    uint64_t t = MGX_REF_EXP2_AT(ki % 32);
    t += (ki + sign_bias) << 47;
    const double s = mgx_fm_from_bits(t);
    const double z = MGX_FMA(0x1.c6af84b912394p-5, r, 0x1.ebfce50fac4f3p-3);
    const double r2 = r * r;
    double y = MGX_FMA(0x1.62e42ff0c52d6p-1, r, 1.0);
    y = MGX_FMA(z, r2, y);
    y = y * s;
    return (float)y;
}

/* 0 if iy is not an integer, 1 if odd, 2 if even (as glibc powf's checkint) */
MGX_FM_HD int mgx_ref_powf_checkint(uint32_t iy)
{
    const int e = (int)((iy >> 23) & 0xff);
    if (e < 0x7f) return 0;
    if (e > 0x7f + 23) return 2;
    if (iy & ((1u << (0x7f + 23 - e)) - 1)) return 0;
    if (iy & (1u << (0x7f + 23 - e))) return 1;
    return 2;
}

MGX_FM_HD int mgx_ref_powf_zeroinfnan(uint32_t i) { return 2 * i - 1 >= 2u * 0x7f800000u - 1; }

/* glibc powf (the FMA build) */
MGX_FM_HD float mgx_ref_powf(float x, float y)
{
    uint64_t sign_bias = 0;
    uint32_t ix = mgx_rf_asuint(x);
    const uint32_t iy = mgx_rf_asuint(y);
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u || mgx_ref_powf_zeroinfnan(iy)) {
        /* x < 0x1p-126 or inf or nan, or y is 0, inf or nan */
        if (mgx_ref_powf_zeroinfnan(iy)) {
            if (2 * iy == 0) return 1.0f;
            if (ix == 0x3f800000u) return 1.0f;
            if (2 * ix > 2u * 0x7f800000u || 2 * iy > 2u * 0x7f800000u) return x + y;
            if (2 * ix == 2u * 0x3f800000u) return 1.0f;
            if ((2 * ix < 2u * 0x3f800000u) == !(iy & 0x80000000u)) return 0.0f;
            return y * y;
        }
        if (mgx_ref_powf_zeroinfnan(ix)) {
            float x2 = x * x;
            if ((ix & 0x80000000u) && mgx_ref_powf_checkint(iy) == 1) x2 = -x2;
            return (iy & 0x80000000u) ? 1.0f / x2 : x2;
        }
        /* x and y are non-zero finite */
        if (ix & 0x80000000u) {
            const int yint = mgx_ref_powf_checkint(iy);
            if (yint == 0) return mgx_rf_asfloat(0x7fc00000u);
            if (yint == 1) sign_bias = 1u << 16;
            ix &= 0x7fffffffu;
        }
        if (ix < 0x00800000u) { /* subnormal x: normalise */
            ix = mgx_rf_asuint(mgx_rf_asfloat(ix) * 0x1p23f);
            ix &= 0x7fffffffu;
            ix -= 23u << 23;
        }
    }
    const double logx = mgx_ref_powf_log2(ix);
    const double ylogx = (double)y * logx;
    if (((mgx_fm_bits(ylogx) >> 47) & 0xffff) >= (mgx_fm_bits(126.0) >> 47)) {
        /* |y log2(x)| >= 126 */
        if (ylogx > 0x1.fffffffd1d571p+6)
            return (sign_bias ? -0x1p97f : 0x1p97f) * 0x1p97f;
        if (ylogx <= -150.0)
            return (sign_bias ? -0x1p-95f : 0x1p-95f) * 0x1p-95f;
    }
    return mgx_ref_powf_exp2(ylogx, sign_bias);
}

/* The restated functions by number, for the exhaustive comparisons (powf takes
 * the exponent y; NaN results are canonicalised so that payloads do not count). */
enum { MGX_FN_EXPF = 0, MGX_FN_EXPM1F = 1, MGX_FN_ERFF = 2, MGX_FN_LOGF = 3, MGX_FN_POWF = 4 };

MGX_FM_HD uint32_t mgx_ref_eval_bits(int fn, uint32_t xbits, float y)
{
    const float x = mgx_rf_asfloat(xbits);
    float r;
    switch (fn) {
    case MGX_FN_EXPF: r = mgx_ref_expf(x); break;
    case MGX_FN_EXPM1F: r = mgx_ref_expm1f(x); break;
    case MGX_FN_ERFF: r = mgx_ref_erff(x); break;
    case MGX_FN_LOGF: r = mgx_ref_logf(x); break;
    default: r = mgx_ref_powf(x, y); break;
    }
    return (r != r) ? 0x7fc00000u : mgx_rf_asuint(r);
}

/* Order-independent digest of one (argument, result) pair: the 64-bit
 * finaliser of SplitMix64 applied to (argument bits << 32 | result bits); a
 * chunk's digest is the wrapping sum over its arguments. */
MGX_FM_HD uint64_t mgx_fm_digest(uint32_t xbits, uint32_t rbits)
{
    uint64_t z = ((uint64_t)xbits << 32) | rbits;
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
