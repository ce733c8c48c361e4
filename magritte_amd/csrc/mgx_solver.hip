/*
 * gfx950 kernels of the plain-method Feautrier formal solver, and the C-ABI
 * entry points that launch them.
 *
 * Work decomposition (one 64-lane wavefront per origin point o):
 *   for each ray pair rr (sequential, so J and Lambda accumulate in the
 *   reference's order without atomics):
 *     1. walk the neighbour lists from o along ray rr and along its antipod.
 *        The 64 lanes evaluate the neighbours of the current point in
 *        parallel and a lexicographic (d2, list index) arg-min selects the
 *        next point; that is the first-wins strict "<" of
 *        Geometry::get_next_general_geometry (geometry.tpp:14-44).  The walk
 *        (point, dZ, Doppler shift per step) is staged in LDS.
 *     2. lanes take the quadrature frequencies of up to two lines (32 lanes
 *        per line).  Each lane expands its line's interpolation points on the
 *        fly (set_data_for_line, solver.tpp:1057-1097) and solves the
 *        second-order Feautrier system of its frequency with a two-sided
 *        elimination: backward from the antipod-side boundary to the origin,
 *        forward from the rr-side boundary to the origin, then joined at the
 *        origin.  This gives u(o) and the Lambda diagonal with O(1) state per
 *        lane; it is algebraically the reference's forward elimination plus
 *        back substitution (solver.tpp:2194-2318), with the same source,
 *        optical-depth and boundary terms.
 *     3. u(rr,o,f) is stored coalesced, J(o,f) and Lambda(o,l) updated by the
 *        owning wave.
 *
 * Numerics: FP64 throughout, no FMA contraction (-ffp-contract=off), and the
 * reference's float transcendentals (the C library's expf, expm1f, logf,
 * powf and erff, restated in mgx_fmath_ref.h), so the traced lists and the
 * source / optical-depth terms are those the reference computes.
 */
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "mgx.h"
#include "mgx_constants.h"
#include "mgx_error.h"
#include "mgx_fmath.h"
#include "mgx_fmath_ref.h"

#define MGX_WAVE 64
/* minimum waves per SIMD requested for the solve kernel (register budget) */
#ifndef MGX_SOLVE_WAVES_PER_SIMD
#define MGX_SOLVE_WAVES_PER_SIMD 3
#endif
/* the same for the adjoint kernels */
#ifndef MGX_ADJ_WAVES_PER_SIMD
#define MGX_ADJ_WAVES_PER_SIMD 2
#endif

namespace {

/* kernel status bits: walk buffer full (relaunch with a larger one), no valid next point,
   output list full, adjoint tape full, Lambda list full, interpolation count above 0xffff */
enum { ERR_CAP = 1, ERR_NO_NEXT = 2, ERR_OUT = 4, ERR_CAPE = 8, ERR_LAMCAP = 16, ERR_INTERP = 32 };

/* global-address-space load (the pointers live in a struct, so the compiler would
   otherwise emit flat loads, which also count against the LDS wait counter and so
   serialise the gathers with LDS traffic).  Records are read as 16-, 8- or 4-byte
   words through address-space-1 pointers: a whole-struct copy through such a
   pointer still becomes a flat load. */
#define MGX_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T gld(const T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    static_assert(sizeof(T) % 4 == 0, "gld reads whole 32-bit words");
    if constexpr (sizeof(T) % 16 == 0) {
        typedef unsigned int w4 __attribute__((ext_vector_type(4)));
        const MGX_GLOBAL w4 *q = (const MGX_GLOBAL w4 *)p;
        w4 tmp[sizeof(T) / 16];
#pragma unroll
        for (unsigned i = 0; i < sizeof(T) / 16; i++) tmp[i] = q[i];
        T r;
        __builtin_memcpy(&r, tmp, sizeof(T));
        return r;
    } else if constexpr (sizeof(T) % 8 == 0) {
        const MGX_GLOBAL unsigned long long *q = (const MGX_GLOBAL unsigned long long *)p;
        unsigned long long tmp[sizeof(T) / 8];
#pragma unroll
        for (unsigned i = 0; i < sizeof(T) / 8; i++) tmp[i] = q[i];
        T r;
        __builtin_memcpy(&r, tmp, sizeof(T));
        return r;
    } else {
        const MGX_GLOBAL unsigned *q = (const MGX_GLOBAL unsigned *)p;
        unsigned tmp[sizeof(T) / 4];
#pragma unroll
        for (unsigned i = 0; i < sizeof(T) / 4; i++) tmp[i] = q[i];
        T r;
        __builtin_memcpy(&r, tmp, sizeof(T));
        return r;
    }
#else
    return *p;
#endif
}

/* Per-point record read by the walk and the segments: one 64-byte line. */
struct PointRec {
    double x, y, z;    /* position */
    double vx, vy, vz; /* velocity (fraction of c) */
    double dshift_max; /* Model::dshift_max */
    double ub;         /* sqrt(2kT/(amu c^2) * max_inverse_mass + vturb2) */
};

/* The leading 32 bytes of a PointRec (the walk reads only the position), as two
   16-byte words */
struct PointPos {
    double x, y, z, vx;
};
static_assert(sizeof(PointRec) % 16 == 0, "PointRec is read in 16-byte words");

/* Per-point integers: neighbour count and offset, boundary index (P = interior). */
struct PointIdx {
    uint32_t n_nbs, cum, p2b, pad;
};

/* Device copy of the model view plus the per-call derived arrays. */
struct KParams {
    mgx_model m;
    mgx_options op;
    const double4 *lpack; /* (P, L): inverse_width, emissivity, opacity, emissivity/opacity */
    const PointRec *pts;  /* (P) */
    const PointIdx *pidx; /* (P) */
    double log_max_interp; /* (double)logf((float)max_interpolation_diff) */
    uint32_t cap;         /* LDS walk capacity (steps per half ray) */
    uint32_t nslots;      /* number of 32-lane line slots = L * ceil(Q/32) */
    uint32_t slots_per_line;
    uint32_t j_lds;       /* forward solve: 1 = J accumulated in LDS, 0 = in the output row */
    uint32_t *status;     /* [0] error bits, [1] walk capacity needed, [2] adjoint segments needed */
};

__device__ inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ inline const double *ray_dir(const mgx_model &m, uint32_t o, uint32_t r)
{
    return m.adaptive_directions ? &m.ray_direction[3 * ((size_t)o * m.nrays + r)]
                                 : &m.ray_direction[3 * (size_t)r];
}

__device__ inline double ray_weight(const mgx_model &m, uint32_t o, uint32_t r)
{
    return m.adaptive_directions ? gld(&m.ray_weight[(size_t)o * m.nrays + r]) : gld(&m.ray_weight[r]);
}

/* ------------------------------------------------------------------------- */
/* Doppler shift (Geometry::get_shift<CoMoving>, geometry.tpp:204-207, 379-396) */

__device__ inline double shift_comoving(const KParams &kp, const double *origin,
                                        const double *ovel, const double *dir, uint32_t c,
                                        double Z)
{
    const PointRec pc = gld(&kp.pts[c]);
    if (kp.m.spherical_symmetry) {
        if (dot3(origin, origin) == 0.0) return 1.0;
        if (pc.x == 0.0) return 1.0 - dot3(ovel, dir);
        const double Rcos_plus_Z = dot3(origin, dir) + Z;
        return 1.0 + (pc.vx * Rcos_plus_Z / pc.x - dot3(ovel, dir));
    }
    double dv[3];
    dv[0] = pc.vx - ovel[0];
    dv[1] = pc.vy - ovel[1];
    dv[2] = pc.vz - ovel[2];
    return 1.0 + dot3(dv, dir);
}

/* Geometry::get_shift<Rest, false>(o, r, c, Z) (geometry.tpp:213-221, 405-415,
   429-465): the shift the old imager walks with, reversed (2 - shift) on the
   second half of the ray set */
__device__ inline double shift_rest_ray(const KParams &kp, const double *origin, const double *dir,
                                        uint32_t c, double Z, bool reverse)
{
    const PointRec pc = gld(&kp.pts[c]);
    double sh;
    if (kp.m.spherical_symmetry) {
        if (pc.x == 0.0) {
            sh = 1.0;
        } else {
            const double Rcos_plus_Z = dot3(origin, dir) + Z;
            sh = 1.0 + pc.vx * Rcos_plus_Z / pc.x;
        }
    } else {
        const double v[3] = {pc.vx, pc.vy, pc.vz};
        sh = 1.0 + dot3(v, dir);
    }
    return reverse ? 2.0 - sh : sh;
}

/* ------------------------------------------------------------------------- */
/* next point on the ray, wave-cooperative                                    */

/* Cross-lane moves within a row of 16 lanes through DPP (every lane of the row
   active): the value of the lane that the pattern CTRL selects.  The patterns
   used here: 0xb1 quad_perm [1, 0, 3, 2] (lane ^ 1), 0x4e quad_perm [2, 3, 0, 1]
   (lane ^ 2), 0x141 row_half_mirror (lane i of 8 takes lane 7 - i), 0x140
   row_mirror (lane i of 16 takes lane 15 - i).  After the first two steps every
   lane of a quad holds the quad's result, so the two mirrors pair quads and then
   half-rows like lane ^ 4 and lane ^ 8 would. */
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_get(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_get(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = dpp_get<CTRL>((uint32_t)b), hi = dpp_get<CTRL>((uint32_t)(b >> 32));
    return __builtin_bit_cast(double, (hi << 32) | lo);
}

__device__ __forceinline__ double read_lane(double v, uint32_t l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)l);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)l);
    return __builtin_bit_cast(double, (hi << 32) | lo);
}

/* lexicographic arg-min step: smaller d2 wins, ties go to the smaller list index */
__device__ __forceinline__ void argmin_take(double &d2, uint32_t &idx, double &zn, double od2,
                                            uint32_t oidx, double ozn)
{
    const bool take = (od2 < d2) || (od2 == d2 && oidx < idx);
    if (take) {
        d2 = od2;
        idx = oidx;
        zn = ozn;
    }
}

template <int CTRL>
__device__ __forceinline__ void argmin_dpp(double &d2, uint32_t &idx, double &zn)
{
    argmin_take(d2, idx, zn, dpp_get<CTRL>(d2), dpp_get<CTRL>(idx), dpp_get<CTRL>(zn));
}

/* Lexicographic arg-min over the wave, in every lane; only lanes below n_cand
   hold candidates (the others hold d2 = inf).  The arg-min is unique (the list
   indices differ), so the order of the pairings does not change it. */
__device__ inline void wave_argmin(double &d2, uint32_t &idx, double &zn, uint32_t n_cand)
{
    argmin_dpp<0xb1>(d2, idx, zn);
    argmin_dpp<0x4e>(d2, idx, zn);
    argmin_dpp<0x141>(d2, idx, zn);
    argmin_dpp<0x140>(d2, idx, zn);
    /* every lane of a row holds the row's arg-min: combine the rows with candidates */
    const uint32_t rows = n_cand > 48 ? 4 : n_cand > 32 ? 3 : n_cand > 16 ? 2 : 1;
    double b_d2 = read_lane(d2, 0), b_zn = read_lane(zn, 0);
    uint32_t b_idx = (uint32_t)__builtin_amdgcn_readlane((int)idx, 0);
    for (uint32_t r = 1; r < rows; r++)
        argmin_take(b_d2, b_idx, b_zn, read_lane(d2, 16 * r),
                    (uint32_t)__builtin_amdgcn_readlane((int)idx, (int)(16 * r)), read_lane(zn, 16 * r));
    d2 = b_d2;
    idx = b_idx;
    zn = b_zn;
}

/* get_next_general_geometry<Defaulttracer> (geometry.tpp:14-44), all lanes return the same;
   pi is the current point's neighbour-list record */
__device__ inline uint32_t next_general(const KParams &kp, const double *origin,
                                        const double *dir, const PointIdx &pi, double &Z, double &dZ)
{
    const uint32_t lane = threadIdx.x & (MGX_WAVE - 1);
    double best_d2 = __builtin_inf();
    uint32_t best_i = 0xffffffffu, best_n = 0;
    double best_zn = 0.0;
    for (uint32_t i = lane; i < pi.n_nbs; i += MGX_WAVE) {
        const uint32_t n = gld(&kp.m.neighbors[pi.cum + i]);
        const PointPos pn = gld(reinterpret_cast<const PointPos *>(&kp.pts[n]));
        double R[3];
        R[0] = pn.x - origin[0];
        R[1] = pn.y - origin[1];
        R[2] = pn.z - origin[2];
        const double Z_new = dot3(R, dir);
        if (Z_new > Z) {
            const double d2 = dot3(R, R) - Z_new * Z_new;
            if (d2 < DBL_MAX && d2 < best_d2) {
                best_d2 = d2;
                best_i = i;
                best_zn = Z_new;
                best_n = n;
            }
        }
    }
    wave_argmin(best_d2, best_i, best_zn, pi.n_nbs);
    if (best_i == 0xffffffffu) return kp.m.npoints;
    dZ = best_zn - Z;
    Z += dZ;
    /* list entry i was examined by lane i % 64, whose own best it is: its point
       index comes from that lane instead of a second read of the list */
    return (uint32_t)__builtin_amdgcn_readlane((int)best_n, (int)(best_i & (MGX_WAVE - 1)));
}

__device__ inline double sqnorm_point(const KParams &kp, uint32_t p)
{
    const PointRec r = gld(&kp.pts[p]);
    return r.x * r.x + r.y * r.y + r.z * r.z;
}

/* get_next_spherical_symmetry (geometry.tpp:118-159), evaluated redundantly by every lane */
__device__ inline uint32_t next_spherical(const KParams &kp, const double *origin,
                                          const double *dir, uint32_t c, double &Z, double &dZ)
{
    uint32_t next;
    const double Rcos = dot3(origin, dir);
    const double R2 = dot3(origin, origin);
    const double Rsin2 = R2 - Rcos * Rcos;
    const double Rcos_plus_Z = Rcos + Z;
    if (Z < -Rcos) {
        if (c <= 0) return kp.m.npoints;
        if (sqnorm_point(kp, c - 1) >= Rsin2) {
            next = c - 1;
            dZ = -sqrt(sqnorm_point(kp, next) - Rsin2) - Rcos_plus_Z;
        } else {
            next = c;
            dZ = -2.0 * Rcos_plus_Z;
        }
    } else {
        if (c >= kp.m.npoints - 1) return kp.m.npoints;
        next = c + 1;
        dZ = +sqrt(sqnorm_point(kp, next) - Rsin2) - Rcos_plus_Z;
    }
    Z += dZ;
    return next;
}

__device__ inline bool on_boundary(const KParams &kp, uint32_t p)
{
    return gld(&kp.pidx[p]).p2b != kp.m.npoints;
}

/*
 * Walk one half ray from o along ray r (trace_ray_for_line's geometry,
 * solver.tpp:774-800) into LDS: step s goes from crt_s (o for s = 0) to
 * s_nxt[s] over distance s_dZ[s]; s_shift[s] is the shift of s_nxt[s].
 * Returns 0 or an ERR_* code; nsteps and shift0 (shift of o) are set.
 * REST selects the old imager's trace_ray<Rest, false> shifts (:742-767).
 */
template <bool REST = false>
__device__ int walk_half(const KParams &kp, uint32_t o, uint32_t r, uint32_t *s_nxt, double *s_dZ,
                         double *s_shift, uint32_t &nsteps, double &shift0)
{
    const mgx_model &m = kp.m;
    const uint32_t lane = threadIdx.x & (MGX_WAVE - 1);
    const PointRec po = gld(&kp.pts[o]);
    const double origin[3] = {po.x, po.y, po.z};
    const double ovel[3] = {po.vx, po.vy, po.vz};
    const double *dp = ray_dir(m, o, r);
    const double dir[3] = {gld(&dp[0]), gld(&dp[1]), gld(&dp[2])};
    double Z = 0.0, dZ = 0.0;
    uint32_t crt = o;
    nsteps = 0;
    const bool reverse = r >= m.hnrays;
    shift0 = REST ? shift_rest_ray(kp, origin, dir, o, 0.0, reverse)
                  : shift_comoving(kp, origin, ovel, dir, o, 0.0);
    PointIdx pc = gld(&kp.pidx[o]);
    for (;;) {
        const uint32_t nxt = m.spherical_symmetry ? next_spherical(kp, origin, dir, crt, Z, dZ)
                                                  : next_general(kp, origin, dir, pc, Z, dZ);
        if (nxt >= m.npoints) {
            if (nsteps == 0) break; /* nothing in this direction */
            return ERR_NO_NEXT;
        }
        if (nsteps >= kp.cap) return ERR_CAP;
        /* the next point's record serves the boundary test and the next step */
        const PointIdx pn = gld(&kp.pidx[nxt]);
        const double sh = REST ? shift_rest_ray(kp, origin, dir, nxt, Z, reverse)
                               : shift_comoving(kp, origin, ovel, dir, nxt, Z);
        if (lane == 0) {
            s_nxt[nsteps] = nxt;
            s_dZ[nsteps] = dZ;
            s_shift[nsteps] = sh;
        }
        nsteps++;
        if (pn.p2b != m.npoints) break; /* on the boundary */
        crt = nxt;
        pc = pn;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* per-line interpolation count (InterpHelper::get_n_interp_for_line, :55-75)  */

__device__ inline uint32_t n_interp_for_line(const KParams &kp, uint32_t l, uint32_t c, uint32_t n)
{
    const uint32_t L = kp.m.nlines;
    const double curr_source = gld(&kp.m.interp_crit[(size_t)c * L + l]);
    const double next_source = gld(&kp.m.interp_crit[(size_t)n * L + l]);
    const double source_diff = (double)fabsf(mgx_ref_logf((float)(next_source / curr_source)));
    const uint32_t k = (uint32_t)ceil(source_diff / kp.log_max_interp);
    return (k > 1) ? k : 1;
}

/* float transcendentals: in FP64 mode the reference's own (the C library's
   expf and powf, restated in mgx_fmath_ref.h), in mixed mode the hardware
   single-precision functions */
template <int PREC>
__device__ inline float f_exp(float x)
{
    if (PREC == MGX_PREC_FP64) return mgx_ref_expf(x);
    return __expf(x);
}

template <int PREC>
__device__ inline float f_pow(float x, float y)
{
    if (PREC == MGX_PREC_FP64) return mgx_ref_powf(x, y);
    return __powf(x, y);
}

/* erf is the reference's erff in both modes: the reference takes the
   difference of two float erf values (compute_dtau_single_line), which cancels
   in the line wings, so a one-ulp change of either value moves u by up to
   ~1e-3 relative there */
template <int PREC>
__device__ inline float f_erf(float x, const double *)
{
    return mgx_ref_erff(x);
}

__device__ inline double interp_linear(double a, double b, double t)
{
    if (t == 0.0) return a;
    if (t == 1.0) return b;
    return (b - a) * t + a;
}

/* the powf path of interpolate_log, taken only by interpolated entries: out of line,
   so that its registers do not count against the solve's loop (with the reference's
   double-internal powf inline, the FP64 forward kernel spilled 40 VGPRs) */
template <int PREC>
__device__ __attribute__((noinline)) double interp_log_pow(double a, double b, double t)
{
    return a * (double)f_pow<PREC>((float)(b / a), (float)t);
}

template <int PREC>
__device__ inline double interp_log(double a, double b, double t)
{
    if (t == 0.0) return a;
    if (t == 1.0) return b;
    return interp_log_pow<PREC>(a, b, t);
}

/* ------------------------------------------------------------------------- */
/* emissivity, opacity and optical depth (solver.tpp:1106-2455)               */

/* 1/x to within one unit in the last place for finite normal x (v_rcp_f64 and
   two Newton steps); used only where the result does not feed a float
   conversion, so it cannot change which float argument a transcendental sees */
__device__ inline double rcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

template <int PREC>
__device__ inline double gaussian(double inverse_width, double diff)
{
    const double sqrt_exp = inverse_width * diff;
    return inverse_width * MGX_INVERSE_SQRT_PI * (double)f_exp<PREC>((float)(-sqrt_exp * sqrt_exp));
}

__device__ inline double planck(double temp, double freq)
{
    return MGX_TWO_HH_OVER_CC_SQUARED * (freq * freq * freq) /
           (double)mgx_ref_expm1f((float)(MGX_HH_OVER_KB * freq / temp));
}

__device__ inline double boundary_intensity(const KParams &kp, uint32_t p, double freq)
{
    const mgx_model &m = kp.m;
    const uint32_t b = gld(&kp.pidx[p]).p2b;
    const int32_t bc = (b < m.nboundary) ? gld(&m.boundary_condition[b]) : MGX_BC_CMB;
    if (bc == MGX_BC_ZERO) return 0.0;
    if (bc == MGX_BC_THERMAL) return planck(gld(&m.boundary_temperature[b]), freq);
    return planck(MGX_T_CMB, freq);
}

__device__ inline uint32_t lower_bound_d(const double *a, uint32_t n, double v)
{
    uint32_t lo = 0, len = n;
    while (len > 0) {
        const uint32_t h = len / 2;
        if (gld(&a[lo + h]) < v) {
            lo = lo + h + 1;
            len = len - h - 1;
        } else {
            len = h;
        }
    }
    return lo;
}

__device__ inline uint32_t upper_bound_d(const double *a, uint32_t n, double v)
{
    uint32_t lo = 0, len = n;
    while (len > 0) {
        const uint32_t h = len / 2;
        if (!(v < gld(&a[lo + h]))) {
            lo = lo + h + 1;
            len = len - h - 1;
        } else {
            len = h;
        }
    }
    return lo;
}

/*
 * The lane's own line in the sorted line list: its position g and the
 * neighbouring frequencies.  A CloseLines window [left, right] almost always
 * starts and ends at g; checking that against a[g-1], a[g], a[g+1] gives the
 * same indices as std::lower_bound / std::upper_bound without a search.
 */
struct OwnLine {
    uint32_t g;
    uint32_t lidx;   /* sorted_line_map[g]: the line at position g */
    double a_lo, a_g, a_hi;
    double f_own;    /* line[lidx] */
};

/* Whether the line-integrated segments take the own line's frequency from OwnLine
   (a register) or read line[lidx] (a load that needs no other load first).  The
   adjoint unit, at 256 VGPRs, keeps it in a register; the solve unit, at 168,
   reads it: measured on the 32^3 wind, the register cost the forward 4-6% in
   spills, the load cost the adjoint 8%. */
#ifdef MGX_ADJOINT_TU
#define MGX_OWN_FREQ_IN_REGS 1
#else
#define MGX_OWN_FREQ_IN_REGS 0
#endif

__device__ inline OwnLine own_line(const mgx_model &m, uint32_t l)
{
    OwnLine r;
    uint32_t g = 0;
    for (uint32_t s = 0; s < m.nlines; s++)
        if (gld(&m.sorted_line_map[s]) == l) g = s;
    r.g = g;
    r.lidx = gld(&m.sorted_line_map[g]);
    r.f_own = gld(&m.line[r.lidx]);
    r.a_g = gld(&m.sorted_line[g]);
    r.a_lo = (g > 0) ? gld(&m.sorted_line[g - 1]) : -__builtin_inf();
    r.a_hi = (g + 1 < m.nlines) ? gld(&m.sorted_line[g + 1]) : __builtin_inf();
    return r;
}

__device__ inline void line_window(const mgx_model &m, const OwnLine &ol, double left, double right,
                                   uint32_t &lo, uint32_t &hi)
{
    lo = (!(ol.a_g < left) && ol.a_lo < left) ? ol.g : lower_bound_d(m.sorted_line, m.nlines, left);
    hi = (!(right < ol.a_g) && right < ol.a_hi) ? ol.g + 1
                                                 : upper_bound_d(m.sorted_line, m.nlines, right);
}

__device__ inline double4 lp(const KParams &kp, uint32_t p, uint32_t l)
{
    return gld(&kp.lpack[(size_t)p * kp.m.nlines + l]);
}

/* get_eta_and_chi_interpolated<approx> (solver.tpp:1177-1202, 1260-1302, 1345-1351) */
template <int APPROX, int PREC>
__device__ void eta_chi_interp(const KParams &kp, const OwnLine &ol, uint32_t p1, uint32_t p2,
                               double t, uint32_t ll, double freq, double ub1, double &eta,
                               double &chi)
{
    const mgx_model &m = kp.m;
    if (APPROX == MGX_APPROX_ONE_LINE) {
        const double4 a = lp(kp, p1, ll), b = lp(kp, p2, ll);
        const double diff = freq - m.line[ll];
        const double prof = gaussian<PREC>(interp_linear(a.x, b.x, t), diff);
        eta = prof * interp_log<PREC>(a.y, b.y, t);
        chi = prof * interp_log<PREC>(a.z, b.z, t) + kp.op.min_opacity;
        return;
    }
    eta = 0.0;
    chi = kp.op.min_opacity;
    uint32_t lo = 0, hi = m.nlines;
    if (APPROX == MGX_APPROX_CLOSE_LINES) {
        const double W = kp.op.max_distance_opacity_contribution * (freq * ub1);
        line_window(m, ol, freq - W, freq + W, lo, hi);
    }
    for (uint32_t s = lo; s < hi; s++) {
        uint32_t l;
        double diff;
        if (APPROX == MGX_APPROX_CLOSE_LINES) {
            if (s == ol.g) {
                l = ol.lidx;
                diff = freq - ol.a_g;
            } else {
                l = gld(&m.sorted_line_map[s]);
                diff = freq - gld(&m.sorted_line[s]);
            }
        } else {
            l = s;
            diff = freq - m.line[l];
        }
        const double4 a = lp(kp, p1, l), b = lp(kp, p2, l);
        const double prof = gaussian<PREC>(interp_linear(a.x, b.x, t), diff);
        eta += prof * interp_log<PREC>(a.y, b.y, t);
        chi += prof * interp_log<PREC>(a.z, b.z, t);
    }
}

/* compute_dtau_single_line (solver.tpp:2374-2455).  Everything up to the
   float conversions follows the reference's operation order. */
template <int PREC>
__device__ double dtau_single_line(const KParams &kp, const double *erf_tab, double4 c, double4 cI,
                                   double4 n, double4 nI, double linefreq, double curr_freq,
                                   double next_freq, double curr_interp, double next_interp,
                                   double dz)
{
    const double iw_c = interp_linear(c.x, cI.x, curr_interp);
    const double iw_n = interp_linear(n.x, nI.x, next_interp);
    const double avg_iw = (iw_c + iw_n) / 2.0;
    const double op_c = interp_log<PREC>(c.z, cI.z, curr_interp);
    const double op_n = interp_log<PREC>(n.z, nI.z, next_interp);
    if (curr_freq == next_freq) {
        const double diff = curr_freq - linefreq;
        const double prof = gaussian<PREC>(avg_iw, diff);
        const double avg_op = (op_c + op_n) / 2.0;
        return dz * (prof * avg_op + kp.op.min_opacity);
    }
    const double next_pos = (linefreq - next_freq) * avg_iw;
    const double curr_pos = (linefreq - curr_freq) * avg_iw;
    const double diff_pos = next_pos - curr_pos;
    const double avg_op = (op_n + op_c) / 2.0;
    const float erf_diff = f_erf<PREC>((float)next_pos, erf_tab) - f_erf<PREC>((float)curr_pos, erf_tab);
    const double erfterm = avg_op * (double)erf_diff * (0.5 * rcp(diff_pos));
    return dz * fma(avg_iw, erfterm, kp.op.min_opacity);
}

/* compute_S_dtau_line_integrated<approx> (solver.tpp:1879-2111) */
template <int APPROX, int PREC>
__device__ void S_dtau_line_integrated(const KParams &kp, const double *erf_tab, const OwnLine &ol,
                                       uint32_t cp, uint32_t np, uint32_t cpi, uint32_t npi,
                                       uint32_t lineidx, double currfreq, double nextfreq,
                                       double curr_interp, double next_interp, double dZ,
                                       double ub_c, double ub_n, double &dtau, double &Scurr,
                                       double &Snext)
{
    const mgx_model &m = kp.m;
    if (APPROX == MGX_APPROX_ONE_LINE) {
        const double4 c = lp(kp, cp, lineidx), cI = lp(kp, cpi, lineidx);
        const double4 n = lp(kp, np, lineidx), nI = lp(kp, npi, lineidx);
        dtau = dtau_single_line<PREC>(kp, erf_tab, c, cI, n, nI, m.line[lineidx], currfreq,
                                      nextfreq, curr_interp, next_interp, dZ);
        Scurr = interp_log<PREC>(c.w, cI.w, curr_interp);
        Snext = interp_log<PREC>(n.w, nI.w, next_interp);
        return;
    }
    double sum_dtau = 0.0, sum_Sc = 0.0, sum_Sn = 0.0;
    uint32_t lo = 0, hi = m.nlines;
    if (APPROX == MGX_APPROX_CLOSE_LINES) {
        double left_freq, right_freq;
        if (currfreq < nextfreq) {
            left_freq = currfreq;
            right_freq = nextfreq;
        } else {
            right_freq = currfreq;
            left_freq = nextfreq;
        }
        const double cb = kp.op.max_distance_opacity_contribution * (right_freq * ub_c);
        const double nb = kp.op.max_distance_opacity_contribution * (right_freq * ub_n);
        const double W = (cb < nb) ? nb : cb;
        line_window(m, ol, left_freq - W, right_freq + W, lo, hi);
    }
    for (uint32_t s = lo; s < hi; s++) {
        /* the own line's index and frequency come from OwnLine, so the usual
           one-line window waits on no table load before its gathers */
        uint32_t l = s;
        double linefreq;
        if (APPROX == MGX_APPROX_CLOSE_LINES && s == ol.g) {
            l = ol.lidx;
            linefreq = MGX_OWN_FREQ_IN_REGS ? ol.f_own : gld(&m.line[l]);
        } else {
            if (APPROX == MGX_APPROX_CLOSE_LINES) l = gld(&m.sorted_line_map[s]);
            linefreq = gld(&m.line[l]);
        }
        const double4 c = lp(kp, cp, l), cI = lp(kp, cpi, l);
        const double4 n = lp(kp, np, l), nI = lp(kp, npi, l);
        const double line_dtau = dtau_single_line<PREC>(kp, erf_tab, c, cI, n, nI, linefreq,
                                                        currfreq, nextfreq, curr_interp,
                                                        next_interp, dZ);
        const double Sc = interp_log<PREC>(c.w, cI.w, curr_interp);
        const double Sn = interp_log<PREC>(n.w, nI.w, next_interp);
        sum_dtau += line_dtau;
        sum_Sc = fma(line_dtau, Sc, sum_Sc);
        sum_Sn = fma(line_dtau, Sn, sum_Sn);
    }
    double d = sum_dtau;
    if (APPROX == MGX_APPROX_CLOSE_LINES) {
        const double bound_min_dtau = kp.op.min_opacity * dZ;
        if (-bound_min_dtau < d) d = (bound_min_dtau < d) ? d : bound_min_dtau;
    }
    dtau = d;
    const double rd = rcp(d);
    Scurr = sum_Sc * rd;
    Snext = sum_Sn * rd;
}

/* ------------------------------------------------------------------------- */
/* traced entries                                                             */

/* One entry of the expanded list: point, interpolation partner, factor, shift. */
struct Entry {
    uint32_t nr, nri;
    double ifac, shift;
};

/* Cached emissivity/opacity at an entry (valid = computed). */
struct PointEC {
    double eta, chi;
    bool valid;
};

/*
 * One segment (compute_source_dtau, solver.tpp:2140-2186) between list
 * entries cu (closer to the rr-side boundary) and nx.  Returns dtau and the
 * source values of the two rows; uses/updates the caches of the two entries.
 */
/* (dshift_max, ub) of point p: adjacent fields of its record, read as one 16-byte word */
__device__ __forceinline__ double2 dsm_ub(const KParams &kp, uint32_t p)
{
    static_assert(offsetof(PointRec, ub) == offsetof(PointRec, dshift_max) + 8 &&
                      offsetof(PointRec, dshift_max) % 16 == 0 && sizeof(PointRec) % 16 == 0,
                  "dshift_max and ub form one aligned 16-byte word");
    return gld(reinterpret_cast<const double2 *>(&kp.pts[p].dshift_max));
}

template <int APPROX, int PREC>
__device__ inline bool segment(const KParams &kp, const double *erf_tab, const OwnLine &ol,
                               const Entry &cu, const Entry &nx, double dZ, double freq,
                               uint32_t line, PointEC &ec_cu, PointEC &ec_nx, double &dtau,
                               double &S_cu, double &S_nx)
{
    const double dshift = nx.shift - cu.shift;
    const double dshift_abs = fabs(dshift);
    /* dshift_max and ub of both points, each pair in one 16-byte load */
    const double2 du_c = dsm_ub(kp, cu.nr), du_n = dsm_ub(kp, nx.nr);
    const double a = du_c.x, b = du_n.x;
    const double dshift_max = (b < a) ? b : a;
    const double cf = freq * cu.shift, nf = freq * nx.shift;
    if (dshift_abs > dshift_max) {
        S_dtau_line_integrated<APPROX, PREC>(kp, erf_tab, ol, cu.nr, nx.nr, cu.nri, nx.nri, line,
                                             cf, nf, cu.ifac, nx.ifac, dZ, du_c.y, du_n.y, dtau,
                                             S_cu, S_nx);
        return false;
    } else {
        if (!ec_cu.valid) {
            eta_chi_interp<APPROX, PREC>(kp, ol, cu.nr, cu.nri, cu.ifac, line, cf, du_c.y,
                                         ec_cu.eta, ec_cu.chi);
            ec_cu.valid = true;
        }
        if (!ec_nx.valid) {
            eta_chi_interp<APPROX, PREC>(kp, ol, nx.nr, nx.nri, nx.ifac, line, nf, du_n.y,
                                         ec_nx.eta, ec_nx.chi);
            ec_nx.valid = true;
        }
        S_cu = ec_cu.eta * rcp(ec_cu.chi);
        S_nx = ec_nx.eta * rcp(ec_nx.chi);
        dtau = 0.5 * (ec_cu.chi + ec_nx.chi) * dZ;
        return true;
    }
}

/*
 * Iterator over one half ray's expanded entries from the outer end inward.
 * Step s of the walk (crt_s -> nxt_s) expands into m_s entries i = m_s..1;
 * entry (s, i) is the i-th interpolation point (i = m_s at nxt_s) and the
 * segment from it to the next entry inward has length dZ_s / m_s.  The counts
 * m_s come from mtab when given, else from n_interp_for_line.
 */
struct HalfIter {
    int s;     /* current step, -1 = at the origin */
    uint32_t i, m_s;
    uint32_t crt, nxt;
    double shift_crt, shift_nxt, dZ, dshift_interpl;
};

__device__ inline void hi_load_step(const KParams &kp, HalfIter &it, uint32_t l, uint32_t o,
                                    double shift0, const uint32_t *s_nxt, const double *s_dZ,
                                    const double *s_shift, const uint16_t *mtab)
{
    const int s = it.s;
    it.nxt = s_nxt[s];
    it.crt = (s == 0) ? o : s_nxt[s - 1];
    it.shift_nxt = s_shift[s];
    it.shift_crt = (s == 0) ? shift0 : s_shift[s - 1];
    it.m_s = mtab ? (uint32_t)mtab[s] : n_interp_for_line(kp, l, it.crt, it.nxt);
    it.i = it.m_s;
    if (it.m_s > 1) {
        const double dshift = it.shift_nxt - it.shift_crt;
        it.dZ = s_dZ[s] / it.m_s;
        it.dshift_interpl = dshift / it.m_s;
    } else {
        it.dZ = s_dZ[s];
        it.dshift_interpl = 0.0;
    }
}

__device__ inline Entry hi_entry(const HalfIter &it, uint32_t o)
{
    Entry e;
    if (it.s < 0) {
        e.nr = o;
        e.nri = o;
        e.ifac = 0.0;
        e.shift = 1.0;
    } else if (it.m_s > 1) {
        e.nr = it.nxt;
        e.nri = it.crt;
        e.ifac = 1.0 - ((double)it.i) / it.m_s;
        e.shift = it.shift_crt + it.i * it.dshift_interpl;
    } else {
        e.nr = it.nxt;
        e.nri = it.crt;
        e.ifac = 0.0;
        e.shift = it.shift_nxt;
    }
    return e;
}

/* advance inward; returns the length of the segment just crossed */
__device__ inline double hi_advance(const KParams &kp, HalfIter &it, uint32_t l, uint32_t o,
                                    double shift0, const uint32_t *s_nxt, const double *s_dZ,
                                    const double *s_shift, const uint16_t *mtab)
{
    const double d = it.dZ;
    if (it.i > 1) {
        it.i--;
    } else {
        it.s--;
        if (it.s >= 0) hi_load_step(kp, it, l, o, shift0, s_nxt, s_dZ, s_shift, mtab);
    }
    return d;
}

/* ------------------------------------------------------------------------- */
/* the forward-solve kernel                                                   */

/* LDS layout of the walk buffers of both half rays plus the per-pass
   interpolation-count tables (2 lines x 2 halves x cap u16) */
__host__ __device__ inline size_t walk_lds_bytes(uint32_t cap)
{
    return 8 * ((2 * (size_t)cap * 4 + 7) / 8) + 4 * (size_t)cap * 8 + 4 * (size_t)cap * 2;
}

/* dynamic LDS of the solve kernel: walk buffers, the Lambda (L) accumulator and, when
   j_lds, the J (F) accumulator */
__host__ __device__ inline size_t solve_lds_bytes(uint32_t cap, uint32_t F, uint32_t L, bool j_lds)
{
    return ((walk_lds_bytes(cap) + 7) / 8) * 8 + 8 * (size_t)L + (j_lds ? 8 * (size_t)F : 0);
}

/* J stays in LDS while the whole LDS of a workgroup fits the share of a CU's
   160 KiB that the register-limited occupancy (MGX_SOLVE_WAVES_PER_SIMD waves on
   each of 4 SIMDs) leaves it; beyond that (many lines, F ~ 10^3) LDS would lower
   the occupancy, and J is accumulated in the origin's row of the output instead
   (only that origin's wave writes it) */
/* room for the static LDS of a kernel: the float functions' tables
   (MGX_REF_LDS_TABLE_BYTES) and a few words */
constexpr size_t kStaticLds = 2048;
static_assert(MGX_REF_LDS_TABLE_BYTES + 64 <= kStaticLds, "static LDS estimate");

inline bool j_in_lds(uint32_t cap, uint32_t F, uint32_t L)
{
    const size_t share = (160 * 1024) / (4 * MGX_SOLVE_WAVES_PER_SIMD);
    return solve_lds_bytes(cap, F, L, true) + kStaticLds <= share;
}

/* J(o, f) += c for one ray pair: the first ray pair of a full pass stores (the sum
   then starts from c, as 0 + c), later ones add; ray pairs of an origin split over
   several waves add atomically into a cleared row */
__device__ inline void j_add(double *Jrow, uint32_t f, double c, bool first, bool atomic)
{
    if (atomic)
        unsafeAtomicAdd(&Jrow[f], c);
    else
        Jrow[f] = first ? c : Jrow[f] + c;
}

/*
 * One half ray of the two-sided elimination.  Walking from the boundary end
 * toward the origin, both halves run the same recurrence with the sides of
 * each row swapped: forward (rr half) X = F, XI = 1/(1+F), Y = Su; backward
 * (antipod half) X = G, XI = 1/(1+G), Y = w.  Boundary row:
 *   Y = (S + 2 I_bdy/dtau) / (1 + 2/dtau + 2/dtau^2),  X = (1 + 2/dtau) dtau^2 / 2;
 * interior row with the outer and inner segment depths d_out, d_in:
 *   q_out = (d_out + d_in)/2 d_out, q_in = (d_out + d_in)/2 d_in,
 *   X <- ((X XI)/q_out + 1) q_in,  XI <- 1/(1+X),  Y <- (Y/q_out + S) XI q_in
 * (solver.tpp:2234-2318 forward; the G recurrence of :2296-2305 backward).
 */
struct Sweep {
    double X, XI, Y;
    double prev_dtau;  /* depth of the last segment crossed (the innermost one at the end) */
    double S_inner;    /* source of the innermost entry's row from the last segment */
    double S_prev_row; /* source value carried to the next row (backward half) */
    double chi_outer;  /* opacity at the outer entry of the last segment (when trapezoid) */
    double dz_last;    /* length of the last segment */
    bool last_trap;    /* the last segment took the trapezoid branch */
    bool started;
};

/*
 * The last NB rows a half sweep passed (index 0 = the row next to the origin):
 * h = dY/dS of the row and g = dY/dY_outer (the elimination factors of the
 * row's source and of the row before it), with the row's point and shift.
 * The off-diagonal Lambda bands are products of these: the entry of the
 * inverse of the Feautrier system between the origin and the j-th row on a
 * side is (du_c/dY_1) g_1 ... g_(j-1) h_j.
 */
template <int NB>
struct BandRing {
    double h[NB > 0 ? NB : 1], g[NB > 0 ? NB : 1], sh[NB > 0 ? NB : 1];
    uint32_t nr[NB > 0 ? NB : 1];
    uint32_t n;
};

template <int NB>
__device__ inline void ring_push(BandRing<NB> &r, double h, double g, const Entry &e)
{
#pragma unroll
    for (int i = NB - 1; i > 0; i--) {
        r.h[i] = r.h[i - 1];
        r.g[i] = r.g[i - 1];
        r.sh[i] = r.sh[i - 1];
        r.nr[i] = r.nr[i - 1];
    }
    r.h[0] = h;
    r.g[0] = g;
    r.sh[0] = e.shift;
    r.nr[0] = e.nr;
    r.n = (r.n < (uint32_t)NB) ? r.n + 1 : (uint32_t)NB;
}

template <int APPROX, int PREC, int NB = 0>
__device__ inline void sweep_half(const KParams &kp, const double *s_erf, const OwnLine &ol,
                                  bool backward, uint32_t S, uint32_t o, double shift0,
                                  const uint32_t *s_nxt, const double *s_dZ, const double *s_sh,
                                  const uint16_t *mtab, uint32_t l, double freq, Sweep &sw,
                                  PointEC &ec_centre, BandRing<NB> *ring = nullptr)
{
    HalfIter it;
    it.s = (int)S - 1;
    hi_load_step(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
    Entry outer = hi_entry(it, o);
    PointEC ec_out = {0.0, 0.0, false};
    sw.started = false;
    for (;;) {
        const double dz = hi_advance(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
        Entry inner = hi_entry(it, o);
        PointEC ec_in = (it.s < 0) ? ec_centre : PointEC{0.0, 0.0, false};
        /* the segment runs from the rr-side entry to the antipod-side entry */
        const Entry &cu = backward ? inner : outer;
        const Entry &nx = backward ? outer : inner;
        PointEC &ec_cu = backward ? ec_in : ec_out;
        PointEC &ec_nx = backward ? ec_out : ec_in;
        double dtau, S_cu, S_nx;
        sw.last_trap =
            segment<APPROX, PREC>(kp, s_erf, ol, cu, nx, dz, freq, l, ec_cu, ec_nx, dtau, S_cu, S_nx);
        sw.chi_outer = ec_out.chi;
        sw.dz_last = dz;
        const double S_out = backward ? S_nx : S_cu; /* row "outer" from this segment */
        const double S_in = backward ? S_cu : S_nx;
        /* row "outer": the boundary row first, then interior rows; the backward
           half's interior row takes its source from the segment beyond it */
        const double S_row = backward ? (sw.started ? sw.S_prev_row : S_out) : S_out;
        if (!sw.started) {
            const double I_b = boundary_intensity(kp, outer.nr, freq * outer.shift);
            const double idt = rcp(dtau);
            const double B_min = fma(2.0, idt, 1.0);
            const double B = fma(2.0 * idt, idt, B_min);
            const double iB = rcp(B);
            sw.Y = fma(2.0 * I_b, idt, S_row) * iB;
            sw.X = 0.5 * B_min * dtau * dtau;
            sw.XI = rcp(1.0 + sw.X);
            sw.started = true;
            if constexpr (NB > 0) ring_push<NB>(*ring, iB, 0.0, outer);
        } else {
            const double avg = 0.5 * (sw.prev_dtau + dtau);
            const double q_out = avg * sw.prev_dtau;
            const double q_in = avg * dtau;
            const double r = rcp(q_out);
            sw.X = fma(r * sw.X, sw.XI, 1.0) * q_in;
            sw.XI = rcp(1.0 + sw.X);
            sw.Y = fma(r, sw.Y, S_row) * sw.XI * q_in;
            if constexpr (NB > 0) {
                const double hq = sw.XI * q_in;
                ring_push<NB>(*ring, hq, r * hq, outer);
            }
        }
        sw.prev_dtau = dtau;
        sw.S_inner = S_in;
        sw.S_prev_row = S_in;
        outer = inner;
        ec_out = ec_in;
        if (it.s < 0) break;
    }
    ec_centre = ec_out;
}

/*
 * One direction of the order-0 short-characteristics solve
 * (solve_shortchar_order_0_ray_forward / _backward, solver.tpp:1565-1757): from
 * the outer end of a half ray to the origin, I <- e^-dt I + expm1(-dt)/dt
 * (S_c - S_n (1 + dt)) + S_c - S_n per segment with dt = max(min_dtau, dtau)
 * and the reference's float expf / expm1f.  Segments run from the outer entry
 * (current) to the inner one (next) on both halves, as the reference calls
 * compute_source_dtau.  dt_last receives the last segment's dt (0 when the half
 * has no segment: I is then the boundary intensity at the origin).
 */
template <int APPROX, int PREC>
__device__ inline double shortchar_half(const KParams &kp, const double *s_erf, const OwnLine &ol,
                                        uint32_t S, uint32_t o, double shift0,
                                        const uint32_t *s_nxt, const double *s_dZ,
                                        const double *s_sh, const uint16_t *mtab, uint32_t l,
                                        double freq, double min_dtau, double &dt_last,
                                        PointEC &ec_centre)
{
    dt_last = 0.0;
    if (S == 0) return boundary_intensity(kp, o, freq);
    HalfIter it;
    it.s = (int)S - 1;
    hi_load_step(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
    Entry outer = hi_entry(it, o);
    PointEC ec_out = {0.0, 0.0, false};
    double I = boundary_intensity(kp, outer.nr, freq * outer.shift);
    double dt = 0.0;
    for (;;) {
        const double dz = hi_advance(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
        Entry inner = hi_entry(it, o);
        PointEC ec_in = (it.s < 0) ? ec_centre : PointEC{0.0, 0.0, false};
        double dtau, S_c, S_n;
        segment<APPROX, PREC>(kp, s_erf, ol, outer, inner, dz, freq, l, ec_out, ec_in, dtau, S_c,
                              S_n);
        dt = (min_dtau < dtau) ? dtau : min_dtau; /* std::max(min_dtau, dtau) */
        const float mdt = (float)(-dt);
        I = (double)mgx_ref_expf(mdt) * I + (double)mgx_ref_expm1f(mdt) / dt * (S_c - S_n * (1.0 + dt)) +
            S_c - S_n;
        outer = inner;
        ec_out = ec_in;
        if (it.s < 0) break;
    }
    ec_centre = ec_out;
    dt_last = dt;
    return I;
}

/* The factors of a Lambda element at the origin that do not depend on the
   solve (update_Lambda, solver.tpp:1777-1803, and the short-characteristics
   rays, :1618-1637): the element is constante * frq * phi * X * inverse_chi. */
struct LamPoint {
    double constante, frq, phi, inverse_chi;
};

template <int APPROX, int PREC>
__device__ inline LamPoint lambda_point(const KParams &kp, const OwnLine &ol, uint32_t o,
                                        uint32_t l, uint32_t f, double freq, double w_ang)
{
    const mgx_model &m = kp.m;
    const uint32_t Q = m.nquads;
    const uint32_t spec = gld(&m.f_species[f]);
    const uint32_t k = gld(&m.f_k[f]);
    const double freq_line = gld(&m.line[l]);
    LamPoint r;
    r.constante = gld(&m.line_A[l]) * gld(&m.quad_weights[(size_t)spec * Q + gld(&m.f_z[f])]) * w_ang;
    double eta, chi;
    eta_chi_interp<APPROX, PREC>(kp, ol, o, o, 0.0, k, freq_line, gld(&kp.pts[o].ub), eta, chi);
    r.inverse_chi = rcp(chi);
    r.frq = freq * 1.0;
    const double inv_mass = gld(&m.species_inverse_mass[spec]);
    const double width = freq_line * sqrt(MGX_TWO_KB_OVER_AMU_CC_SQUARED * inv_mass *
                                              gld(&m.temperature[o]) +
                                          gld(&m.vturb2[o]));
    const double inverse_width = 1.0 / width;
    const double se = inverse_width * (r.frq - freq_line);
    const double ex = se * se;
    r.phi = inverse_width * MGX_INVERSE_SQRT_PI * (double)f_exp<PREC>((float)(-ex));
    return r;
}

/* which formal solution solve_origin runs */
enum { METHOD_FEAUTRIER = 0, METHOD_SHORTCHAR = 1, METHOD_UV = 2 };

/* extra outputs of the short-characteristics and u/v solves */
struct ShortcharOut {
    double *I;       /* (nrays, P, F) or NULL (short characteristics) */
    double *v;       /* (hnrays, P, F) or NULL */
    double min_dtau; /* Parameters::min_dtau (short characteristics) */
};

/*
 * Ray pairs rr0..rr1-1 of origin o (the body of feautrier_kernel): walk, per-
 * pass interpolation counts, two-sided elimination per lane, u stored when
 * given, J added into the origin's row Jrow (F; atomically when j_atomic) and
 * the Lambda diagonal into the LDS accumulator s_L (L).  Returns 0, or an
 * ERR_* code after flagging it in kp.status.
 */
/* the general Lambda operator (n_off_diag > 0): for each receiving point and line a list
   of at most cap (emitting point, value) entries, the diagonal element first */
struct LamBands {
    uint32_t *count; /* (P, L) */
    uint32_t *nr;    /* (P, L, cap) */
    double *val;     /* (P, L, cap) */
    uint32_t cap, noff;
};

/* Lambda::add_element (lambda.tpp) by the whole wave (q, v uniform): the lanes compare 64
   entries of the list at a time; the first entry of point q gets v added, else lane 0
   appends.  Only this wave writes the list; the fence orders lane 0's writes before the
   next merge's reads by the other lanes. */
__device__ inline void lam_merge(const KParams &kp, const LamBands &lb, size_t pl, uint32_t q,
                                 double v, uint32_t lane)
{
    const size_t base = pl * lb.cap;
    const uint32_t n = lb.count[pl];
    for (uint32_t i0 = 0; i0 < n; i0 += MGX_WAVE) {
        const uint32_t i = i0 + lane;
        const uint64_t hit = __ballot(i < n && lb.nr[base + i] == q);
        if (hit) {
            if (lane == (uint32_t)__ffsll((unsigned long long)hit) - 1) lb.val[base + i] += v;
            __threadfence_block();
            return;
        }
    }
    if (lane == 0) {
        if (n >= lb.cap) {
            atomicOr(&kp.status[0], (uint32_t)ERR_LAMCAP);
        } else {
            lb.nr[base + n] = q;
            lb.val[base + n] = v;
            lb.count[pl] = n + 1;
        }
    }
    __threadfence_block();
}

template <int APPROX, int PREC, int METHOD = METHOD_FEAUTRIER, int NB = 0>
__device__ __forceinline__ int solve_origin(const KParams &kp, uint32_t o, uint32_t rr0,
                                            uint32_t rr1, unsigned char *smem, const double *s_erf,
                                            double *u, bool with_lam, double *s_J, double *Jrow,
                                            bool j_atomic, double *s_L, ShortcharOut sc = {},
                                            LamBands lb = {})
{
    const mgx_model &m = kp.m;
    const uint32_t lane = threadIdx.x;
    const uint32_t cap = kp.cap;
    const uint32_t P = m.npoints, F = m.nfreqs, Q = m.nquads;
    uint32_t *s_nxt0 = reinterpret_cast<uint32_t *>(smem);
    uint32_t *s_nxt1 = s_nxt0 + cap;
    double *s_dZ0 = reinterpret_cast<double *>(smem + 8 * ((2 * cap * 4 + 7) / 8));
    double *s_dZ1 = s_dZ0 + cap;
    double *s_sh0 = s_dZ1 + cap;
    double *s_sh1 = s_sh0 + cap;
    uint16_t *s_m = reinterpret_cast<uint16_t *>(s_sh1 + cap); /* [line j][half h][step] */
    const size_t PF = (size_t)P * F;

    for (uint32_t rr = rr0; rr < rr1; rr++) {
        const uint32_t ar = gld(&m.antipod[rr]);
        uint32_t S0, S1;
        double sh00, sh01;
        int e = walk_half(kp, o, rr, s_nxt0, s_dZ0, s_sh0, S0, sh00);
        if (!e) e = walk_half(kp, o, ar, s_nxt1, s_dZ1, s_sh1, S1, sh01);
        if (e) {
            if (lane == 0) {
                atomicOr(&kp.status[0], (uint32_t)e);
                if (e == ERR_CAP) atomicMax(&kp.status[1], kp.cap + 1);
            }
            return e;
        }
        __syncthreads();
        const double w = ray_weight(m, o, rr);
        const uint32_t Sall = S0 + S1;

        for (uint32_t slot0 = 0; slot0 < kp.nslots; slot0 += 2) {
            /* interpolation counts of this pass's two lines, one per (line, step) */
            const uint32_t lA = slot0 / kp.slots_per_line;
            const uint32_t lB = (slot0 + 1 < kp.nslots) ? (slot0 + 1) / kp.slots_per_line : lA;
            for (uint32_t i = lane; i < 2 * Sall; i += MGX_WAVE) {
                const uint32_t j = i / Sall, r = i % Sall;
                const uint32_t h = (r < S0) ? 0 : 1;
                const uint32_t st = h ? r - S0 : r;
                const uint32_t *nx = h ? s_nxt1 : s_nxt0;
                const uint32_t crt = (st == 0) ? o : nx[st - 1];
                const uint32_t mm = n_interp_for_line(kp, j ? lB : lA, crt, nx[st]);
                if (mm > 0xffffu) atomicOr(&kp.status[0], (uint32_t)ERR_INTERP);
                s_m[(j * 2 + h) * cap + st] = (uint16_t)(mm > 0xffffu ? 0xffffu : mm);
            }
            __syncthreads();

            const uint32_t jj = lane >> 5;
            const uint32_t slot = slot0 + jj;
            const uint32_t l = jj ? lB : lA;
            const uint32_t z = (slot % kp.slots_per_line) * 32 + (lane & 31);
            const bool active = (slot < kp.nslots) && (z < Q);
            double Lval = 0.0;
            /* band elements of this lane, rr side (lo) and antipod side (up), j = 0 .. NB-1 */
            double Llo[NB > 0 ? NB : 1], Lup[NB > 0 ? NB : 1];
            uint32_t qlo[NB > 0 ? NB : 1], qup[NB > 0 ? NB : 1];
            uint32_t nlo = 0, nup = 0;
            bool lam_added = false; /* update_Lambda ran for this lane (n_tot > 1, in the line) */
#pragma unroll
            for (int j = 0; j < (NB > 0 ? NB : 1); j++) {
                Llo[j] = Lup[j] = 0.0;
                qlo[j] = qup[j] = 0;
            }
            if (active) {
                const OwnLine ol = own_line(m, l);
                const uint32_t f = l * Q + z;
                const size_t of = (size_t)o * F + f;
                const double freq = gld(&m.nu[of]);
                if constexpr (METHOD == METHOD_SHORTCHAR) {
                    /* solve_shortchar_order_0 (solver.tpp:186-224): both directions,
                       u, v, J and a Lambda element per direction */
                    double If, Ib, uc, vc;
                    if (S0 == 0 && S1 == 0) {
                        If = Ib = uc = boundary_intensity(kp, o, freq);
                        vc = 0.0;
                    } else {
                        PointEC ec_cent = {0.0, 0.0, false};
                        double dtf, dtb;
                        If = shortchar_half<APPROX, PREC>(kp, s_erf, ol, S0, o, sh00, s_nxt0, s_dZ0,
                                                          s_sh0, s_m + (jj * 2 + 0) * cap, l, freq,
                                                          sc.min_dtau, dtf, ec_cent);
                        Ib = shortchar_half<APPROX, PREC>(kp, s_erf, ol, S1, o, sh01, s_nxt1, s_dZ1,
                                                          s_sh1, s_m + (jj * 2 + 1) * cap, l, freq,
                                                          sc.min_dtau, dtb, ec_cent);
                        uc = 0.5 * (If + Ib);
                        vc = 0.5 * (If - Ib);
                        if (with_lam) {
                            /* a direction without a segment adds nothing (the reference
                               reads an uninitialised dtau there) */
                            const LamPoint lq = lambda_point<APPROX, PREC>(kp, ol, o, l, f, freq, w);
                            if (S0 > 0) {
                                const double factor = (double)mgx_ref_expm1f((float)(-dtf)) / dtf;
                                Lval += lq.constante * lq.frq * lq.phi * (factor + 1.0) * lq.inverse_chi;
                            }
                            if (S1 > 0) {
                                const double factor = (double)mgx_ref_expm1f((float)(-dtb)) / dtb;
                                Lval += lq.constante * lq.frq * lq.phi * (factor + 1.0) * lq.inverse_chi;
                            }
                        }
                    }
                    const double wt = w * 2.0;
                    if (s_J)
                        s_J[f] = s_J[f] + wt * uc;
                    else
                        j_add(Jrow, f, wt * uc, rr == rr0, j_atomic);
                    if (sc.I) {
                        sc.I[(size_t)rr * PF + of] = If;
                        sc.I[(size_t)ar * PF + of] = Ib;
                    }
                    if (u) u[(size_t)rr * PF + of] = uc;
                    if (sc.v) sc.v[(size_t)rr * PF + of] = vc;
                } else {
                double uc;
                double vc = 0.0; /* v at the centre (METHOD_UV) */
                if (S0 == 0 && S1 == 0) {
                    uc = boundary_intensity(kp, o, freq);
                    if constexpr (METHOD != METHOD_UV) {
                        if (s_J)
                            s_J[f] = s_J[f] + 2.0 * w * uc;
                        else
                            j_add(Jrow, f, 2.0 * w * uc, rr == rr0, j_atomic);
                    }
                } else {
                    PointEC ec_cent = {0.0, 0.0, false};
                    Sweep bw, fw;
                    bw.started = fw.started = false;
                    BandRing<NB> ring_b, ring_f;
                    ring_b.n = ring_f.n = 0;
                    if (S1 > 0)
                        sweep_half<APPROX, PREC, NB>(kp, s_erf, ol, true, S1, o, sh01, s_nxt1,
                                                     s_dZ1, s_sh1, s_m + (jj * 2 + 1) * cap, l,
                                                     freq, bw, ec_cent, &ring_b);
                    if (S0 > 0)
                        sweep_half<APPROX, PREC, NB>(kp, s_erf, ol, false, S0, o, sh00, s_nxt0,
                                                     s_dZ0, s_sh0, s_m + (jj * 2 + 0) * cap, l,
                                                     freq, fw, ec_cent, &ring_f);
                    double Lc;
                    double du_fw = 0.0, du_bw = 0.0; /* du_c / dY of each half's last row */
                    if (S1 > 0) {
                        /* row of the origin: forward quantities at the origin, joined with
                           the backward ones of its antipod-side neighbour */
                        const double dtau_c = bw.prev_dtau;
                        const double S_c = bw.S_inner;
                        double FF, Su, iC_c;
                        if (S0 == 0) { /* the origin is the first row */
                            const double I_f = boundary_intensity(kp, o, freq);
                            const double idt = rcp(dtau_c);
                            const double B_min = fma(2.0, idt, 1.0);
                            const double B = fma(2.0 * idt, idt, B_min);
                            iC_c = 0.5 * dtau_c * dtau_c;
                            Su = fma(2.0 * I_f, idt, S_c) * rcp(B);
                            FF = 0.5 * B_min * dtau_c * dtau_c;
                        } else {
                            const double avg = 0.5 * (fw.prev_dtau + dtau_c);
                            const double r = rcp(avg * fw.prev_dtau);
                            iC_c = avg * dtau_c;
                            FF = fma(r * fw.X, fw.XI, 1.0) * iC_c;
                            const double FI = rcp(1.0 + FF);
                            Su = fma(r, fw.Y, S_c) * FI * iC_c;
                            du_fw = r * iC_c; /* times rden below */
                        }
                        const double GP = bw.X * bw.XI;
                        const double rden = rcp(FF + GP);
                        uc = fma(Su, 1.0 + FF, bw.Y) * rden;
                        Lc = iC_c * rden;
                        du_fw *= rden;
                        du_bw = rden;
                        if constexpr (METHOD == METHOD_UV) {
                            if (S0 == 0) {
                                /* the origin is the first row: v = I_bdy - u there */
                                vc = boundary_intensity(kp, o, freq) - uc;
                            } else {
                                /* solve_feautrier_order_2_uv (solver.tpp:2882-2919): u one row to
                                   either side (the back substitution u_n = Y_n + XI_n u_c of
                                   each half) and -du/dtau at the centre by the second-order
                                   difference over the two optical depths */
                                const double um = fw.Y + uc * fw.XI;
                                const double up = bw.Y + uc * bw.XI;
                                const double dtaumin = fw.prev_dtau;
                                /* the reference recomputes both depths with one pair of
                                   opacity variables: when both segments take the trapezoid
                                   branch, the second reuses the opacity of row c-1 for its
                                   current point (solver.tpp:2898-2908) */
                                const double dtauplus =
                                    (fw.last_trap && bw.last_trap)
                                        ? 0.5 * (fw.chi_outer + bw.chi_outer) * bw.dz_last
                                        : dtau_c;
                                const double coeffmin =
                                    -dtauplus / (dtaumin * dtaumin + dtaumin * dtauplus);
                                const double coeffplus =
                                    dtaumin / (dtaumin * dtauplus + dtauplus * dtauplus);
                                const double coeffzero = -coeffmin - coeffplus;
                                vc = -(coeffmin * um + coeffplus * up + coeffzero * uc);
                            }
                        }
                    } else { /* the origin is the last row */
                        const double I_l = boundary_intensity(kp, o, freq);
                        const double idt = rcp(fw.prev_dtau);
                        const double A_l = 2.0 * idt * idt;
                        const double B_min = fma(2.0, idt, 1.0);
                        const double B = B_min + A_l;
                        const double den = rcp(fma(B, fw.X, B_min));
                        const double Sl = fma(2.0 * I_l, idt, fw.S_inner);
                        uc = fma(A_l, fw.Y, Sl) * (1.0 + fw.X) * den;
                        Lc = (1.0 + fw.X) * den;
                        du_fw = A_l * Lc;
                        if constexpr (METHOD == METHOD_UV) vc = uc - I_l; /* v = u - I_bdy at the last row */
                    }
                    if constexpr (METHOD != METHOD_UV) {
                        if (s_J)
                            s_J[f] = s_J[f] + uc * 2.0 * w;
                        else
                            j_add(Jrow, f, uc * 2.0 * w, rr == rr0, j_atomic);
                    }
                    /* ---- off-diagonal bands (update_Lambda, solver.tpp:1808-1850): the
                       entry of the inverse between the origin and the j-th row on each
                       side, an element at that row's point, its opacity at the line
                       centre and its profile at nu(point, f) times the row's shift ---- */
                    if constexpr (NB > 0) {
                        if (with_lam && m.f_in_line[f]) {
                            lam_added = true;
                            nlo = ring_f.n < lb.noff ? ring_f.n : lb.noff;
                            nup = ring_b.n < lb.noff ? ring_b.n : lb.noff;
                            double af = du_fw, ab = du_bw;
#pragma unroll
                            for (int j = 0; j < NB; j++) {
                                if ((uint32_t)j < nlo) {
                                    const uint32_t q = ring_f.nr[j];
                                    const double fq = gld(&m.nu[(size_t)q * F + f]) * ring_f.sh[j];
                                    const LamPoint lq =
                                        lambda_point<APPROX, PREC>(kp, ol, q, l, f, fq, 2.0 * w);
                                    Llo[j] = lq.constante * lq.frq * lq.phi * (af * ring_f.h[j]) *
                                             lq.inverse_chi;
                                    qlo[j] = q;
                                }
                                if ((uint32_t)j < nup) {
                                    const uint32_t q = ring_b.nr[j];
                                    const double fq = gld(&m.nu[(size_t)q * F + f]) * ring_b.sh[j];
                                    const LamPoint lq =
                                        lambda_point<APPROX, PREC>(kp, ol, q, l, f, fq, 2.0 * w);
                                    Lup[j] = lq.constante * lq.frq * lq.phi * (ab * ring_b.h[j]) *
                                             lq.inverse_chi;
                                    qup[j] = q;
                                }
                                af *= ring_f.g[j];
                                ab *= ring_b.g[j];
                            }
                        }
                    }
                    /* ---- Lambda diagonal (update_Lambda, solver.tpp:1761-1800) ---- */
                    if (METHOD != METHOD_UV && with_lam && m.f_in_line[f]) {
                        const LamPoint lq = lambda_point<APPROX, PREC>(kp, ol, o, l, f, freq, 2.0 * w);
                        Lval = lq.constante * lq.frq * lq.phi * Lc * lq.inverse_chi;
                    }
                }
                if (u) u[(size_t)rr * PF + of] = uc;
                if (METHOD == METHOD_UV && sc.v) sc.v[(size_t)rr * PF + of] = vc;
                }
            }
            /* per-line sum of the Lambda contributions over the 32 lanes of a slot;
               lane 0 then adds both slots of the pass in order (they may share a line) */
            if (with_lam) {
                double v = Lval;
#pragma unroll
                for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, MGX_WAVE);
                const double v_hi = __shfl(v, 32, MGX_WAVE);
                if (lane == 0) {
                    for (uint32_t h = 0; h < 2; h++) {
                        const uint32_t slot_h = slot0 + h;
                        if (slot_h >= kp.nslots) break;
                        const uint32_t ll = slot_h / kp.slots_per_line;
                        s_L[ll] = s_L[ll] + (h ? v_hi : v);
                    }
                }
                /* band elements: the same sums, then lane 0 merges them into the lists of
                   (o, line) in the reference's order (lower then upper, band by band) */
                if constexpr (NB > 0) {
                    const uint64_t added = __ballot(lam_added);
                    /* rows and points are the same on every lane of a slot that ran
                       update_Lambda, zero elsewhere: take the largest */
#pragma unroll
                    for (int off = 16; off >= 1; off >>= 1) {
                        const uint32_t a = __shfl_xor(nlo, off, MGX_WAVE), b = __shfl_xor(nup, off, MGX_WAVE);
                        nlo = a > nlo ? a : nlo;
                        nup = b > nup ? b : nup;
                    }
                    const uint32_t nlo_hi = __shfl(nlo, 32, MGX_WAVE), nup_hi = __shfl(nup, 32, MGX_WAVE);
#pragma unroll
                    for (int j = 0; j < NB; j++) {
                        double a = Llo[j], b = Lup[j];
                        uint32_t qa = qlo[j], qb = qup[j];
#pragma unroll
                        for (int off = 16; off >= 1; off >>= 1) {
                            a += __shfl_xor(a, off, MGX_WAVE);
                            b += __shfl_xor(b, off, MGX_WAVE);
                            const uint32_t ta = __shfl_xor(qa, off, MGX_WAVE), tb = __shfl_xor(qb, off, MGX_WAVE);
                            qa = ta > qa ? ta : qa;
                            qb = tb > qb ? tb : qb;
                        }
                        Llo[j] = a;
                        Lup[j] = b;
                        qlo[j] = qa;
                        qup[j] = qb;
                    }
                    /* every lane takes slot h's values from its first lane (0 or 32) */
                    for (uint32_t h = 0; h < 2; h++) {
                        const uint32_t slot_h = slot0 + h;
                        if (slot_h >= kp.nslots) break;
                        if (((added >> (32 * h)) & 0xffffffffull) == 0) continue;
                        const int src = 32 * (int)h;
                        const size_t pl = (size_t)o * m.nlines + slot_h / kp.slots_per_line;
                        const uint32_t n_lo = __shfl(nlo, src, MGX_WAVE), n_up = __shfl(nup, src, MGX_WAVE);
                        lam_merge(kp, lb, pl, o, __shfl(v, src, MGX_WAVE), lane); /* diagonal first */
#pragma unroll
                        for (int j = 0; j < NB; j++) {
                            if ((uint32_t)j < n_lo)
                                lam_merge(kp, lb, pl, __shfl(qlo[j], src, MGX_WAVE),
                                          __shfl(Llo[j], src, MGX_WAVE), lane);
                            if ((uint32_t)j < n_up)
                                lam_merge(kp, lb, pl, __shfl(qup[j], src, MGX_WAVE),
                                          __shfl(Lup[j], src, MGX_WAVE), lane);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    return 0;
}

/* One origin per wave.  With Jline set (the sparse solve) the origin's J stays in
   LDS (or in its row of J, a scratch) and only its line integrals are written:
   Jline(o, l) = sum_z w_z J(o, l Q + z) with the quadrature weights of line l's
   species (lspec.J of Solver::solve_feautrier_order_2_sparse, solver.tpp:340-357). */
template <int APPROX, int PREC>
__global__ __launch_bounds__(MGX_WAVE, MGX_SOLVE_WAVES_PER_SIMD) void feautrier_kernel(
    const KParams *__restrict__ kpp, double *u, double *J, double *lam, double *Jline = nullptr)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    const mgx_model &m = kp.m;
    const uint32_t o = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t F = m.nfreqs, L = m.nlines;
    /* a walk of this launch has outgrown the walk buffer: the host relaunches the whole
       solve with a larger one, so the origins not started yet skip their work */
    if (__builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&kp.status[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) &
        ERR_CAP)
        return;
    double *s_L = reinterpret_cast<double *>(smem + ((walk_lds_bytes(kp.cap) + 7) / 8) * 8);
    double *s_J = kp.j_lds ? s_L + L : nullptr;
    for (uint32_t i = lane; i < L + (s_J ? F : 0); i += MGX_WAVE) s_L[i] = 0.0;
    if (solve_origin<APPROX, PREC>(kp, o, 0, m.hnrays, smem, s_erf, u, lam != nullptr, s_J,
                                   s_J ? nullptr : J + (size_t)o * F, false, s_L))
        return;
    if (Jline) {
        __syncthreads();
        const double *Jo = s_J ? s_J : J + (size_t)o * F;
        const uint32_t Q = m.nquads;
        for (uint32_t l = lane; l < L; l += MGX_WAVE) {
            const double *w = &m.quad_weights[(size_t)m.line_species[l] * Q];
            double acc = 0.0;
            for (uint32_t z = 0; z < Q; z++) acc += w[z] * Jo[(size_t)l * Q + z];
            Jline[(size_t)o * L + l] = acc;
        }
    } else if (s_J) {
        for (uint32_t i = lane; i < F; i += MGX_WAVE) J[(size_t)o * F + i] = s_J[i];
    }
    if (lam)
        for (uint32_t i = lane; i < L; i += MGX_WAVE) lam[(size_t)o * L + i] = s_L[i];
}

/* The forward solve with the general Lambda operator: the bands of up to MGX_MAX_OFF_DIAG
   rows on either side of each origin merged into per-(point, line) lists.  Origins are
   not split over waves, so each list is written by one lane of one wave. */
template <int APPROX, int PREC>
__global__ __launch_bounds__(MGX_WAVE, 2) void feautrier_bands_kernel(
    const KParams *__restrict__ kpp, double *u, double *J, LamBands lb)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    const mgx_model &m = kp.m;
    const uint32_t o = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t F = m.nfreqs, L = m.nlines;
    /* a walk of this launch has outgrown the walk buffer: the host relaunches the whole
       solve with a larger one, so the origins not started yet skip their work */
    if (__builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&kp.status[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) &
        ERR_CAP)
        return;
    double *s_L = reinterpret_cast<double *>(smem + ((walk_lds_bytes(kp.cap) + 7) / 8) * 8);
    double *s_J = kp.j_lds ? s_L + L : nullptr;
    for (uint32_t i = lane; i < L + (s_J ? F : 0); i += MGX_WAVE) s_L[i] = 0.0;
    if (lane == 0)
        for (uint32_t l = 0; l < L; l++) lb.count[(size_t)o * L + l] = 0;
    if (solve_origin<APPROX, PREC, METHOD_FEAUTRIER, MGX_MAX_OFF_DIAG>(
            kp, o, 0, m.hnrays, smem, s_erf, u, true, s_J, J + (size_t)o * F, false, s_L, {}, lb))
        return;
    if (s_J)
        for (uint32_t i = lane; i < F; i += MGX_WAVE) J[(size_t)o * F + i] = s_J[i];
}

/* Order-0 short characteristics (Solver::solve_shortchar_order_0, solver.tpp:151-235):
   the same grid, walk and segments as feautrier_kernel */
template <int APPROX, int PREC>
__global__ __launch_bounds__(MGX_WAVE, MGX_SOLVE_WAVES_PER_SIMD) void shortchar_kernel(
    const KParams *__restrict__ kpp, ShortcharOut sc, double *u, double *J, double *lam)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    const mgx_model &m = kp.m;
    const uint32_t o = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t F = m.nfreqs, L = m.nlines;
    /* a walk of this launch has outgrown the walk buffer: the host relaunches the whole
       solve with a larger one, so the origins not started yet skip their work */
    if (__builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&kp.status[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) &
        ERR_CAP)
        return;
    double *s_L = reinterpret_cast<double *>(smem + ((walk_lds_bytes(kp.cap) + 7) / 8) * 8);
    double *s_J = kp.j_lds ? s_L + L : nullptr;
    for (uint32_t i = lane; i < L + (s_J ? F : 0); i += MGX_WAVE) s_L[i] = 0.0;
    if (solve_origin<APPROX, PREC, METHOD_SHORTCHAR>(kp, o, 0, m.hnrays, smem, s_erf, u,
                                                     lam != nullptr, s_J, J + (size_t)o * F,
                                                     false, s_L, sc))
        return;
    if (s_J)
        for (uint32_t i = lane; i < F; i += MGX_WAVE) J[(size_t)o * F + i] = s_J[i];
    if (lam)
        for (uint32_t i = lane; i < L; i += MGX_WAVE) lam[(size_t)o * L + i] = s_L[i];
}

/* Feautrier order 2 with v (Solver::solve_feautrier_order_2_uv, solver.tpp:238-295,
   2767-2921): u and v only */
template <int APPROX, int PREC>
__global__ __launch_bounds__(MGX_WAVE, MGX_SOLVE_WAVES_PER_SIMD) void feautrier_uv_kernel(
    const KParams *__restrict__ kpp, double *u, double *v)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    const uint32_t o = blockIdx.x;
    double *s_L = reinterpret_cast<double *>(smem + ((walk_lds_bytes(kp.cap) + 7) / 8) * 8);
    ShortcharOut sc{nullptr, v, 0.0};
    solve_origin<APPROX, PREC, METHOD_UV>(kp, o, 0, kp.m.hnrays, smem, s_erf, u, false, nullptr,
                                          nullptr, false, s_L, sc);
}

/* ------------------------------------------------------------------------- */
/* batched forward solve: many models in one persistent launch                */

/* one work item: ray pairs rr0..rr1-1 of origin `origin` of model `model` */
struct BatchItem {
    uint32_t model, origin, rr0, rr1;
};

/*
 * Persistent kernel over a list of work items from several models (each with
 * its own KParams: geometry, line tables, sizes).  A wave takes the next item
 * from an atomic counter, runs solve_origin for it and writes J and the Lambda
 * diagonal of that origin: stored when the item covers all ray pairs of the
 * origin, added atomically when the origin's ray pairs are split over items
 * (the outputs are then cleared before the launch).  WPS is the register
 * budget in waves per SIMD, one of the autotuner's knobs.
 */
template <int APPROX, int PREC, int WPS>
__global__ __launch_bounds__(MGX_WAVE, WPS) void feautrier_batch_kernel(
    const KParams *__restrict__ kps, const BatchItem *__restrict__ items, uint32_t n_items,
    uint32_t *counter, double *const *__restrict__ J, double *const *__restrict__ lam)
{
    mgx_ref_tables_init();
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    __shared__ uint32_t s_item;
    const uint32_t lane = threadIdx.x;
    for (;;) {
        __syncthreads();
        if (lane == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t w = s_item;
        if (w >= n_items) break;
        const BatchItem it = items[w];
        const KParams &kp = kps[it.model];
        const uint32_t F = kp.m.nfreqs, L = kp.m.nlines;
        double *s_L = reinterpret_cast<double *>(smem + ((walk_lds_bytes(kp.cap) + 7) / 8) * 8);
        double *s_J = kp.j_lds ? s_L + L : nullptr;
        for (uint32_t i = lane; i < L + (s_J ? F : 0); i += MGX_WAVE) s_L[i] = 0.0;
        double *Jm = J[it.model];
        double *Lm = lam ? lam[it.model] : nullptr;
        const size_t o = it.origin;
        const bool whole = it.rr0 == 0 && it.rr1 == kp.m.hnrays;
        if (solve_origin<APPROX, PREC>(kp, it.origin, it.rr0, it.rr1, smem, s_erf, nullptr,
                                       Lm != nullptr, s_J, Jm + o * F, !whole, s_L))
            return;
        __syncthreads();
        if (whole) {
            if (s_J)
                for (uint32_t i = lane; i < F; i += MGX_WAVE) Jm[o * F + i] = s_J[i];
            if (Lm)
                for (uint32_t i = lane; i < L; i += MGX_WAVE) Lm[o * L + i] = s_L[i];
        } else {
            if (s_J)
                for (uint32_t i = lane; i < F; i += MGX_WAVE)
                    unsafeAtomicAdd(&Jm[o * F + i], s_J[i]);
            if (Lm)
                for (uint32_t i = lane; i < L; i += MGX_WAVE)
                    unsafeAtomicAdd(&Lm[o * L + i], s_L[i]);
        }
    }
}

/* clears n words (graph-safe: a kernel, not a memset node) */
__global__ void fill_u32_kernel(uint32_t *p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_f64_kernel(double *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

/* ------------------------------------------------------------------------- */
/* trace kernels: counts and dumps of expanded lists                          */

/* one wave per (rr, o): n_tot for every line */
__global__ __launch_bounds__(MGX_WAVE) void trace_counts_kernel(const KParams *__restrict__ kpp, uint32_t *n_tot)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const mgx_model &m = kp.m;
    const uint32_t o = blockIdx.x, rr = blockIdx.y, lane = threadIdx.x;
    const uint32_t cap = kp.cap;
    uint32_t *s_nxt0 = reinterpret_cast<uint32_t *>(smem);
    uint32_t *s_nxt1 = s_nxt0 + cap;
    double *s_dZ0 = reinterpret_cast<double *>(smem + 8 * ((2 * cap * 4 + 7) / 8));
    double *s_dZ1 = s_dZ0 + cap;
    double *s_sh0 = s_dZ1 + cap;
    double *s_sh1 = s_sh0 + cap;
    uint32_t S0, S1;
    double sh00, sh01;
    int e = walk_half(kp, o, rr, s_nxt0, s_dZ0, s_sh0, S0, sh00);
    if (!e) e = walk_half(kp, o, m.antipod[rr], s_nxt1, s_dZ1, s_sh1, S1, sh01);
    if (e) {
        if (lane == 0) {
            atomicOr(&kp.status[0], (uint32_t)e);
            if (e == ERR_CAP) atomicMax(&kp.status[1], kp.cap + 1);
        }
        return;
    }
    __syncthreads();
    for (uint32_t l = lane; l < m.nlines; l += MGX_WAVE) {
        uint32_t n = 1;
        for (uint32_t s = 0; s < S0; s++)
            n += n_interp_for_line(kp, l, s == 0 ? o : s_nxt0[s - 1], s_nxt0[s]);
        for (uint32_t s = 0; s < S1; s++)
            n += n_interp_for_line(kp, l, s == 0 ? o : s_nxt1[s - 1], s_nxt1[s]);
        n_tot[((size_t)rr * m.npoints + o) * m.nlines + l] = n;
    }
}

/* one wave per (rr, o): walk and trace statistics (S, neighbours, T, T_interp) */
__global__ __launch_bounds__(MGX_WAVE) void trace_stats_kernel(const KParams *__restrict__ kpp, unsigned long long *stats)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const mgx_model &m = kp.m;
    const uint32_t o = blockIdx.x, rr = blockIdx.y, lane = threadIdx.x;
    const uint32_t cap = kp.cap;
    uint32_t *s_nxt0 = reinterpret_cast<uint32_t *>(smem);
    uint32_t *s_nxt1 = s_nxt0 + cap;
    double *s_dZ0 = reinterpret_cast<double *>(smem + 8 * ((2 * cap * 4 + 7) / 8));
    double *s_dZ1 = s_dZ0 + cap;
    double *s_sh0 = s_dZ1 + cap;
    double *s_sh1 = s_sh0 + cap;
    uint32_t S0, S1;
    double sh00, sh01;
    int e = walk_half(kp, o, rr, s_nxt0, s_dZ0, s_sh0, S0, sh00);
    if (!e) e = walk_half(kp, o, m.antipod[rr], s_nxt1, s_dZ1, s_sh1, S1, sh01);
    if (e) {
        if (lane == 0) {
            atomicOr(&kp.status[0], (uint32_t)e);
            if (e == ERR_CAP) atomicMax(&kp.status[1], kp.cap + 1);
        }
        return;
    }
    __syncthreads();
    unsigned long long nb = 0, T = 0, Ti = 0;
    for (uint32_t l = lane; l < m.nlines; l += MGX_WAVE) {
        uint32_t n = 1, ni = 0;
        for (uint32_t s = 0; s < S0; s++) {
            const uint32_t k = n_interp_for_line(kp, l, s == 0 ? o : s_nxt0[s - 1], s_nxt0[s]);
            n += k;
            ni += k - 1;
        }
        for (uint32_t s = 0; s < S1; s++) {
            const uint32_t k = n_interp_for_line(kp, l, s == 0 ? o : s_nxt1[s - 1], s_nxt1[s]);
            n += k;
            ni += k - 1;
        }
        T += n;
        Ti += ni;
    }
    if (lane == 0) {
        for (uint32_t s = 0; s < S0; s++) nb += m.n_neighbors[s == 0 ? o : s_nxt0[s - 1]];
        for (uint32_t s = 0; s < S1; s++) nb += m.n_neighbors[s == 0 ? o : s_nxt1[s - 1]];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        T += __shfl_xor(T, off, MGX_WAVE);
        Ti += __shfl_xor(Ti, off, MGX_WAVE);
    }
    if (lane == 0) {
        atomicAdd(&stats[0], (unsigned long long)(S0 + S1));
        atomicAdd(&stats[1], nb);
        atomicAdd(&stats[2], T);
        atomicAdd(&stats[3], Ti);
    }
}

/* The walk-buffer size a solve needs, before the solve (the reference sizes its
   scratch from get_ray_lengths_max, solver.tpp:106-126): one wave per (origin,
   ray pair) walks both halves without storing them and the longest half's step
   count goes into *max_steps.  Z grows along a walk, so a general-geometry walk
   visits no point twice, and a spherically symmetric one each shell at most twice
   (in and out): 2 npoints + 2 steps bound it, and a longer walk flags ERR_NO_NEXT. */
__global__ __launch_bounds__(MGX_WAVE) void walk_max_kernel(const KParams *__restrict__ kpp,
                                                            uint32_t *max_steps)
{
    const KParams &kp = *kpp;
    const mgx_model &m = kp.m;
    const uint32_t o = blockIdx.x, rr = blockIdx.y, lane = threadIdx.x;
    uint32_t longest = 0;
    for (int h = 0; h < 2; h++) {
        const uint32_t r = h ? m.antipod[rr] : rr;
        const PointRec po = gld(&kp.pts[o]);
        const double origin[3] = {po.x, po.y, po.z};
        const double *dp = ray_dir(m, o, r);
        const double dir[3] = {gld(&dp[0]), gld(&dp[1]), gld(&dp[2])};
        double Z = 0.0, dZ = 0.0;
        uint32_t crt = o, n = 0;
        PointIdx pc = gld(&kp.pidx[o]);
        for (;;) {
            const uint32_t nxt = m.spherical_symmetry ? next_spherical(kp, origin, dir, crt, Z, dZ)
                                                      : next_general(kp, origin, dir, pc, Z, dZ);
            if (nxt >= m.npoints) {
                if (n != 0 && lane == 0) atomicOr(&kp.status[0], (uint32_t)ERR_NO_NEXT);
                break;
            }
            if (++n > 2 * m.npoints + 2) {
                if (lane == 0) atomicOr(&kp.status[0], (uint32_t)ERR_NO_NEXT);
                break;
            }
            const PointIdx pn = gld(&kp.pidx[nxt]);
            if (pn.p2b != m.npoints) break; /* on the boundary */
            crt = nxt;
            pc = pn;
        }
        longest = n > longest ? n : longest;
    }
    if (lane == 0) atomicMax(max_steps, longest);
}

/* one wave per selected (rr, o, l); lane 0 writes the expanded list */
__global__ __launch_bounds__(MGX_WAVE) void trace_rays_kernel(
    const KParams *__restrict__ kpp, const uint32_t *sel_rr, const uint32_t *sel_o, const uint32_t *sel_l,
    uint32_t capacity, uint32_t *count, uint32_t *centre_idx, uint32_t *nr, uint32_t *nr_interp,
    double *shift, double *interp_factor, double *dZ)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const mgx_model &m = kp.m;
    const uint32_t i = blockIdx.x, lane = threadIdx.x;
    const uint32_t rr = sel_rr[i], o = sel_o[i], l = sel_l[i];
    const uint32_t cap = kp.cap;
    uint32_t *s_nxt0 = reinterpret_cast<uint32_t *>(smem);
    uint32_t *s_nxt1 = s_nxt0 + cap;
    double *s_dZ0 = reinterpret_cast<double *>(smem + 8 * ((2 * cap * 4 + 7) / 8));
    double *s_dZ1 = s_dZ0 + cap;
    double *s_sh0 = s_dZ1 + cap;
    double *s_sh1 = s_sh0 + cap;
    uint32_t S0, S1;
    double sh00, sh01;
    int e = walk_half(kp, o, rr, s_nxt0, s_dZ0, s_sh0, S0, sh00);
    if (!e) e = walk_half(kp, o, m.antipod[rr], s_nxt1, s_dZ1, s_sh1, S1, sh01);
    if (e) {
        if (lane == 0) {
            atomicOr(&kp.status[0], (uint32_t)e);
            if (e == ERR_CAP) atomicMax(&kp.status[1], kp.cap + 1);
        }
        return;
    }
    __syncthreads();
    if (lane != 0) return;
    /* count entries of each half */
    uint32_t N0 = 0, N1 = 0;
    for (uint32_t s = 0; s < S0; s++) N0 += n_interp_for_line(kp, l, s == 0 ? o : s_nxt0[s - 1], s_nxt0[s]);
    for (uint32_t s = 0; s < S1; s++) N1 += n_interp_for_line(kp, l, s == 0 ? o : s_nxt1[s - 1], s_nxt1[s]);
    const uint32_t n = N0 + N1 + 1;
    count[i] = n;
    centre_idx[i] = N0;
    if (n > capacity) {
        atomicOr(&kp.status[0], (uint32_t)ERR_OUT);
        return;
    }
    const size_t row = (size_t)i * capacity;
    /* rr half, outermost first: list index j = 0..N0-1 */
    {
        HalfIter it;
        uint32_t j = 0;
        if (S0 > 0) {
            it.s = (int)S0 - 1;
            hi_load_step(kp, it, l, o, sh00, s_nxt0, s_dZ0, s_sh0, nullptr);
            while (it.s >= 0) {
                const Entry en = hi_entry(it, o);
                nr[row + j] = en.nr;
                nr_interp[row + j] = en.nri;
                interp_factor[row + j] = en.ifac;
                shift[row + j] = en.shift;
                dZ[row + j] = hi_advance(kp, it, l, o, sh00, s_nxt0, s_dZ0, s_sh0, nullptr);
                j++;
            }
        }
        /* centre */
        nr[row + j] = o;
        nr_interp[row + j] = 0; /* never written by the reference's trace (zero scratch) */
        interp_factor[row + j] = 0.0;
        shift[row + j] = 1.0;
    }
    /* antipod half, innermost first: list index N0+1 .. N0+N1; dZ[j-1] for entry j */
    if (S1 > 0) {
        /* iterate outward by collecting per-step counts, then entries in order */
        uint32_t j = N0 + 1;
        double shift_crt = sh01;
        uint32_t crt = o;
        for (uint32_t s = 0; s < S1; s++) {
            const uint32_t nxt = s_nxt1[s];
            const uint32_t ms = n_interp_for_line(kp, l, crt, nxt);
            const double shift_nxt = s_sh1[s];
            const double dshift = shift_nxt - shift_crt;
            if (ms > 1) {
                const double dZ_interpl = s_dZ1[s] / ms;
                const double dshift_interpl = dshift / ms;
                for (uint32_t q = 1; q <= ms; q++, j++) {
                    nr[row + j] = nxt;
                    shift[row + j] = shift_crt + q * dshift_interpl;
                    nr_interp[row + j] = crt;
                    interp_factor[row + j] = 1.0 - ((double)q) / ms;
                    dZ[row + j - 1] = dZ_interpl;
                }
            } else {
                nr[row + j] = nxt;
                shift[row + j] = shift_nxt;
                nr_interp[row + j] = crt;
                interp_factor[row + j] = 0.0;
                dZ[row + j - 1] = s_dZ1[s];
                j++;
            }
            crt = nxt;
            shift_crt = shift_nxt;
        }
    }
    dZ[row + n - 1] = 0.0;
}

/* ------------------------------------------------------------------------- */
/* projection-surface imager (image_feautrier_order_2_new_imager,             */
/* solver.tpp:577-627; Image::set_coordinates_projection_surface,             */
/* image.cpp:221-398)                                                         */

struct ImageFrame {
    double rd[3];          /* normalised ray direction */
    double ix, iy, jx, jy, jz, denominator;
    double cpt[3];         /* surface_center_point */
    double min_x, min_y, deltax, deltay;
    uint32_t start_bdy;    /* get_closest_bdy_point_in_custom_raydir */
    uint32_t nx, ny;
};

/* lexicographic min over the block of (value, key) pairs */
__device__ inline bool lex_less(double a, uint32_t ka, double b, uint32_t kb)
{
    return (a < b) || (a == b && ka < kb);
}

/* one block of 256 threads: projected bounding box of the boundary points and
   the closest boundary point along the ray direction */
__global__ __launch_bounds__(256) void image_frame_kernel(mgx_model m, ImageFrame *out)
{
    mgx_ref_tables_init();
    __shared__ double s_min_x[256], s_max_x[256], s_min_y[256], s_max_y[256], s_pv[256];
    __shared__ uint32_t s_pk[256], s_pp[256];
    ImageFrame fr = *out; /* rd, nx, ny set by the host */
    const double rx = fr.rd[0], ry = fr.rd[1], rz = fr.rd[2];
    fr.denominator = sqrt(rx * rx + ry * ry);
    const double inverse_denominator = 1.0 / fr.denominator;
    fr.ix = ry * inverse_denominator;
    fr.iy = -rx * inverse_denominator;
    fr.jx = rx * rz * inverse_denominator;
    fr.jy = ry * rz * inverse_denominator;
    fr.jz = -fr.denominator;
    double mnx = DBL_MAX, mxx = -DBL_MAX, mny = DBL_MAX, mxy = -DBL_MAX;
    /* closest: initial guess point2boundary[0] used as a point index (key 0),
       then boundary points b >= 1 (key b), strict "<" */
    double pv = __builtin_inf();
    uint32_t pk = 0xffffffffu, pp = 0;
    const uint32_t t = threadIdx.x;
    if (t == 0 && !m.spherical_symmetry) {
        pp = m.point2boundary[0];
        pv = dot3(fr.rd, &m.position[3 * (size_t)pp]);
        pk = 0;
    }
    for (uint32_t b = t; b < m.nboundary; b += 256) {
        const double *p = &m.position[3 * (size_t)m.boundary2point[b]];
        double X, Y;
        if (m.dimension == 1) {
            X = p[0];
            Y = 0.0;
        } else if (fr.denominator >= 1.0e-9) {
            X = fr.ix * p[0] + fr.iy * p[1];
            Y = fr.jx * p[0] + fr.jy * p[1] + fr.jz * p[2];
        } else {
            X = p[0];
            Y = p[1];
        }
        mnx = (X < mnx) ? X : mnx;
        mxx = (mxx < X) ? X : mxx;
        mny = (Y < mny) ? Y : mny;
        mxy = (mxy < Y) ? Y : mxy;
        if (b >= 1 && !m.spherical_symmetry) {
            const double pd = dot3(fr.rd, p);
            if (lex_less(pd, b, pv, pk)) {
                pv = pd;
                pk = b;
                pp = m.boundary2point[b];
            }
        }
    }
    s_min_x[t] = mnx;
    s_max_x[t] = mxx;
    s_min_y[t] = mny;
    s_max_y[t] = mxy;
    s_pv[t] = pv;
    s_pk[t] = pk;
    s_pp[t] = pp;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (t < h) {
            s_min_x[t] = (s_min_x[t + h] < s_min_x[t]) ? s_min_x[t + h] : s_min_x[t];
            s_max_x[t] = (s_max_x[t] < s_max_x[t + h]) ? s_max_x[t + h] : s_max_x[t];
            s_min_y[t] = (s_min_y[t + h] < s_min_y[t]) ? s_min_y[t + h] : s_min_y[t];
            s_max_y[t] = (s_max_y[t] < s_max_y[t + h]) ? s_max_y[t + h] : s_max_y[t];
            if (lex_less(s_pv[t + h], s_pk[t + h], s_pv[t], s_pk[t])) {
                s_pv[t] = s_pv[t + h];
                s_pk[t] = s_pk[t + h];
                s_pp[t] = s_pp[t + h];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        double min_x = s_min_x[0], max_x = s_max_x[0], min_y = s_min_y[0], max_y = s_max_y[0];
        if (m.dimension == 1) {
            const double ax = fabs(min_x);
            max_x = (ax < max_x) ? max_x : ax;
            min_x = -max_x;
            min_y = min_x;
            max_y = max_x;
        }
        fr.start_bdy = m.spherical_symmetry ? m.npoints - 1 : s_pp[0];
        const double *cp = &m.position[3 * (size_t)fr.start_bdy];
        double dist = dot3(cp, fr.rd);
        if (m.spherical_symmetry) dist = -cp[0];
        for (int i = 0; i < 3; i++) fr.cpt[i] = fr.rd[i] * dist;
        fr.min_x = min_x;
        fr.min_y = min_y;
        fr.deltax = (max_x - min_x) / (fr.nx);
        fr.deltay = (max_y - min_y) / (fr.ny);
        *out = fr;
    }
}

/* get_boundary_point_closer_to_custom_ray (geometry.tpp:310-341), wave-cooperative */
__device__ inline uint32_t bdy_closer(const mgx_model &m, const double *origin, const double *rd,
                                      uint32_t crt)
{
    const uint32_t lane = threadIdx.x & (MGX_WAVE - 1);
    const uint32_t n_nbs = m.n_neighbors[crt];
    const uint32_t cum = m.cum_n_neighbors[crt];
    double Rc[3];
    for (int i = 0; i < 3; i++) Rc[i] = m.position[3 * (size_t)crt + i] - origin[i];
    const double Zc = dot3(Rc, rd);
    /* the current point is the candidate to beat (key 0); neighbour i has key i + 1 */
    double best = (lane == 0) ? dot3(Rc, Rc) - Zc * Zc : __builtin_inf();
    uint32_t key = (lane == 0) ? 0u : 0xffffffffu;
    double dummy = 0.0;
    for (uint32_t i = lane; i < n_nbs; i += MGX_WAVE) {
        const uint32_t n = m.neighbors[cum + i];
        if (m.point2boundary[n] != m.npoints) {
            double R[3];
            for (int j = 0; j < 3; j++) R[j] = m.position[3 * (size_t)n + j] - origin[j];
            const double Zn = dot3(R, rd);
            const double d2 = dot3(R, R) - Zn * Zn;
            if (lex_less(d2, i + 1, best, key)) {
                best = d2;
                key = i + 1;
            }
        }
    }
    wave_argmin(best, key, dummy, n_nbs);
    return (key == 0 || key == 0xffffffffu) ? crt : m.neighbors[cum + key - 1];
}

__device__ inline double shift_rest(const mgx_model &m, const double *origin, const double *rd,
                                    uint32_t c, double Z)
{
    if (m.spherical_symmetry) {
        const double *pc = &m.position[3 * (size_t)c];
        if (pc[0] == 0.0) return 1.0;
        const double Rcos_plus_Z = dot3(origin, rd) + Z;
        return 1.0 + m.velocity[3 * (size_t)c] * Rcos_plus_Z / pc[0];
    }
    return 1.0 + dot3(&m.velocity[3 * (size_t)c], rd);
}

/* image kinds: the intensity (2 Su[last] - I_bdy) or the optical depth along the ray */
enum { IMG_INTENSITY = 0, IMG_OPTICAL_DEPTH = 1 };

/* Forward elimination of a whole ray, one row at a time
   (image_feautrier_order_2, solver.tpp:2463-2560).  fe_row takes each row from
   the first on, with the depth of the segment leaving it and its source there;
   fe_last closes the last row and returns Su[last]. */
struct FwdElim {
    double FF, FI, Su, prev_dtau;
    bool started;
};

__device__ inline void fe_row(const KParams &kp, FwdElim &fe, double dtau, double S_row,
                              const Entry &row, double freq)
{
    if (!fe.started) {
        const double I_f = boundary_intensity(kp, row.nr, freq * row.shift);
        const double idt = rcp(dtau);
        const double Bf_min_Cf = fma(2.0, idt, 1.0);
        const double Bf = fma(2.0 * idt, idt, Bf_min_Cf);
        fe.Su = fma(2.0 * I_f, idt, S_row) * rcp(Bf);
        fe.FF = 0.5 * Bf_min_Cf * dtau * dtau;
        fe.FI = rcp(1.0 + fe.FF);
        fe.started = true;
    } else {
        const double dtau_avg = 0.5 * (fe.prev_dtau + dtau);
        const double iA = dtau_avg * fe.prev_dtau;
        const double iC = dtau_avg * dtau;
        const double A = rcp(iA);
        fe.FF = fma(A * fe.FF, fe.FI, 1.0) * iC;
        fe.FI = rcp(1.0 + fe.FF);
        fe.Su = fma(A, fe.Su, S_row) * fe.FI * iC;
    }
    fe.prev_dtau = dtau;
}

__device__ inline double fe_last(const FwdElim &fe, double S_last, double I_l)
{
    const double idt = rcp(fe.prev_dtau);
    const double A_l = 2.0 * idt * idt;
    const double Bl_min_Al = fma(2.0, idt, 1.0);
    const double Bl = Bl_min_Al + A_l;
    const double denominator = rcp(fma(Bl, fe.FF, Bl_min_Al));
    const double Sl = fma(2.0 * I_l, idt, S_last);
    return fma(A_l, fe.Su, Sl) * (1.0 + fe.FF) * denominator;
}

/* all-lines interpolation count (InterpHelper::get_n_interp, :10-46) */
__device__ inline uint32_t n_interp_all(const KParams &kp, uint32_t c, uint32_t n)
{
    const uint32_t L = kp.m.nlines;
    double mx = 0.0;
    for (uint32_t l = 0; l < L; l++) {
        const double d = (double)fabsf(mgx_ref_logf((float)(kp.m.interp_crit[(size_t)n * L + l] /
                                                        kp.m.interp_crit[(size_t)c * L + l])));
        if (l == 0 || mx < d) mx = d;
    }
    const uint32_t k = (uint32_t)ceil(mx / kp.log_max_interp);
    return (k > 1) ? k : 1;
}

/* The projection-surface imager, one pixel at a time (the pixel loop bodies of
   Solver::image_feautrier_order_2_new_imager and image_optical_depth_new_imager,
   solver.tpp:590-620, 701-729), in two parts: image_walk, where the whole wave
   walks the pixel's ray into its lists in the LDS walk buffer, and image_freqs,
   where lanes take the pixel's frequencies.  The walk buffer of capacity cap
   (walk_lds_bytes(cap) bytes) holds two pixels' lists of up to cap steps each:
   slot 0 and slot 1. */
struct PixelWalk {
    uint32_t *nxt;
    double *dZ, *sh;
    uint16_t *m;
    uint32_t ip, S;
    double shift0, shift_centre;
};

__device__ inline PixelWalk pixel_slot(unsigned char *smem, uint32_t cap, uint32_t slot)
{
    PixelWalk w;
    double *dZ = reinterpret_cast<double *>(smem + 8 * ((2 * cap * 4 + 7) / 8));
    w.nxt = reinterpret_cast<uint32_t *>(smem) + slot * cap;
    w.dZ = dZ + slot * cap;
    w.sh = dZ + 2 * cap + slot * cap;
    w.m = reinterpret_cast<uint16_t *>(dZ + 4 * cap) + slot * cap;
    w.ip = 0;
    w.S = 0;
    w.shift0 = w.shift_centre = 0.0;
    return w;
}

/* Walks pixel pix's ray into w's lists (the whole wave; lane 0 writes ImX, ImY and
   the lists).  Returns false when the walk needs more than kp.cap steps (the status
   words then hold ERR_CAP and the capacity needed). */
__device__ bool image_walk(const KParams &kp, const ImageFrame &fr, uint32_t pix, PixelWalk &w,
                           double *ImX, double *ImY)
{
    const mgx_model &m = kp.m;
    const uint32_t lane = threadIdx.x & (MGX_WAVE - 1);
    const uint32_t cap = kp.cap;
    uint32_t *s_nxt = w.nxt;
    double *s_dZ = w.dZ;
    double *s_sh = w.sh;
    uint16_t *s_m = w.m;

    /* pixel position on the projection surface and its 3D origin */
    const uint32_t x_idx = pix / fr.ny, y_idx = pix % fr.ny;
    const double xloc = fr.min_x + fr.deltax / 2.0 + x_idx * fr.deltax;
    const double yloc = fr.min_y + fr.deltay / 2.0 + y_idx * fr.deltay;
    if (lane == 0 && ImX) {
        ImX[pix] = xloc;
        ImY[pix] = yloc;
    }
    double origin[3];
    if (fr.denominator >= 1.0e-9) {
        origin[0] = (xloc * fr.ix + yloc * fr.jx) + fr.cpt[0];
        origin[1] = (xloc * fr.iy + yloc * fr.jy) + fr.cpt[1];
        origin[2] = (yloc * fr.jz) + fr.cpt[2];
    } else {
        origin[0] = xloc + fr.cpt[0];
        origin[1] = yloc + fr.cpt[1];
        origin[2] = 0.0 + fr.cpt[2];
    }
    const double *rd = fr.rd;

    /* trace_ray_imaging_get_start (solver.tpp:808-830) */
    uint32_t ip = fr.start_bdy;
    for (uint32_t it = 0; it <= m.nboundary; it++) {
        const uint32_t na = bdy_closer(m, origin, rd, ip);
        if (na == ip) break;
        ip = na;
    }
    double Z;
    {
        double R[3];
        for (int i = 0; i < 3; i++) R[i] = m.position[3 * (size_t)ip + i] - origin[i];
        Z = dot3(rd, R);
        if (m.spherical_symmetry) {
            const double Rcos = dot3(origin, rd);
            const double R2 = dot3(origin, origin);
            const double Rsin2 = R2 - Rcos * Rcos;
            const double *pb = &m.position[3 * (size_t)ip];
            const double h2 = dot3(pb, pb) - Rsin2;
            Z = (h2 > 0) ? (-sqrt(h2) - Rcos) : -Rcos;
        }
    }
    const double shift_centre = shift_rest(m, origin, rd, ip, Z);

    /* trace_ray_imaging<Rest> (solver.tpp:842-897): walk into LDS */
    uint32_t S = 0;
    double shift0 = 0.0;
    {
        double dZ = 0.0;
        uint32_t crt = ip;
        uint32_t nxt = m.spherical_symmetry ? next_spherical(kp, origin, rd, crt, Z, dZ)
                                            : next_general(kp, origin, rd, gld(&kp.pidx[crt]), Z, dZ);
        bool accessed_non_boundary = false;
        if (nxt < m.npoints) {
            shift0 = shift_rest(m, origin, rd, crt, Z - dZ);
            double shift_nxt = shift_rest(m, origin, rd, nxt, Z);
            for (;;) {
                /* past the buffer the walk goes on uncounted in LDS, so that the
                   relaunch gets the ray's exact length; a chord crosses every point
                   at most twice (each shell, under spherical symmetry) */
                if (S > 2 * m.npoints + 2) {
                    if (lane == 0) atomicOr(&kp.status[0], (uint32_t)ERR_NO_NEXT);
                    return false;
                }
                if (S < cap && lane == 0) {
                    s_nxt[S] = nxt;
                    s_dZ[S] = dZ;
                    s_sh[S] = shift_nxt;
                }
                S++;
                double rel[3];
                for (int i = 0; i < 3; i++) rel[i] = m.position[3 * (size_t)nxt + i] - m.center[i];
                /* the record of nxt serves the boundary test and the next step */
                const PointIdx pn = gld(&kp.pidx[nxt]);
                const bool interior = pn.p2b == m.npoints;
                if (!(interior || dot3(rel, rd) < 0)) break;
                crt = nxt;
                if (interior) accessed_non_boundary = true;
                nxt = m.spherical_symmetry ? next_spherical(kp, origin, rd, crt, Z, dZ)
                                           : next_general(kp, origin, rd, pn, Z, dZ);
                if (nxt >= m.npoints) break;
                shift_nxt = shift_rest(m, origin, rd, nxt, Z);
            }
        }
        if (!accessed_non_boundary && !m.spherical_symmetry) S = 0;
    }
    if (S > cap) {
        if (lane == 0) {
            atomicOr(&kp.status[0], (uint32_t)ERR_CAP);
            atomicMax(&kp.status[1], S);
        }
        return false;
    }
    __syncthreads();
    for (uint32_t i = lane; i < S; i += MGX_WAVE) {
        const uint32_t crt = (i == 0) ? ip : s_nxt[i - 1];
        const uint32_t mm = n_interp_all(kp, crt, s_nxt[i]);
        if (mm > 0xffffu) atomicOr(&kp.status[0], (uint32_t)ERR_INTERP);
        s_m[i] = (uint16_t)(mm > 0xffffu ? 0xffffu : mm);
    }
    __syncthreads();
    w.ip = ip;
    w.S = S;
    w.shift0 = shift0;
    w.shift_centre = shift_centre;
    return true;
}

/* The frequencies f0, f0 + fstep, ... of the walked pixel pix. */
template <int APPROX, int PREC, int KIND>
__device__ void image_freqs(const KParams &kp, const PixelWalk &w, uint32_t pix, uint32_t f0,
                            uint32_t fstep, double *I)
{
    const double *s_erf = nullptr; /* erf needs no table */
    const mgx_model &m = kp.m;
    const uint32_t F = m.nfreqs;
    const uint32_t ip = w.ip, S = w.S;
    const double shift0 = w.shift0, shift_centre = w.shift_centre;
    const uint32_t *s_nxt = w.nxt;
    const double *s_dZ = w.dZ, *s_sh = w.sh;
    const uint16_t *s_m = w.m;

    for (uint32_t f = f0; f < F; f += fstep) {
        const double nu_ip = m.nu[(size_t)ip * F + f];
        double Ival;
        if (S == 0) {
            /* n_tot = 1: the boundary intensity, or no optical depth
               (image_optical_depth_new_imager, solver.tpp:703-729) */
            Ival = (KIND == IMG_OPTICAL_DEPTH) ? 0.0 : boundary_intensity(kp, ip, nu_ip);
        } else {
            const double freq = nu_ip;
            const uint32_t line = m.f_line[f];
            const uint32_t ub_idx = upper_bound_d(m.sorted_line, m.nlines, freq);
            const uint32_t own_l = m.sorted_line_map[ub_idx > 0 ? ub_idx - 1 : 0];
            const OwnLine ol = own_line(m, own_l);
            HalfIter it;
            it.s = (int)S - 1;
            hi_load_step(kp, it, line, ip, shift0, s_nxt, s_dZ, s_sh, s_m);
            Entry outer = hi_entry(it, ip);
            PointEC ec_out = {0.0, 0.0, false};
            FwdElim fe;
            fe.started = false;
            double prev_Snext = 0.0, tau = 0.0;
            for (;;) {
                const double dz = hi_advance(kp, it, line, ip, shift0, s_nxt, s_dZ, s_sh, s_m);
                Entry inner = hi_entry(it, ip);
                if (it.s < 0) inner.shift = shift_centre;
                PointEC ec_in = {0.0, 0.0, false};
                double dtau, S_out, S_in;
                segment<APPROX, PREC>(kp, s_erf, ol, outer, inner, dz, freq, line, ec_out, ec_in,
                                      dtau, S_out, S_in);
                if constexpr (KIND == IMG_OPTICAL_DEPTH)
                    tau += dtau; /* image_optical_depth (solver.tpp:2718-2758) */
                else
                    fe_row(kp, fe, dtau, S_out, outer, freq);
                prev_Snext = S_in;
                outer = inner;
                ec_out = ec_in;
                if (it.s < 0) break;
            }
            if constexpr (KIND == IMG_OPTICAL_DEPTH) {
                Ival = tau;
            } else {
                /* last row at the start point (image_feautrier_order_2, solver.tpp:2546-2559) */
                const double I_l = boundary_intensity(kp, ip, freq * shift_centre);
                Ival = 2.0 * fe_last(fe, prev_Snext, I_l) - boundary_intensity(kp, ip, nu_ip);
            }
        }
        I[(size_t)pix * F + f] = Ival;
    }
}

/* One pixel per wave, or two when the image has at most 32 frequencies: then the
   wave walks both pixels' rays and each half-wave takes one pixel's frequencies.
   Returns false when a walk outgrew the walk buffer. */
template <int APPROX, int PREC, int KIND>
__device__ bool image_pixels(const KParams &kp, const ImageFrame &fr, uint32_t pix0, uint32_t npix,
                             unsigned char *smem, double *I, double *ImX, double *ImY)
{
    const uint32_t lane = threadIdx.x & (MGX_WAVE - 1);
    PixelWalk w0 = pixel_slot(smem, kp.cap, 0);
    const bool ok0 = image_walk(kp, fr, pix0, w0, ImX, ImY);
    if (npix == 1) {
        if (!ok0) return false;
        image_freqs<APPROX, PREC, KIND>(kp, w0, pix0, lane, MGX_WAVE, I);
        return true;
    }
    /* the second ray is walked even when the first outgrew the buffer, so that
       its length is reported too */
    PixelWalk w1 = pixel_slot(smem, kp.cap, 1);
    const bool ok1 = image_walk(kp, fr, pix0 + 1, w1, ImX, ImY);
    if (!ok0 || !ok1) return false;
    const uint32_t half = lane >> 5;
    image_freqs<APPROX, PREC, KIND>(kp, half ? w1 : w0, pix0 + half, lane & 31, 32, I);
    return true;
}


/* one wave per ppw (1 or 2) consecutive pixels of one model's image */
template <int APPROX, int PREC, int KIND>
__global__ __launch_bounds__(MGX_WAVE) void image_kernel(const KParams *__restrict__ kpp, const ImageFrame *frp,
                                                         double *I, double *ImX, double *ImY, uint32_t ppw)
{
    mgx_ref_tables_init();
    extern __shared__ unsigned char smem[];
    const ImageFrame fr = *frp;
    const uint32_t npix_all = fr.nx * fr.ny, p0 = blockIdx.x * ppw;
    const uint32_t n = (p0 + ppw <= npix_all) ? ppw : npix_all - p0;
    (void)image_pixels<APPROX, PREC, KIND>(*kpp, fr, p0, n, smem, I, ImX, ImY);
}

/*
 * The images of a batch of models in one persistent launch: model i's pixels are
 * cut into chunks of kImageChunk consecutive pixels, the chunks of all models form
 * one work list (models by decreasing walk length, so the long rays start first),
 * and a wave takes the next chunk from an atomic counter and images its pixels
 * with image_pixel.  chunk_start[s] is the first chunk of the s-th model of that
 * order, slot_model[s] the model (n_slots + 1 entries of chunk_start).
 */
constexpr uint32_t kImageChunk = 4;

template <int APPROX, int PREC>
__global__ __launch_bounds__(MGX_WAVE) void image_batch_kernel(
    const KParams *__restrict__ kps, const ImageFrame *__restrict__ frames,
    const uint32_t *__restrict__ chunk_start, const uint32_t *__restrict__ slot_model,
    uint32_t n_slots, uint32_t *counter, double *const *__restrict__ I,
    double *const *__restrict__ ImX, double *const *__restrict__ ImY)
{
    mgx_ref_tables_init();
    extern __shared__ unsigned char smem[];
    __shared__ uint32_t s_item;
    const uint32_t lane = threadIdx.x;
    const uint32_t n_chunks = chunk_start[n_slots];
    for (;;) {
        __syncthreads();
        if (lane == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t w = s_item;
        if (w >= n_chunks) break;
        /* the slot s with chunk_start[s] <= w < chunk_start[s + 1] */
        uint32_t lo = 0, hi = n_slots;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (chunk_start[mid] <= w) lo = mid;
            else hi = mid;
        }
        const uint32_t mi = slot_model[lo];
        const KParams &kp = kps[mi];
        const ImageFrame fr = frames[mi];
        const uint32_t npix = fr.nx * fr.ny;
        const uint32_t p0 = (w - chunk_start[lo]) * kImageChunk;
        const uint32_t p1 = (p0 + kImageChunk < npix) ? p0 + kImageChunk : npix;
        const uint32_t ppw = kp.m.nfreqs <= 32 ? 2 : 1;
        for (uint32_t pix = p0; pix < p1; pix += ppw) {
            const uint32_t n = (pix + ppw <= p1) ? ppw : p1 - pix;
            if (!image_pixels<APPROX, PREC, IMG_INTENSITY>(kp, fr, pix, n, smem, I[mi], ImX[mi], ImY[mi]))
                return;
            __syncthreads();
        }
    }
}

/* The walked ray pair of one model point for the old imager: the rr half and the
   antipod half (steps, lengths, shifts, all-lines interpolation counts) in one
   slot of walk_lds_bytes(cap) bytes of LDS. */
struct RayPairWalk {
    uint32_t *nxt0, *nxt1;
    double *dZ0, *dZ1, *sh0, *sh1;
    uint16_t *m0, *m1;
    uint32_t o, S0, S1;
    double sh00, sh01;
};

__device__ inline RayPairWalk ray_pair_slot(unsigned char *smem, uint32_t cap, uint32_t slot)
{
    unsigned char *base = smem + (size_t)slot * walk_lds_bytes(cap);
    RayPairWalk w;
    w.nxt0 = reinterpret_cast<uint32_t *>(base);
    w.nxt1 = w.nxt0 + cap;
    w.dZ0 = reinterpret_cast<double *>(base + 8 * ((2 * (size_t)cap * 4 + 7) / 8));
    w.dZ1 = w.dZ0 + cap;
    w.sh0 = w.dZ1 + cap;
    w.sh1 = w.sh0 + cap;
    w.m0 = reinterpret_cast<uint16_t *>(w.sh1 + cap);
    w.m1 = w.m0 + cap;
    w.o = 0;
    w.S0 = w.S1 = 0;
    w.sh00 = w.sh01 = 0.0;
    return w;
}

/* The image-plane coordinates of point o (lane 0 writes them), then the walk of
   its ray pair into w and the interpolation counts; the whole wave takes part.
   Returns 0 or an ERR_* code. */
__device__ inline int ray_pair_walk(const KParams &kp, uint32_t rr, uint32_t ar, uint32_t o,
                                    RayPairWalk &w, double *ImX, double *ImY)
{
    const mgx_model &m = kp.m;
    const uint32_t lane = threadIdx.x & (MGX_WAVE - 1);
    if (lane == 0) {
        const PointRec po = gld(&kp.pts[o]);
        double X = po.x, Y = 0.0;
        if (m.dimension == 3) {
            const double rx = m.ray_direction[3 * (size_t)rr + 0];
            const double ry = m.ray_direction[3 * (size_t)rr + 1];
            const double rz = m.ray_direction[3 * (size_t)rr + 2];
            const double denominator = sqrt(rx * rx + ry * ry);
            if (denominator >= 1.0e-9) {
                const double inverse_denominator = 1.0 / denominator;
                const double ix = ry * inverse_denominator, iy = -rx * inverse_denominator;
                const double jx = rx * rz * inverse_denominator, jy = ry * rz * inverse_denominator;
                const double jz = -denominator;
                X = ix * po.x + iy * po.y;
                Y = jx * po.x + jy * po.y + jz * po.z;
            } else {
                Y = po.y;
            }
        } else if (m.dimension != 1) {
            X = 0.0;
        }
        ImX[o] = X;
        ImY[o] = Y;
    }
    w.o = o;
    int e = walk_half<true>(kp, o, rr, w.nxt0, w.dZ0, w.sh0, w.S0, w.sh00);
    if (!e) e = walk_half<true>(kp, o, ar, w.nxt1, w.dZ1, w.sh1, w.S1, w.sh01);
    if (e) return e;
    __syncthreads();
    for (uint32_t i = lane; i < w.S0 + w.S1; i += MGX_WAVE) {
        const uint32_t h = (i < w.S0) ? 0 : 1;
        const uint32_t st = h ? i - w.S0 : i;
        const uint32_t *nx = h ? w.nxt1 : w.nxt0;
        const uint32_t crt = (st == 0) ? o : nx[st - 1];
        const uint32_t mm = n_interp_all(kp, crt, nx[st]);
        if (mm > 0xffffu) atomicOr(&kp.status[0], (uint32_t)ERR_INTERP);
        (h ? w.m1 : w.m0)[st] = (uint16_t)(mm > 0xffffu ? 0xffffu : mm);
    }
    __syncthreads();
    return 0;
}

/*
 * The old imager at every model point along ray rr (Solver::image_feautrier_order_2
 * and Solver::image_optical_depth, solver.tpp:543-574, 656-686).  One wave per
 * ppw (1 or 2) consecutive points: each point's ray pair (rr, antipod) is walked
 * with the Rest-frame shifts of trace_ray<Rest, false> (reversed on rays >=
 * hnrays) and the all-lines interpolation count of every step, into its own LDS
 * slot; lanes take frequencies (with two points, each half-wave takes one
 * point's) and run over the whole ray from the rr end (first) to the antipod
 * end (last): the rr half inward, the origin, then the antipod half outward.
 * KIND selects I(o, f) = 2 Su[last] - B(nr[last], nu(o, f)) (the boundary term
 * at the unshifted frequency, as the reference) or the summed optical depth.
 * ImX, ImY: the point's coordinates on the image plane
 * (Image::set_coordinates_all_model_points, image.cpp:154-212).
 */
template <int APPROX, int PREC, int KIND>
__global__ __launch_bounds__(MGX_WAVE) void image_ray_kernel(const KParams *__restrict__ kpp,
                                                             uint32_t rr, double *I, double *ImX,
                                                             double *ImY, uint32_t ppw)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    const mgx_model &m = kp.m;
    const uint32_t lane = threadIdx.x;
    const uint32_t F = m.nfreqs;
    const uint32_t o0 = blockIdx.x * ppw;
    const uint32_t npt = (o0 + ppw <= m.npoints) ? ppw : m.npoints - o0;
    const uint32_t ar = gld(&m.antipod[rr]);

    RayPairWalk w0 = ray_pair_slot(smem, kp.cap, 0);
    RayPairWalk w1 = w0;
    int e = ray_pair_walk(kp, rr, ar, o0, w0, ImX, ImY);
    if (!e && npt == 2) {
        w1 = ray_pair_slot(smem, kp.cap, 1);
        e = ray_pair_walk(kp, rr, ar, o0 + 1, w1, ImX, ImY);
    }
    if (e) {
        if (lane == 0) {
            atomicOr(&kp.status[0], (uint32_t)e);
            if (e == ERR_CAP) atomicMax(&kp.status[1], kp.cap + 1);
        }
        return;
    }
    /* with two points, each half-wave takes one point's frequencies */
    const uint32_t half = (npt == 2) ? (lane >> 5) : 0;
    const uint32_t f0 = (npt == 2) ? (lane & 31) : lane;
    const uint32_t fstep = (npt == 2) ? 32 : MGX_WAVE;
    const RayPairWalk &w = half ? w1 : w0;
    const uint32_t o = w.o, S0 = w.S0, S1 = w.S1;
    const double sh00 = w.sh00, sh01 = w.sh01;
    const uint32_t *s_nxt0 = w.nxt0, *s_nxt1 = w.nxt1;
    const double *s_dZ0 = w.dZ0, *s_dZ1 = w.dZ1, *s_sh0 = w.sh0, *s_sh1 = w.sh1;
    const uint16_t *s_m0 = w.m0, *s_m1 = w.m1;

    const uint32_t last_nr = (S1 > 0) ? s_nxt1[S1 - 1] : o;
    for (uint32_t f = f0; f < F; f += fstep) {
        const double freq = m.nu[(size_t)o * F + f];
        double val;
        if (S0 + S1 == 0) {
            /* n_tot = 1 */
            val = (KIND == IMG_OPTICAL_DEPTH) ? 0.0 : boundary_intensity(kp, o, freq);
        } else {
            const uint32_t line = m.f_line[f];
            const uint32_t ub_idx = upper_bound_d(m.sorted_line, m.nlines, freq);
            const OwnLine ol = own_line(m, m.sorted_line_map[ub_idx > 0 ? ub_idx - 1 : 0]);
            FwdElim fe;
            fe.started = false;
            double tau = 0.0, S_last = 0.0;
            Entry prev;
            PointEC ec_prev = {0.0, 0.0, false};
            /* one segment from prev to nx: row prev of the elimination, or its depth */
            auto cross = [&](const Entry &nx, double dz) {
                PointEC ec_nx = {0.0, 0.0, false};
                double dtau, S_cu, S_nx;
                segment<APPROX, PREC>(kp, s_erf, ol, prev, nx, dz, freq, line, ec_prev, ec_nx,
                                      dtau, S_cu, S_nx);
                if constexpr (KIND == IMG_OPTICAL_DEPTH)
                    tau += dtau;
                else
                    fe_row(kp, fe, dtau, S_cu, prev, freq);
                S_last = S_nx;
                prev = nx;
                ec_prev = ec_nx;
            };
            if (S0 > 0) { /* the rr half from its far end inward to the origin */
                HalfIter it;
                it.s = (int)S0 - 1;
                hi_load_step(kp, it, line, o, sh00, s_nxt0, s_dZ0, s_sh0, s_m0);
                prev = hi_entry(it, o);
                for (;;) {
                    const double dz = hi_advance(kp, it, line, o, sh00, s_nxt0, s_dZ0, s_sh0, s_m0);
                    Entry nx = hi_entry(it, o);
                    if (it.s < 0) nx.shift = sh00;
                    cross(nx, dz);
                    if (it.s < 0) break;
                }
            } else {
                prev.nr = o;
                prev.nri = o;
                prev.ifac = 0.0;
                prev.shift = sh00;
            }
            if (S1 > 0) { /* the antipod half from the origin outward */
                HalfIter it;
                it.s = 0;
                hi_load_step(kp, it, line, o, sh01, s_nxt1, s_dZ1, s_sh1, s_m1);
                it.i = 1;
                for (;;) {
                    cross(hi_entry(it, o), it.dZ);
                    if (it.i < it.m_s) {
                        it.i++;
                    } else {
                        it.s++;
                        if (it.s >= (int)S1) break;
                        hi_load_step(kp, it, line, o, sh01, s_nxt1, s_dZ1, s_sh1, s_m1);
                        it.i = 1;
                    }
                }
            }
            if constexpr (KIND == IMG_OPTICAL_DEPTH) {
                val = tau;
            } else {
                const double I_l = boundary_intensity(kp, last_nr, freq * prev.shift);
                val = 2.0 * fe_last(fe, S_last, I_l) - boundary_intensity(kp, last_nr, freq);
            }
        }
        I[(size_t)o * F + f] = val;
    }
}

/* ------------------------------------------------------------------------- */
/* adjoint: vector-Jacobian products of J and images with respect to the line */
/* emissivity and opacity (P, L).  No reference counterpart.  The trace, the  */
/* branch of every segment, the CloseLines windows and the interpolation      */
/* counts are held fixed.                                                     */

/* Emissivity and opacity of an entry; phi is the line profile value when the
   entry has only its own line in the CloseLines window and no interpolation
   partner (then eta = phi eps and chi = chi_min + phi kappa), NaN otherwise. */
struct PointECT {
    double eta, chi, phi;
    bool valid;
};

template <int APPROX, int PREC>
__device__ inline void eta_chi_tape(const KParams &kp, const OwnLine &ol, const Entry &e,
                                    uint32_t line, double freq, PointECT &ec)
{
    ec.valid = true;
    ec.phi = __builtin_nan("");
    if (APPROX == MGX_APPROX_CLOSE_LINES && e.ifac == 0.0) {
        const double W = kp.op.max_distance_opacity_contribution * (freq * gld(&kp.pts[e.nr].ub));
        const double left = freq - W, right = freq + W;
        if (!(ol.a_g < left) && ol.a_lo < left && !(right < ol.a_g) && right < ol.a_hi) {
            /* the arithmetic of eta_chi_interp for a one-line window at t = 0 */
            const double4 a = lp(kp, e.nr, kp.m.sorted_line_map[ol.g]);
            const double prof = gaussian<PREC>(a.x, freq - ol.a_g);
            ec.eta = 0.0 + prof * a.y;
            ec.chi = kp.op.min_opacity + prof * a.z;
            ec.phi = prof;
            return;
        }
        if ((ol.a_g < left && right < ol.a_hi) || (right < ol.a_g && ol.a_lo < left)) {
            /* no line in the window: eta = 0, chi = chi_min (phi = 0 reproduces both) */
            ec.eta = 0.0;
            ec.chi = kp.op.min_opacity;
            ec.phi = 0.0;
            return;
        }
    }
    eta_chi_interp<APPROX, PREC>(kp, ol, e.nr, e.nri, e.ifac, line, freq, gld(&kp.pts[e.nr].ub),
                                 ec.eta, ec.chi);
}

/* Segment kinds recorded for the reverse pass */
#define MGX_SEG_ERF_GENERIC (-1.0)
#define MGX_SEG_TRAPEZOID (-2.0)
#define MGX_SEG_ERF_EMPTY (-3.0) /* no line in the window: no dependence on the lines */

/*
 * segment() for the adjoint's forward sweep: the same values, plus the kind of
 * the segment for the reverse pass in g: g >= 0 for a line-integrated segment
 * with only the own line in the window, no interpolation partner and no
 * clamp (g = d dtau / d mean-opacity), MGX_SEG_ERF_GENERIC for the other
 * line-integrated segments, MGX_SEG_TRAPEZOID for the trapezoid rule.
 */
template <int APPROX, int PREC>
__device__ inline void segment_tape(const KParams &kp, const double *erf_tab, const OwnLine &ol,
                                    const Entry &cu, const Entry &nx, double dZ, double freq,
                                    uint32_t line, PointECT &ec_cu, PointECT &ec_nx,
                                    double &dtau, double &S_cu, double &S_nx, double &g)
{
    const mgx_model &m = kp.m;
    const double dshift = nx.shift - cu.shift;
    const double dshift_abs = fabs(dshift);
    const double2 du_c = dsm_ub(kp, cu.nr), du_n = dsm_ub(kp, nx.nr);
    const double a = du_c.x, b = du_n.x;
    const double dshift_max = (b < a) ? b : a;
    const double cf = freq * cu.shift, nf = freq * nx.shift;
    if (dshift_abs > dshift_max) {
        if (APPROX == MGX_APPROX_CLOSE_LINES && cu.ifac == 0.0 && nx.ifac == 0.0) {
            double left_freq, right_freq;
            if (cf < nf) {
                left_freq = cf;
                right_freq = nf;
            } else {
                right_freq = cf;
                left_freq = nf;
            }
            const double cb = kp.op.max_distance_opacity_contribution * (right_freq * du_c.y);
            const double nb = kp.op.max_distance_opacity_contribution * (right_freq * du_n.y);
            const double W = (cb < nb) ? nb : cb;
            const double left = left_freq - W, right = right_freq + W;
            if (!(ol.a_g < left) && ol.a_lo < left && !(right < ol.a_g) && right < ol.a_hi) {
                /* the arithmetic of S_dtau_line_integrated for one line at t = 0 */
                const uint32_t l = m.sorted_line_map[ol.g];
                const double4 c = lp(kp, cu.nr, l), n = lp(kp, nx.nr, l);
                const double linefreq = m.line[l];
                const double avg_iw = (c.x + n.x) / 2.0;
                double line_dtau, dop;
                if (cf == nf) {
                    const double prof = gaussian<PREC>(avg_iw, cf - linefreq);
                    line_dtau = dZ * (prof * ((c.z + n.z) / 2.0) + kp.op.min_opacity);
                    dop = dZ * prof;
                } else {
                    const double next_pos = (linefreq - nf) * avg_iw;
                    const double curr_pos = (linefreq - cf) * avg_iw;
                    const double diff_pos = next_pos - curr_pos;
                    const double avg_op = (n.z + c.z) / 2.0;
                    const float erf_diff = f_erf<PREC>((float)next_pos, erf_tab) -
                                           f_erf<PREC>((float)curr_pos, erf_tab);
                    const double h = 0.5 * rcp(diff_pos);
                    const double erfterm = avg_op * (double)erf_diff * h;
                    line_dtau = dZ * fma(avg_iw, erfterm, kp.op.min_opacity);
                    dop = dZ * (avg_iw * ((double)erf_diff * h));
                }
                double d = 0.0 + line_dtau;
                const double sum_Sc = fma(line_dtau, c.w, 0.0), sum_Sn = fma(line_dtau, n.w, 0.0);
                const double bound_min_dtau = kp.op.min_opacity * dZ;
                const bool clamped = (-bound_min_dtau < d) && (d < bound_min_dtau);
                if (-bound_min_dtau < d) d = (bound_min_dtau < d) ? d : bound_min_dtau;
                dtau = d;
                const double rd = rcp(d);
                S_cu = sum_Sc * rd;
                S_nx = sum_Sn * rd;
                g = clamped ? MGX_SEG_ERF_GENERIC : dop;
                return;
            }
            if ((ol.a_g < left && right < ol.a_hi) || (right < ol.a_g && ol.a_lo < left)) {
                /* empty window: the sums are zero and the depth is clamped to its bound */
                double d = 0.0;
                const double bound_min_dtau = kp.op.min_opacity * dZ;
                if (-bound_min_dtau < d) d = (bound_min_dtau < d) ? d : bound_min_dtau;
                dtau = d;
                const double rd = rcp(d);
                S_cu = 0.0 * rd;
                S_nx = 0.0 * rd;
                g = MGX_SEG_ERF_EMPTY;
                return;
            }
        }
        S_dtau_line_integrated<APPROX, PREC>(kp, erf_tab, ol, cu.nr, nx.nr, cu.nri, nx.nri, line,
                                             cf, nf, cu.ifac, nx.ifac, dZ, du_c.y, du_n.y, dtau, S_cu,
                                             S_nx);
        g = MGX_SEG_ERF_GENERIC;
        return;
    }
    if (!ec_cu.valid) eta_chi_tape<APPROX, PREC>(kp, ol, cu, line, cf, ec_cu);
    if (!ec_nx.valid) eta_chi_tape<APPROX, PREC>(kp, ol, nx, line, nf, ec_nx);
    S_cu = ec_cu.eta * rcp(ec_cu.chi);
    S_nx = ec_nx.eta * rcp(ec_nx.chi);
    dtau = 0.5 * (ec_cu.chi + ec_nx.chi) * dZ;
    g = MGX_SEG_TRAPEZOID;
}

/* Tape of one half ray: segment k (counted from the outer end) stores five
   doubles, Y, X, dtau, the segment kind g and the profile value phi of its
   outer entry, each lane-interleaved (64 lanes side by side) within the
   segment's record.  The words stay doubles in mixed mode too: the reverse
   forms u_k - u_{k+1} from Y and X, which cancels where u is smooth, and the
   inverse-width gradient sums terms of both signs over the profile, so a float
   tape moved it by up to 6e-3 relative. */
#define MGX_TAPE_WORDS 5
struct Tape {
    /* bytes of one segment's record (all 64 lanes) */
    static constexpr size_t kRecord = (size_t)MGX_WAVE * 8 * MGX_TAPE_WORDS;
    static __device__ __forceinline__ size_t offset(uint32_t k, uint32_t c, uint32_t lane)
    {
        return kRecord * k + (size_t)MGX_WAVE * 8 * c + 8 * lane;
    }
    static __device__ __forceinline__ void put(unsigned char *base, uint32_t k, uint32_t c,
                                               uint32_t lane, double v)
    {
        *reinterpret_cast<double *>(base + offset(k, c, lane)) = v;
    }
    static __device__ __forceinline__ double get(const unsigned char *base, uint32_t k, uint32_t c,
                                                 uint32_t lane)
    {
        return *reinterpret_cast<const double *>(base + offset(k, c, lane));
    }
};

template <int PREC>
struct SweepStore {
    unsigned char *st; /* the half's tape (records of all 64 lanes) */
    uint32_t lane;
    uint32_t cap; /* segments that fit */
    uint32_t n;   /* segments of the half (may exceed cap: nothing is stored past it) */
    double I_b;   /* boundary intensity of the outer boundary row */
};

/* sweep_half of the forward kernel, recording the tape */
template <int APPROX, int PREC>
__device__ inline void sweep_half_tape(const KParams &kp, const double *s_erf, const OwnLine &ol,
                                       bool backward, uint32_t S, uint32_t o, double shift0,
                                       double shift_centre, const uint32_t *s_nxt,
                                       const double *s_dZ, const double *s_sh,
                                       const uint16_t *mtab, uint32_t l, double freq, Sweep &sw,
                                       PointECT &ec_centre, SweepStore<PREC> &store)
{
    HalfIter it;
    it.s = (int)S - 1;
    hi_load_step(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
    Entry outer = hi_entry(it, o);
    PointECT ec_out = {0.0, 0.0, 0.0, false};
    sw.started = false;
    store.n = 0;
    for (;;) {
        const double dz = hi_advance(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
        Entry inner = hi_entry(it, o);
        if (it.s < 0) inner.shift = shift_centre;
        PointECT ec_in = (it.s < 0) ? ec_centre : PointECT{0.0, 0.0, 0.0, false};
        const Entry &cu = backward ? inner : outer;
        const Entry &nx = backward ? outer : inner;
        PointECT &ec_cu = backward ? ec_in : ec_out;
        PointECT &ec_nx = backward ? ec_out : ec_in;
        double dtau, S_cu, S_nx, g;
        segment_tape<APPROX, PREC>(kp, s_erf, ol, cu, nx, dz, freq, l, ec_cu, ec_nx, dtau, S_cu,
                                   S_nx, g);
        const double S_out = backward ? S_nx : S_cu;
        const double S_in = backward ? S_cu : S_nx;
        const double S_row = backward ? (sw.started ? sw.S_prev_row : S_out) : S_out;
        if (!sw.started) {
            const double I_b = boundary_intensity(kp, outer.nr, freq * outer.shift);
            store.I_b = I_b;
            const double idt = rcp(dtau);
            const double B_min = fma(2.0, idt, 1.0);
            const double B = fma(2.0 * idt, idt, B_min);
            sw.Y = fma(2.0 * I_b, idt, S_row) * rcp(B);
            sw.X = 0.5 * B_min * dtau * dtau;
            sw.XI = rcp(1.0 + sw.X);
            sw.started = true;
        } else {
            const double avg = 0.5 * (sw.prev_dtau + dtau);
            const double q_out = avg * sw.prev_dtau;
            const double q_in = avg * dtau;
            const double r = rcp(q_out);
            sw.X = fma(r * sw.X, sw.XI, 1.0) * q_in;
            sw.XI = rcp(1.0 + sw.X);
            sw.Y = fma(r, sw.Y, S_row) * sw.XI * q_in;
        }
        sw.prev_dtau = dtau;
        sw.S_inner = S_in;
        sw.S_prev_row = S_in;
        if (store.n < store.cap) {
            typedef Tape T;
            T::put(store.st, store.n, 0, store.lane, sw.Y);
            T::put(store.st, store.n, 1, store.lane, sw.X);
            T::put(store.st, store.n, 2, store.lane, dtau);
            T::put(store.st, store.n, 3, store.lane, g);
            T::put(store.st, store.n, 4, store.lane, ec_out.valid ? ec_out.phi : __builtin_nan(""));
        }
        store.n++;
        outer = inner;
        ec_out = ec_in;
        if (it.s < 0) break;
    }
    ec_centre = ec_out;
}

/*
 * Gradient sink of one lane.  The outward walk of the reverse pass visits the
 * segments of walk step s (outer entry in step s) together, and all of them
 * touch only the two points crt_s and nxt_s; the own-line contributions for
 * those two points are summed in registers, and when the walk moves to step
 * s+1 the finished point (crt_s) is summed over the lanes that share the line
 * (width lanes, aligned) and added by one lane.  Contributions to any other
 * line are added by the lane directly.
 */
struct GradSink {
    double *g_eps, *g_kap;
    uint32_t L;
    uint32_t line;   /* own line of the reducing lanes */
    uint32_t width;  /* lanes summed together (32 or 64), 0: no reduction */
    uint32_t p_crt, p_nxt;
    double e_crt, k_crt, e_nxt, k_nxt;
    double *g_iw;    /* inverse-width gradient (IW instantiations only) */
    double w_crt, w_nxt;
};

/* IW: the sink also carries the gradient with respect to the inverse line widths */
template <bool IW = false>
__device__ inline void sink_put(const GradSink &g, uint32_t p, uint32_t l, double de, double dk,
                                double dw = 0.0)
{
    const size_t i = (size_t)p * g.L + l;
    if (de != 0.0) unsafeAtomicAdd(&g.g_eps[i], de);
    if (dk != 0.0) unsafeAtomicAdd(&g.g_kap[i], dk);
    if (IW && dw != 0.0) unsafeAtomicAdd(&g.g_iw[i], dw);
}

template <bool IW = false>
__device__ inline void sink_add(GradSink &g, uint32_t p, uint32_t l, double de, double dk,
                                double dw = 0.0)
{
    if (l != g.line) {
        sink_put<IW>(g, p, l, de, dk, dw);
    } else if (p == g.p_nxt) {
        g.e_nxt += de;
        g.k_nxt += dk;
        if (IW) g.w_nxt += dw;
    } else {
        g.e_crt += de;
        g.k_crt += dk;
        if (IW) g.w_crt += dw;
    }
}

/* The sum of v over the aligned group of width lanes (2, 4, ..., 64) that holds
   the lane, valid in the group's last lane; all lanes of the group active.
   Within a row of 16 the partners come through DPP (lane ^ 1, lane ^ 2, then the
   mirror images within 8 and 16 lanes, which pair the partial sums the same
   way); rows are then joined by DPP broadcasts, row_bcast15 (0x142: lane 15 of
   row r into row r + 1, here rows 1 and 3) and row_bcast31 (0x143: lane 31 into
   rows 2 and 3), which leave the sum in the group's last row. */
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_rows(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWS, 0xf, false);
    const uint64_t hi =
        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, (hi << 32) | lo);
}

__device__ __forceinline__ double group_sum(double v, uint32_t width)
{
    v += dpp_get<0xb1>(v); /* quad_perm [1, 0, 3, 2] */
    if (width > 2) v += dpp_get<0x4e>(v); /* quad_perm [2, 3, 0, 1] */
    if (width > 4) v += dpp_get<0x141>(v); /* row_half_mirror */
    if (width > 8) v += dpp_get<0x140>(v); /* row_mirror */
    if (width > 16) v += dpp_rows<0x142, 0xa>(v); /* rows 1, 3 += lane 15, 47 */
    if (width > 32) v += dpp_rows<0x143, 0xc>(v); /* rows 2, 3 += lane 31 */
    return v;
}

/* all lanes of a reduction group call this together */
template <bool IW = false>
__device__ inline void sink_flush(const GradSink &g, uint32_t p, double e, double k, double w)
{
    if (g.width == 0) {
        sink_put<IW>(g, p, g.line, e, k, w);
        return;
    }
    e = group_sum(e, g.width);
    k = group_sum(k, g.width);
    if (IW) w = group_sum(w, g.width);
    if ((threadIdx.x & (g.width - 1)) == g.width - 1) sink_put<IW>(g, p, g.line, e, k, w);
}

/* the walk moves outward to a step whose points are (crt = old nxt, nxt) */
template <bool IW = false>
__device__ inline void sink_step(GradSink &g, uint32_t nxt)
{
    sink_flush<IW>(g, g.p_crt, g.e_crt, g.k_crt, g.w_crt);
    g.p_crt = g.p_nxt;
    g.e_crt = g.e_nxt;
    g.k_crt = g.k_nxt;
    g.w_crt = g.w_nxt;
    g.p_nxt = nxt;
    g.e_nxt = 0.0;
    g.k_nxt = 0.0;
    g.w_nxt = 0.0;
}

template <bool IW = false>
__device__ inline void sink_finish(GradSink &g)
{
    sink_flush<IW>(g, g.p_crt, g.e_crt, g.k_crt, g.w_crt);
    sink_flush<IW>(g, g.p_nxt, g.e_nxt, g.k_nxt, g.w_nxt);
}

/* partial derivatives of interpolate_log(a, b, t) = a^(1-t) b^t with value r */
__device__ inline void interp_log_partials(double a, double b, double t, double r, double &da,
                                           double &db)
{
    da = (t == 1.0) ? 0.0 : (1.0 - t) * r / a;
    db = (t == 0.0) ? 0.0 : t * r / b;
}

/* adds (de, dk) to the gradients of (p, l) directly (the rare paths of the reverse
   pass, which do not go through a lane's sink) */
__device__ inline void grad_put(double *g_eps, double *g_kap, uint32_t L, uint32_t p, uint32_t l,
                                double de, double dk)
{
    const size_t i = (size_t)p * L + l;
    if (de != 0.0) unsafeAtomicAdd(&g_eps[i], de);
    if (dk != 0.0) unsafeAtomicAdd(&g_kap[i], dk);
}

/* reverse of interpolate_linear(iw(p), iw(pi), t): partials 1 - t and t */
__device__ inline void iw_put(double *g_iw, uint32_t L, uint32_t p, uint32_t pi, uint32_t l,
                              double t, double bar)
{
    const double a = (1.0 - t) * bar;
    if (a != 0.0) unsafeAtomicAdd(&g_iw[(size_t)p * L + l], a);
    if (t != 0.0) {
        const double b = t * bar;
        if (b != 0.0) unsafeAtomicAdd(&g_iw[(size_t)pi * L + l], b);
    }
}

/* d(gaussian(w, diff))/dw = exp(-x^2) (1 - 2 x^2) / sqrt(pi), x = w diff: the
   derivative of the smooth profile, in double */
__device__ inline double gaussian_dw(double w, double diff)
{
    const double x = w * diff;
    return MGX_INVERSE_SQRT_PI * exp(-x * x) * (1.0 - 2.0 * x * x);
}

/* d(compute_dtau_single_line)/d(avg_iw) at fixed frequencies, of the smooth
   function: dz avg_op d(profile)/dw when the two frequencies are equal, else the
   derivative of dz avg_op (erf(np) - erf(cp)) / (2 (cf - nf)), np = (lf - nf) w,
   cp = (lf - cf) w: dz avg_op (np exp(-np^2) - cp exp(-cp^2)) / (sqrt(pi) (np - cp)) */
__device__ inline double dtau_line_dw(double avg_iw, double avg_op, double linefreq, double cf,
                                      double nf, double dz)
{
    if (cf == nf) return dz * avg_op * gaussian_dw(avg_iw, cf - linefreq);
    const double np = (linefreq - nf) * avg_iw, cp = (linefreq - cf) * avg_iw;
    return dz * avg_op * MGX_INVERSE_SQRT_PI * (np * exp(-np * np) - cp * exp(-cp * cp)) /
           (np - cp);
}

/* the inverse-width part of the reverse of compute_dtau_single_line; bar is the
   adjoint of the line's optical depth */
template <int PREC>
__device__ inline void dtau_iw_bwd(double *g_iw, uint32_t L, uint32_t cp, uint32_t cpi,
                                   uint32_t np, uint32_t npi, uint32_t l, double4 c, double4 cI,
                                   double4 n, double4 nI, double op_c, double op_n, double linefreq,
                                   double cf, double nf, double ct, double nt, double dz, double bar)
{
    if (bar == 0.0) return;
    const double avg_iw = (interp_linear(c.x, cI.x, ct) + interp_linear(n.x, nI.x, nt)) / 2.0;
    const double avg_op = (op_c + op_n) / 2.0;
    const double wb = 0.5 * bar * dtau_line_dw(avg_iw, avg_op, linefreq, cf, nf, dz);
    iw_put(g_iw, L, cp, cpi, l, ct, wb);
    iw_put(g_iw, L, np, npi, l, nt, wb);
}

/* reverse of eta_chi_interp: eta_bar d(eta) + chi_bar d(chi), added directly
   (and, with IW, the inverse widths' part into g_iw).
   A rare path of the reverse pass (interpolated entries, several lines in the
   window): kept out of line, with every argument by value so that the
   caller's registers stay out of memory. */
template <int APPROX, int PREC, bool IW = false>
__device__ __attribute__((noinline)) void eta_chi_interp_bwd(
    const KParams *kpp, OwnLine ol, uint32_t p1, uint32_t p2, double t, uint32_t ll, double freq,
    double eta_bar, double chi_bar, double *g_eps, double *g_kap, double *g_iw)
{
    const KParams &kp = *kpp;
    const mgx_model &m = kp.m;
    uint32_t lo = 0, hi = m.nlines;
    if (APPROX == MGX_APPROX_ONE_LINE) {
        lo = ll;
        hi = ll + 1;
    } else if (APPROX == MGX_APPROX_CLOSE_LINES) {
        const double W = kp.op.max_distance_opacity_contribution * (freq * gld(&kp.pts[p1].ub));
        line_window(m, ol, freq - W, freq + W, lo, hi);
    }
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = (APPROX == MGX_APPROX_CLOSE_LINES) ? m.sorted_line_map[s] : s;
        const double diff = freq - m.line[l];
        const double4 a = lp(kp, p1, l), b = lp(kp, p2, l);
        const double iw = interp_linear(a.x, b.x, t);
        const double prof = gaussian<PREC>(iw, diff);
        const double ei = interp_log<PREC>(a.y, b.y, t), ki = interp_log<PREC>(a.z, b.z, t);
        double de1, de2, dk1, dk2;
        interp_log_partials(a.y, b.y, t, ei, de1, de2);
        interp_log_partials(a.z, b.z, t, ki, dk1, dk2);
        const double pe = prof * eta_bar, pk = prof * chi_bar;
        grad_put(g_eps, g_kap, m.nlines, p1, l, pe * de1, pk * dk1);
        if (t != 0.0) grad_put(g_eps, g_kap, m.nlines, p2, l, pe * de2, pk * dk2);
        /* eta = prof eps_i, chi = chi_min + prof kap_i */
        if (IW) iw_put(g_iw, m.nlines, p1, p2, l, t, gaussian_dw(iw, diff) * (eta_bar * ei + chi_bar * ki));
    }
}

/* d(compute_dtau_single_line)/d(avg_op) */
template <int PREC>
__device__ double dtau_single_line_dop(const double *erf_tab, double4 c, double4 cI, double4 n,
                                       double4 nI, double linefreq, double curr_freq,
                                       double next_freq, double curr_interp, double next_interp,
                                       double dz)
{
    const double iw_c = interp_linear(c.x, cI.x, curr_interp);
    const double iw_n = interp_linear(n.x, nI.x, next_interp);
    const double avg_iw = (iw_c + iw_n) / 2.0;
    if (curr_freq == next_freq) return dz * gaussian<PREC>(avg_iw, curr_freq - linefreq);
    const double next_pos = (linefreq - next_freq) * avg_iw;
    const double curr_pos = (linefreq - curr_freq) * avg_iw;
    const double diff_pos = next_pos - curr_pos;
    const float erf_diff = f_erf<PREC>((float)next_pos, erf_tab) - f_erf<PREC>((float)curr_pos, erf_tab);
    return dz * (avg_iw * ((double)erf_diff * (0.5 * rcp(diff_pos))));
}

/* reverse of the line terms at one end of a segment: opacity (kappa interpolation,
   value op) and line source (interpolation of eps/kappa, value S), added directly */
__device__ inline void line_end_bwd(double *g_eps, double *g_kap, uint32_t L, uint32_t p,
                                    uint32_t pi, uint32_t l, double t, double4 a, double4 b,
                                    double op, double S, double op_bar, double S_bar)
{
    double dk1, dk2, dR1, dR2;
    interp_log_partials(a.z, b.z, t, op, dk1, dk2);
    interp_log_partials(a.w, b.w, t, S, dR1, dR2);
    /* R = eps/kappa: dR/deps = 1/kappa, dR/dkappa = -R/kappa */
    const double r1 = S_bar * dR1 * rcp(a.z);
    grad_put(g_eps, g_kap, L, p, l, r1, op_bar * dk1 - r1 * a.w);
    if (t != 0.0) {
        const double r2 = S_bar * dR2 * rcp(b.z);
        grad_put(g_eps, g_kap, L, pi, l, r2, op_bar * dk2 - r2 * b.w);
    }
}

/* reverse of S_dtau_line_integrated (the general case: out of line, arguments by
   value, gradients added directly) */
template <int APPROX, int PREC, bool IW = false>
__device__ __attribute__((noinline)) void S_dtau_line_integrated_bwd(
    const KParams *kpp, const double *erf_tab, OwnLine ol, uint32_t cp, uint32_t np, uint32_t cpi,
    uint32_t npi, uint32_t lineidx, double currfreq, double nextfreq, double curr_interp,
    double next_interp, double dZ, double dtau_bar, double Sc_bar, double Sn_bar, double *g_eps,
    double *g_kap, double *g_iw)
{
    const KParams &kp = *kpp;
    const mgx_model &m = kp.m;
    const uint32_t L = m.nlines;
    if (APPROX == MGX_APPROX_ONE_LINE) {
        const double4 c = lp(kp, cp, lineidx), cI = lp(kp, cpi, lineidx);
        const double4 n = lp(kp, np, lineidx), nI = lp(kp, npi, lineidx);
        const double dop = dtau_single_line_dop<PREC>(erf_tab, c, cI, n, nI, m.line[lineidx],
                                                      currfreq, nextfreq, curr_interp,
                                                      next_interp, dZ);
        const double op_c = interp_log<PREC>(c.z, cI.z, curr_interp);
        const double op_n = interp_log<PREC>(n.z, nI.z, next_interp);
        line_end_bwd(g_eps, g_kap, L, cp, cpi, lineidx, curr_interp, c, cI, op_c,
                     interp_log<PREC>(c.w, cI.w, curr_interp), 0.5 * dop * dtau_bar, Sc_bar);
        line_end_bwd(g_eps, g_kap, L, np, npi, lineidx, next_interp, n, nI, op_n,
                     interp_log<PREC>(n.w, nI.w, next_interp), 0.5 * dop * dtau_bar, Sn_bar);
        if (IW)
            dtau_iw_bwd<PREC>(g_iw, L, cp, cpi, np, npi, lineidx, c, cI, n, nI, op_c, op_n,
                              m.line[lineidx], currfreq, nextfreq, curr_interp, next_interp, dZ,
                              dtau_bar);
        return;
    }
    uint32_t lo = 0, hi = m.nlines;
    if (APPROX == MGX_APPROX_CLOSE_LINES) {
        double left_freq, right_freq;
        if (currfreq < nextfreq) {
            left_freq = currfreq;
            right_freq = nextfreq;
        } else {
            right_freq = currfreq;
            left_freq = nextfreq;
        }
        const double cb = kp.op.max_distance_opacity_contribution * (right_freq * gld(&kp.pts[cp].ub));
        const double nb = kp.op.max_distance_opacity_contribution * (right_freq * gld(&kp.pts[np].ub));
        const double W = (cb < nb) ? nb : cb;
        line_window(m, ol, left_freq - W, right_freq + W, lo, hi);
    }
    if (hi == lo + 1) {
        /* one line in the window: a single pass (the erf difference serves the
           optical depth and its derivative) */
        const uint32_t l = (APPROX == MGX_APPROX_CLOSE_LINES) ? m.sorted_line_map[lo] : lo;
        const double4 c = lp(kp, cp, l), cI = lp(kp, cpi, l);
        const double4 n = lp(kp, np, l), nI = lp(kp, npi, l);
        const double linefreq = m.line[l];
        const double iw_c = interp_linear(c.x, cI.x, curr_interp);
        const double iw_n = interp_linear(n.x, nI.x, next_interp);
        const double avg_iw = (iw_c + iw_n) / 2.0;
        const double op_c = interp_log<PREC>(c.z, cI.z, curr_interp);
        const double op_n = interp_log<PREC>(n.z, nI.z, next_interp);
        double line_dtau, dop;
        if (currfreq == nextfreq) {
            const double prof = gaussian<PREC>(avg_iw, currfreq - linefreq);
            line_dtau = dZ * (prof * ((op_c + op_n) / 2.0) + kp.op.min_opacity);
            dop = dZ * prof;
        } else {
            const double next_pos = (linefreq - nextfreq) * avg_iw;
            const double curr_pos = (linefreq - currfreq) * avg_iw;
            const double diff_pos = next_pos - curr_pos;
            const double avg_op = (op_n + op_c) / 2.0;
            const float erf_diff =
                f_erf<PREC>((float)next_pos, erf_tab) - f_erf<PREC>((float)curr_pos, erf_tab);
            const double ed = (double)erf_diff * (0.5 * rcp(diff_pos));
            line_dtau = dZ * fma(avg_iw, avg_op * ed, kp.op.min_opacity);
            dop = dZ * (avg_iw * ed);
        }
        const double Sc = interp_log<PREC>(c.w, cI.w, curr_interp);
        const double Sn = interp_log<PREC>(n.w, nI.w, next_interp);
        double d = line_dtau;
        bool clamped = false;
        if (APPROX == MGX_APPROX_CLOSE_LINES) {
            const double bound_min_dtau = kp.op.min_opacity * dZ;
            if (-bound_min_dtau < d && d < bound_min_dtau) {
                d = bound_min_dtau;
                clamped = true;
            }
        }
        const double rd = rcp(d);
        const double sSc_bar = Sc_bar * rd, sSn_bar = Sn_bar * rd;
        /* one line: unclamped, S = S_l exactly and the adjoint of the line depth is
           dtau_bar; clamped, the depth is the bound and S scales with the line depth */
        const double ld_bar = clamped ? sSc_bar * Sc + sSn_bar * Sn : dtau_bar;
        line_end_bwd(g_eps, g_kap, L, cp, cpi, l, curr_interp, c, cI, op_c, Sc, 0.5 * dop * ld_bar,
                     sSc_bar * line_dtau);
        line_end_bwd(g_eps, g_kap, L, np, npi, l, next_interp, n, nI, op_n, Sn, 0.5 * dop * ld_bar,
                     sSn_bar * line_dtau);
        if (IW)
            dtau_iw_bwd<PREC>(g_iw, L, cp, cpi, np, npi, l, c, cI, n, nI, op_c, op_n, linefreq,
                              currfreq, nextfreq, curr_interp, next_interp, dZ, ld_bar);
        return;
    }
    /* forward sums (the same operations as S_dtau_line_integrated) */
    double sum_dtau = 0.0, sum_Sc = 0.0, sum_Sn = 0.0;
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = (APPROX == MGX_APPROX_CLOSE_LINES) ? m.sorted_line_map[s] : s;
        const double4 c = lp(kp, cp, l), cI = lp(kp, cpi, l);
        const double4 n = lp(kp, np, l), nI = lp(kp, npi, l);
        const double line_dtau = dtau_single_line<PREC>(kp, erf_tab, c, cI, n, nI, m.line[l],
                                                        currfreq, nextfreq, curr_interp,
                                                        next_interp, dZ);
        sum_dtau += line_dtau;
        sum_Sc = fma(line_dtau, interp_log<PREC>(c.w, cI.w, curr_interp), sum_Sc);
        sum_Sn = fma(line_dtau, interp_log<PREC>(n.w, nI.w, next_interp), sum_Sn);
    }
    double d = sum_dtau;
    bool clamped = false;
    if (APPROX == MGX_APPROX_CLOSE_LINES) {
        const double bound_min_dtau = kp.op.min_opacity * dZ;
        if (-bound_min_dtau < d && d < bound_min_dtau) {
            d = bound_min_dtau;
            clamped = true;
        }
    }
    const double rd = rcp(d);
    const double sSc_bar = Sc_bar * rd, sSn_bar = Sn_bar * rd;
    const double sdt_bar =
        clamped ? 0.0 : dtau_bar - sSc_bar * (sum_Sc * rd) - sSn_bar * (sum_Sn * rd);
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = (APPROX == MGX_APPROX_CLOSE_LINES) ? m.sorted_line_map[s] : s;
        const double4 c = lp(kp, cp, l), cI = lp(kp, cpi, l);
        const double4 n = lp(kp, np, l), nI = lp(kp, npi, l);
        const double line_dtau = dtau_single_line<PREC>(kp, erf_tab, c, cI, n, nI, m.line[l],
                                                        currfreq, nextfreq, curr_interp,
                                                        next_interp, dZ);
        const double Sc = interp_log<PREC>(c.w, cI.w, curr_interp);
        const double Sn = interp_log<PREC>(n.w, nI.w, next_interp);
        const double op_c = interp_log<PREC>(c.z, cI.z, curr_interp);
        const double op_n = interp_log<PREC>(n.z, nI.z, next_interp);
        const double ld_bar = sdt_bar + sSc_bar * Sc + sSn_bar * Sn;
        const double dop = dtau_single_line_dop<PREC>(erf_tab, c, cI, n, nI, m.line[l], currfreq,
                                                      nextfreq, curr_interp, next_interp, dZ);
        line_end_bwd(g_eps, g_kap, L, cp, cpi, l, curr_interp, c, cI, op_c, Sc, 0.5 * dop * ld_bar,
                     sSc_bar * line_dtau);
        line_end_bwd(g_eps, g_kap, L, np, npi, l, next_interp, n, nI, op_n, Sn, 0.5 * dop * ld_bar,
                     sSn_bar * line_dtau);
        if (IW)
            dtau_iw_bwd<PREC>(g_iw, L, cp, cpi, np, npi, l, c, cI, n, nI, op_c, op_n, m.line[l],
                              currfreq, nextfreq, curr_interp, next_interp, dZ, ld_bar);
    }
}

/* emissivity/opacity of an entry, its profile value (NaN: not of the simple
   kind) and the adjoints collected for them from the trapezoid segments on
   either side (one back-propagation per entry) */
struct EntryBar {
    double eta, chi, phi, eta_bar, chi_bar;
    bool have;
};

template <int APPROX, int PREC>
__device__ __attribute__((noinline)) double2 eta_chi_generic(const KParams *kpp, OwnLine ol,
                                                             uint32_t nr, uint32_t nri,
                                                             double ifac, double f,
                                                             uint32_t line)
{
    double2 r;
    eta_chi_interp<APPROX, PREC>(*kpp, ol, nr, nri, ifac, line, f, gld(&kpp->pts[nr].ub), r.x, r.y);
    return r;
}

/* eta and chi of an entry for the reverse pass: from the recorded profile value
   when there is one (the forward arithmetic of eta_chi_tape), else recomputed */
template <int APPROX, int PREC>
__device__ inline void entry_eta_chi(const KParams &kp, const OwnLine &ol, const Entry &e,
                                     double freq, uint32_t line, uint32_t lown, EntryBar &eb)
{
    if (eb.have) return;
    if (eb.phi == eb.phi) {
        const double4 a = lp(kp, e.nr, lown);
        eb.eta = 0.0 + eb.phi * a.y;
        eb.chi = kp.op.min_opacity + eb.phi * a.z;
    } else {
        const double2 r =
            eta_chi_generic<APPROX, PREC>(&kp, ol, e.nr, e.nri, e.ifac, freq * e.shift, line);
        eb.eta = r.x;
        eb.chi = r.y;
    }
    eb.have = true;
}

/* back-propagates an entry's collected (eta, chi) adjoints */
template <int APPROX, int PREC, bool IW = false>
__device__ inline void entry_bwd(const KParams &kp, const OwnLine &ol, const Entry &e, double freq,
                                 uint32_t line, uint32_t lown, const EntryBar &eb, GradSink &g)
{
    if (eb.eta_bar == 0.0 && eb.chi_bar == 0.0) return;
    if (eb.phi == eb.phi) {
        /* eta = phi eps, chi = chi_min + phi kappa at this entry's own point, phi the
           profile of its inverse width (phi = 0: no line in the window) */
        double dw = 0.0;
        if (IW && eb.phi != 0.0) {
            const double4 a = lp(kp, e.nr, lown);
            dw = gaussian_dw(a.x, freq * e.shift - ol.a_g) * (a.y * eb.eta_bar + a.z * eb.chi_bar);
        }
        sink_add<IW>(g, e.nr, lown, eb.phi * eb.eta_bar, eb.phi * eb.chi_bar, dw);
    } else {
        eta_chi_interp_bwd<APPROX, PREC, IW>(&kp, ol, e.nr, e.nri, e.ifac, line, freq * e.shift,
                                             eb.eta_bar, eb.chi_bar, g.g_eps, g.g_kap, g.g_iw);
    }
}

/*
 * Reverse of one segment.  dtau_bar, Sc_bar, Sn_bar are the adjoints of its
 * optical depth and of its two source values; kind is the tape's g.
 * Line-integrated segments scatter their gradients directly (from the recorded
 * d dtau / d mean-opacity when g >= 0, else by recomputation); trapezoid
 * segments add the adjoints of (eta, chi) at their two entries to ecu / enx.
 */
template <int APPROX, int PREC, bool IW = false>
__device__ void segment_bwd(const KParams &kp, const double *erf_tab, const OwnLine &ol,
                            const Entry &cu, const Entry &nx, double dZ, double freq,
                            uint32_t line, uint32_t lown, double d, double kind, double dtau_bar,
                            double Sc_bar, double Sn_bar, GradSink &g, EntryBar &ecu,
                            EntryBar &enx)
{
    if (kind >= 0.0) {
        /* one line, no interpolation partner, not clamped: dtau = line_dtau = d */
        const double4 c = lp(kp, cu.nr, lown), n = lp(kp, nx.nr, lown);
        const double rd = rcp(d);
        const double sSc_bar = Sc_bar * rd, sSn_bar = Sn_bar * rd;
        /* the segment's sources are the line's own (S = d S_l / d): the adjoint of the
           line depth is dtau_bar (the source terms cancel exactly) */
        const double ld_bar = dtau_bar;
        const double op_bar = 0.5 * kind * ld_bar;
        const double rc = sSc_bar * d * rcp(c.z), rn = sSn_bar * d * rcp(n.z);
        double wb = 0.0;
        if (IW && ld_bar != 0.0)
            wb = 0.5 * ld_bar * dtau_line_dw((c.x + n.x) / 2.0, (n.z + c.z) / 2.0, ol.a_g,
                                             freq * cu.shift, freq * nx.shift, dZ);
        sink_add<IW>(g, cu.nr, lown, rc, op_bar - rc * c.w, wb);
        sink_add<IW>(g, nx.nr, lown, rn, op_bar - rn * n.w, wb);
        return;
    }
    if (kind == MGX_SEG_ERF_EMPTY) return;
    const double cf = freq * cu.shift, nf = freq * nx.shift;
    if (kind == MGX_SEG_ERF_GENERIC) {
        S_dtau_line_integrated_bwd<APPROX, PREC, IW>(&kp, erf_tab, ol, cu.nr, nx.nr, cu.nri,
                                                     nx.nri, line, cf, nf, cu.ifac, nx.ifac, dZ,
                                                     dtau_bar, Sc_bar, Sn_bar, g.g_eps, g.g_kap,
                                                     g.g_iw);
        return;
    }
    entry_eta_chi<APPROX, PREC>(kp, ol, cu, freq, line, lown, ecu);
    entry_eta_chi<APPROX, PREC>(kp, ol, nx, freq, line, lown, enx);
    const double ric = rcp(ecu.chi), rin = rcp(enx.chi);
    ecu.chi_bar += 0.5 * dZ * dtau_bar - Sc_bar * ecu.eta * ric * ric;
    enx.chi_bar += 0.5 * dZ * dtau_bar - Sn_bar * enx.eta * rin * rin;
    ecu.eta_bar += Sc_bar * ric;
    enx.eta_bar += Sn_bar * rin;
}

/* derivative of a row  u + A_a (u - u_a) + A_b (u - u_b) = S  with respect to d_a,
   A_a = 2/((d_a + d_b) d_a), A_b = 2/((d_a + d_b) d_b); Da = u - u_a, Db = u - u_b */
__device__ inline double row_dd(double da, double db, double Da, double Db)
{
    const double s = da + db;
    const double rs = rcp(s);
    const double Aa = 2.0 * rs * rcp(da), Ab = 2.0 * rs * rcp(db);
    return -Aa * (rs + rcp(da)) * Da - Ab * rs * Db;
}

/* derivative of a boundary row  (1 + 2/d) u + (2/d^2)(u - u_n) = S + 2 I/d  w.r.t. d */
__device__ inline double bdy_row_dd(double d, double I, double u, double Dn)
{
    const double r = rcp(d);
    return 2.0 * r * r * (I - u) - 4.0 * r * r * r * Dn;
}

/*
 * Reverse pass over one half ray, outward from the centre.  With the system
 * T u = r of the ray (rows as in the reference's elimination) and the seed
 * ubar at the centre, lam = T^-T e_c ubar = D T^-1 D^-1 e_c ubar (D T is
 * symmetric, D = half the sum of a row's two segment depths, or half the one
 * depth of a boundary row), so lam_k = D_k mu_k with mu_k = XI_k mu_{k+1}, the
 * same factors that give u_k = Y_k + XI_k u_{k+1}.  dL/dS_k = lam_k and
 * dL/dd = -sum_rows lam_row dR_row/dd; the differences u_k - u_{k+1} are taken
 * as Y_k - X_k XI_k u_{k+1} (optically thin steps leave u constant to ~13
 * digits, so subtracting two u values would lose them).
 *
 * Segment sources as the reference uses them (row n takes Scurr of segment n,
 * the last row Snext of the last segment): rr-side half (backward = false):
 * segment k supplies its outer row (Sc), and its Sn feeds the centre row when
 * the centre is the last row; antipod-side half: segment k supplies the row of
 * its inner entry (Sc) and segment 0 the boundary row (Sn).
 */
struct RevCentre {
    double u, mu, lam; /* u, D^-1-scaled adjoint and adjoint of the centre row */
    double dbar;       /* the centre row's part of dL/d(depth of the innermost segment) */
    bool centre_last;  /* rr-side half only: the centre is the system's last row */
    PointECT ec;       /* the centre entry's cache from the forward sweeps */
};

template <int APPROX, int PREC, bool IW = false>
__device__ __forceinline__ void reverse_half(const KParams &kp, const double *s_erf, const OwnLine &ol,
                             bool backward, uint32_t S, uint32_t o, double shift0,
                             double shift_centre, const uint32_t *s_nxt, const double *s_dZ,
                             const double *s_sh, const uint16_t *mtab, uint32_t l, double freq,
                             const SweepStore<PREC> store, const RevCentre rc, GradSink g)
{
    const uint32_t K = store.n;
    const uint32_t lown = kp.m.sorted_line_map[ol.g];
    /* outward entry iterator: step s, entry i = 1..m_s */
    HalfIter it;
    it.s = 0;
    hi_load_step(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
    it.i = 1;
    Entry inner;
    inner.nr = o;
    inner.nri = o;
    inner.ifac = 0.0;
    inner.shift = shift_centre;
    g.p_crt = o;
    g.p_nxt = it.nxt;
    g.e_crt = g.k_crt = g.e_nxt = g.k_nxt = 0.0;
    g.w_crt = g.w_nxt = 0.0;

    double u_in = rc.u, mu_in = rc.mu, lam_in = rc.lam, dbar = rc.dbar;
    EntryBar ein = {rc.ec.eta, rc.ec.chi, rc.ec.valid ? rc.ec.phi : __builtin_nan(""), 0.0, 0.0,
                    rc.ec.valid};
    typedef Tape T;
    const uint32_t ln = store.lane;
    double Y = T::get(store.st, K - 1, 0, ln), X = T::get(store.st, K - 1, 1, ln),
           d = T::get(store.st, K - 1, 2, ln), kind = T::get(store.st, K - 1, 3, ln),
           phi = T::get(store.st, K - 1, 4, ln);
    /* the record of the segment after next, loaded one iteration ahead of its use so
       that the global load's latency overlaps a segment's reverse */
    double fY = 0.0, fX = 0.0, fd = 0.0, fk = 0.0, fp = 0.0;
    if (K >= 2) {
        fY = T::get(store.st, K - 2, 0, ln);
        fX = T::get(store.st, K - 2, 1, ln);
        fd = T::get(store.st, K - 2, 2, ln);
        fk = T::get(store.st, K - 2, 3, ln);
        fp = T::get(store.st, K - 2, 4, ln);
    }
    for (uint32_t k = K; k-- > 0;) {
        const Entry outer = hi_entry(it, o);
        const double dz = it.dZ;
        const double XI = rcp(1.0 + X);
        const double u_k = fma(XI, u_in, Y);
        const double mu_k = XI * mu_in;
        const double Din = fma(-X * XI, u_in, Y); /* u_k - u_in */
        double lam_k, dbar_next = 0.0, Yn = 0.0, Xn = 0.0, dn = 0.0, kn = 0.0, pn = 0.0;
        if (k > 0) {
            Yn = fY;
            Xn = fX;
            dn = fd;
            kn = fk;
            pn = fp;
            if (k >= 2) {
                fY = T::get(store.st, k - 2, 0, ln);
                fX = T::get(store.st, k - 2, 1, ln);
                fd = T::get(store.st, k - 2, 2, ln);
                fk = T::get(store.st, k - 2, 3, ln);
                fp = T::get(store.st, k - 2, 4, ln);
            }
            lam_k = 0.5 * (dn + d) * mu_k;
            /* u_k - u_{k-1} = X_{k-1} XI_{k-1} u_k - Y_{k-1} */
            const double XIn = rcp(1.0 + Xn);
            const double Dout = fma(Xn * XIn, u_k, -Yn);
            dbar -= lam_k * row_dd(d, dn, Din, Dout);
            dbar_next = -lam_k * row_dd(dn, d, Dout, Din);
        } else {
            lam_k = 0.5 * d * mu_k;
            dbar -= lam_k * bdy_row_dd(d, store.I_b, u_k, Din);
        }
        EntryBar eout = {0.0, 0.0, phi, 0.0, 0.0, false};
        if (!backward) {
            const double Sn_bar = (k == K - 1 && rc.centre_last) ? rc.lam : 0.0;
            segment_bwd<APPROX, PREC, IW>(kp, s_erf, ol, outer, inner, dz, freq, l, lown, d, kind,
                                      dbar, lam_k, Sn_bar, g, eout, ein);
        } else {
            const double Sn_bar = (k == 0) ? lam_k : 0.0;
            segment_bwd<APPROX, PREC, IW>(kp, s_erf, ol, inner, outer, dz, freq, l, lown, d, kind,
                                      dbar, lam_in, Sn_bar, g, ein, eout);
        }
        /* the inner entry has all its contributions now */
        entry_bwd<APPROX, PREC, IW>(kp, ol, inner, freq, l, lown, ein, g);
        if (k == 0) {
            entry_bwd<APPROX, PREC, IW>(kp, ol, outer, freq, l, lown, eout, g);
            break;
        }
        /* move outward */
        ein = eout;
        inner = outer;
        u_in = u_k;
        mu_in = mu_k;
        lam_in = lam_k;
        dbar = dbar_next;
        Y = Yn;
        X = Xn;
        d = dn;
        kind = kn;
        phi = pn;
        if (it.i < it.m_s) {
            it.i++;
        } else {
            it.s++;
            hi_load_step(kp, it, l, o, shift0, s_nxt, s_dZ, s_sh, mtab);
            it.i = 1;
            sink_step<IW>(g, it.nxt);
        }
    }
    sink_finish<IW>(g);
}

/*
 * Centre row of a ray pair after both sweeps: u at the centre, x_c = (T^-1)_cc
 * (the Lambda diagonal), D_c and the centre row's derivative with respect to
 * the depths of the innermost segments of the two halves.
 */
struct CentreRow {
    double u, xc, Dc;
    double dd_fw, dd_bw; /* dR_c/d(depth), rr side and antipod side */
};

template <int PREC>
__device__ inline CentreRow centre_row(const KParams &kp, uint32_t o, double freq, double shift_c,
                                       bool has_fw, bool has_bw, const Sweep &fw, const Sweep &bw)
{
    CentreRow c;
    c.dd_fw = c.dd_bw = 0.0;
    if (has_bw) {
        const double dtau_c = bw.prev_dtau;
        const double S_c = bw.S_inner;
        double FF, Su, iC_c;
        if (!has_fw) {
            const double I_f = boundary_intensity(kp, o, freq * shift_c);
            const double idt = rcp(dtau_c);
            const double B_min = fma(2.0, idt, 1.0);
            const double B = fma(2.0 * idt, idt, B_min);
            iC_c = 0.5 * dtau_c * dtau_c;
            Su = fma(2.0 * I_f, idt, S_c) * rcp(B);
            FF = 0.5 * B_min * dtau_c * dtau_c;
        } else {
            const double avg = 0.5 * (fw.prev_dtau + dtau_c);
            const double r = rcp(avg * fw.prev_dtau);
            iC_c = avg * dtau_c;
            FF = fma(r * fw.X, fw.XI, 1.0) * iC_c;
            const double FI = rcp(1.0 + FF);
            Su = fma(r, fw.Y, S_c) * FI * iC_c;
        }
        const double GP = bw.X * bw.XI;
        const double rden = rcp(FF + GP);
        c.u = fma(Su, 1.0 + FF, bw.Y) * rden;
        c.xc = iC_c * rden;
        const double Dbw = fma(GP, c.u, -bw.Y); /* u_c - u_{c+1} */
        if (has_fw) {
            const double Dfw = fma(fw.X * fw.XI, c.u, -fw.Y); /* u_c - u_{c-1} */
            c.Dc = 0.5 * (fw.prev_dtau + dtau_c);
            c.dd_fw = row_dd(fw.prev_dtau, dtau_c, Dfw, Dbw);
            c.dd_bw = row_dd(dtau_c, fw.prev_dtau, Dbw, Dfw);
        } else {
            const double I_f = boundary_intensity(kp, o, freq * shift_c);
            c.Dc = 0.5 * dtau_c;
            c.dd_bw = bdy_row_dd(dtau_c, I_f, c.u, Dbw);
        }
    } else {
        const double I_l = boundary_intensity(kp, o, freq * shift_c);
        const double idt = rcp(fw.prev_dtau);
        const double A_l = 2.0 * idt * idt;
        const double B_min = fma(2.0, idt, 1.0);
        const double B = B_min + A_l;
        const double den = rcp(fma(B, fw.X, B_min));
        const double Sl = fma(2.0 * I_l, idt, fw.S_inner);
        c.u = fma(A_l, fw.Y, Sl) * (1.0 + fw.X) * den;
        c.xc = (1.0 + fw.X) * den;
        const double Dfw = fma(fw.X * fw.XI, c.u, -fw.Y);
        c.Dc = 0.5 * fw.prev_dtau;
        c.dd_fw = bdy_row_dd(fw.prev_dtau, I_l, c.u, Dfw);
    }
    return c;
}

/*
 * Adjoint of the formal solve: given J_bar (P, F), adds (dJ/d eps)^T J_bar and
 * (dJ/d kappa)^T J_bar into g_eps, g_kap (P, L).  Persistent grid: each wave takes
 * the next origin from a counter; it keeps the elimination state of its current ray
 * pair in its own slice of the scratch (2 halves x capE tape records).
 * The lane layout is the forward kernel's (two 32-lane line slots); lanes past
 * the last quadrature bin repeat it with a zero seed so every slot reduces its
 * gradients over 32 lanes.
 */
template <int APPROX, int PREC, bool IW>
__global__ __launch_bounds__(MGX_WAVE, MGX_ADJ_WAVES_PER_SIMD) void feautrier_adjoint_kernel(
    const KParams *__restrict__ kpp, const double *__restrict__ J_bar, double *g_eps,
    double *g_kap, double *g_iw, double *scratch, uint32_t capE)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    const mgx_model &m = kp.m;
    const uint32_t lane = threadIdx.x;
    const uint32_t cap = kp.cap;
    const uint32_t P = m.npoints, F = m.nfreqs, Q = m.nquads;
    uint32_t *s_nxt0 = reinterpret_cast<uint32_t *>(smem);
    uint32_t *s_nxt1 = s_nxt0 + cap;
    double *s_dZ0 = reinterpret_cast<double *>(smem + 8 * ((2 * cap * 4 + 7) / 8));
    double *s_dZ1 = s_dZ0 + cap;
    double *s_sh0 = s_dZ1 + cap;
    double *s_sh1 = s_sh0 + cap;
    uint16_t *s_m = reinterpret_cast<uint16_t *>(s_sh1 + cap);
    unsigned char *st_wave = reinterpret_cast<unsigned char *>(scratch) +
                             (size_t)blockIdx.x * 2 * capE * Tape::kRecord;

    /* origins are handed out by an atomic counter (status word 3, cleared before
       the launch), so the waves finish together however the origins' costs vary */
    for (;;) {
        uint32_t onext = 0;
        if (lane == 0) onext = atomicAdd(&kp.status[3], 1u);
        const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)onext, 0);
        if (o >= P) break;
        /* a buffer of this launch overflowed: the host clears the gradients and
           relaunches, so the remaining origins are skipped */
        if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(&kp.status[0], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) &
            (ERR_CAP | ERR_CAPE))
            return;
        for (uint32_t rr = 0; rr < m.hnrays; rr++) {
            const uint32_t ar = gld(&m.antipod[rr]);
            uint32_t S0, S1;
            double sh00, sh01;
            __syncthreads();
            int e = walk_half(kp, o, rr, s_nxt0, s_dZ0, s_sh0, S0, sh00);
            if (!e) e = walk_half(kp, o, ar, s_nxt1, s_dZ1, s_sh1, S1, sh01);
            if (e) {
                if (lane == 0) {
                    atomicOr(&kp.status[0], (uint32_t)e);
                    if (e == ERR_CAP) atomicMax(&kp.status[1], kp.cap + 1);
                }
                return;
            }
            __syncthreads();
            if (S0 == 0 && S1 == 0) continue; /* u = I_bdy: no dependence */
            const double w = ray_weight(m, o, rr);
            const uint32_t Sall = S0 + S1;
            for (uint32_t slot0 = 0; slot0 < kp.nslots; slot0 += 2) {
                const uint32_t lA = slot0 / kp.slots_per_line;
                const uint32_t lB = (slot0 + 1 < kp.nslots) ? (slot0 + 1) / kp.slots_per_line : lA;
                __syncthreads();
                for (uint32_t i = lane; i < 2 * Sall; i += MGX_WAVE) {
                    const uint32_t j = i / Sall, r = i % Sall;
                    const uint32_t h = (r < S0) ? 0 : 1;
                    const uint32_t stp = h ? r - S0 : r;
                    const uint32_t *nx = h ? s_nxt1 : s_nxt0;
                    const uint32_t crt = (stp == 0) ? o : nx[stp - 1];
                    const uint32_t mm = n_interp_for_line(kp, j ? lB : lA, crt, nx[stp]);
                    if (mm > 0xffffu) atomicOr(&kp.status[0], (uint32_t)ERR_INTERP);
                    s_m[(j * 2 + h) * cap + stp] = (uint16_t)(mm > 0xffffu ? 0xffffu : mm);
                }
                __syncthreads();
                const uint32_t jj = lane >> 5;
                const uint32_t slot = slot0 + jj;
                if (slot >= kp.nslots) continue; /* whole 32-lane half idle */
                const uint32_t l = jj ? lB : lA;
                const uint32_t z0 = (slot % kp.slots_per_line) * 32 + (lane & 31);
                const uint32_t z = (z0 < Q) ? z0 : Q - 1;
                const OwnLine ol = own_line(m, l);
                const uint32_t f = l * Q + z;
                const size_t of = (size_t)o * F + f;
                const double freq = gld(&m.nu[of]);
                const double ubar = (z0 < Q) ? 2.0 * w * J_bar[of] : 0.0;
                PointECT ec_cent = {0.0, 0.0, 0.0, false};
                Sweep bw, fw;
                bw.started = fw.started = false;
                SweepStore<PREC> st0{st_wave, lane, capE, 0, 0.0},
                    st1{st_wave + (size_t)capE * Tape::kRecord, lane, capE, 0, 0.0};
                if (S1 > 0)
                    sweep_half_tape<APPROX, PREC>(kp, s_erf, ol, true, S1, o, sh01, 1.0, s_nxt1,
                                                  s_dZ1, s_sh1, s_m + (jj * 2 + 1) * cap, l, freq,
                                                  bw, ec_cent, st1);
                if (S0 > 0)
                    sweep_half_tape<APPROX, PREC>(kp, s_erf, ol, false, S0, o, sh00, 1.0, s_nxt0,
                                                  s_dZ0, s_sh0, s_m + (jj * 2 + 0) * cap, l, freq,
                                                  fw, ec_cent, st0);
                if (st0.n > capE || st1.n > capE) {
                    if ((lane & 31) == 0) {
                        atomicOr(&kp.status[0], (uint32_t)ERR_CAPE);
                        atomicMax(&kp.status[2], st0.n > st1.n ? st0.n : st1.n);
                    }
                    continue;
                }
                const CentreRow c = centre_row<PREC>(kp, o, freq, 1.0, S0 > 0, S1 > 0, fw, bw);
                const double mu_c = ubar * c.xc * rcp(c.Dc);
                const double lam_c = ubar * c.xc;
                GradSink g{g_eps, g_kap, m.nlines, l, 32, 0, 0, 0.0, 0.0, 0.0, 0.0, g_iw, 0.0, 0.0};
                if (S0 > 0) {
                    const RevCentre rc{c.u, mu_c, lam_c, -lam_c * c.dd_fw, S1 == 0, ec_cent};
                    reverse_half<APPROX, PREC, IW>(kp, s_erf, ol, false, S0, o, sh00, 1.0, s_nxt0,
                                                   s_dZ0, s_sh0, s_m + (jj * 2 + 0) * cap, l,
                                                   freq, st0, rc, g);
                }
                if (S1 > 0) {
                    const RevCentre rc{c.u, mu_c, lam_c, -lam_c * c.dd_bw, false, ec_cent};
                    reverse_half<APPROX, PREC, IW>(kp, s_erf, ol, true, S1, o, sh01, 1.0, s_nxt1,
                                                   s_dZ1, s_sh1, s_m + (jj * 2 + 1) * cap, l,
                                                   freq, st1, rc, g);
                }
            }
        }
    }
}

/*
 * Adjoint of the projection-surface image: given I_bar (nx*ny, F), adds
 * (dI/d eps)^T I_bar and (dI/d kappa)^T I_bar into g_eps, g_kap.  One wave per
 * pixel (persistent grid), lanes across image frequencies in chunks of 64
 * (lanes past F repeat the last frequency with a zero seed).  I = 2 u_last -
 * I_bdy, so the seed of u at the start point is 2 I_bar.
 */
template <int APPROX, int PREC, bool IW>
__global__ __launch_bounds__(MGX_WAVE, MGX_ADJ_WAVES_PER_SIMD) void image_adjoint_kernel(
    const KParams *__restrict__ kpp, const ImageFrame *frp, const double *__restrict__ I_bar,
    double *g_eps, double *g_kap, double *g_iw, double *scratch, uint32_t capE)
{
    mgx_ref_tables_init();
    const KParams &kp = *kpp;
    extern __shared__ unsigned char smem[];
    const double *s_erf = nullptr; /* erf needs no table */
    const mgx_model &m = kp.m;
    const ImageFrame fr = *frp;
    const uint32_t lane = threadIdx.x;
    const uint32_t F = m.nfreqs;
    const uint32_t npix = fr.nx * fr.ny;
    /* two pixels per wave when a half-wave covers the frequencies (image_pixels) */
    const uint32_t ppw = F <= 32 ? 2 : 1;
    const uint32_t nitems = (npix + ppw - 1) / ppw;
    unsigned char *st_wave = reinterpret_cast<unsigned char *>(scratch) +
                             (size_t)blockIdx.x * capE * Tape::kRecord;

    for (uint32_t q = blockIdx.x; q < nitems; q += gridDim.x) {
        /* a buffer of this launch overflowed: the host clears the gradients and
           relaunches, so the remaining pixels are skipped */
        if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(&kp.status[0], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) &
            (ERR_CAP | ERR_CAPE))
            return;
        const uint32_t p0 = q * ppw;
        const uint32_t n = (p0 + ppw <= npix) ? ppw : npix - p0;
        __syncthreads();
        PixelWalk w0 = pixel_slot(smem, kp.cap, 0), w1 = pixel_slot(smem, kp.cap, 1);
        const bool ok0 = image_walk(kp, fr, p0, w0, nullptr, nullptr);
        const bool ok1 = n < 2 || image_walk(kp, fr, p0 + 1, w1, nullptr, nullptr);
        if (!ok0 || !ok1) return;
        /* the lanes of a group (the wave, or a half-wave per pixel) take one pixel */
        const uint32_t gw = (n == 2) ? 32 : 64;
        const uint32_t half = (n == 2) ? (lane >> 5) : 0;
        const PixelWalk &w = half ? w1 : w0;
        const uint32_t pix = p0 + half;
        const uint32_t S = w.S, ip = w.ip;
        if (S == 0) continue; /* I = I_bdy: no dependence (for the whole group) */
        for (uint32_t fb = 0; fb < F; fb += gw) {
            const uint32_t fl = fb + (lane & (gw - 1));
            const uint32_t f = (fl < F) ? fl : F - 1;
            const double freq = m.nu[(size_t)ip * F + f];
            const uint32_t line = m.f_line[f];
            const uint32_t ub_idx = upper_bound_d(m.sorted_line, m.nlines, freq);
            const uint32_t own_l = m.sorted_line_map[ub_idx > 0 ? ub_idx - 1 : 0];
            const OwnLine ol = own_line(m, own_l);
            const double ubar = (fl < F) ? 2.0 * I_bar[(size_t)pix * F + f] : 0.0;
            Sweep fw;
            fw.started = false;
            PointECT ec = {0.0, 0.0, 0.0, false};
            SweepStore<PREC> st{st_wave, lane, capE, 0, 0.0};
            sweep_half_tape<APPROX, PREC>(kp, s_erf, ol, false, S, ip, w.shift0, w.shift_centre,
                                          w.nxt, w.dZ, w.sh, w.m, line, freq, fw, ec, st);
            if (st.n > capE) {
                if (lane == 0 || lane == 32) {
                    atomicOr(&kp.status[0], (uint32_t)ERR_CAPE);
                    atomicMax(&kp.status[2], st.n);
                }
                continue;
            }
            const CentreRow c = centre_row<PREC>(kp, ip, freq, w.shift_centre, true, false, fw, fw);
            const double mu_c = ubar * c.xc * rcp(c.Dc);
            const double lam_c = ubar * c.xc;
            /* one reduction per group when the group's lanes share their own line */
            const uint32_t l0 = __shfl(own_l, (int)(lane & ~(gw - 1)), MGX_WAVE);
            const uint64_t same = __ballot(own_l == l0);
            const uint64_t gmask = (gw == 64) ? ~0ull : (0xffffffffull << (32 * half));
            const bool uniform = (same & gmask) == gmask;
            GradSink g{g_eps, g_kap, m.nlines, own_l, uniform ? gw : 0u, 0, 0, 0.0, 0.0, 0.0, 0.0,
                       g_iw, 0.0, 0.0};
            const RevCentre rc{c.u, mu_c, lam_c, -lam_c * c.dd_fw, true, ec};
            reverse_half<APPROX, PREC, IW>(kp, s_erf, ol, false, S, ip, w.shift0, w.shift_centre,
                                           w.nxt, w.dZ, w.sh, w.m, line, freq, st, rc, g);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* preparation kernels                                                       */

__global__ void pack_kernel(const mgx_model m, double4 *lpack, PointRec *pts, PointIdx *pidx)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t PL = (size_t)m.npoints * m.nlines;
    if (i < PL) {
        const double e = m.emissivity[i], k = m.opacity[i];
        lpack[i] = make_double4(m.inverse_width[i], e, k, e / k);
    }
    if (i < m.npoints) {
        PointRec r;
        r.x = m.position[3 * i + 0];
        r.y = m.position[3 * i + 1];
        r.z = m.position[3 * i + 2];
        r.vx = m.velocity[3 * i + 0];
        r.vy = m.velocity[3 * i + 1];
        r.vz = m.velocity[3 * i + 2];
        r.dshift_max = m.dshift_max[i];
        r.ub = sqrt(MGX_TWO_KB_OVER_AMU_CC_SQUARED * m.max_inverse_mass * m.temperature[i] +
                    m.vturb2[i]);
        pts[i] = r;
        PointIdx q;
        q.n_nbs = m.n_neighbors[i];
        q.cum = m.cum_n_neighbors[i];
        q.p2b = m.point2boundary[i];
        q.pad = 0;
        pidx[i] = q;
    }
}

/* ------------------------------------------------------------------------- */
/* host side                                                                 */

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return mgx::fail(MGX_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

struct DevBuffers {
    double4 *lpack = nullptr;
    PointRec *pts = nullptr;
    PointIdx *pidx = nullptr;
    uint32_t *status = nullptr;
    KParams *kp_dev = nullptr; /* device copy of the launch parameters */
    ImageFrame *frame = nullptr; /* the imager's frame (carved with the rest) */
    bool owned = true;         /* false: carved out of a shared buffer */
    std::unique_lock<std::mutex> ws_lock; /* held while carved from the call workspace */
    ~DevBuffers()
    {
        if (!owned) return;
        if (kp_dev) (void)hipFree(kp_dev);
        if (lpack) (void)hipFree(lpack);
        if (pts) (void)hipFree(pts);
        if (pidx) (void)hipFree(pidx);
        if (status) (void)hipFree(status);
    }
};

/* entry points whose Lambda output is the diagonal only: with n_off_diag > 0
   the reference also stores the bands, which mgx_feautrier_forward_lambda_device
   returns */
int check_lambda_diagonal(const mgx_options *op)
{
    if (op && op->n_off_diag != 0)
        return mgx::fail(MGX_E_UNSUPPORTED,
                         "n_off_diag > 0: the Lambda bands come from mgx_feautrier_forward_lambda_device");
    return 0;
}

int check_view(const mgx_model *m, const mgx_options *op)
{
    if (!m || !op) return mgx::fail(MGX_E_ARG, "null model or options");
    if (op->approx < 0 || op->approx > 2) return mgx::fail(MGX_E_ARG, "unknown approximation");
    if (op->precision != MGX_PREC_FP64 && op->precision != MGX_PREC_MIXED)
        return mgx::fail(MGX_E_ARG, "unknown precision mode");
    if (m->npoints == 0 || m->nlines == 0 || m->nquads == 0 || m->nfreqs == 0)
        return mgx::fail(MGX_E_ARG, "empty model");
    if (m->nfreqs != m->nlines * m->nquads)
        return mgx::fail(MGX_E_UNSUPPORTED,
                         "the formal solve needs the line spectral discretisation (nfreqs = nlines*nquads)");
    if (m->hnrays == 0 || m->hnrays > m->nrays) return mgx::fail(MGX_E_ARG, "bad hnrays");
    if (!m->position || !m->velocity || !m->n_neighbors || !m->cum_n_neighbors ||
        !m->neighbors || !m->point2boundary || !m->ray_direction || !m->ray_weight ||
        !m->antipod || !m->temperature || !m->vturb2 || !m->line || !m->sorted_line ||
        !m->sorted_line_map || !m->emissivity || !m->opacity || !m->inverse_width ||
        !m->interp_crit || !m->dshift_max || !m->nu || !m->f_line || !m->f_species ||
        !m->f_k || !m->f_z || !m->f_in_line || !m->line_A || !m->species_inverse_mass ||
        !m->quad_weights)
        return mgx::fail(MGX_E_ARG, "null array in model view");
    return MGX_OK;
}


const uint32_t kMaxCap = 8192; /* walk steps per half ray: a ceiling above any LDS budget */

} // namespace

/*
 * State shared by the solver's two translation units: this file, and
 * mgx_adjoint.hip, which compiles it again with MGX_ADJOINT_TU for the adjoint
 * entry points (so that the adjoint kernels get code-generation options of their
 * own).  Kernel timing, the per-call workspace and the walk-capacity hints are
 * defined here and declared in the adjoint unit.
 */
#ifdef MGX_ADJOINT_TU
#define MGX_SHARED(decl, init) extern decl
#else
#define MGX_SHARED(decl, init) decl init
#endif
#define MGX_COMMA ,
namespace mgx_shared {
MGX_SHARED(bool g_timing, = false);
MGX_SHARED(double g_last_kernel_ms, = 0.0);
MGX_SHARED(thread_local bool t_no_events, = false); /* set by the no-allocation entry point */
} // namespace mgx_shared

namespace {
using namespace mgx_shared;

/* launches fn() between two events on st when timing is enabled */
template <typename Fn>
void timed(hipStream_t st, Fn fn)
{
    if (!g_timing || t_no_events) {
        fn();
        return;
    }
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, st);
    fn();
    (void)hipEventRecord(b, st);
    (void)hipEventSynchronize(b);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    g_last_kernel_ms = ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
}

/* fills the packed tables of b (already allocated) and the launch parameters */
int prepare_tables(const mgx_model *m, const mgx_options *op, hipStream_t st, DevBuffers &b,
                   KParams &kp)
{
    const size_t PL = (size_t)m->npoints * m->nlines;
    HIP_TRY(hipMemsetAsync(b.status, 0, 4 * sizeof(uint32_t), st));
    const size_t n = PL > m->npoints ? PL : m->npoints;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, *m,
                       b.lpack, b.pts, b.pidx);
    HIP_TRY(hipGetLastError());
    kp.m = *m;
    kp.op = *op;
    kp.lpack = b.lpack;
    kp.pts = b.pts;
    kp.pidx = b.pidx;
    kp.log_max_interp = (double)mgx_ref_logf((float)op->max_interpolation_diff);
    kp.slots_per_line = (m->nquads + 31) / 32;
    kp.nslots = m->nlines * kp.slots_per_line;
    kp.status = b.status;
    return MGX_OK;
}

/* Device memory of a call's packed tables, status words, launch parameters and
   image frame: one grow-only buffer per device, reused by every call, so a call
   allocates nothing once the buffer has grown to its model.  A call holds the
   buffer (b.ws_lock) until it returns, so calls that use it run one at a time. */
} // namespace
namespace mgx_shared {
MGX_SHARED(std::mutex g_prep_mutex, );
MGX_SHARED(unsigned char *g_prep[64], = {});
MGX_SHARED(size_t g_prep_bytes[64], = {});
} // namespace mgx_shared
namespace {

/* the tables of model m in b: from the call workspace, or (persistent, for objects
   that keep them, such as a batch) in allocations of their own */
int prepare(const mgx_model *m, const mgx_options *op, hipStream_t st, DevBuffers &b, KParams &kp,
            bool persistent = false)
{
    const size_t P = m->npoints, PL = (size_t)m->npoints * m->nlines;
    if (persistent) {
        HIP_TRY(hipMalloc(&b.lpack, PL * sizeof(double4)));
        HIP_TRY(hipMalloc(&b.pts, P * sizeof(PointRec)));
        HIP_TRY(hipMalloc(&b.pidx, P * sizeof(PointIdx)));
        HIP_TRY(hipMalloc(&b.status, 4 * sizeof(uint32_t)));
        HIP_TRY(hipMalloc(&b.kp_dev, sizeof(KParams)));
        return prepare_tables(m, op, st, b, kp);
    }
    const auto rnd = [](size_t n) { return (n + 255) & ~(size_t)255; };
    const size_t need = rnd(PL * sizeof(double4)) + rnd(P * sizeof(PointRec)) +
                        rnd(P * sizeof(PointIdx)) + rnd(4 * sizeof(uint32_t)) +
                        rnd(sizeof(KParams)) + rnd(sizeof(ImageFrame));
    b.ws_lock = std::unique_lock<std::mutex>(g_prep_mutex);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return mgx::fail(MGX_E_ARG, "device index %d out of range", dev);
    if (g_prep_bytes[dev] < need) {
        /* the previous holder has returned, so nothing uses the old buffer */
        if (g_prep[dev]) (void)hipFree(g_prep[dev]);
        g_prep[dev] = nullptr;
        g_prep_bytes[dev] = 0;
        HIP_TRY(hipMalloc(&g_prep[dev], need));
        g_prep_bytes[dev] = need;
    }
    unsigned char *q = g_prep[dev];
    b.owned = false;
    b.lpack = reinterpret_cast<double4 *>(q);
    q += rnd(PL * sizeof(double4));
    b.pts = reinterpret_cast<PointRec *>(q);
    q += rnd(P * sizeof(PointRec));
    b.pidx = reinterpret_cast<PointIdx *>(q);
    q += rnd(P * sizeof(PointIdx));
    b.status = reinterpret_cast<uint32_t *>(q);
    q += rnd(4 * sizeof(uint32_t));
    b.kp_dev = reinterpret_cast<KParams *>(q);
    q += rnd(sizeof(KParams));
    b.frame = reinterpret_cast<ImageFrame *>(q);
    return prepare_tables(m, op, st, b, kp);
}

/* ------------------------------------------------------------------------- */
/* the no-allocation scratch                                                 */

/* Device memory of mgx_feautrier_forward_noalloc: the status words, the launch
   parameters and the packed point and line tables live here instead of in
   per-call allocations.  Each GPU holds its own copy for the life of the
   process; host code reaches it only through the address hipGetSymbolAddress
   returns on that GPU. */
#ifndef MGX_ADJOINT_TU
static __device__ __align__(256) unsigned char scratch[MGX_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mutex; /* guards g_scratch_ptr and serialises the calls that use it */
constexpr int kMaxDevices = 64;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

/* the current GPU's copy of scratch (looked up once per GPU); g_scratch_mutex held */
int scratch_base(unsigned char **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return mgx::fail(MGX_E_ARG, "device index %d out of range", dev);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return mgx::fail(MGX_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return MGX_OK;
}
#endif

/* reads the status words; returns MGX_OK, or an error (needed capacity in *need;
   segments needed by the adjoint's scratch in *need_e) */
int read_status(hipStream_t st, const DevBuffers &b, uint32_t *need, uint32_t *need_e = nullptr)
{
    uint32_t h[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(h, b.status, sizeof h, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *need = h[1];
    if (need_e) *need_e = (h[0] & ERR_CAPE) ? (h[2] ? h[2] : 1) : 0;
    if (h[0] & ERR_NO_NEXT)
        return mgx::fail(MGX_E_NO_NEXT, "a ray from an interior point found no valid next point");
    if (h[0] & ERR_INTERP)
        return mgx::fail(MGX_E_UNSUPPORTED,
                         "a segment needs more than 65535 interpolation points (the interpolation "
                         "criterion changes by more than max_interpolation_diff^65535 between neighbours)");
    if (h[0] & ERR_OUT) return mgx::fail(MGX_E_RAY_LENGTH, "a traced list exceeds the output capacity");
    if (h[0] & ERR_LAMCAP) return mgx::fail(MGX_E_ARG, "a Lambda list exceeds its capacity");
    if (h[0] & ERR_CAP) return MGX_E_RAY_LENGTH;
    return MGX_OK;
}

/* LDS per workgroup available to dynamic allocations: the device's limit (160 KiB
   on gfx950) less the static LDS, queried once per device */
size_t lds_budget()
{
    static size_t cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (!cached[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            v <= 0)
            v = 65536;
        cached[dev] = (size_t)v - kStaticLds;
    }
    return cached[dev];
}

/* the largest walk capacity whose buffers fit the LDS budget beside extra bytes */
uint32_t max_walk_cap(size_t extra)
{
    const size_t budget = lds_budget();
    uint32_t c = kMaxCap;
    while (c > 16 && walk_lds_bytes(c) + extra > budget) c -= 16;
    return c;
}

/* Walk capacities that solved without a relaunch, by geometry and ray set: a hint
   for the next solve of the same model (a model changed in place under the same
   arrays costs at most a relaunch). */
} // namespace
namespace mgx_shared {
struct WalkKey {
    const void *pos, *nbs, *dirs;
    uint32_t P, nrays, adaptive;
    double rd[3];    /* image: ray direction (zero for the ray pairs of a solve) */
    uint32_t nx, ny; /* image: pixels (zero for a solve) */
    bool operator<(const WalkKey &o) const
    {
        return std::tie(pos, nbs, dirs, P, nrays, adaptive, rd[0], rd[1], rd[2], nx, ny) <
               std::tie(o.pos, o.nbs, o.dirs, o.P, o.nrays, o.adaptive, o.rd[0], o.rd[1], o.rd[2],
                        o.nx, o.ny);
    }
};
MGX_SHARED(std::mutex g_cap_mutex, );
MGX_SHARED(std::map<WalkKey MGX_COMMA uint32_t> g_cap_hint, );
} // namespace mgx_shared
namespace {

WalkKey walk_key(const mgx_model &m)
{
    return WalkKey{m.position, m.neighbors, m.ray_direction, m.npoints, m.nrays,
                   (uint32_t)m.adaptive_directions, {0.0, 0.0, 0.0}, 0, 0};
}

/* the key of a projection-surface image's pixel rays */
WalkKey image_key(const mgx_model &m, const double *raydir, uint32_t nx, uint32_t ny)
{
    WalkKey k = walk_key(m);
    k.dirs = nullptr;
    k.nrays = 0;
    k.adaptive = 0;
    for (int i = 0; i < 3; i++) k.rd[i] = raydir[i];
    k.nx = nx;
    k.ny = ny;
    return k;
}

bool cap_hint(const WalkKey &key, uint32_t &cap)
{
    std::lock_guard<std::mutex> lk(g_cap_mutex);
    auto it = g_cap_hint.find(key);
    if (it == g_cap_hint.end()) return false;
    cap = it->second;
    return true;
}

void set_cap_hint(const WalkKey &key, uint32_t cap)
{
    std::lock_guard<std::mutex> lk(g_cap_mutex);
    if (g_cap_hint.size() > 4096) g_cap_hint.clear();
    g_cap_hint[key] = cap;
}

/* The walk capacity after a launch whose walks outgrew cap: need is the largest
   length a walk reported.  The pixel walks of the imager report their exact
   length, the ray-pair walks cap + 1 (they stop at the buffer's end); the
   capacity then doubles.  0 when no capacity up to max_cap can hold the walk. */
uint32_t grown_cap(uint32_t cap, uint32_t need, uint32_t max_cap)
{
    if (need > max_cap || cap >= max_cap) return 0;
    const uint32_t want = (need > cap + 1) ? (need + 15) / 16 * 16 : cap * 2;
    return want > max_cap ? max_cap : want;
}

/*
 * Launch loop of the solves: sets the LDS walk capacity, launches, and relaunches
 * with a larger capacity when a walk outgrew it.  With pair_walks (the launch walks
 * the model's ray pairs from every origin) the first capacity is the one that
 * served this geometry and ray set before, or else the longest half-ray that
 * walk_max_kernel finds, so no launch is aborted for want of walk buffer.
 */
template <typename Launch>
int launch_with_capacity(hipStream_t st, DevBuffers &b, KParams &kp, Launch launch,
                         size_t extra_lds = 0, bool pair_walks = false, const WalkKey *img = nullptr)
{
    uint32_t cap = 128;
    double total_ms = 0.0; /* kernel time summed over the launches */
    struct Report {
        double &t;
        ~Report() { g_last_kernel_ms = t; }
    } report{total_ms};
    const uint32_t max_cap = max_walk_cap(8 + extra_lds);
    const WalkKey key = img ? *img : walk_key(kp.m);
    const bool hinted = pair_walks || img;
    if (hinted) {
        const bool have = cap_hint(key, cap);
        if (!have && pair_walks) {
            HIP_TRY(hipMemsetAsync(b.status, 0, 4 * sizeof(uint32_t), st));
            HIP_TRY(hipMemcpyAsync(b.kp_dev, &kp, sizeof(KParams), hipMemcpyHostToDevice, st));
            g_last_kernel_ms = 0.0;
            timed(st, [&] {
                hipLaunchKernelGGL(walk_max_kernel, dim3(kp.m.npoints, kp.m.hnrays), dim3(MGX_WAVE), 0,
                                   st, b.kp_dev, b.status + 3);
            });
            total_ms += g_last_kernel_ms;
            HIP_TRY(hipGetLastError());
            uint32_t h[4] = {0, 0, 0, 0};
            HIP_TRY(hipMemcpyAsync(h, b.status, sizeof h, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (h[0] & ERR_NO_NEXT)
                return mgx::fail(MGX_E_NO_NEXT, "a ray from an interior point found no valid next point");
            if (h[3] > max_cap)
                return mgx::fail(MGX_E_RAY_LENGTH, "a ray needs more than %u walk steps", max_cap);
            cap = (h[3] + 15) / 16 * 16;
            cap = cap < 16 ? 16 : cap;
        }
        cap = cap > max_cap ? max_cap : cap;
    }
    for (;;) {
        kp.cap = cap;
        kp.j_lds = j_in_lds(cap, kp.m.nfreqs, kp.m.nlines) ? 1 : 0;
        HIP_TRY(hipMemsetAsync(b.status, 0, 4 * sizeof(uint32_t), st));
        HIP_TRY(hipMemcpyAsync(b.kp_dev, &kp, sizeof(KParams), hipMemcpyHostToDevice, st));
        g_last_kernel_ms = 0.0;
        launch(walk_lds_bytes(cap));
        total_ms += g_last_kernel_ms;
        HIP_TRY(hipGetLastError());
        uint32_t need = 0;
        const int e = read_status(st, b, &need);
        if (e == MGX_OK && hinted && !t_no_events) set_cap_hint(key, cap);
        if (e != MGX_E_RAY_LENGTH) return e;
        cap = grown_cap(cap, need, max_cap);
        if (!cap) return mgx::fail(MGX_E_RAY_LENGTH, "a ray needs more than %u walk steps", max_cap);
    }
}

/* argument checks, model preparation and the frame of the projection-surface
   imager (Model::compute_image_new, model.cpp:1008-1035; the frame reduction
   runs on the device); dfr is the device ImageFrame of b */
int image_setup(const mgx_model *model, const mgx_options *opts, const double *raydir, uint32_t nx,
                uint32_t ny, hipStream_t st, DevBuffers &b, KParams &kp, ImageFrame *&dfr)
{
    if (!model || !opts || !raydir) return mgx::fail(MGX_E_ARG, "null argument");
    if (nx == 0 || ny == 0) return mgx::fail(MGX_E_ARG, "the image needs at least one pixel");
    const double sq = raydir[0] * raydir[0] + raydir[1] * raydir[1] + raydir[2] * raydir[2];
    if (sq == 0.0)
        return mgx::fail(MGX_E_ARG, "The given ray direction vector does not point in a "
                                    "direction. Please use a non-zero (normed) direction vector "
                                    "to generate an image.");
    if (opts->approx == MGX_APPROX_ONE_LINE)
        return mgx::fail(MGX_E_UNSUPPORTED, "One line approximation is not supported for imaging.");
    if (!model->position || !model->nu || !model->boundary2point || !model->point2boundary ||
        !model->interp_crit || !model->sorted_line || !model->emissivity)
        return mgx::fail(MGX_E_ARG, "null array in model view");
    if (model->npoints == 0 || model->nfreqs == 0 || model->nlines == 0 || model->nboundary == 0)
        return mgx::fail(MGX_E_ARG, "empty model");
    const double inv = 1 / std::sqrt(sq);
    ImageFrame fr;
    std::memset(&fr, 0, sizeof fr);
    for (int i = 0; i < 3; i++) fr.rd[i] = raydir[i] * inv;
    if (model->dimension == 1 && (fr.rd[0] != 0.0 || fr.rd[1] != 1.0 || fr.rd[2] != 0.0))
        return mgx::fail(MGX_E_ARG, "In 1D, the image ray has to be (0,1,0)");
    fr.nx = nx;
    fr.ny = ny;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    dfr = b.frame;
    HIP_TRY(hipMemcpyAsync(dfr, &fr, sizeof fr, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(image_frame_kernel, dim3(1), dim3(256), 0, st, *model, dfr);
    HIP_TRY(hipGetLastError());
    return MGX_OK;
}

/* Device scratch of the adjoint kernels, kept between calls (grow-only; freed by
   mgx_release_workspace).  Calls are serialised on it. */
std::mutex g_ws_mutex;
double *g_ws = nullptr;
size_t g_ws_bytes = 0;
uint32_t g_capE_hint = 128;

int ensure_workspace(size_t bytes)
{
    if (bytes <= g_ws_bytes) return MGX_OK;
    if (g_ws) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_ws);
        g_ws = nullptr;
        g_ws_bytes = 0;
    }
    HIP_TRY(hipMalloc(&g_ws, bytes));
    g_ws_bytes = bytes;
    return MGX_OK;
}

/* waves of the persistent adjoint grid: resident waves per CU x CUs */
template <typename K>
uint32_t persistent_waves(K kernel, size_t lds, uint32_t work)
{
    int dev = 0, ncu = 0, nb = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, MGX_WAVE, lds) != hipSuccess ||
        nb <= 0)
        nb = 4;
    const uint64_t g = (uint64_t)(ncu > 0 ? ncu : 256) * (uint64_t)nb;
    return (uint32_t)(g < work ? g : work);
}

/*
 * Launch loop of the adjoint kernels: clears the gradients and the status,
 * launches, and retries with a larger LDS walk buffer (ERR_CAP) or a larger
 * scratch (ERR_CAPE) until the launch completes.  launch(lds, grid, capE)
 * receives the dynamic LDS size; grid_of(lds) the persistent grid size.
 */
template <typename Launch, typename Grid>
int adjoint_launch(hipStream_t st, DevBuffers &b, KParams &kp, const WalkKey &key, double *d_eps,
                   double *d_kap, double *d_iw, size_t PL, size_t per_wave_segments_factor,
                   Grid grid_of, Launch launch)
{
    const uint32_t max_cap = max_walk_cap(8);
    uint32_t cap = 128;
    (void)cap_hint(key, cap); /* the walks of the forward solve or image */
    if (cap > max_cap) cap = max_cap;
    uint32_t capE = g_capE_hint;
    double total_ms = 0.0; /* kernel time summed over the launches */
    struct Report {
        double &t;
        ~Report() { g_last_kernel_ms = t; }
    } report{total_ms};
    for (;;) {
        kp.cap = cap;
        const size_t lds = walk_lds_bytes(cap);
        const uint32_t grid = grid_of(lds);
        const size_t bytes = (size_t)grid * per_wave_segments_factor * capE * Tape::kRecord;
        if (int e = ensure_workspace(bytes)) return e;
        HIP_TRY(hipMemsetAsync(b.status, 0, 4 * sizeof(uint32_t), st));
        HIP_TRY(hipMemsetAsync(d_eps, 0, PL * sizeof(double), st));
        HIP_TRY(hipMemsetAsync(d_kap, 0, PL * sizeof(double), st));
        if (d_iw) HIP_TRY(hipMemsetAsync(d_iw, 0, PL * sizeof(double), st));
        HIP_TRY(hipMemcpyAsync(b.kp_dev, &kp, sizeof(KParams), hipMemcpyHostToDevice, st));
        g_last_kernel_ms = 0.0;
        launch(lds, grid, capE);
        total_ms += g_last_kernel_ms;
        HIP_TRY(hipGetLastError());
        uint32_t need = 0, need_e = 0;
        const int e = read_status(st, b, &need, &need_e);
        if (e != MGX_OK && e != MGX_E_RAY_LENGTH) return e;
        bool retry = false;
        if (e == MGX_E_RAY_LENGTH) {
            cap = grown_cap(cap, need, max_cap);
            if (!cap) return mgx::fail(MGX_E_RAY_LENGTH, "a ray needs more than %u walk steps", max_cap);
            retry = true;
        }
        if (need_e) {
            uint32_t c = capE * 2;
            while (c < need_e) c *= 2;
            if (c > (1u << 20))
                return mgx::fail(MGX_E_RAY_LENGTH, "a ray has more than %u traced entries", 1u << 20);
            capE = c;
            g_capE_hint = c;
            retry = true;
        }
        if (!retry) {
            if (!t_no_events) set_cap_hint(key, cap);
            return MGX_OK;
        }
    }
}

template <int A, int PR, bool IW>
int run_J_adjoint_k(hipStream_t st, DevBuffers &b, KParams &kp, const double *J_bar,
                    double *d_eps, double *d_kap, double *d_iw)
{
    auto kern = feautrier_adjoint_kernel<A, PR, IW>;
    const uint32_t P = kp.m.npoints;
    return adjoint_launch(
        st, b, kp, walk_key(kp.m), d_eps, d_kap, d_iw, (size_t)P * kp.m.nlines, 2,
        [&](size_t lds) { return persistent_waves(kern, lds, P); },
        [&](size_t lds, uint32_t grid, uint32_t capE) {
            timed(st, [&] {
                hipLaunchKernelGGL(kern, dim3(grid), dim3(MGX_WAVE), lds, st, b.kp_dev, J_bar, d_eps,
                                   d_kap, d_iw, g_ws, capE);
            });
        });
}

template <int A, int PR>
int run_J_adjoint(hipStream_t st, DevBuffers &b, KParams &kp, const double *J_bar, double *d_eps,
                  double *d_kap, double *d_iw)
{
    return d_iw ? run_J_adjoint_k<A, PR, true>(st, b, kp, J_bar, d_eps, d_kap, d_iw)
                : run_J_adjoint_k<A, PR, false>(st, b, kp, J_bar, d_eps, d_kap, nullptr);
}


/* ------------------------------------------------------------------------- */
/* batched forward solve: plan, autotuner, graph                              */

struct BatchConfig {
    uint32_t rays_per_item; /* 0: all ray pairs of an origin in one item */
    uint32_t order;         /* 0: items by decreasing cost; 1: model by model */
    uint32_t wps;           /* register budget of the kernel: 2, 3 or 4 waves/SIMD */
};

} // namespace

struct mgx_batch {
    uint32_t n = 0;
    mgx_options op{};
    std::vector<mgx_model> views;
    std::vector<DevBuffers *> bufs;
    std::vector<KParams> kps;
    KParams *kps_dev = nullptr;
    BatchItem *items_dev = nullptr;
    uint32_t n_items = 0, items_cap = 0;
    uint32_t *counter = nullptr, *status = nullptr;
    double **J_dev = nullptr, **lam_dev = nullptr;
    std::vector<double *> J, lam;
    bool with_lam = false;
    BatchConfig cfg{0, 0, MGX_SOLVE_WAVES_PER_SIMD};
    uint32_t cap = 128;
    uint32_t maxF = 0, maxL = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int use_graph = 1;
    double last_ms = 0.0;
    /* images (mgx_batch_set_images): one per model when n_img == n */
    uint32_t n_img = 0, n_chunks = 0;
    bool images_on = true; /* off while the tuner times the solve alone */
    std::vector<KParams> img_kps;
    KParams *img_kps_dev = nullptr;
    ImageFrame *frames_dev = nullptr;
    uint32_t *chunk_start_dev = nullptr, *slot_model_dev = nullptr, *img_counter = nullptr;
    double **I_dev = nullptr, **ImX_dev = nullptr, **ImY_dev = nullptr;
    void free_images()
    {
        for (void *p : {(void *)img_kps_dev, (void *)frames_dev, (void *)chunk_start_dev,
                        (void *)slot_model_dev, (void *)img_counter, (void *)I_dev, (void *)ImX_dev,
                        (void *)ImY_dev})
            if (p) (void)hipFree(p);
        img_kps_dev = nullptr;
        frames_dev = nullptr;
        chunk_start_dev = slot_model_dev = img_counter = nullptr;
        I_dev = ImX_dev = ImY_dev = nullptr;
        img_kps.clear();
        n_img = n_chunks = 0;
    }
    ~mgx_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        free_images();
        for (auto *b : bufs) delete b;
        if (kps_dev) (void)hipFree(kps_dev);
        if (items_dev) (void)hipFree(items_dev);
        if (counter) (void)hipFree(counter);
        if (status) (void)hipFree(status);
        if (J_dev) (void)hipFree(J_dev);
        if (lam_dev) (void)hipFree(lam_dev);
    }
};

namespace {

/* work list of a configuration (host) */
std::vector<BatchItem> batch_items(const mgx_batch &B, const BatchConfig &c)
{
    std::vector<BatchItem> it;
    std::vector<double> cost;
    for (uint32_t mi = 0; mi < B.n; mi++) {
        const mgx_model &m = B.views[mi];
        const uint32_t R = m.hnrays;
        const uint32_t step = (c.rays_per_item == 0 || c.rays_per_item > R) ? R : c.rays_per_item;
        /* walk length grows like the model's linear size */
        const double len = std::cbrt((double)m.npoints);
        for (uint32_t o = 0; o < m.npoints; o++)
            for (uint32_t r0 = 0; r0 < R; r0 += step) {
                const uint32_t r1 = (r0 + step < R) ? r0 + step : R;
                it.push_back(BatchItem{mi, o, r0, r1});
                cost.push_back((double)(r1 - r0) * len * m.nlines);
            }
    }
    if (c.order == 0) {
        std::vector<uint32_t> idx(it.size());
        for (uint32_t i = 0; i < idx.size(); i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
        std::vector<BatchItem> s(it.size());
        for (uint32_t i = 0; i < idx.size(); i++) s[i] = it[idx[i]];
        it.swap(s);
    }
    return it;
}

template <int A, int PR>
void launch_batch_kernel(const mgx_batch &B, uint32_t wps, dim3 grid, size_t lds, hipStream_t st,
                         uint32_t n_items)
{
    double *const *lam = B.with_lam ? B.lam_dev : nullptr;
    if (wps == 2)
        hipLaunchKernelGGL((feautrier_batch_kernel<A, PR, 2>), grid, dim3(MGX_WAVE), lds, st,
                           B.kps_dev, B.items_dev, n_items, B.counter, B.J_dev, lam);
    else if (wps == 4)
        hipLaunchKernelGGL((feautrier_batch_kernel<A, PR, 4>), grid, dim3(MGX_WAVE), lds, st,
                           B.kps_dev, B.items_dev, n_items, B.counter, B.J_dev, lam);
    else
        hipLaunchKernelGGL((feautrier_batch_kernel<A, PR, 3>), grid, dim3(MGX_WAVE), lds, st,
                           B.kps_dev, B.items_dev, n_items, B.counter, B.J_dev, lam);
}

template <int A, int PR>
uint32_t batch_grid(uint32_t wps, size_t lds)
{
    if (wps == 2) return persistent_waves(feautrier_batch_kernel<A, PR, 2>, lds, 0xffffffffu);
    if (wps == 4) return persistent_waves(feautrier_batch_kernel<A, PR, 4>, lds, 0xffffffffu);
    return persistent_waves(feautrier_batch_kernel<A, PR, 3>, lds, 0xffffffffu);
}

bool batch_j_lds(const mgx_batch &B) { return j_in_lds(B.cap, B.maxF, B.maxL); }

size_t batch_lds(const mgx_batch &B)
{
    return walk_lds_bytes(B.cap) + 8 + 8 * (size_t)B.maxL + (batch_j_lds(B) ? 8 * (size_t)B.maxF : 0);
}

/* persistent grid of the current configuration (resident waves, at most n_items) */
uint32_t batch_grid_size(const mgx_batch &B, uint32_t n_items)
{
    const size_t lds = batch_lds(B);
    const bool fp64 = B.op.precision == MGX_PREC_FP64;
    uint32_t grid = 0;
    switch (B.op.approx) {
    case MGX_APPROX_NONE:
        grid = fp64 ? batch_grid<MGX_APPROX_NONE, MGX_PREC_FP64>(B.cfg.wps, lds)
                    : batch_grid<MGX_APPROX_NONE, MGX_PREC_MIXED>(B.cfg.wps, lds);
        break;
    case MGX_APPROX_ONE_LINE:
        grid = fp64 ? batch_grid<MGX_APPROX_ONE_LINE, MGX_PREC_FP64>(B.cfg.wps, lds)
                    : batch_grid<MGX_APPROX_ONE_LINE, MGX_PREC_MIXED>(B.cfg.wps, lds);
        break;
    default:
        grid = fp64 ? batch_grid<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>(B.cfg.wps, lds)
                    : batch_grid<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>(B.cfg.wps, lds);
    }
    if (grid > n_items) grid = n_items ? n_items : 1;
    return grid;
}

bool batch_images(const mgx_batch &B) { return B.n_img == B.n && B.images_on; }

/* persistent grid of the image kernel (resident waves, at most the chunk count) */
uint32_t image_grid_size(const mgx_batch &B)
{
    const size_t lds = walk_lds_bytes(B.cap);
    const uint32_t work = B.n_chunks ? B.n_chunks : 1;
    const bool fp64 = B.op.precision == MGX_PREC_FP64;
    if (B.op.approx == MGX_APPROX_NONE)
        return fp64 ? persistent_waves(image_batch_kernel<MGX_APPROX_NONE, MGX_PREC_FP64>, lds, work)
                    : persistent_waves(image_batch_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED>, lds, work);
    return fp64 ? persistent_waves(image_batch_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>, lds, work)
                : persistent_waves(image_batch_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>, lds, work);
}

void launch_image_batch(const mgx_batch &B, hipStream_t st, uint32_t grid)
{
    const size_t lds = walk_lds_bytes(B.cap);
    const bool fp64 = B.op.precision == MGX_PREC_FP64;
    const dim3 g(grid), blk(MGX_WAVE);
#define MGX_IMG_ARGS B.img_kps_dev, B.frames_dev, B.chunk_start_dev, B.slot_model_dev, B.n_img, \
                     B.img_counter, B.I_dev, B.ImX_dev, B.ImY_dev
    if (B.op.approx == MGX_APPROX_NONE) {
        if (fp64)
            hipLaunchKernelGGL((image_batch_kernel<MGX_APPROX_NONE, MGX_PREC_FP64>), g, blk, lds, st, MGX_IMG_ARGS);
        else
            hipLaunchKernelGGL((image_batch_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED>), g, blk, lds, st, MGX_IMG_ARGS);
    } else {
        if (fp64)
            hipLaunchKernelGGL((image_batch_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>), g, blk, lds, st, MGX_IMG_ARGS);
        else
            hipLaunchKernelGGL((image_batch_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>), g, blk, lds, st, MGX_IMG_ARGS);
    }
#undef MGX_IMG_ARGS
}

/* the launches of one execution: line tables, output clearing (split items),
   counter and status reset, the persistent kernel over n_items items on `grid`
   waves, then (when images are set) the persistent image kernel on `igrid` waves
   (no device queries: the sequence is also captured into a graph) */
void batch_enqueue(const mgx_batch &B, hipStream_t st, uint32_t n_items, bool clear_outputs,
                   uint32_t grid, uint32_t igrid)
{
    for (uint32_t mi = 0; mi < B.n; mi++) {
        const mgx_model &m = B.views[mi];
        const size_t PL = (size_t)m.npoints * m.nlines;
        const size_t n = PL > m.npoints ? PL : m.npoints;
        hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, m,
                           B.bufs[mi]->lpack, B.bufs[mi]->pts, B.bufs[mi]->pidx);
        if (clear_outputs) {
            const size_t PF = (size_t)m.npoints * m.nfreqs;
            hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((PF + 255) / 256)), dim3(256), 0,
                               st, B.J[mi], PF);
            if (B.with_lam)
                hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((PL + 255) / 256)), dim3(256),
                                   0, st, B.lam[mi], PL);
        }
    }
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, B.counter, 1u, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, B.status, 4u, 0u);
    const size_t lds = batch_lds(B);
    const bool fp64 = B.op.precision == MGX_PREC_FP64;
    const dim3 g(grid);
    switch (B.op.approx) {
    case MGX_APPROX_NONE:
        fp64 ? launch_batch_kernel<MGX_APPROX_NONE, MGX_PREC_FP64>(B, B.cfg.wps, g, lds, st, n_items)
             : launch_batch_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED>(B, B.cfg.wps, g, lds, st, n_items);
        break;
    case MGX_APPROX_ONE_LINE:
        fp64 ? launch_batch_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_FP64>(B, B.cfg.wps, g, lds, st, n_items)
             : launch_batch_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_MIXED>(B, B.cfg.wps, g, lds, st, n_items);
        break;
    default:
        fp64 ? launch_batch_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>(B, B.cfg.wps, g, lds, st, n_items)
             : launch_batch_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>(B, B.cfg.wps, g, lds, st, n_items);
    }
    if (batch_images(B)) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, B.img_counter, 1u, 0u);
        launch_image_batch(B, st, igrid);
    }
}

/* uploads the work list of the current configuration and the per-model
   parameters with the current walk capacity */
int batch_upload(mgx_batch &B, hipStream_t st)
{
    const std::vector<BatchItem> it = batch_items(B, B.cfg);
    if (it.size() > B.items_cap) {
        if (B.items_dev) (void)hipFree(B.items_dev);
        B.items_dev = nullptr;
        HIP_TRY(hipMalloc(&B.items_dev, it.size() * sizeof(BatchItem)));
        B.items_cap = (uint32_t)it.size();
    }
    B.n_items = (uint32_t)it.size();
    HIP_TRY(hipMemcpyAsync(B.items_dev, it.data(), it.size() * sizeof(BatchItem),
                           hipMemcpyHostToDevice, st));
    for (auto &k : B.kps) {
        k.cap = B.cap;
        k.j_lds = batch_j_lds(B) ? 1 : 0;
    }
    HIP_TRY(hipMemcpyAsync(B.kps_dev, B.kps.data(), B.kps.size() * sizeof(KParams),
                           hipMemcpyHostToDevice, st));
    if (B.n_img) {
        for (auto &k : B.img_kps) k.cap = B.cap;
        HIP_TRY(hipMemcpyAsync(B.img_kps_dev, B.img_kps.data(), B.img_kps.size() * sizeof(KParams),
                               hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return MGX_OK;
}

bool batch_split(const mgx_batch &B)
{
    if (B.cfg.rays_per_item == 0) return false;
    for (const auto &v : B.views)
        if (B.cfg.rays_per_item < v.hnrays) return true;
    return false;
}

void batch_drop_graph(mgx_batch &B)
{
    if (B.exec) (void)hipGraphExecDestroy(B.exec);
    if (B.graph) (void)hipGraphDestroy(B.graph);
    B.exec = nullptr;
    B.graph = nullptr;
}

/* one full execution with the walk-capacity retry; the launches go through a
   captured graph when use_graph is set */
int batch_execute(mgx_batch &B, hipStream_t st)
{
    for (;;) {
        const bool clear = batch_split(B);
        const uint32_t grid = batch_grid_size(B, B.n_items);
        const uint32_t igrid = batch_images(B) ? image_grid_size(B) : 0;
        if (B.use_graph) {
            if (!B.exec) {
                /* capture on a private stream (the caller's may be the null stream) */
                hipStream_t cs;
                HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
                hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    batch_enqueue(B, cs, B.n_items, clear, grid, igrid);
                    e = hipStreamEndCapture(cs, &B.graph);
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0);
                (void)hipStreamDestroy(cs);
                if (e != hipSuccess) {
                    batch_drop_graph(B);
                    return mgx::fail(MGX_E_HIP, "graph capture of the batch: %s", hipGetErrorString(e));
                }
            }
            timed(st, [&] { (void)hipGraphLaunch(B.exec, st); });
        } else {
            timed(st, [&] { batch_enqueue(B, st, B.n_items, clear, grid, igrid); });
        }
        HIP_TRY(hipGetLastError());
        B.last_ms = g_last_kernel_ms;
        uint32_t h[4] = {0, 0, 0, 0};
        HIP_TRY(hipMemcpyAsync(h, B.status, sizeof h, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (h[0] & ERR_NO_NEXT)
            return mgx::fail(MGX_E_NO_NEXT, "a ray from an interior point found no valid next point");
        if (h[0] & ERR_INTERP)
            return mgx::fail(MGX_E_UNSUPPORTED,
                             "a segment needs more than 65535 interpolation points");
        if (!(h[0] & ERR_CAP)) return MGX_OK;
        const uint32_t max_cap = max_walk_cap(8 + 8 * (size_t)B.maxL);
        const uint32_t cap = grown_cap(B.cap, h[1], max_cap);
        if (!cap) return mgx::fail(MGX_E_RAY_LENGTH, "a ray needs more than %u walk steps", max_cap);
        B.cap = cap;
        batch_drop_graph(B);
        if (int e = batch_upload(B, st)) return e;
    }
}

} // namespace

namespace {

/* argument checks of the old imager (Model::compute_image / compute_image_optical_depth,
   model.cpp:957-976, 1061-1076; the 1D rule of the Image constructor, image.cpp:12-26) and
   the launch of image_ray_kernel */
template <int KIND>
int run_image_ray(const mgx_model *model, const mgx_options *opts, uint32_t rr, double *out,
                  double *ImX, double *ImY, hipStream_t st)
{
    if (!model || !opts || !out || !ImX || !ImY) return mgx::fail(MGX_E_ARG, "null argument");
    if (opts->approx == MGX_APPROX_ONE_LINE)
        return mgx::fail(MGX_E_UNSUPPORTED, "One line approximation is not supported for imaging.");
    if (opts->approx < 0 || opts->approx > 2) return mgx::fail(MGX_E_ARG, "unknown approximation");
    if (opts->precision != MGX_PREC_FP64 && opts->precision != MGX_PREC_MIXED)
        return mgx::fail(MGX_E_ARG, "unknown precision mode");
    if (!model->position || !model->velocity || !model->nu || !model->ray_direction ||
        !model->antipod || !model->point2boundary || !model->interp_crit || !model->sorted_line ||
        !model->emissivity)
        return mgx::fail(MGX_E_ARG, "null array in model view");
    if (model->npoints == 0 || model->nfreqs == 0 || model->nlines == 0)
        return mgx::fail(MGX_E_ARG, "empty model");
    if (rr >= model->nrays) return mgx::fail(MGX_E_ARG, "ray %u out of range", rr);
    if (model->adaptive_directions)
        return mgx::fail(MGX_E_UNSUPPORTED,
                         "the imager walks the model's global ray directions (no adaptive directions)");
    if (model->dimension == 1) {
        double d[3];
        HIP_TRY(hipMemcpyAsync(d, model->ray_direction + 3 * (size_t)rr, sizeof d,
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (d[0] != 0.0 || d[1] != 1.0 || d[2] != 0.0)
            return mgx::fail(MGX_E_ARG, "In 1D, the image ray has to be (0,1,0)");
    }
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    return launch_with_capacity(st, b, kp, [&](size_t lds) {
        /* two points per wave when a half-wave covers the frequencies and two walk
           slots fit the LDS */
        const uint32_t ppw = (model->nfreqs <= 32 && 2 * lds <= lds_budget()) ? 2 : 1;
        const dim3 grid((model->npoints + ppw - 1) / ppw), block(MGX_WAVE);
        const size_t l2 = ppw * lds;
        timed(st, [&] {
            const bool fp64 = opts->precision == MGX_PREC_FP64;
            if (opts->approx == MGX_APPROX_NONE) {
                if (fp64)
                    hipLaunchKernelGGL((image_ray_kernel<MGX_APPROX_NONE, MGX_PREC_FP64, KIND>), grid,
                                       block, l2, st, b.kp_dev, rr, out, ImX, ImY, ppw);
                else
                    hipLaunchKernelGGL((image_ray_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED, KIND>), grid,
                                       block, l2, st, b.kp_dev, rr, out, ImX, ImY, ppw);
            } else {
                if (fp64)
                    hipLaunchKernelGGL((image_ray_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64, KIND>),
                                       grid, block, l2, st, b.kp_dev, rr, out, ImX, ImY, ppw);
                else
                    hipLaunchKernelGGL((image_ray_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED, KIND>),
                                       grid, block, l2, st, b.kp_dev, rr, out, ImX, ImY, ppw);
            }
        });
    });
}

} // namespace

extern "C" {

#ifndef MGX_ADJOINT_TU
const char *mgx_last_error(void) { return mgx::error_buffer(); }

const char *mgx_version(void) { return "magritte_amd 0.1.0 (gfx950)"; }

namespace {

/* the plain-method solve on prepared buffers (shared by the allocating and the
   no-allocation entry points) */
int forward_launch(const mgx_model *model, const mgx_options *opts, double *u, double *J_out,
                   double *lambda_diag, hipStream_t st, DevBuffers &b, KParams &kp,
                   double *Jline = nullptr, double **J_scratch = nullptr)
{
    bool scratch_failed = false;
    const int e = launch_with_capacity(st, b, kp, [&](size_t) {
        /* the line-integrated solve keeps J in LDS, or else in a scratch J (P, F) */
        double *J = J_out;
        if (Jline && !kp.j_lds) {
            if (!*J_scratch &&
                hipMalloc(J_scratch, (size_t)model->npoints * model->nfreqs * sizeof(double)) !=
                    hipSuccess) {
                scratch_failed = true;
                return;
            }
            J = *J_scratch;
        }
        const dim3 grid(model->npoints), block(MGX_WAVE);
        const size_t lds = solve_lds_bytes(kp.cap, model->nfreqs, model->nlines, kp.j_lds != 0);
        timed(st, [&] {
            const bool fp64 = opts->precision == MGX_PREC_FP64;
            switch (opts->approx) {
            case MGX_APPROX_NONE:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_kernel<MGX_APPROX_NONE, MGX_PREC_FP64>), grid,
                                       block, lds, st, b.kp_dev, u, J, lambda_diag, Jline);
                else
                    hipLaunchKernelGGL((feautrier_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED>), grid,
                                       block, lds, st, b.kp_dev, u, J, lambda_diag, Jline);
                break;
            case MGX_APPROX_ONE_LINE:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_FP64>), grid,
                                       block, lds, st, b.kp_dev, u, J, lambda_diag, Jline);
                else
                    hipLaunchKernelGGL((feautrier_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, u, J, lambda_diag, Jline);
                break;
            default:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>),
                                       grid, block, lds, st, b.kp_dev, u, J, lambda_diag, Jline);
                else
                    hipLaunchKernelGGL((feautrier_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, u, J, lambda_diag, Jline);
            }
        });
    }, 8 * (size_t)model->nlines, true);
    if (scratch_failed) return mgx::fail(MGX_E_HIP, "allocation of the J scratch failed");
    return e;
}

} // namespace

int mgx_feautrier_forward_device(const mgx_model *model, const mgx_options *opts, double *u,
                                 double *J, double *lambda_diag, void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (int e = check_lambda_diagonal(opts)) return e;
    if (!J) return mgx::fail(MGX_E_ARG, "J is required");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    return forward_launch(model, opts, u, J, lambda_diag, st, b, kp);
}

int mgx_feautrier_sparse_device(const mgx_model *model, const mgx_options *opts, double *J_line,
                                double *lambda_diag, void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (int e = check_lambda_diagonal(opts)) return e;
    if (!J_line) return mgx::fail(MGX_E_ARG, "J_line is required");
    if (!model->line_species) return mgx::fail(MGX_E_ARG, "null array in model view");
    hipStream_t st = (hipStream_t)stream;
    /* the line sums read bin z of line l at l nquads + z: the line grid's order */
    {
        std::vector<uint32_t> fl(model->nfreqs), fz(model->nfreqs);
        if (hipMemcpyAsync(fl.data(), model->f_line, fl.size() * 4, hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipMemcpyAsync(fz.data(), model->f_z, fz.size() * 4, hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return mgx::fail(MGX_E_HIP, "reading the frequency map failed");
        for (uint32_t f = 0; f < model->nfreqs; f++)
            if (fl[f] != f / model->nquads || fz[f] != f % model->nquads)
                return mgx::fail(MGX_E_UNSUPPORTED,
                                 "the sparse solve needs the line frequency grid (bin z of line l at l*nquads+z)");
    }
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    /* J of an origin stays in LDS; when it does not fit (long walks), its row of a
       scratch J, allocated by the launch that needs it */
    double *Jtmp = nullptr;
    struct Free {
        double *&p;
        ~Free() { if (p) (void)hipFree(p); }
    } free_tmp{Jtmp};
    return forward_launch(model, opts, nullptr, nullptr, lambda_diag, st, b, kp, J_line, &Jtmp);
}

static size_t noalloc_scratch_bytes(const mgx_model *model)
{
    if (!model) return 0;
    const size_t P = model->npoints, PL = (size_t)model->npoints * model->nlines;
    const auto rnd = [](size_t n) { return (n + 255) & ~(size_t)255; };
    return rnd(4 * sizeof(uint32_t)) + rnd(sizeof(KParams)) + rnd(P * sizeof(PointRec)) +
           rnd(P * sizeof(PointIdx)) + rnd(PL * sizeof(double4));
}

int mgx_feautrier_forward_noalloc(const mgx_model *model, const mgx_options *opts, double *u,
                                  double *J, double *lambda_diag, void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (int e = check_lambda_diagonal(opts)) return e;
    if (!J) return mgx::fail(MGX_E_ARG, "J is required");
    const size_t need = noalloc_scratch_bytes(model);
    if (need > (size_t)MGX_NOALLOC_SCRATCH_BYTES)
        return mgx::fail(MGX_E_ARG,
                         "the model needs %zu bytes of scratch (npoints %u, nlines %u); the "
                         "no-allocation solve holds %zu (use mgx_feautrier_forward_device)",
                         need, model->npoints, model->nlines, (size_t)MGX_NOALLOC_SCRATCH_BYTES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    unsigned char *base = nullptr;
    if (int e = scratch_base(&base)) return e;
    const size_t P = model->npoints, PL = (size_t)model->npoints * model->nlines;
    size_t off = 0;
    const auto take = [&](size_t bytes) {
        unsigned char *p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    DevBuffers b;
    b.owned = false;
    b.status = reinterpret_cast<uint32_t *>(take(4 * sizeof(uint32_t)));
    b.kp_dev = reinterpret_cast<KParams *>(take(sizeof(KParams)));
    b.pts = reinterpret_cast<PointRec *>(take(P * sizeof(PointRec)));
    b.pidx = reinterpret_cast<PointIdx *>(take(P * sizeof(PointIdx)));
    b.lpack = reinterpret_cast<double4 *>(take(PL * sizeof(double4)));
    struct NoEvents {
        NoEvents() { t_no_events = true; }
        ~NoEvents() { t_no_events = false; }
    } no_events;
    hipStream_t st = (hipStream_t)stream;
    KParams kp;
    if (int e = prepare_tables(model, opts, st, b, kp)) return e;
    return forward_launch(model, opts, u, J, lambda_diag, st, b, kp);
}

int mgx_feautrier_forward_lambda_device(const mgx_model *model, const mgx_options *opts,
                                        double *u, double *J, uint32_t cap, uint32_t *lam_count,
                                        uint32_t *lam_nr, double *lam_val, void *stream)
{
    mgx::clear_error();
    if (!opts) return mgx::fail(MGX_E_ARG, "null options");
    if (opts->n_off_diag > MGX_MAX_OFF_DIAG)
        return mgx::fail(MGX_E_UNSUPPORTED, "n_off_diag > %d is not supported", MGX_MAX_OFF_DIAG);
    mgx_options op = *opts;
    op.n_off_diag = 0; /* the rest of the checks as the diagonal solve */
    if (int e = check_view(model, &op)) return e;
    if (!J || !lam_count || !lam_nr || !lam_val) return mgx::fail(MGX_E_ARG, "null argument");
    if (cap < 1) return mgx::fail(MGX_E_ARG, "the Lambda lists need a capacity of at least 1");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, &op, st, b, kp)) return e;
    const LamBands lb{lam_count, lam_nr, lam_val, cap, opts->n_off_diag};
    return launch_with_capacity(st, b, kp, [&](size_t) {
        const dim3 grid(model->npoints), block(MGX_WAVE);
        const size_t lds = solve_lds_bytes(kp.cap, model->nfreqs, model->nlines, kp.j_lds != 0);
        timed(st, [&] {
            const bool fp64 = op.precision == MGX_PREC_FP64;
            switch (op.approx) {
            case MGX_APPROX_NONE:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_bands_kernel<MGX_APPROX_NONE, MGX_PREC_FP64>), grid,
                                       block, lds, st, b.kp_dev, u, J, lb);
                else
                    hipLaunchKernelGGL((feautrier_bands_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED>), grid,
                                       block, lds, st, b.kp_dev, u, J, lb);
                break;
            case MGX_APPROX_ONE_LINE:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_bands_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_FP64>),
                                       grid, block, lds, st, b.kp_dev, u, J, lb);
                else
                    hipLaunchKernelGGL((feautrier_bands_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, u, J, lb);
                break;
            default:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_bands_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>),
                                       grid, block, lds, st, b.kp_dev, u, J, lb);
                else
                    hipLaunchKernelGGL((feautrier_bands_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, u, J, lb);
            }
        });
    }, 8 * (size_t)model->nlines, true);
}

int mgx_shortchar_order_0_device(const mgx_model *model, const mgx_options *opts, double min_dtau,
                                 double *I, double *u, double *v, double *J, double *lambda_diag,
                                 void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (!J) return mgx::fail(MGX_E_ARG, "J is required");
    if (!(min_dtau > 0.0)) return mgx::fail(MGX_E_ARG, "min_dtau must be positive");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    const ShortcharOut sc{I, v, min_dtau};
    return launch_with_capacity(st, b, kp, [&](size_t) {
        const dim3 grid(model->npoints), block(MGX_WAVE);
        const size_t lds = solve_lds_bytes(kp.cap, model->nfreqs, model->nlines, kp.j_lds != 0);
        timed(st, [&] {
            const bool fp64 = opts->precision == MGX_PREC_FP64;
            switch (opts->approx) {
            case MGX_APPROX_NONE:
                if (fp64)
                    hipLaunchKernelGGL((shortchar_kernel<MGX_APPROX_NONE, MGX_PREC_FP64>), grid,
                                       block, lds, st, b.kp_dev, sc, u, J, lambda_diag);
                else
                    hipLaunchKernelGGL((shortchar_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED>), grid,
                                       block, lds, st, b.kp_dev, sc, u, J, lambda_diag);
                break;
            case MGX_APPROX_ONE_LINE:
                if (fp64)
                    hipLaunchKernelGGL((shortchar_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_FP64>), grid,
                                       block, lds, st, b.kp_dev, sc, u, J, lambda_diag);
                else
                    hipLaunchKernelGGL((shortchar_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, sc, u, J, lambda_diag);
                break;
            default:
                if (fp64)
                    hipLaunchKernelGGL((shortchar_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>),
                                       grid, block, lds, st, b.kp_dev, sc, u, J, lambda_diag);
                else
                    hipLaunchKernelGGL((shortchar_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, sc, u, J, lambda_diag);
            }
        });
    }, 8 * (size_t)model->nlines, true);
}

int mgx_feautrier_order_2_uv_device(const mgx_model *model, const mgx_options *opts, double *u,
                                    double *v, void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (!u || !v) return mgx::fail(MGX_E_ARG, "u and v are required");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    return launch_with_capacity(st, b, kp, [&](size_t) {
        const dim3 grid(model->npoints), block(MGX_WAVE);
        const size_t lds = solve_lds_bytes(kp.cap, model->nfreqs, model->nlines, false);
        timed(st, [&] {
            const bool fp64 = opts->precision == MGX_PREC_FP64;
            switch (opts->approx) {
            case MGX_APPROX_NONE:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_uv_kernel<MGX_APPROX_NONE, MGX_PREC_FP64>), grid,
                                       block, lds, st, b.kp_dev, u, v);
                else
                    hipLaunchKernelGGL((feautrier_uv_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED>), grid,
                                       block, lds, st, b.kp_dev, u, v);
                break;
            case MGX_APPROX_ONE_LINE:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_uv_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_FP64>),
                                       grid, block, lds, st, b.kp_dev, u, v);
                else
                    hipLaunchKernelGGL((feautrier_uv_kernel<MGX_APPROX_ONE_LINE, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, u, v);
                break;
            default:
                if (fp64)
                    hipLaunchKernelGGL((feautrier_uv_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>),
                                       grid, block, lds, st, b.kp_dev, u, v);
                else
                    hipLaunchKernelGGL((feautrier_uv_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>),
                                       grid, block, lds, st, b.kp_dev, u, v);
            }
        });
    }, 8 * (size_t)model->nlines, true);
}


int mgx_image_device(const mgx_model *model, const mgx_options *opts, const double *raydir,
                     uint32_t nx, uint32_t ny, double *I, double *ImX, double *ImY, void *stream)
{
    mgx::clear_error();
    if (!I || !ImX || !ImY) return mgx::fail(MGX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    ImageFrame *dfr = nullptr;
    if (int e = image_setup(model, opts, raydir, nx, ny, st, b, kp, dfr)) return e;
    const WalkKey ikey = image_key(*model, raydir, nx, ny);
    return launch_with_capacity(st, b, kp, [&](size_t lds) {
        /* two pixels per wave when a half-wave covers the frequencies */
        const uint32_t ppw = model->nfreqs <= 32 ? 2 : 1;
        const dim3 grid((nx * ny + ppw - 1) / ppw), block(MGX_WAVE);
        timed(st, [&] {
            const bool fp64 = opts->precision == MGX_PREC_FP64;
            if (opts->approx == MGX_APPROX_NONE) {
                if (fp64)
                    hipLaunchKernelGGL((image_kernel<MGX_APPROX_NONE, MGX_PREC_FP64, IMG_INTENSITY>), grid, block,
                                       lds, st, b.kp_dev, dfr, I, ImX, ImY, ppw);
                else
                    hipLaunchKernelGGL((image_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED, IMG_INTENSITY>), grid, block,
                                       lds, st, b.kp_dev, dfr, I, ImX, ImY, ppw);
            } else {
                if (fp64)
                    hipLaunchKernelGGL((image_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64, IMG_INTENSITY>), grid,
                                       block, lds, st, b.kp_dev, dfr, I, ImX, ImY, ppw);
                else
                    hipLaunchKernelGGL((image_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED, IMG_INTENSITY>), grid,
                                       block, lds, st, b.kp_dev, dfr, I, ImX, ImY, ppw);
            }
        });
    }, 0, false, &ikey);
}

int mgx_image_optical_depth_device(const mgx_model *model, const mgx_options *opts,
                                   const double *raydir, uint32_t nx, uint32_t ny, double *tau,
                                   double *ImX, double *ImY, void *stream)
{
    mgx::clear_error();
    if (!tau || !ImX || !ImY) return mgx::fail(MGX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    ImageFrame *dfr = nullptr;
    if (int e = image_setup(model, opts, raydir, nx, ny, st, b, kp, dfr)) return e;
    const WalkKey ikey = image_key(*model, raydir, nx, ny);
    return launch_with_capacity(st, b, kp, [&](size_t lds) {
        /* two pixels per wave when a half-wave covers the frequencies */
        const uint32_t ppw = model->nfreqs <= 32 ? 2 : 1;
        const dim3 grid((nx * ny + ppw - 1) / ppw), block(MGX_WAVE);
        timed(st, [&] {
            const bool fp64 = opts->precision == MGX_PREC_FP64;
            if (opts->approx == MGX_APPROX_NONE) {
                if (fp64)
                    hipLaunchKernelGGL((image_kernel<MGX_APPROX_NONE, MGX_PREC_FP64, IMG_OPTICAL_DEPTH>),
                                       grid, block, lds, st, b.kp_dev, dfr, tau, ImX, ImY, ppw);
                else
                    hipLaunchKernelGGL((image_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED, IMG_OPTICAL_DEPTH>),
                                       grid, block, lds, st, b.kp_dev, dfr, tau, ImX, ImY, ppw);
            } else {
                if (fp64)
                    hipLaunchKernelGGL(
                        (image_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64, IMG_OPTICAL_DEPTH>), grid,
                        block, lds, st, b.kp_dev, dfr, tau, ImX, ImY, ppw);
                else
                    hipLaunchKernelGGL(
                        (image_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED, IMG_OPTICAL_DEPTH>), grid,
                        block, lds, st, b.kp_dev, dfr, tau, ImX, ImY, ppw);
            }
        });
    }, 0, false, &ikey);
}

int mgx_image_ray_device(const mgx_model *model, const mgx_options *opts, uint32_t rr, double *I,
                         double *ImX, double *ImY, void *stream)
{
    mgx::clear_error();
    return run_image_ray<IMG_INTENSITY>(model, opts, rr, I, ImX, ImY, (hipStream_t)stream);
}

int mgx_image_optical_depth_ray_device(const mgx_model *model, const mgx_options *opts,
                                       uint32_t rr, double *tau, double *ImX, double *ImY,
                                       void *stream)
{
    mgx::clear_error();
    return run_image_ray<IMG_OPTICAL_DEPTH>(model, opts, rr, tau, ImX, ImY, (hipStream_t)stream);
}

#endif /* !MGX_ADJOINT_TU */

#ifdef MGX_ADJOINT_TU
static int J_adjoint_entry(const mgx_model *model, const mgx_options *opts, const double *J_bar,
                           double *d_eps, double *d_kap, double *d_iw, void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (!J_bar || !d_eps || !d_kap) return mgx::fail(MGX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    const bool fp64 = opts->precision == MGX_PREC_FP64;
    switch (opts->approx) {
    case MGX_APPROX_NONE:
        return fp64 ? run_J_adjoint<MGX_APPROX_NONE, MGX_PREC_FP64>(st, b, kp, J_bar, d_eps, d_kap, d_iw)
                    : run_J_adjoint<MGX_APPROX_NONE, MGX_PREC_MIXED>(st, b, kp, J_bar, d_eps, d_kap, d_iw);
    case MGX_APPROX_ONE_LINE:
        return fp64 ? run_J_adjoint<MGX_APPROX_ONE_LINE, MGX_PREC_FP64>(st, b, kp, J_bar, d_eps, d_kap, d_iw)
                    : run_J_adjoint<MGX_APPROX_ONE_LINE, MGX_PREC_MIXED>(st, b, kp, J_bar, d_eps, d_kap, d_iw);
    default:
        return fp64 ? run_J_adjoint<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64>(st, b, kp, J_bar, d_eps, d_kap, d_iw)
                    : run_J_adjoint<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED>(st, b, kp, J_bar, d_eps, d_kap, d_iw);
    }
}

int mgx_feautrier_adjoint_device(const mgx_model *model, const mgx_options *opts,
                                 const double *J_bar, double *d_eps, double *d_kap, void *stream)
{
    return J_adjoint_entry(model, opts, J_bar, d_eps, d_kap, nullptr, stream);
}

int mgx_feautrier_adjoint_width_device(const mgx_model *model, const mgx_options *opts,
                                       const double *J_bar, double *d_eps, double *d_kap,
                                       double *d_inverse_width, void *stream)
{
    if (!d_inverse_width) {
        mgx::clear_error();
        return mgx::fail(MGX_E_ARG, "null argument");
    }
    return J_adjoint_entry(model, opts, J_bar, d_eps, d_kap, d_inverse_width, stream);
}

static int image_adjoint_entry(const mgx_model *model, const mgx_options *opts,
                               const double *raydir, uint32_t nx, uint32_t ny,
                               const double *I_bar, double *d_eps, double *d_kap, double *d_iw,
                               void *stream)
{
    mgx::clear_error();
    if (!I_bar || !d_eps || !d_kap) return mgx::fail(MGX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    DevBuffers b;
    KParams kp;
    ImageFrame *dfr = nullptr;
    if (int e = image_setup(model, opts, raydir, nx, ny, st, b, kp, dfr)) return e;
    const size_t PL = (size_t)model->npoints * model->nlines;
    const bool fp64 = opts->precision == MGX_PREC_FP64;
    const WalkKey ikey = image_key(*model, raydir, nx, ny);
    auto run = [&](auto kern) {
        return adjoint_launch(
            st, b, kp, ikey, d_eps, d_kap, d_iw, PL, 1,
            [&](size_t lds) {
                const uint32_t ppw = model->nfreqs <= 32 ? 2 : 1; /* pixels per wave */
                return persistent_waves(kern, lds, (nx * ny + ppw - 1) / ppw);
            },
            [&](size_t lds, uint32_t grid, uint32_t capE) {
                timed(st, [&] {
                    hipLaunchKernelGGL(kern, dim3(grid), dim3(MGX_WAVE), lds, st, b.kp_dev, dfr,
                                       I_bar, d_eps, d_kap, d_iw, g_ws, capE);
                });
            });
    };
    if (d_iw) {
        if (opts->approx == MGX_APPROX_NONE)
            return fp64 ? run(image_adjoint_kernel<MGX_APPROX_NONE, MGX_PREC_FP64, true>)
                        : run(image_adjoint_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED, true>);
        return fp64 ? run(image_adjoint_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64, true>)
                    : run(image_adjoint_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED, true>);
    }
    if (opts->approx == MGX_APPROX_NONE)
        return fp64 ? run(image_adjoint_kernel<MGX_APPROX_NONE, MGX_PREC_FP64, false>)
                    : run(image_adjoint_kernel<MGX_APPROX_NONE, MGX_PREC_MIXED, false>);
    return fp64 ? run(image_adjoint_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_FP64, false>)
                : run(image_adjoint_kernel<MGX_APPROX_CLOSE_LINES, MGX_PREC_MIXED, false>);
}

int mgx_image_adjoint_device(const mgx_model *model, const mgx_options *opts, const double *raydir,
                             uint32_t nx, uint32_t ny, const double *I_bar, double *d_eps,
                             double *d_kap, void *stream)
{
    return image_adjoint_entry(model, opts, raydir, nx, ny, I_bar, d_eps, d_kap, nullptr, stream);
}

int mgx_image_adjoint_width_device(const mgx_model *model, const mgx_options *opts,
                                   const double *raydir, uint32_t nx, uint32_t ny,
                                   const double *I_bar, double *d_eps, double *d_kap,
                                   double *d_inverse_width, void *stream)
{
    if (!d_inverse_width) {
        mgx::clear_error();
        return mgx::fail(MGX_E_ARG, "null argument");
    }
    return image_adjoint_entry(model, opts, raydir, nx, ny, I_bar, d_eps, d_kap, d_inverse_width,
                               stream);
}

void mgx_release_workspace(void)
{
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (g_ws) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_ws);
    }
    g_ws = nullptr;
    g_ws_bytes = 0;
}

#endif /* MGX_ADJOINT_TU */

#ifndef MGX_ADJOINT_TU
void mgx_set_kernel_timing(int enable) { g_timing = enable != 0; }

double mgx_last_kernel_ms(void) { return g_last_kernel_ms; }

int mgx_trace_stats_device(const mgx_model *model, const mgx_options *opts, uint64_t *stats,
                           void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (!stats) return mgx::fail(MGX_E_ARG, "stats is required");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    unsigned long long *d = nullptr;
    HIP_TRY(hipMalloc(&d, 4 * sizeof(unsigned long long)));
    int e = launch_with_capacity(st, b, kp, [&](size_t lds) {
        (void)hipMemsetAsync(d, 0, 4 * sizeof(unsigned long long), st);
        hipLaunchKernelGGL(trace_stats_kernel, dim3(model->npoints, model->hnrays),
                           dim3(MGX_WAVE), lds, st, b.kp_dev, d);
    });
    if (!e) {
        hipError_t he = hipMemcpy(stats, d, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
        if (he != hipSuccess) e = mgx::fail(MGX_E_HIP, "stats copy: %s", hipGetErrorString(he));
    }
    (void)hipFree(d);
    return e;
}

int mgx_trace_counts_device(const mgx_model *model, const mgx_options *opts, uint32_t *n_tot,
                            void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (!n_tot) return mgx::fail(MGX_E_ARG, "n_tot is required");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    return launch_with_capacity(st, b, kp, [&](size_t lds) {
        hipLaunchKernelGGL(trace_counts_kernel, dim3(model->npoints, model->hnrays),
                           dim3(MGX_WAVE), lds, st, b.kp_dev, n_tot);
    });
}

int mgx_trace_rays_device(const mgx_model *model, const mgx_options *opts, uint32_t nsel,
                          const uint32_t *sel_rr, const uint32_t *sel_o, const uint32_t *sel_l,
                          uint32_t capacity, uint32_t *count, uint32_t *centre_idx, uint32_t *nr,
                          uint32_t *nr_interp, double *shift, double *interp_factor, double *dZ,
                          void *stream)
{
    mgx::clear_error();
    if (int e = check_view(model, opts)) return e;
    if (nsel == 0) return MGX_OK;
    if (!sel_rr || !sel_o || !sel_l || !count || !centre_idx || !nr || !nr_interp || !shift ||
        !interp_factor || !dZ || capacity == 0)
        return mgx::fail(MGX_E_ARG, "null output or zero capacity");
    hipStream_t st = (hipStream_t)stream;
    DevBuffers b;
    KParams kp;
    if (int e = prepare(model, opts, st, b, kp)) return e;
    const int e = launch_with_capacity(st, b, kp, [&](size_t lds) {
        hipLaunchKernelGGL(trace_rays_kernel, dim3(nsel), dim3(MGX_WAVE), lds, st, b.kp_dev, sel_rr,
                           sel_o, sel_l, capacity, count, centre_idx, nr, nr_interp, shift,
                           interp_factor, dZ);
    });
    return e;
}

int mgx_batch_create(const mgx_model *models, uint32_t n, const mgx_options *opts,
                     double *const *J, double *const *lambda_diag, mgx_batch **out, void *stream)
{
    mgx::clear_error();
    if (!models || !opts || !J || !out || n == 0) return mgx::fail(MGX_E_ARG, "null argument");
    *out = nullptr;
    for (uint32_t i = 0; i < n; i++) {
        if (int e = check_view(&models[i], opts)) return e;
        if (int e = check_lambda_diagonal(opts)) return e;
        if (!J[i]) return mgx::fail(MGX_E_ARG, "J of model %u is null", i);
        if (lambda_diag && !lambda_diag[i]) return mgx::fail(MGX_E_ARG, "lambda of model %u is null", i);
    }
    hipStream_t st = (hipStream_t)stream;
    mgx_batch *B = new mgx_batch();
    B->n = n;
    B->op = *opts;
    B->with_lam = lambda_diag != nullptr;
    auto fail = [&](int e) {
        delete B;
        return e;
    };
    if (hipMalloc(&B->status, 4 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&B->counter, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&B->kps_dev, n * sizeof(KParams)) != hipSuccess ||
        hipMalloc(&B->J_dev, n * sizeof(double *)) != hipSuccess ||
        (B->with_lam && hipMalloc(&B->lam_dev, n * sizeof(double *)) != hipSuccess))
        return fail(mgx::fail(MGX_E_HIP, "batch allocation failed"));
    for (uint32_t i = 0; i < n; i++) {
        B->views.push_back(models[i]);
        B->bufs.push_back(new DevBuffers());
        KParams kp;
        if (int e = prepare(&models[i], opts, st, *B->bufs.back(), kp, true)) return fail(e);
        kp.status = B->status;
        B->kps.push_back(kp);
        B->J.push_back(J[i]);
        B->lam.push_back(lambda_diag ? lambda_diag[i] : nullptr);
        B->maxF = std::max(B->maxF, models[i].nfreqs);
        B->maxL = std::max(B->maxL, models[i].nlines);
    }
    if (hipMemcpyAsync(B->J_dev, B->J.data(), n * sizeof(double *), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        (B->with_lam && hipMemcpyAsync(B->lam_dev, B->lam.data(), n * sizeof(double *),
                                       hipMemcpyHostToDevice, st) != hipSuccess))
        return fail(mgx::fail(MGX_E_HIP, "batch upload failed"));
    if (int e = batch_upload(*B, st)) return fail(e);
    /* the walk buffer from the longest half-ray of any model (walk_max_kernel) */
    if (hipMemsetAsync(B->status, 0, 4 * sizeof(uint32_t), st) != hipSuccess)
        return fail(mgx::fail(MGX_E_HIP, "batch status reset failed"));
    for (uint32_t i = 0; i < n; i++)
        hipLaunchKernelGGL(walk_max_kernel, dim3(models[i].npoints, models[i].hnrays), dim3(MGX_WAVE),
                           0, st, B->kps_dev + i, B->status + 3);
    uint32_t h[4] = {0, 0, 0, 0};
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h, B->status, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(mgx::fail(MGX_E_HIP, "batch walk-length pass failed"));
    if (h[0] & ERR_NO_NEXT)
        return fail(mgx::fail(MGX_E_NO_NEXT, "a ray from an interior point found no valid next point"));
    const uint32_t max_cap = max_walk_cap(8 + 8 * (size_t)B->maxL);
    if (h[3] > max_cap)
        return fail(mgx::fail(MGX_E_RAY_LENGTH, "a ray needs more than %u walk steps", max_cap));
    B->cap = std::max<uint32_t>(16, (h[3] + 15) / 16 * 16);
    if (int e = batch_upload(*B, st)) return fail(e);
    *out = B;
    return MGX_OK;
}

int mgx_batch_configure(mgx_batch *B, uint32_t rays_per_item, uint32_t order, uint32_t wps,
                        int use_graph, void *stream)
{
    mgx::clear_error();
    if (!B) return mgx::fail(MGX_E_ARG, "null batch");
    if (order > 1 || (wps != 2 && wps != 3 && wps != 4))
        return mgx::fail(MGX_E_ARG, "order must be 0 or 1 and wps 2, 3 or 4");
    B->cfg = BatchConfig{rays_per_item, order, wps};
    B->use_graph = use_graph != 0;
    batch_drop_graph(*B);
    return batch_upload(*B, (hipStream_t)stream);
}

int mgx_batch_run(mgx_batch *B, void *stream)
{
    mgx::clear_error();
    if (!B) return mgx::fail(MGX_E_ARG, "null batch");
    return batch_execute(*B, (hipStream_t)stream);
}

int mgx_batch_set_images(mgx_batch *B, const mgx_model *image_views, const double *raydirs,
                         uint32_t nx, uint32_t ny, double *const *I, double *const *ImX,
                         double *const *ImY, void *stream)
{
    mgx::clear_error();
    if (!B || !image_views || !raydirs || !I || !ImX || !ImY) return mgx::fail(MGX_E_ARG, "null argument");
    if (nx == 0 || ny == 0) return mgx::fail(MGX_E_ARG, "the image needs at least one pixel");
    if (B->op.approx == MGX_APPROX_ONE_LINE)
        return mgx::fail(MGX_E_UNSUPPORTED, "One line approximation is not supported for imaging.");
    const uint64_t npix = (uint64_t)nx * ny;
    const uint64_t chunks_per_model = (npix + kImageChunk - 1) / kImageChunk;
    if (npix > 0xffffffffull || chunks_per_model * B->n > 0xffffffffull)
        return mgx::fail(MGX_E_ARG, "too many pixels in the batch's images");
    hipStream_t st = (hipStream_t)stream;
    std::vector<ImageFrame> frames(B->n);
    for (uint32_t i = 0; i < B->n; i++) {
        const mgx_model &v = image_views[i], &m = B->views[i];
        if (!I[i] || !ImX[i] || !ImY[i]) return mgx::fail(MGX_E_ARG, "image outputs of model %u are null", i);
        if (v.npoints != m.npoints || v.nlines != m.nlines || v.nboundary != m.nboundary ||
            v.dimension != m.dimension || v.spherical_symmetry != m.spherical_symmetry ||
            v.max_inverse_mass != m.max_inverse_mass)
            return mgx::fail(MGX_E_ARG, "image view %u does not describe batch model %u", i, i);
        /* the image shares the solve's packed point and line tables */
        if (v.position != m.position || v.velocity != m.velocity || v.n_neighbors != m.n_neighbors ||
            v.cum_n_neighbors != m.cum_n_neighbors || v.neighbors != m.neighbors ||
            v.point2boundary != m.point2boundary || v.boundary2point != m.boundary2point ||
            v.emissivity != m.emissivity || v.opacity != m.opacity ||
            v.inverse_width != m.inverse_width || v.dshift_max != m.dshift_max ||
            v.temperature != m.temperature || v.vturb2 != m.vturb2 || v.interp_crit != m.interp_crit)
            return mgx::fail(MGX_E_ARG, "image view %u must share the geometry and line arrays of batch model %u", i, i);
        if (v.nfreqs == 0 || !v.nu || !v.f_line || !v.sorted_line || !v.sorted_line_map ||
            !v.boundary_condition || !v.boundary_temperature)
            return mgx::fail(MGX_E_ARG, "image view %u has no frequencies or boundary data", i);
        const double *rd = &raydirs[3 * (size_t)i];
        const double sq = rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2];
        if (sq == 0.0)
            return mgx::fail(MGX_E_ARG, "The given ray direction vector does not point in a "
                                        "direction. Please use a non-zero (normed) direction vector "
                                        "to generate an image.");
        ImageFrame fr;
        std::memset(&fr, 0, sizeof fr);
        const double inv = 1 / std::sqrt(sq);
        for (int k = 0; k < 3; k++) fr.rd[k] = rd[k] * inv;
        if (v.dimension == 1 && (fr.rd[0] != 0.0 || fr.rd[1] != 1.0 || fr.rd[2] != 0.0))
            return mgx::fail(MGX_E_ARG, "In 1D, the image ray has to be (0,1,0)");
        fr.nx = nx;
        fr.ny = ny;
        frames[i] = fr;
    }
    batch_drop_graph(*B);
    B->free_images();
    const uint32_t n = B->n;
    if (hipMalloc(&B->img_kps_dev, n * sizeof(KParams)) != hipSuccess ||
        hipMalloc(&B->frames_dev, n * sizeof(ImageFrame)) != hipSuccess ||
        hipMalloc(&B->chunk_start_dev, (n + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&B->slot_model_dev, n * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&B->img_counter, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&B->I_dev, n * sizeof(double *)) != hipSuccess ||
        hipMalloc(&B->ImX_dev, n * sizeof(double *)) != hipSuccess ||
        hipMalloc(&B->ImY_dev, n * sizeof(double *)) != hipSuccess) {
        B->free_images();
        return mgx::fail(MGX_E_HIP, "batch image allocation failed");
    }
    /* models by decreasing walk length (the cube root of the point count) */
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return B->views[a].npoints > B->views[b].npoints;
    });
    std::vector<uint32_t> start(n + 1);
    for (uint32_t s_ = 0; s_ <= n; s_++) start[s_] = (uint32_t)(chunks_per_model * s_);
    B->img_kps.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        B->img_kps[i] = B->kps[i];
        B->img_kps[i].m = image_views[i];
        B->img_kps[i].cap = B->cap;
    }
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    hipError_t e = up(B->frames_dev, frames.data(), n * sizeof(ImageFrame));
    if (e == hipSuccess) e = up(B->chunk_start_dev, start.data(), (n + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = up(B->slot_model_dev, order.data(), n * sizeof(uint32_t));
    if (e == hipSuccess) e = up(B->img_kps_dev, B->img_kps.data(), n * sizeof(KParams));
    if (e == hipSuccess) e = up(B->I_dev, I, n * sizeof(double *));
    if (e == hipSuccess) e = up(B->ImX_dev, ImX, n * sizeof(double *));
    if (e == hipSuccess) e = up(B->ImY_dev, ImY, n * sizeof(double *));
    for (uint32_t i = 0; i < n && e == hipSuccess; i++) {
        hipLaunchKernelGGL(image_frame_kernel, dim3(1), dim3(256), 0, st, image_views[i],
                           B->frames_dev + i);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        B->free_images();
        return mgx::fail(MGX_E_HIP, "batch image set-up: %s", hipGetErrorString(e));
    }
    B->n_chunks = (uint32_t)(chunks_per_model * n);
    B->n_img = n;
    return MGX_OK;
}

int mgx_batch_tune(mgx_batch *B, uint32_t sample_stride, uint32_t *best, void *stream)
{
    mgx::clear_error();
    if (!B) return mgx::fail(MGX_E_ARG, "null batch");
    hipStream_t st = (hipStream_t)stream;
    if (sample_stride == 0) sample_stride = 1;
    const BatchConfig cands[] = {{0, 0, 3}, {0, 0, 2}, {0, 0, 4}, {0, 1, 3},
                                 {4, 0, 3}, {2, 0, 3}, {1, 0, 3}, {4, 0, 2}};
    const bool saved_timing = g_timing;
    g_timing = true;
    B->images_on = false; /* the candidates tile the solve; the images are timed apart */
    struct Restore {
        mgx_batch *b;
        ~Restore() { b->images_on = true; }
    } restore{B};
    double best_ms = 1e300;
    BatchConfig best_cfg = B->cfg;
    const int saved_graph = B->use_graph;
    for (const BatchConfig &c : cands) {
        B->cfg = c;
        B->use_graph = 0;
        batch_drop_graph(*B);
        /* time a strided sample of the work list (every sample_stride-th item) */
        std::vector<BatchItem> all = batch_items(*B, c);
        std::vector<BatchItem> sample;
        for (size_t i = 0; i < all.size(); i += sample_stride) sample.push_back(all[i]);
        B->n_items = (uint32_t)sample.size();
        if (B->n_items > B->items_cap) {
            if (int e = batch_upload(*B, st)) {
                g_timing = saved_timing;
                return e;
            }
            B->n_items = (uint32_t)sample.size();
        }
        if (hipMemcpyAsync(B->items_dev, sample.data(), sample.size() * sizeof(BatchItem),
                           hipMemcpyHostToDevice, st) != hipSuccess) {
            g_timing = saved_timing;
            return mgx::fail(MGX_E_HIP, "tuning upload failed");
        }
        double ms = 0.0;
        for (int rep = 0; rep < 2; rep++) {
            if (int e = batch_execute(*B, st)) {
                g_timing = saved_timing;
                return e;
            }
            ms = B->last_ms; /* the second run: caches warm */
        }
        if (ms < best_ms) {
            best_ms = ms;
            best_cfg = c;
        }
    }
    g_timing = saved_timing;
    B->cfg = best_cfg;
    B->use_graph = saved_graph;
    batch_drop_graph(*B);
    if (best) {
        best[0] = best_cfg.rays_per_item;
        best[1] = best_cfg.order;
        best[2] = best_cfg.wps;
    }
    return batch_upload(*B, st);
}

double mgx_batch_last_ms(const mgx_batch *B) { return B ? B->last_ms : 0.0; }

void mgx_batch_destroy(mgx_batch *B) { delete B; }

#endif /* !MGX_ADJOINT_TU */

} // extern "C"
