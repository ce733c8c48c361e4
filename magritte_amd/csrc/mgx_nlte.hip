/*
 * Level populations (non-LTE), sparse variant, on gfx950.
 *
 * Replaces, for the sparse NLTE loop of the reference:
 *   Model::compute_Jeff_sparse                       src/model/model.cpp:549-570
 *   Lines::iteration_using_statistical_equilibrium_sparse   src/model/lines/lines.cpp:139-150
 *   LineProducingSpecies::update_using_statistical_equilibrium_sparse
 *                                                    lineProducingSpecies.tpp:589-683
 *   LineProducingSpecies::correct_negative_populations       :711-802
 *   LineProducingSpecies::check_for_convergence(_trial)      :127-201
 *   LineProducingSpecies::update_using_acceleration(_trial)  :266-369
 *   LineProducingSpecies::update_using_LTE                   :52-75
 *   Lines::set_emissivity_and_opacity                lines.tpp:29-41
 *   Model::compute_level_populations_sparse          model.cpp:756-850
 *   Model::ng_acceleration_criterion<Default/Adaptive>       model.cpp:580-645
 *
 * Layout in HBM: a population vector holds every species' block of (P, nlev)
 * FP64 values, species after species (LineProducingSpecies::index(p, i) =
 * i + p nlev within a block).  The handle keeps the current vector, a ring
 * of the LineProducingSpecies::populations history (residuals are the
 * differences of consecutive entries, computed where they are used) and the
 * trial vectors of the adaptive Ng criterion.
 *
 * Statistical equilibrium: one wavefront per (point, species).  Lane r owns
 * row r of the nlev x nlev rate matrix, held column-major in LDS next to its
 * LU factors, so every accumulation into an entry happens in one lane in the
 * reference's order (radiative, approximated Lambda, collisional).  The
 * solve is Gaussian elimination with partial pivoting (pivot search by a
 * wave reduction, pivot row broadcast from LDS), then two steps of iterative
 * refinement whose residuals are compensated (double-double) dot products.
 * The reference solves the same system by long-double column-pivoted QR.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "mgx.h"
#include "mgx_constants.h"
#include "mgx_error.h"
#include "mgx_fmath.h"
#include "mgx_fmath_ref.h"

#define NL_TRY(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return mgx::fail(MGX_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

namespace {

constexpr int NL_WAVE = 64;
constexpr uint32_t NL_MAX_NLEV = 64;
constexpr uint32_t NL_MAX_ORDER = 12;
constexpr int NL_RED_THREADS = 256;
constexpr int NL_GRAM_BLOCKS = 512;

/* device tables */
struct NCol {
    uint32_t num, op, ntmp, ncol;
    const uint32_t *icol, *jcol;
    const double *tmp, *Cd, *Ce;
};

struct NSpec {
    uint32_t num, nlev, nrad, lid0;
    uint32_t ncolpar, col0;
    uint64_t pop_off;
    const double *energy, *weight, *A, *Bs, *Ba, *freq;
    const uint32_t *irad, *jrad;
};

/* kernel arguments shared by the per-point kernels */
struct NView {
    uint32_t P, L, Q, F, nspec, maxnlev, maxnrad, reserved;
    const NSpec *spec;
    const NCol *col;
    const double *temperature, *abundance, *quad_weights;
    double min_e, min_o, inv_frac, min_rel, pop_prec;
};

/* ------------------------------------------------------------------------ */
/* device helpers                                                            */
/* ------------------------------------------------------------------------ */

__device__ inline double line_emissivity(const NSpec &S, uint32_t k, double ni, double mn)
{
    const double e = S.freq[k] * MGX_HH_OVER_FOUR_PI * S.A[k] * ni;
    return e > mn ? e : mn;
}

__device__ inline double line_opacity(const NSpec &S, uint32_t k, double ni, double nj,
                                      double mn)
{
    const double o = S.freq[k] * MGX_HH_OVER_FOUR_PI * (nj * S.Ba[k] - ni * S.Bs[k]);
    return o > mn ? o : mn;
}

/* tools/interpolation.hpp search(): index of the first tabulated value >= v */
__device__ inline uint32_t search_tmp(const double *x, uint32_t n, double v)
{
    uint32_t start = 0, stop = n - 1;
    if (v >= x[stop]) return stop;
    if (v <= x[start]) return start;
    while (stop > start + 1) {
        const uint32_t mid = (stop + start) / 2;
        if (v > x[mid]) start = mid;
        else if (v < x[mid]) stop = mid;
        else return mid;
    }
    return stop;
}

/* wave-uniform argmax of val with the smallest lane winning ties */
__device__ inline int wave_argmax(double val, int idx)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(val, o);
        const int oi = __shfl_xor(idx, o);
        if (ov > val || (ov == val && oi < idx)) {
            val = ov;
            idx = oi;
        }
    }
    return __shfl(idx, 0);
}

__device__ inline int lane_of_step(int my_step, int c)
{
    const unsigned long long b = __ballot(my_step == c);
    return b ? (int)__ffsll((long long)b) - 1 : 0;
}

/* exp(-E / (k T)) weighted level population (update_using_LTE, set_LTE) */
__device__ inline double boltzmann(const NSpec &S, uint32_t i, double T)
{
    return S.weight[i] * exp(-S.energy[i] / (MGX_KB * T));
}

/*
 * Negative-population correction, inversion check, convergence statistics
 * and the line emissivity / opacity / interpolation criterion of point p of
 * species S from its new populations x (lane r holds level r) and the
 * previous ones old[r].  pv is an nlev-long LDS scratch.  Returns x after
 * the correction.
 */
__device__ double finish_point(const NView &v, const NSpec &S, uint32_t p, uint32_t s, double x,
                               const double *old, double *cur_p, double *pv, double *eps,
                               double *kap, double *ic, double *part)
{
    const uint32_t lane = threadIdx.x, n = S.nlev;
    const double ptot = v.abundance[(size_t)p * v.nspec + S.num];
    const double T = v.temperature[p];
    /* correct_negative_populations, first pass */
    if (ptot > 0.0) {
        if (lane < n && x < 0.0) x = 0.0;
        if (lane < n) pv[lane] = x;
        __syncthreads();
        double tpp = 0.0;
        for (uint32_t i = 0; i < n; i++) tpp += pv[i]; /* the reference's sequential sum */
        x = x * ptot / tpp;
        __syncthreads();
    }
    /* population inversions -> LTE of the levels involved (second pass) */
    if (lane < n) pv[lane] = x;
    __syncthreads();
    bool any = false;
    for (uint32_t k = lane; k < S.nrad; k += NL_WAVE) {
        const double ni = pv[S.irad[k]], nj = pv[S.jrad[k]];
        any = any || (nj < S.Bs[k] / S.Ba[k] * ni * v.inv_frac);
    }
    if (__ballot(any)) {
        if (lane == 0) {
            char mod[NL_MAX_NLEV];
            for (uint32_t i = 0; i < n; i++) mod[i] = 0;
            uint32_t next = 0, curr;
            do {
                curr = next;
                bool inv = false;
                for (uint32_t pass = 0; pass < 2; pass++)
                    for (uint32_t q = 0; q < S.nrad; q++) {
                        const uint32_t k = pass == 0 ? S.nrad - 1 - q : q;
                        const uint32_t I = S.irad[k], J = S.jrad[k];
                        if (pv[J] < S.Bs[k] / S.Ba[k] * pv[I] * v.inv_frac) {
                            const double tot = pv[I] + pv[J];
                            pv[I] = boltzmann(S, I, T);
                            pv[J] = boltzmann(S, J, T);
                            const double pf = pv[I] + pv[J];
                            pv[I] *= tot / pf;
                            pv[J] *= tot / pf;
                            mod[I] = mod[J] = 1;
                            inv = true;
                        }
                    }
                if (inv) {
                    double tot = 0.0, pf = 0.0;
                    for (uint32_t i = 0; i < n; i++)
                        if (mod[i]) {
                            tot += pv[i];
                            pv[i] = boltzmann(S, i, T);
                            pf += pv[i];
                        }
                    for (uint32_t i = 0; i < n; i++)
                        if (mod[i]) pv[i] *= tot / pf;
                }
                next = 0;
                for (uint32_t i = 0; i < n; i++) next += mod[i];
            } while (next != curr);
        }
        __syncthreads();
        if (lane < n) x = pv[lane];
    }
    if (lane < n) cur_p[lane] = x;
    /* check_for_convergence: per-point partial sums (levels in order) */
    const double weight = 1.0 / (double)((uint64_t)v.P * n);
    double rc = 0.0;
    bool counted = false;
    if (lane < n) {
        const double o = old[lane];
        counted = x > v.min_rel * ptot;
        if (counted) rc = 2.0 * fabs((x - o) / (x + o));
    }
    double *rcs = pv + n; /* n more doubles of scratch */
    if (lane < n) rcs[lane] = counted ? rc : -1.0;
    __syncthreads();
    if (lane == 0) {
        double cnt = 0.0, rcm = 0.0, rcmax = 0.0;
        for (uint32_t i = 0; i < n; i++) {
            const double r = rcs[i];
            if (r < 0.0) continue; /* below min_rel_pop_for_convergence */
            if (r > v.pop_prec) cnt += 1.0;
            rcm += weight * r;
            rcmax = r < rcmax ? rcmax : r;
        }
        const size_t P = v.P;
        part[((size_t)s * 3 + 0) * P + p] = cnt;
        part[((size_t)s * 3 + 1) * P + p] = rcm;
        part[((size_t)s * 3 + 2) * P + p] = rcmax;
    }
    /* Lines::set_emissivity_and_opacity + interpolation criterion */
    if (lane < n) pv[lane] = x;
    __syncthreads();
    for (uint32_t k = lane; k < S.nrad; k += NL_WAVE) {
        const double ni = pv[S.irad[k]], nj = pv[S.jrad[k]];
        const size_t idx = (size_t)p * v.L + S.lid0 + k;
        const double o = line_opacity(S, k, ni, nj, v.min_o);
        eps[idx] = line_emissivity(S, k, ni, v.min_e);
        kap[idx] = o;
        ic[idx] = o / ptot;
    }
    return x;
}

/* ------------------------------------------------------------------------ */
/* statistical equilibrium                                                   */
/* ------------------------------------------------------------------------ */

__device__ inline void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

/* LDS per wave (doubles): A0[n n], LU[n n], xs[n], jeff[maxnrad], kold[maxnrad], pv[2 n] */
__host__ __device__ inline size_t stateq_lds_doubles(uint32_t n, uint32_t maxnrad)
{
    return 2 * (size_t)n * n + n + 2 * (size_t)maxnrad + 2 * (size_t)n;
}

/*
 * Inputs of the global statistical equilibrium (update_using_statistical_equilibrium,
 * lineProducingSpecies.tpp:377-523): the general Lambda lists, the block-Jacobi iterate
 * (population layout) and the diagonal terms the reference's approximated-Lambda elements
 * put on the rows of other points (lambda_extra_kernel).  Each sweep solves every
 * point's block with the other points' populations on the right-hand side.
 */
struct GlobalIn {
    const uint32_t *count, *nr; /* (P, L), (P, L, cap) */
    const double *val;          /* (P, L, cap) */
    uint32_t cap;
    const double *x_it;         /* iterate */
    const double *dextra;       /* diagonal terms from other points' lists */
    double *x_out;              /* this sweep's solution */
    unsigned long long *dmax;   /* largest |x_out - x_it| / n_tot (bits of a positive double) */
};

template <bool GLOBAL>
__global__ __launch_bounds__(NL_WAVE) void stateq_kernel(NView v, const double *__restrict__ J,
                                                         const double *__restrict__ lam,
                                                         const double *__restrict__ old_all,
                                                         double *cur_all, double *eps,
                                                         double *kap, double *ic,
                                                         double *Jlin_out, double *Jeff_out,
                                                         double *part, GlobalIn g)
{
    mgx_ref_tables_init();
    extern __shared__ double sm[];
    const uint32_t p = blockIdx.x, s = blockIdx.y, lane = threadIdx.x;
    const NSpec S = v.spec[s];
    const uint32_t n = S.nlev, nrad = S.nrad;
    double *A0 = sm;
    double *LU = A0 + (size_t)n * n;
    double *xs = LU + (size_t)n * n;
    double *jeff = xs + n;
    double *kold = jeff + v.maxnrad;
    double *pv = kold + v.maxnrad;
    const double *old = old_all + S.pop_off + (size_t)p * n;
    double *cur_p = cur_all + S.pop_off + (size_t)p * n;
    const double ptot = v.abundance[(size_t)p * v.nspec + S.num];

    /* compute_Jeff_sparse, and the opacity of the populations before the update */
    for (uint32_t k = lane; k < nrad; k += NL_WAVE) {
        const uint32_t lid = S.lid0 + k;
        const double *Jp = J + (size_t)p * v.F + (size_t)lid * v.Q;
        const double *w = v.quad_weights + (size_t)s * v.Q;
        double jl = 0.0;
        for (uint32_t z = 0; z < v.Q; z++) jl += w[z] * Jp[z];
        const double ni = old[S.irad[k]], nj = old[S.jrad[k]];
        double diff;
        if constexpr (GLOBAL) {
            /* Model::compute_Jeff (model.cpp:514-545): the list's elements times the upper
               population at their emitting points */
            const size_t pl = (size_t)p * v.L + lid;
            diff = 0.0;
            for (uint32_t m = 0; m < g.count[pl]; m++) {
                const uint32_t q = g.nr[pl * g.cap + m];
                diff += g.val[pl * g.cap + m] * old_all[S.pop_off + (size_t)q * n + S.irad[k]];
            }
        } else {
            diff = lam[(size_t)p * v.L + lid] * ni;
        }
        const double je = jl - MGX_HH_OVER_FOUR_PI * diff;
        jeff[k] = je;
        kold[k] = line_opacity(S, k, ni, nj, v.min_o);
        if (Jlin_out) Jlin_out[(size_t)p * v.L + lid] = jl;
        if (Jeff_out) Jeff_out[(size_t)p * v.L + lid] = je;
    }
    for (uint32_t i = lane; i < n * n; i += NL_WAVE) A0[i] = 0.0;
    __syncthreads();

    /* the rate matrix; lane r accumulates row r (column-major A0[c n + r]) */
    const uint32_t r = lane;
    double yoff = 0.0; /* right-hand side of row r from other points (global mode) */
    if (r < n) {
        for (uint32_t k = 0; k < nrad; k++) {
            const uint32_t I = S.irad[k], Jl = S.jrad[k];
            const double je = jeff[k];
            const double v_IJ = S.A[k] + S.Bs[k] * je;
            const double v_JI = S.Ba[k] * je;
            if (r == Jl) {
                A0[(size_t)I * n + r] += v_IJ;
                A0[(size_t)Jl * n + r] -= v_JI;
            }
            if (r == I) {
                A0[(size_t)Jl * n + r] += v_JI;
                A0[(size_t)I * n + r] -= v_IJ;
            }
        }
        for (uint32_t k = 0; k < nrad; k++) {
            const uint32_t I = S.irad[k], Jl = S.jrad[k];
            if constexpr (GLOBAL) {
                /* elements at this point stay in the block; one at another point q couples
                   row (p, lower) to that point's upper level: moved to the right-hand side
                   with the iterate (its diagonal term on q's row is in dextra) */
                const size_t pl = (size_t)p * v.L + S.lid0 + k;
                for (uint32_t m = 0; m < g.count[pl]; m++) {
                    const uint32_t q = g.nr[pl * g.cap + m];
                    const double v_IJ = -g.val[pl * g.cap + m] * kold[k] / S.freq[k];
                    if (q == p) {
                        if (r == Jl) A0[(size_t)I * n + r] += v_IJ;
                        if (r == I) A0[(size_t)I * n + r] -= v_IJ;
                    } else if (r == Jl) {
                        yoff -= v_IJ * g.x_it[S.pop_off + (size_t)q * n + I];
                    }
                }
            } else {
                const double v_IJ = -lam[(size_t)p * v.L + S.lid0 + k] * kold[k] / S.freq[k];
                if (r == Jl) A0[(size_t)I * n + r] += v_IJ;
                if (r == I) A0[(size_t)I * n + r] -= v_IJ;
            }
        }
        if constexpr (GLOBAL) A0[(size_t)r * n + r] += g.dextra[S.pop_off + (size_t)p * n + r];
        const double T = v.temperature[p];
        for (uint32_t c = 0; c < S.ncolpar; c++) {
            const NCol C = v.col[S.col0 + c];
            double abn = v.abundance[(size_t)p * v.nspec + C.num];
            if (C.op != 'n') {
                const double frac_H2_para = 1.0 / (1.0 + 9.0 * mgx_ref_expf((float)(-170.5 / T)));
                if (C.op == 'o') abn *= (1.0 - frac_H2_para);
                if (C.op == 'p') abn *= frac_H2_para;
            }
            const uint32_t t = search_tmp(C.tmp, C.ntmp, T);
            const bool edge = (t == 0) || (t == C.ntmp - 1);
            const double step = edge ? 0.0 : (T - C.tmp[t - 1]) / (C.tmp[t] - C.tmp[t - 1]);
            const double *Cd1 = C.Cd + (size_t)t * C.ncol, *Ce1 = C.Ce + (size_t)t * C.ncol;
            const double *Cd0 = edge ? Cd1 : Cd1 - C.ncol, *Ce0 = edge ? Ce1 : Ce1 - C.ncol;
            for (uint32_t k = 0; k < C.ncol; k++) {
                const uint32_t I = C.icol[k], Jl = C.jcol[k];
                if (r != I && r != Jl) continue;
                const double cd = edge ? Cd1[k] : Cd0[k] + (Cd1[k] - Cd0[k]) * step;
                const double ce = edge ? Ce1[k] : Ce0[k] + (Ce1[k] - Ce0[k]) * step;
                const double v_IJ = cd * abn, v_JI = ce * abn;
                if (r == Jl) {
                    A0[(size_t)I * n + r] += v_IJ;
                    A0[(size_t)Jl * n + r] -= v_JI;
                }
                if (r == I) {
                    A0[(size_t)Jl * n + r] += v_JI;
                    A0[(size_t)I * n + r] -= v_IJ;
                }
            }
        }
        if (r == n - 1)
            for (uint32_t c = 0; c < n; c++) A0[(size_t)c * n + r] = 1.0;
    }
    __syncthreads();
    for (uint32_t i = lane; i < n * n; i += NL_WAVE) LU[i] = A0[i];
    const double y0 = (r == n - 1) ? ptot : yoff;
    double yv = y0;
    __syncthreads();

    /* Gaussian elimination with partial pivoting; my_step = the column this
       lane's row became the pivot of (n while not yet) */
    int my_step = (int)n;
    for (uint32_t c = 0; c < n; c++) {
        double val = (r < n && my_step == (int)n) ? fabs(LU[(size_t)c * n + r]) : -1.0;
        if (val != val) val = 0.0; /* a NaN row is picked only when nothing else is left */
        const int pr = wave_argmax(val, (int)lane);
        if ((int)lane == pr) my_step = (int)c;
        const double pivot = LU[(size_t)c * n + pr];
        const double ypiv = __shfl(yv, pr);
        if (r < n && my_step > (int)c) {
            const double f = pivot != 0.0 ? LU[(size_t)c * n + r] / pivot : 0.0;
            LU[(size_t)c * n + r] = f;
            for (uint32_t j = c + 1; j < n; j++)
                LU[(size_t)j * n + r] = fma(-f, LU[(size_t)j * n + pr], LU[(size_t)j * n + r]);
            yv = fma(-f, ypiv, yv);
        }
        __syncthreads();
    }
    /* back substitution by columns; lane c ends with x_c */
    double x = 0.0;
    for (int c = (int)n - 1; c >= 0; c--) {
        const int pr = lane_of_step(my_step, c);
        const double d = LU[(size_t)c * n + pr];
        double xc = ((int)lane == pr && d != 0.0) ? yv / d : 0.0;
        xc = __shfl(xc, pr);
        if ((int)lane == c) x = xc;
        if (r < n && my_step < c) yv = fma(-LU[(size_t)c * n + r], xc, yv);
    }
    /* iterative refinement with compensated residuals */
    for (int it = 0; it < 2; it++) {
        if (r < n) xs[r] = x;
        __syncthreads();
        double res = 0.0;
        if (r < n) {
            double hi = y0, lo = 0.0;
            for (uint32_t j = 0; j < n; j++) {
                const double a = A0[(size_t)j * n + r], b = xs[j];
                const double pr_ = a * b;
                const double pe = fma(a, b, -pr_);
                double s_, e_;
                two_sum(hi, -pr_, s_, e_);
                hi = s_;
                lo += e_ - pe;
            }
            res = hi + lo;
        }
        __syncthreads();
        for (uint32_t c = 0; c < n; c++) {
            const int pr = lane_of_step(my_step, (int)c);
            const double dp = __shfl(res, pr);
            if (r < n && my_step > (int)c) res = fma(-LU[(size_t)c * n + r], dp, res);
        }
        double dx = 0.0;
        for (int c = (int)n - 1; c >= 0; c--) {
            const int pr = lane_of_step(my_step, c);
            const double d = LU[(size_t)c * n + pr];
            double xc = ((int)lane == pr && d != 0.0) ? res / d : 0.0;
            xc = __shfl(xc, pr);
            if ((int)lane == c) dx = xc;
            if (r < n && my_step < c) res = fma(-LU[(size_t)c * n + r], xc, res);
        }
        x += dx;
    }
    __syncthreads();
    if constexpr (GLOBAL) {
        double d = 0.0;
        if (lane < n) {
            const size_t i = S.pop_off + (size_t)p * n + lane;
            g.x_out[i] = x;
            d = fabs(x - g.x_it[i]) / ptot;
            if (!(d == d)) d = __builtin_inf();
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const double o2 = __shfl_xor(d, off);
            d = o2 > d ? o2 : d;
        }
        if (lane == 0) atomicMax(g.dmax, (unsigned long long)__double_as_longlong(d));
        return;
    }
    finish_point(v, S, p, s, x, old, cur_p, pv, eps, kap, ic, part);
}

/* the diagonal terms an approximated-Lambda element of the list of (p, k) at another point
   q puts on row (q, upper) (lineProducingSpecies.tpp:455-458): + Ls opacity(p, k) / nu */
__global__ void lambda_extra_kernel(NView v, GlobalIn g, const double *__restrict__ old_all,
                                    double *dextra)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
    if (p >= v.P) return;
    const NSpec S = v.spec[s];
    const uint32_t n = S.nlev;
    const double *old = old_all + S.pop_off + (size_t)p * n;
    for (uint32_t k = 0; k < S.nrad; k++) {
        const uint32_t I = S.irad[k];
        if (I == n - 1) continue;
        const double kold = line_opacity(S, k, old[I], old[S.jrad[k]], v.min_o);
        const size_t pl = (size_t)p * v.L + S.lid0 + k;
        for (uint32_t m = 0; m < g.count[pl]; m++) {
            const uint32_t q = g.nr[pl * g.cap + m];
            if (q == p) continue;
            atomicAdd(&dextra[S.pop_off + (size_t)q * n + I], g.val[pl * g.cap + m] * kold / S.freq[k]);
        }
    }
}

/* negative correction, convergence and line data after an Ng step */
__global__ __launch_bounds__(NL_WAVE) void finish_kernel(NView v, const double *__restrict__ old_all,
                                                         double *cur_all, double *eps,
                                                         double *kap, double *ic, double *part)
{
    extern __shared__ double sm[];
    const uint32_t p = blockIdx.x, s = blockIdx.y, lane = threadIdx.x;
    const NSpec S = v.spec[s];
    const uint32_t n = S.nlev;
    const double *old = old_all + S.pop_off + (size_t)p * n;
    double *cur_p = cur_all + S.pop_off + (size_t)p * n;
    const double x = lane < n ? cur_p[lane] : 0.0;
    __syncthreads();
    finish_point(v, S, p, s, x, old, cur_p, sm, eps, kap, ic, part);
}

/* update_using_LTE (levels summed in order) and the line data, one thread per (p, s) */
__global__ void lte_kernel(NView v, uint32_t nlspecs, double *cur_all, int set_lte, double *eps,
                           double *kap, double *ic)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)v.P * nlspecs) return;
    const uint32_t p = (uint32_t)(t % v.P), s = (uint32_t)(t / v.P);
    const NSpec S = v.spec[s];
    double *pop = cur_all + S.pop_off + (size_t)p * S.nlev;
    const double ptot = v.abundance[(size_t)p * v.nspec + S.num];
    if (set_lte) {
        const double T = v.temperature[p];
        double pf = 0.0;
        for (uint32_t i = 0; i < S.nlev; i++) {
            pop[i] = boltzmann(S, i, T);
            pf += pop[i];
        }
        for (uint32_t i = 0; i < S.nlev; i++) pop[i] *= ptot / pf;
    }
    for (uint32_t k = 0; k < S.nrad; k++) {
        const double ni = pop[S.irad[k]], nj = pop[S.jrad[k]];
        const size_t idx = (size_t)p * v.L + S.lid0 + k;
        const double o = line_opacity(S, k, ni, nj, v.min_o);
        eps[idx] = line_emissivity(S, k, ni, v.min_e);
        kap[idx] = o;
        ic[idx] = o / ptot;
    }
}

__global__ void pop_tot_kernel(NView v, uint32_t nlspecs, double *out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)v.P * nlspecs) return;
    const uint32_t p = (uint32_t)(t % v.P), s = (uint32_t)(t / v.P);
    out[t] = v.abundance[(size_t)p * v.nspec + v.spec[s].num];
}

/* per-species totals of the per-point partials: out[s*3 + {0,1,2}] = count, mean, max.
   One block per species; fixed summation order. */
__global__ __launch_bounds__(NL_RED_THREADS) void reduce_parts_kernel(const double *part,
                                                                      uint32_t P, double *out)
{
    __shared__ double sc[NL_RED_THREADS], sm_[NL_RED_THREADS], sx[NL_RED_THREADS];
    const uint32_t s = blockIdx.x, t = threadIdx.x;
    const double *c = part + ((size_t)s * 3 + 0) * P;
    const double *m = part + ((size_t)s * 3 + 1) * P;
    const double *x = part + ((size_t)s * 3 + 2) * P;
    double a = 0.0, b = 0.0, mx = 0.0;
    for (uint32_t p = t; p < P; p += NL_RED_THREADS) {
        a += c[p];
        b += m[p];
        mx = x[p] < mx ? mx : x[p];
    }
    sc[t] = a;
    sm_[t] = b;
    sx[t] = mx;
    __syncthreads();
    for (int w = NL_RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)t < w) {
            sc[t] += sc[t + w];
            sm_[t] += sm_[t + w];
            sx[t] = sx[t + w] < sx[t] ? sx[t] : sx[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[3 * s + 0] = sc[0];
        out[3 * s + 1] = sm_[0];
        out[3 * s + 2] = sx[0];
    }
}

/* ------------------------------------------------------------------------ */
/* Ng acceleration                                                           */
/* ------------------------------------------------------------------------ */

struct VecList {
    const double *x[NL_MAX_ORDER + 1]; /* oldest .. newest */
};

/* Gram matrix of the residuals R_i = x[i+1] - x[i], i < order, over n
   elements; per-block partials of the upper triangle (order (order+1)/2). */
template <int MO>
__global__ __launch_bounds__(256) void gram_kernel(VecList xl, uint32_t order, uint64_t off,
                                                   uint64_t n, double *partial)
{
    constexpr int NP = MO * (MO + 1) / 2;
    __shared__ double red[4][NP];
    double acc[NP];
#pragma unroll
    for (int a = 0; a < NP; a++) acc[a] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        double rv[MO];
        double prev = xl.x[0][off + t];
#pragma unroll
        for (int i = 0; i < MO; i++) {
            if (i < (int)order) {
                const double nx = xl.x[i + 1][off + t];
                rv[i] = nx - prev;
                prev = nx;
            } else {
                rv[i] = 0.0;
            }
        }
        int a = 0;
#pragma unroll
        for (int i = 0; i < MO; i++)
#pragma unroll
            for (int j = i; j < MO; j++, a++) acc[a] = fma(rv[i], rv[j], acc[a]);
    }
    const int wave = threadIdx.x / NL_WAVE, lane = threadIdx.x % NL_WAVE;
#pragma unroll
    for (int a = 0; a < NP; a++) {
        double vv = acc[a];
        for (int o = 32; o > 0; o >>= 1) vv += __shfl_xor(vv, o);
        if (lane == 0) red[wave][a] = vv;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < NP; a += blockDim.x)
        partial[(size_t)blockIdx.x * NP + a] = ((red[0][a] + red[1][a]) + red[2][a]) + red[3][a];
}

__global__ void gram_reduce_kernel(const double *partial, uint32_t nblocks, uint32_t np,
                                   double *out)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= np) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nblocks; b++) s += partial[(size_t)b * np + a];
    out[a] = s;
}

struct Coefs {
    double c[NL_MAX_ORDER];
};

/* out = sum_{i < order} x[order - i] * c[order - 1 - i], accumulated from zero in
   that order (the reference's population += populations[popsize-1-i] * coef[order-1-i]) */
__global__ void combine_kernel(VecList xl, Coefs cf, uint32_t order, uint64_t off, uint64_t n,
                               double *out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        double acc = 0.0;
        for (uint32_t i = 0; i < order; i++)
            acc += xl.x[order - i][off + t] * cf.c[order - 1 - i];
        out[off + t] = acc;
    }
}

/* check_for_convergence_trial: relative change of every entry */
__global__ __launch_bounds__(NL_RED_THREADS) void trial_conv_kernel(const double *tr,
                                                                    const double *tp,
                                                                    uint64_t off, uint64_t n,
                                                                    double pop_prec, double weight,
                                                                    double *partial)
{
    __shared__ double sc[NL_RED_THREADS], sm_[NL_RED_THREADS], sx[NL_RED_THREADS];
    double a = 0.0, b = 0.0, mx = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const double x = tr[off + t], y = tp[off + t];
        const double rc = 2.0 * fabs((x - y) / (x + y));
        if (rc > pop_prec) a += 1.0;
        b += weight * rc;
        mx = rc < mx ? mx : rc;
    }
    const uint32_t t = threadIdx.x;
    sc[t] = a;
    sm_[t] = b;
    sx[t] = mx;
    __syncthreads();
    for (int w = NL_RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)t < w) {
            sc[t] += sc[t + w];
            sm_[t] += sm_[t + w];
            sx[t] = sx[t + w] < sx[t] ? sx[t] : sx[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        partial[3 * (size_t)blockIdx.x + 0] = sc[0];
        partial[3 * (size_t)blockIdx.x + 1] = sm_[0];
        partial[3 * (size_t)blockIdx.x + 2] = sx[0];
    }
}

/* ------------------------------------------------------------------------ */
/* host: long-double column-pivoted QR of the small Ng system                */
/* ------------------------------------------------------------------------ */

/* Householder QR with column pivoting (Businger and Golub 1965; column-norm
   downdating of LAPACK xGEQP3) and the rank-revealing solve of
   Eigen::ColPivHouseholderQR, which the reference calls for the Ng
   coefficients.  a: n x n column-major, overwritten. */
void cpqr_solve(int n, std::vector<long double> &a, const std::vector<long double> &b,
                std::vector<long double> &x)
{
    using LD = long double;
    const LD eps = std::numeric_limits<LD>::epsilon();
    auto at = [&](int i, int j) -> LD & { return a[(size_t)j * n + i]; };
    std::vector<LD> tau(n, 0.0L), nu(n), nd(n), c(b);
    std::vector<int> perm(n);
    LD maxnorm = 0.0L;
    for (int j = 0; j < n; j++) {
        LD s = 0.0L;
        for (int i = 0; i < n; i++) s += at(i, j) * at(i, j);
        nu[j] = nd[j] = std::sqrt(s);
        perm[j] = j;
        maxnorm = std::max(maxnorm, nu[j]);
    }
    const LD thr_helper = (maxnorm * eps) * (maxnorm * eps) / (LD)n;
    int nonzero = n;
    LD maxpivot = 0.0L;
    for (int k = 0; k < n; k++) {
        int big = k;
        for (int j = k + 1; j < n; j++)
            if (nu[j] > nu[big]) big = j;
        if (nonzero == n && nu[big] * nu[big] < thr_helper * (LD)(n - k)) nonzero = k;
        if (big != k) {
            for (int i = 0; i < n; i++) std::swap(at(i, k), at(i, big));
            std::swap(nu[k], nu[big]);
            std::swap(nd[k], nd[big]);
            std::swap(perm[k], perm[big]);
        }
        LD tail = 0.0L;
        for (int i = k + 1; i < n; i++) tail += at(i, k) * at(i, k);
        const LD c0 = at(k, k);
        LD beta;
        if (tail <= std::numeric_limits<LD>::min()) {
            tau[k] = 0.0L;
            beta = c0;
            for (int i = k + 1; i < n; i++) at(i, k) = 0.0L;
        } else {
            beta = std::sqrt(c0 * c0 + tail);
            if (c0 >= 0.0L) beta = -beta;
            for (int i = k + 1; i < n; i++) at(i, k) /= (c0 - beta);
            tau[k] = (beta - c0) / beta;
        }
        at(k, k) = beta;
        maxpivot = std::max(maxpivot, std::fabs(beta));
        if (tau[k] != 0.0L)
            for (int j = k + 1; j < n; j++) {
                LD t = at(k, j);
                for (int i = k + 1; i < n; i++) t += at(i, k) * at(i, j);
                at(k, j) -= tau[k] * t;
                for (int i = k + 1; i < n; i++) at(i, j) -= tau[k] * at(i, k) * t;
            }
        for (int j = k + 1; j < n; j++) {
            if (nu[j] == 0.0L) continue;
            LD t = std::fabs(at(k, j)) / nu[j];
            t = std::max((LD)0.0L, (1.0L + t) * (1.0L - t));
            const LD q = nu[j] / nd[j];
            if (t * q * q <= std::sqrt(eps)) {
                LD s = 0.0L;
                for (int i = k + 1; i < n; i++) s += at(i, j) * at(i, j);
                nu[j] = nd[j] = std::sqrt(s);
            } else {
                nu[j] *= std::sqrt(t);
            }
        }
    }
    const LD pthr = maxpivot * eps * (LD)n;
    int rank = 0;
    for (int i = 0; i < nonzero; i++) rank += std::fabs(at(i, i)) > pthr ? 1 : 0;
    x.assign(n, 0.0L);
    if (rank == 0) return;
    for (int k = 0; k < rank; k++) {
        if (tau[k] == 0.0L) continue;
        LD t = c[k];
        for (int i = k + 1; i < n; i++) t += at(i, k) * c[i];
        c[k] -= tau[k] * t;
        for (int i = k + 1; i < n; i++) c[i] -= tau[k] * at(i, k) * t;
    }
    for (int i = rank - 1; i >= 0; i--) {
        LD s = c[i];
        for (int j = i + 1; j < rank; j++) s -= at(i, j) * c[j];
        c[i] = s / at(i, i);
    }
    for (int i = 0; i < rank; i++) x[perm[i]] = c[i];
}

} // namespace

/* ------------------------------------------------------------------------ */
/* handle                                                                    */
/* ------------------------------------------------------------------------ */

struct mgx_nlte {
    mgx_model view{};           /* device view (borrowed pointers) */
    mgx_nlte_params prm{};
    uint32_t P = 0, L = 0, Q = 0, F = 0, nlspecs = 0, nspec = 0, maxnlev = 0, maxnrad = 0;
    const double *abundance = nullptr;
    std::vector<uint32_t> nlev, nrad;
    std::vector<uint64_t> pop_off;
    uint64_t nvec = 0; /* doubles per population vector */
    void *tables = nullptr;
    const NSpec *spec_dev = nullptr;
    const NCol *col_dev = nullptr;
    double *cur = nullptr, *trial = nullptr, *trial_prev = nullptr;
    std::vector<double *> ring; /* slots */
    std::vector<int> hist;      /* slot indices, oldest .. newest */
    double *part = nullptr;     /* (nlspecs, 3, P) per-point partials */
    double *red = nullptr;      /* 3 nlspecs totals + Gram scratch */
    double *gram_partial = nullptr;
    double *J = nullptr, *lambda = nullptr; /* the loop's J and Lambda */
    size_t J_elems = 0, lambda_elems = 0;
    double *gx = nullptr; /* global update: two iterates and the extra diagonal (3 nvec), dmax */
    uint32_t *lam_count = nullptr, *lam_nr = nullptr; /* the loop's Lambda lists */
    double *lam_val = nullptr;
    size_t lam_elems = 0, lam_count_elems = 0;
    int last_loop = -1;     /* kind of the loop whose formal solve filled J and Lambda last */
    uint32_t last_cap = 0;  /* its list capacity (general loop) */
    hipStream_t stream = nullptr;
    std::vector<mgx_nlte_conv> conv;
};

namespace {

NView make_view(const mgx_nlte *h)
{
    NView v{};
    v.P = h->P;
    v.L = h->L;
    v.Q = h->Q;
    v.F = h->F;
    v.nspec = h->nspec;
    v.maxnlev = h->maxnlev;
    v.maxnrad = h->maxnrad;
    v.spec = h->spec_dev;
    v.col = h->col_dev;
    v.temperature = h->view.temperature;
    v.abundance = h->abundance;
    v.quad_weights = h->view.quad_weights;
    v.min_e = h->prm.min_line_emissivity;
    v.min_o = h->prm.min_line_opacity;
    v.inv_frac = h->prm.population_inversion_fraction;
    v.min_rel = h->prm.min_rel_pop_for_convergence;
    v.pop_prec = h->prm.pop_prec;
    return v;
}

size_t point_lds(const mgx_nlte *h, bool stateq)
{
    const size_t d = stateq ? stateq_lds_doubles(h->maxnlev, h->maxnrad) : 2 * (size_t)h->maxnlev;
    return 8 * d;
}

int set_lds_limit(const void *fn, size_t bytes)
{
    if (bytes > 65536)
        NL_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return MGX_OK;
}

/* totals of the per-point partials into h->conv (synchronises) */
int collect_conv(mgx_nlte *h, hipStream_t st)
{
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(h->nlspecs), dim3(NL_RED_THREADS), 0, st,
                       h->part, h->P, h->red);
    NL_TRY(hipGetLastError());
    std::vector<double> t(3 * h->nlspecs);
    NL_TRY(hipMemcpyAsync(t.data(), h->red, t.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    NL_TRY(hipStreamSynchronize(st));
    for (uint32_t s = 0; s < h->nlspecs; s++) {
        const double w = 1.0 / (double)((uint64_t)h->P * h->nlev[s]);
        h->conv[s].fraction_not_converged = t[3 * s + 0] * w;
        h->conv[s].relative_change_mean = t[3 * s + 1];
        h->conv[s].relative_change_max = t[3 * s + 2];
    }
    return MGX_OK;
}

/* append a copy of the current populations to the history */
int push_history(mgx_nlte *h, hipStream_t st)
{
    const int cap = (int)h->ring.size();
    int slot;
    if ((int)h->hist.size() < cap) {
        std::vector<char> used(cap, 0);
        for (int q : h->hist) used[q] = 1;
        slot = (int)(std::find(used.begin(), used.end(), 0) - used.begin());
    } else {
        slot = h->hist.front(); /* the oldest entry is beyond any usable order */
        h->hist.erase(h->hist.begin());
    }
    NL_TRY(hipMemcpyAsync(h->ring[slot], h->cur, h->nvec * sizeof(double),
                          hipMemcpyDeviceToDevice, st));
    h->hist.push_back(slot);
    return MGX_OK;
}

int restart_history(mgx_nlte *h, hipStream_t st)
{
    h->hist.clear();
    return push_history(h, st);
}

int line_data(mgx_nlte *h, int set_lte, double *eps, double *kap, double *ic, hipStream_t st)
{
    if (!eps || !kap || !ic) return mgx::fail(MGX_E_ARG, "emissivity, opacity and interp_crit are required");
    const uint64_t n = (uint64_t)h->P * h->nlspecs;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    if (blocks)
        hipLaunchKernelGGL(lte_kernel, dim3(blocks), dim3(256), 0, st, make_view(h), h->nlspecs,
                           h->cur, set_lte, eps, kap, ic);
    NL_TRY(hipGetLastError());
    return MGX_OK;
}

/* Ng coefficients of every species from the newest order residuals of the
   list xs (oldest .. newest, order + 1 vectors) */
int ng_coefficients(mgx_nlte *h, const std::vector<const double *> &xs, uint32_t order,
                    std::vector<Coefs> &cf, hipStream_t st)
{
    VecList xl{};
    for (uint32_t i = 0; i <= order; i++) xl.x[i] = xs[xs.size() - 1 - order + i];
    const uint32_t np = order * (order + 1) / 2;
    cf.assign(h->nlspecs, Coefs{});
    for (uint32_t s = 0; s < h->nlspecs; s++) {
        const uint64_t off = h->pop_off[s], n = (uint64_t)h->P * h->nlev[s];
        const int MO = order <= 4 ? 4 : order <= 8 ? 8 : 12;
        const int npm = MO * (MO + 1) / 2;
        const dim3 g(NL_GRAM_BLOCKS), b(256);
        if (MO == 4)
            hipLaunchKernelGGL(gram_kernel<4>, g, b, 0, st, xl, order, off, n, h->gram_partial);
        else if (MO == 8)
            hipLaunchKernelGGL(gram_kernel<8>, g, b, 0, st, xl, order, off, n, h->gram_partial);
        else
            hipLaunchKernelGGL(gram_kernel<12>, g, b, 0, st, xl, order, off, n, h->gram_partial);
        NL_TRY(hipGetLastError());
        hipLaunchKernelGGL(gram_reduce_kernel, dim3((npm + 63) / 64), dim3(64), 0, st,
                           h->gram_partial, (uint32_t)NL_GRAM_BLOCKS, (uint32_t)npm, h->red + 3 * h->nlspecs);
        NL_TRY(hipGetLastError());
        std::vector<double> g_(npm);
        NL_TRY(hipMemcpyAsync(g_.data(), h->red + 3 * h->nlspecs, npm * sizeof(double),
                              hipMemcpyDeviceToHost, st));
        NL_TRY(hipStreamSynchronize(st));
        /* RTR(i, j) from the packed upper triangle of the MO x MO accumulator */
        std::vector<long double> A((size_t)order * order), one(order, 1.0L), c;
        for (uint32_t i = 0; i < order; i++)
            for (uint32_t j = 0; j < order; j++) {
                const uint32_t a = std::min(i, j), bb = std::max(i, j);
                const uint32_t idx = a * MO - a * (a - 1) / 2 + (bb - a);
                A[(size_t)j * order + i] = g_[idx];
            }
        (void)np;
        cpqr_solve((int)order, A, one, c);
        long double sum = 0.0L;
        for (uint32_t i = 0; i < order; i++) sum += c[i];
        for (uint32_t i = 0; i < order; i++) cf[s].c[i] = (double)(c[i] / sum);
    }
    return MGX_OK;
}

int combine(mgx_nlte *h, const std::vector<const double *> &xs, uint32_t order,
            const std::vector<Coefs> &cf, double *out, hipStream_t st)
{
    VecList xl{};
    for (uint32_t i = 0; i <= order; i++) xl.x[i] = xs[xs.size() - 1 - order + i];
    for (uint32_t s = 0; s < h->nlspecs; s++) {
        const uint64_t n = (uint64_t)h->P * h->nlev[s];
        hipLaunchKernelGGL(combine_kernel, dim3(1024), dim3(256), 0, st, xl, cf[s], order,
                           h->pop_off[s], n, out);
        NL_TRY(hipGetLastError());
    }
    return MGX_OK;
}

std::vector<const double *> history_ptrs(const mgx_nlte *h)
{
    std::vector<const double *> v;
    for (int q : h->hist) v.push_back(h->ring[q]);
    return v;
}

int launch_finish(mgx_nlte *h, const double *old, double *eps, double *kap, double *ic,
                  hipStream_t st)
{
    const size_t lds = point_lds(h, false);
    if (int e = set_lds_limit((const void *)finish_kernel, lds)) return e;
    hipLaunchKernelGGL(finish_kernel, dim3(h->P, h->nlspecs), dim3(NL_WAVE), lds, st,
                       make_view(h), old, h->cur, eps, kap, ic, h->part);
    NL_TRY(hipGetLastError());
    return MGX_OK;
}

void release(mgx_nlte *h)
{
    (void)hipFree(h->tables);
    (void)hipFree(h->cur);
    (void)hipFree(h->trial);
    (void)hipFree(h->trial_prev);
    for (double *q : h->ring) (void)hipFree(q);
    (void)hipFree(h->part);
    (void)hipFree(h->red);
    (void)hipFree(h->gram_partial);
    (void)hipFree(h->J);
    (void)hipFree(h->lambda);
    (void)hipFree(h->gx);
    (void)hipFree(h->lam_count);
    (void)hipFree(h->lam_nr);
    (void)hipFree(h->lam_val);
}

/* grow-only device buffer */
template <typename T>
int ensure(T *&p, size_t &have, size_t want)
{
    if (have >= want) return MGX_OK;
    (void)hipFree(p);
    p = nullptr;
    NL_TRY(hipMalloc(&p, want * sizeof(T)));
    have = want;
    return MGX_OK;
}

} // namespace

extern "C" {

void mgx_nlte_default_params(mgx_nlte_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->pop_prec = 1.0e-6;
    p->min_rel_pop_for_convergence = 1.0e-10;
    p->population_inversion_fraction = -1.0;
    p->min_line_emissivity = 1.0e-50;
    p->min_line_opacity = 1.0e-13;
    p->convergence_fraction = 0.995;
    p->use_adaptive_Ng_acceleration = 1;
    p->Ng_acceleration_mem_limit = 9;
    p->adaptive_Ng_acceleration_min_order = 2;
    p->adaptive_Ng_acceleration_use_max_criterion = 1;
    p->Ng_acceleration_remove_N_its = 0;
}

int mgx_nlte_create(const mgx_model *m, uint32_t nlspecs, const mgx_linedata *ld,
                    const uint32_t *ncolpar, const mgx_colldata *colpar, uint32_t nspec,
                    const double *abundance_dev, const mgx_nlte_params *params, mgx_nlte **out,
                    void *stream)
{
    mgx::clear_error();
    if (!m || !ld || !params || !out || !abundance_dev)
        return mgx::fail(MGX_E_ARG, "null argument");
    *out = nullptr;
    if (nlspecs == 0 || nlspecs != m->nlspecs)
        return mgx::fail(MGX_E_ARG, "nlspecs (%u) must match the model (%u)", nlspecs, m->nlspecs);
    if (!m->temperature || !m->quad_weights) return mgx::fail(MGX_E_ARG, "model view lacks temperature / quad_weights");
    const uint32_t mem = params->Ng_acceleration_mem_limit + params->Ng_acceleration_remove_N_its;
    if (params->Ng_acceleration_mem_limit > NL_MAX_ORDER)
        return mgx::fail(MGX_E_UNSUPPORTED, "Ng_acceleration_mem_limit above %u", NL_MAX_ORDER);
    uint32_t lsum = 0, ctot = 0;
    for (uint32_t s = 0; s < nlspecs; s++) {
        const mgx_linedata &d = ld[s];
        if (d.nlev < 2 || d.nlev > NL_MAX_NLEV)
            return mgx::fail(MGX_E_UNSUPPORTED, "species %u: %u levels (2..%u supported)", s, d.nlev,
                             NL_MAX_NLEV);
        if (d.num >= nspec) return mgx::fail(MGX_E_ARG, "species %u: abundance column out of range", s);
        for (uint32_t k = 0; k < d.nrad; k++)
            if (d.irad[k] >= d.nlev || d.jrad[k] >= d.nlev)
                return mgx::fail(MGX_E_ARG, "species %u: radiative level index out of range", s);
        const uint32_t nc = ncolpar ? ncolpar[s] : 0;
        for (uint32_t c = 0; c < nc; c++) {
            const mgx_colldata &C = colpar[ctot + c];
            if (C.num_col_partner >= nspec) return mgx::fail(MGX_E_ARG, "collision partner column out of range");
            if (C.ntmp == 0) return mgx::fail(MGX_E_ARG, "collision partner without temperatures");
            if (C.orth_or_para_H2 != 'n' && C.orth_or_para_H2 != 'o' && C.orth_or_para_H2 != 'p')
                return mgx::fail(MGX_E_ARG, "orth_or_para_H2 must be 'n', 'o' or 'p'");
            for (uint32_t k = 0; k < C.ncol; k++)
                if (C.icol[k] >= d.nlev || C.jcol[k] >= d.nlev)
                    return mgx::fail(MGX_E_ARG, "species %u: collisional level index out of range", s);
        }
        ctot += nc;
        lsum += d.nrad;
    }
    if (lsum != m->nlines) return mgx::fail(MGX_E_ARG, "line count of the species (%u) != model nlines (%u)", lsum, m->nlines);
    if (m->nfreqs != m->nlines * m->nquads)
        return mgx::fail(MGX_E_ARG, "the model's frequencies must be the line grid (nfreqs = nlines * nquads)");

    mgx_nlte *h = new mgx_nlte();
    h->view = *m;
    h->prm = *params;
    h->P = m->npoints;
    h->L = m->nlines;
    h->Q = m->nquads;
    h->F = m->nfreqs;
    h->nlspecs = nlspecs;
    h->nspec = nspec;
    h->abundance = abundance_dev;
    h->conv.assign(nlspecs, mgx_nlte_conv{0.0, 0.0, 0.0});
    hipStream_t st = (hipStream_t)stream;
    h->stream = st;

    /* pack the tables: NSpec[nlspecs], NCol[ctot], then the arrays (8-byte aligned) */
    std::vector<unsigned char> img;
    auto align8 = [&]() { img.resize((img.size() + 7) & ~(size_t)7); };
    const size_t spec_at = 0;
    img.resize(sizeof(NSpec) * nlspecs);
    align8();
    const size_t col_at = img.size();
    img.resize(img.size() + sizeof(NCol) * std::max<uint32_t>(ctot, 1));
    std::vector<NSpec> sp(nlspecs);
    std::vector<NCol> cp(std::max<uint32_t>(ctot, 1));
    std::vector<size_t> fix_off;        /* offsets of pointer fields to relocate */
    std::vector<size_t> fix_target;     /* image offsets they point to */
    auto put = [&](const void *src, size_t bytes) {
        align8();
        const size_t at = img.size();
        img.resize(at + bytes);
        if (bytes) std::memcpy(img.data() + at, src, bytes);
        return at;
    };
    struct Rel {
        size_t field_at, target;
    };
    std::vector<Rel> rel;
    uint32_t lid0 = 0, c0 = 0;
    uint64_t off = 0;
    for (uint32_t s = 0; s < nlspecs; s++) {
        const mgx_linedata &d = ld[s];
        NSpec &S = sp[s];
        S.num = d.num;
        S.nlev = d.nlev;
        S.nrad = d.nrad;
        S.lid0 = lid0;
        S.ncolpar = ncolpar ? ncolpar[s] : 0;
        S.col0 = c0;
        S.pop_off = off;
        h->nlev.push_back(d.nlev);
        h->nrad.push_back(d.nrad);
        h->pop_off.push_back(off);
        h->maxnlev = std::max(h->maxnlev, d.nlev);
        h->maxnrad = std::max(h->maxnrad, d.nrad);
        off += (uint64_t)m->npoints * d.nlev;
        lid0 += d.nrad;
        const size_t base = spec_at + sizeof(NSpec) * s;
        rel.push_back({base + offsetof(NSpec, energy), put(d.energy, 8 * (size_t)d.nlev)});
        rel.push_back({base + offsetof(NSpec, weight), put(d.weight, 8 * (size_t)d.nlev)});
        rel.push_back({base + offsetof(NSpec, A), put(d.A, 8 * (size_t)d.nrad)});
        rel.push_back({base + offsetof(NSpec, Bs), put(d.Bs, 8 * (size_t)d.nrad)});
        rel.push_back({base + offsetof(NSpec, Ba), put(d.Ba, 8 * (size_t)d.nrad)});
        rel.push_back({base + offsetof(NSpec, freq), put(d.frequency, 8 * (size_t)d.nrad)});
        rel.push_back({base + offsetof(NSpec, irad), put(d.irad, 4 * (size_t)d.nrad)});
        rel.push_back({base + offsetof(NSpec, jrad), put(d.jrad, 4 * (size_t)d.nrad)});
        for (uint32_t c = 0; c < S.ncolpar; c++, c0++) {
            const mgx_colldata &C = colpar[c0];
            NCol &Cc = cp[c0];
            Cc.num = C.num_col_partner;
            Cc.op = (uint32_t)C.orth_or_para_H2;
            Cc.ntmp = C.ntmp;
            Cc.ncol = C.ncol;
            const size_t cb = col_at + sizeof(NCol) * c0;
            rel.push_back({cb + offsetof(NCol, icol), put(C.icol, 4 * (size_t)C.ncol)});
            rel.push_back({cb + offsetof(NCol, jcol), put(C.jcol, 4 * (size_t)C.ncol)});
            rel.push_back({cb + offsetof(NCol, tmp), put(C.tmp, 8 * (size_t)C.ntmp)});
            rel.push_back({cb + offsetof(NCol, Cd), put(C.Cd, 8 * (size_t)C.ntmp * C.ncol)});
            rel.push_back({cb + offsetof(NCol, Ce), put(C.Ce, 8 * (size_t)C.ntmp * C.ncol)});
        }
    }
    h->nvec = off;
    std::memcpy(img.data() + spec_at, sp.data(), sizeof(NSpec) * nlspecs);
    std::memcpy(img.data() + col_at, cp.data(), sizeof(NCol) * std::max<uint32_t>(ctot, 1));
    auto fail_free = [&](int e) {
        release(h);
        delete h;
        return e;
    };
    if (hipMalloc(&h->tables, img.size()) != hipSuccess)
        return fail_free(mgx::fail(MGX_E_HIP, "hipMalloc of the species tables failed"));
    const uintptr_t dev = (uintptr_t)h->tables;
    for (const Rel &r : rel) {
        const uintptr_t pv = dev + r.target;
        std::memcpy(img.data() + r.field_at, &pv, sizeof(pv));
    }
    h->spec_dev = (const NSpec *)(dev + spec_at);
    h->col_dev = (const NCol *)(dev + col_at);
    const size_t vb = h->nvec * sizeof(double);
    const int cap = (int)mem + 2;
    bool ok = hipMemcpyAsync(h->tables, img.data(), img.size(), hipMemcpyHostToDevice, st) == hipSuccess;
    ok = ok && hipMalloc(&h->cur, vb) == hipSuccess;
    ok = ok && hipMalloc(&h->trial, vb) == hipSuccess;
    ok = ok && hipMalloc(&h->trial_prev, vb) == hipSuccess;
    for (int q = 0; ok && q < cap; q++) {
        double *d = nullptr;
        ok = hipMalloc(&d, vb) == hipSuccess;
        if (ok) h->ring.push_back(d);
    }
    ok = ok && hipMalloc(&h->part, 3 * sizeof(double) * (size_t)nlspecs * h->P) == hipSuccess;
    const size_t npm = (size_t)NL_MAX_ORDER * (NL_MAX_ORDER + 1) / 2;
    ok = ok && hipMalloc(&h->red, sizeof(double) * (3 * (size_t)nlspecs + npm + 3 * 1024)) == hipSuccess;
    ok = ok && hipMalloc(&h->gram_partial, sizeof(double) * npm * NL_GRAM_BLOCKS) == hipSuccess;
    ok = ok && hipMemsetAsync(h->cur, 0, vb, st) == hipSuccess;
    ok = ok && hipMemsetAsync(h->trial, 0, vb, st) == hipSuccess;
    ok = ok && hipMemsetAsync(h->trial_prev, 0, vb, st) == hipSuccess;
    ok = ok && hipStreamSynchronize(st) == hipSuccess;
    if (!ok) return fail_free(mgx::fail(MGX_E_HIP, "device allocation of the population state failed"));
    *out = h;
    return MGX_OK;
}

void mgx_nlte_destroy(mgx_nlte *h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    release(h);
    delete h;
}

int mgx_nlte_set_lte(mgx_nlte *h, double *eps, double *kap, double *ic, void *stream)
{
    mgx::clear_error();
    if (!h) return mgx::fail(MGX_E_ARG, "null handle");
    hipStream_t st = (hipStream_t)stream;
    if (int e = line_data(h, 1, eps, kap, ic, st)) return e;
    if (int e = restart_history(h, st)) return e;
    NL_TRY(hipStreamSynchronize(st));
    return MGX_OK;
}

int mgx_nlte_set_populations(mgx_nlte *h, const double *pop, double *eps, double *kap, double *ic,
                             void *stream)
{
    mgx::clear_error();
    if (!h || !pop) return mgx::fail(MGX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    NL_TRY(hipMemcpyAsync(h->cur, pop, h->nvec * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (int e = line_data(h, 0, eps, kap, ic, st)) return e;
    if (int e = restart_history(h, st)) return e;
    NL_TRY(hipStreamSynchronize(st));
    return MGX_OK;
}

int mgx_nlte_get_populations(const mgx_nlte *h, double *pop, double *ptot, void *stream)
{
    mgx::clear_error();
    if (!h) return mgx::fail(MGX_E_ARG, "null handle");
    hipStream_t st = (hipStream_t)stream;
    if (pop) NL_TRY(hipMemcpyAsync(pop, h->cur, h->nvec * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (ptot) {
        const uint64_t n = (uint64_t)h->P * h->nlspecs;
        hipLaunchKernelGGL(pop_tot_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                           make_view(h), h->nlspecs, ptot);
        NL_TRY(hipGetLastError());
    }
    NL_TRY(hipStreamSynchronize(st));
    return MGX_OK;
}

int mgx_nlte_history(const mgx_nlte *h, uint32_t *n_residuals)
{
    mgx::clear_error();
    if (!h || !n_residuals) return mgx::fail(MGX_E_ARG, "null argument");
    *n_residuals = h->hist.empty() ? 0 : (uint32_t)h->hist.size() - 1;
    return MGX_OK;
}

int mgx_nlte_iteration_stateq(mgx_nlte *h, const double *J, const double *lam, double *eps,
                              double *kap, double *ic, mgx_nlte_conv *conv, double *Jlin,
                              double *Jeff, void *stream)
{
    mgx::clear_error();
    if (!h || !J || !lam || !eps || !kap || !ic) return mgx::fail(MGX_E_ARG, "null argument");
    if (h->hist.empty()) return mgx::fail(MGX_E_ARG, "no populations set (call mgx_nlte_set_lte first)");
    hipStream_t st = (hipStream_t)stream;
    if (int e = push_history(h, st)) return e;
    const double *old = h->ring[h->hist.back()];
    const size_t lds = point_lds(h, true);
    if (int e = set_lds_limit((const void *)stateq_kernel<false>, lds)) return e;
    hipLaunchKernelGGL(stateq_kernel<false>, dim3(h->P, h->nlspecs), dim3(NL_WAVE), lds, st,
                       make_view(h), J, lam, old, h->cur, eps, kap, ic, Jlin, Jeff, h->part,
                       GlobalIn{});
    NL_TRY(hipGetLastError());
    if (int e = collect_conv(h, st)) return e;
    if (conv) std::memcpy(conv, h->conv.data(), sizeof(mgx_nlte_conv) * h->nlspecs);
    return MGX_OK;
}

int mgx_nlte_iteration_stateq_global(mgx_nlte *h, const double *J, uint32_t cap,
                                     const uint32_t *lam_count, const uint32_t *lam_nr,
                                     const double *lam_val, double *eps, double *kap, double *ic,
                                     mgx_nlte_conv *conv, double *Jlin, double *Jeff,
                                     uint32_t *sweeps, void *stream)
{
    mgx::clear_error();
    if (!h || !J || !lam_count || !lam_nr || !lam_val || !eps || !kap || !ic)
        return mgx::fail(MGX_E_ARG, "null argument");
    if (cap < 1) return mgx::fail(MGX_E_ARG, "the Lambda lists need a capacity of at least 1");
    if (h->hist.empty()) return mgx::fail(MGX_E_ARG, "no populations set (call mgx_nlte_set_lte first)");
    hipStream_t st = (hipStream_t)stream;
    if (!h->gx) NL_TRY(hipMalloc(&h->gx, (3 * h->nvec + 1) * sizeof(double)));
    if (int e = push_history(h, st)) return e;
    const double *old = h->ring[h->hist.back()];
    double *xa = h->gx, *xb = h->gx + h->nvec, *dx = h->gx + 2 * h->nvec;
    unsigned long long *dmax = reinterpret_cast<unsigned long long *>(h->gx + 3 * h->nvec);
    NL_TRY(hipMemcpyAsync(xa, old, h->nvec * sizeof(double), hipMemcpyDeviceToDevice, st));
    NL_TRY(hipMemsetAsync(dx, 0, h->nvec * sizeof(double), st));
    GlobalIn g{lam_count, lam_nr, lam_val, cap, nullptr, dx, nullptr, dmax};
    hipLaunchKernelGGL(lambda_extra_kernel, dim3((h->P + 63) / 64, h->nlspecs), dim3(64), 0, st,
                       make_view(h), g, old, dx);
    NL_TRY(hipGetLastError());
    const size_t lds = point_lds(h, true);
    if (int e = set_lds_limit((const void *)stateq_kernel<true>, lds)) return e;
    /* block-Jacobi sweeps until the largest change is at the rounding level of the point
       solves (below 1e-15 of the point's total, or no longer decreasing) */
    double prev = __builtin_inf();
    uint32_t it = 0;
    const uint32_t max_sweeps = 100;
    for (;;) {
        g.x_it = xa;
        g.x_out = xb;
        NL_TRY(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(stateq_kernel<true>, dim3(h->P, h->nlspecs), dim3(NL_WAVE), lds, st,
                           make_view(h), J, nullptr, old, h->cur, eps, kap, ic,
                           it == 0 ? Jlin : nullptr, it == 0 ? Jeff : nullptr, h->part, g);
        NL_TRY(hipGetLastError());
        unsigned long long bits = 0;
        NL_TRY(hipMemcpyAsync(&bits, dmax, sizeof bits, hipMemcpyDeviceToHost, st));
        NL_TRY(hipStreamSynchronize(st));
        double d;
        std::memcpy(&d, &bits, sizeof d);
        std::swap(xa, xb);
        it++;
        if (d <= 1.0e-15 || (it >= 3 && d >= prev && d <= 1.0e-12)) break;
        if (it >= max_sweeps || !(d < 1.0e300))
            return mgx::fail(MGX_E_UNSUPPORTED,
                             "the global statistical equilibrium did not converge by block-Jacobi "
                             "sweeps (largest change %.3e of the total after %u sweeps)", d, it);
        prev = d;
    }
    if (sweeps) *sweeps = it;
    NL_TRY(hipMemcpyAsync(h->cur, xa, h->nvec * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (int e = launch_finish(h, old, eps, kap, ic, st)) return e;
    if (int e = collect_conv(h, st)) return e;
    if (conv) std::memcpy(conv, h->conv.data(), sizeof(mgx_nlte_conv) * h->nlspecs);
    return MGX_OK;
}

int mgx_nlte_iteration_ng(mgx_nlte *h, uint32_t order, double *eps, double *kap, double *ic,
                          mgx_nlte_conv *conv, void *stream)
{
    mgx::clear_error();
    if (!h || !eps || !kap || !ic) return mgx::fail(MGX_E_ARG, "null argument");
    if (h->hist.empty()) return mgx::fail(MGX_E_ARG, "no populations set");
    if (order == 0 || order > NL_MAX_ORDER)
        return mgx::fail(MGX_E_UNSUPPORTED, "Ng order %u (1..%u supported)", order, NL_MAX_ORDER);
    hipStream_t st = (hipStream_t)stream;
    if (int e = push_history(h, st)) return e;
    const uint32_t nres = (uint32_t)h->hist.size() - 1;
    if (nres < order)
        return mgx::fail(MGX_E_ARG, "Error during ng acceleration; order of acceleration (%u) greater than data stored (%u residuals)", order, nres);
    const std::vector<const double *> xs = history_ptrs(h);
    std::vector<Coefs> cf;
    if (int e = ng_coefficients(h, xs, order, cf, st)) return e;
    if (int e = combine(h, xs, order, cf, h->cur, st)) return e;
    /* residuals.clear(); populations.erase(begin, end - 1) */
    h->hist.erase(h->hist.begin(), h->hist.end() - 1);
    if (int e = launch_finish(h, h->ring[h->hist.back()], eps, kap, ic, st)) return e;
    if (int e = collect_conv(h, st)) return e;
    if (conv) std::memcpy(conv, h->conv.data(), sizeof(mgx_nlte_conv) * h->nlspecs);
    return MGX_OK;
}

int mgx_nlte_trial_ng(mgx_nlte *h, uint32_t order, mgx_nlte_conv *conv, void *stream)
{
    mgx::clear_error();
    if (!h) return mgx::fail(MGX_E_ARG, "null handle");
    if (h->hist.empty()) return mgx::fail(MGX_E_ARG, "no populations set");
    if (order == 0 || order > NL_MAX_ORDER)
        return mgx::fail(MGX_E_UNSUPPORTED, "Ng order %u (1..%u supported)", order, NL_MAX_ORDER);
    hipStream_t st = (hipStream_t)stream;
    std::vector<const double *> xs = history_ptrs(h);
    xs.push_back(h->cur); /* the trial's push of the current populations */
    if (xs.size() < (size_t)order + 1)
        return mgx::fail(MGX_E_ARG, "Error during ng acceleration; order of acceleration (%u) greater than data stored", order);
    std::swap(h->trial, h->trial_prev);
    std::vector<Coefs> cf;
    if (int e = ng_coefficients(h, xs, order, cf, st)) return e;
    if (int e = combine(h, xs, order, cf, h->trial, st)) return e;
    double *partial = h->red + 3 * h->nlspecs + NL_MAX_ORDER * (NL_MAX_ORDER + 1) / 2;
    for (uint32_t s = 0; s < h->nlspecs; s++) {
        const uint64_t n = (uint64_t)h->P * h->nlev[s];
        const double w = 1.0 / (double)((uint64_t)h->P * h->nlev[s]);
        hipLaunchKernelGGL(trial_conv_kernel, dim3(1024), dim3(NL_RED_THREADS), 0, st, h->trial,
                           h->trial_prev, h->pop_off[s], n, h->prm.pop_prec, w, partial);
        NL_TRY(hipGetLastError());
        std::vector<double> t(3 * 1024);
        NL_TRY(hipMemcpyAsync(t.data(), partial, t.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        NL_TRY(hipStreamSynchronize(st));
        double c = 0.0, m = 0.0, x = 0.0;
        for (int b = 0; b < 1024; b++) {
            c += t[3 * b];
            m += t[3 * b + 1];
            x = t[3 * b + 2] < x ? x : t[3 * b + 2];
        }
        h->conv[s].fraction_not_converged = c * w;
        h->conv[s].relative_change_mean = m;
        h->conv[s].relative_change_max = x;
    }
    if (conv) std::memcpy(conv, h->conv.data(), sizeof(mgx_nlte_conv) * h->nlspecs);
    return MGX_OK;
}

} // extern "C"

namespace {

/* Model::compute_level_populations_sparse (model.cpp:756-850) or, with global,
   Model::compute_level_populations (model.cpp:653-744) */
/* the formal solve of each iteration: the plain method with the Lambda diagonal (sparse
   update), with the general Lambda lists (global update), or order-0 short characteristics
   (whose Lambda is diagonal, so compute_Jeff and the global update reduce to the per-point
   update) */
enum LoopKind { LOOP_SPARSE = 0, LOOP_GLOBAL = 1, LOOP_SHORTCHAR = 2 };

int level_populations_loop(const mgx_model *m, const mgx_options *opts, mgx_nlte *h, int use_ng,
                           int64_t max_niterations, double *eps, double *kap, double *ic,
                           int32_t *iterations, double *error_mean, double *error_max,
                           hipStream_t st, int kind, double min_dtau = 0.0)
{
    const bool global = kind == LOOP_GLOBAL;
    if (!m || !opts || !h || !eps || !kap || !ic) return mgx::fail(MGX_E_ARG, "null argument");
    if (eps != m->emissivity || kap != m->opacity || ic != m->interp_crit)
        return mgx::fail(MGX_E_ARG, "emissivity, opacity and interp_crit must be the model view's arrays");
    if (m->npoints != h->P || m->nlines != h->L || m->nfreqs != h->F)
        return mgx::fail(MGX_E_ARG, "model view does not match the level-population handle");
    /* the sparse update needs the diagonal Lambda; short characteristics
       ignore n_off_diag (their Lambda is always diagonal, model.cpp:859-935) */
    if (kind == LOOP_SPARSE && opts->n_off_diag != 0)
        return mgx::fail(MGX_E_ARG, "parameters->n_off_diag != 0 so cannot use sparse update.");
    const size_t nJ = (size_t)h->P * h->F, nL = (size_t)h->P * h->L;
    const uint32_t cap = 1 + 2 * opts->n_off_diag * m->hnrays;
    if (int e = ensure(h->J, h->J_elems, nJ)) return e;
    if (global) {
        size_t have_n = h->lam_elems;
        if (int e = ensure(h->lam_val, h->lam_elems, nL * cap)) return e;
        if (int e = ensure(h->lam_nr, have_n, nL * cap)) return e;
        if (int e = ensure(h->lam_count, h->lam_count_elems, nL)) return e;
    } else {
        if (int e = ensure(h->lambda, h->lambda_elems, nL)) return e;
    }
    const mgx_nlte_params &pr = h->prm;
    const uint32_t skip = pr.Ng_acceleration_remove_N_its;
    int32_t iteration = 0;
    uint32_t normal = 0;
    size_t nerr = 0;
    bool some_not_converged = true;
    std::vector<mgx_nlte_conv> cv(h->nlspecs);
    while (some_not_converged && iteration < max_niterations) {
        /* Model::ng_acceleration_criterion */
        bool use_step = false;
        uint32_t order = 0;
        if (use_ng) {
            if (pr.use_adaptive_Ng_acceleration) {
                if (normal >= pr.adaptive_Ng_acceleration_min_order + skip) {
                    double curr = 0.0, trial = 0.0;
                    for (uint32_t s = 0; s < h->nlspecs; s++)
                        curr += pr.adaptive_Ng_acceleration_use_max_criterion
                                    ? h->conv[s].relative_change_max
                                    : h->conv[s].relative_change_mean;
                    if (int e = mgx_nlte_trial_ng(h, normal - skip, cv.data(), st)) return e;
                    for (uint32_t s = 0; s < h->nlspecs; s++)
                        trial += pr.adaptive_Ng_acceleration_use_max_criterion
                                     ? cv[s].relative_change_max
                                     : cv[s].relative_change_mean;
                    if (trial < curr || normal == pr.Ng_acceleration_mem_limit + skip) {
                        use_step = true;
                        order = normal - skip;
                    }
                }
            } else if (normal == pr.Ng_acceleration_mem_limit + skip) {
                use_step = true;
                order = pr.Ng_acceleration_mem_limit;
            }
        }
        if (use_step) {
            if (int e = mgx_nlte_iteration_ng(h, order, eps, kap, ic, cv.data(), st)) return e;
            normal = 0;
            continue;
        }
        iteration++;
        some_not_converged = false;
        if (global) {
            if (int e = mgx_feautrier_forward_lambda_device(m, opts, nullptr, h->J, cap,
                                                            h->lam_count, h->lam_nr, h->lam_val, st))
                return e;
            if (int e = mgx_nlte_iteration_stateq_global(h, h->J, cap, h->lam_count, h->lam_nr,
                                                         h->lam_val, eps, kap, ic, cv.data(),
                                                         nullptr, nullptr, nullptr, st))
                return e;
        } else {
            if (kind == LOOP_SHORTCHAR) {
                if (int e = mgx_shortchar_order_0_device(m, opts, min_dtau, nullptr, nullptr,
                                                         nullptr, h->J, h->lambda, st))
                    return e;
            } else if (int e = mgx_feautrier_forward_device(m, opts, nullptr, h->J, h->lambda, st)) {
                return e;
            }
            if (int e = mgx_nlte_iteration_stateq(h, h->J, h->lambda, eps, kap, ic, cv.data(),
                                                  nullptr, nullptr, st))
                return e;
        }
        h->last_loop = kind;
        h->last_cap = global ? cap : 0;
        normal++;
        for (uint32_t s = 0; s < h->nlspecs; s++) {
            if (error_mean) error_mean[nerr] = cv[s].relative_change_mean;
            if (error_max) error_max[nerr] = cv[s].relative_change_max;
            nerr++;
            if (cv[s].fraction_not_converged > 1.0 - pr.convergence_fraction)
                some_not_converged = true;
        }
    }
    if (iterations) *iterations = iteration;
    return MGX_OK;
}

} // namespace

extern "C" {

int mgx_nlte_last_lambda(const mgx_nlte *h, uint32_t *cap, const uint32_t **count,
                         const uint32_t **nr, const double **val, const double **diag)
{
    mgx::clear_error();
    if (!h || !cap || !count || !nr || !val || !diag) return mgx::fail(MGX_E_ARG, "null argument");
    if (h->last_loop < 0) return mgx::fail(MGX_E_ARG, "no iteration loop has run a formal solve");
    *cap = h->last_cap;
    const bool lists = h->last_loop == LOOP_GLOBAL;
    *count = lists ? h->lam_count : nullptr;
    *nr = lists ? h->lam_nr : nullptr;
    *val = lists ? h->lam_val : nullptr;
    *diag = lists ? nullptr : h->lambda;
    return MGX_OK;
}

int mgx_compute_level_populations_sparse(const mgx_model *m, const mgx_options *opts, mgx_nlte *h,
                                         int use_ng, int64_t max_niterations, double *eps,
                                         double *kap, double *ic, int32_t *iterations,
                                         double *error_mean, double *error_max, void *stream)
{
    mgx::clear_error();
    return level_populations_loop(m, opts, h, use_ng, max_niterations, eps, kap, ic, iterations,
                                  error_mean, error_max, (hipStream_t)stream, LOOP_SPARSE);
}

int mgx_compute_level_populations_shortchar(const mgx_model *m, const mgx_options *opts,
                                            mgx_nlte *h, int use_ng, int64_t max_niterations,
                                            double min_dtau, double *eps, double *kap, double *ic,
                                            int32_t *iterations, double *error_mean,
                                            double *error_max, void *stream)
{
    mgx::clear_error();
    return level_populations_loop(m, opts, h, use_ng, max_niterations, eps, kap, ic, iterations,
                                  error_mean, error_max, (hipStream_t)stream, LOOP_SHORTCHAR,
                                  min_dtau);
}

int mgx_compute_level_populations(const mgx_model *m, const mgx_options *opts, mgx_nlte *h,
                                  int use_ng, int64_t max_niterations, double *eps, double *kap,
                                  double *ic, int32_t *iterations, double *error_mean,
                                  double *error_max, void *stream)
{
    mgx::clear_error();
    if (opts && opts->n_off_diag > MGX_MAX_OFF_DIAG)
        return mgx::fail(MGX_E_UNSUPPORTED, "n_off_diag > %d is not supported", MGX_MAX_OFF_DIAG);
    return level_populations_loop(m, opts, h, use_ng, max_niterations, eps, kap, ic, iterations,
                                  error_mean, error_max, (hipStream_t)stream, LOOP_GLOBAL);
}

} // extern "C"
