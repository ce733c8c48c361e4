/*
 * CPU ORACLE -- test infrastructure only (tests/ and bench.py's cpu_baseline).
 *
 * Long-double restatement of the reference's sparse level-population
 * update, following the reference's types: populations, the residual
 * history and the statistical-equilibrium matrix are long double (Eigen
 * VectorXld / MatrixXld there, x87 80-bit here as with gcc on x86-64),
 * J_lin / J_eff and the line emissivity / opacity are double.
 *
 *   LineProducingSpecies::update_using_LTE        lineProducingSpecies.tpp:52-75
 *   LineProducingSpecies::set_LTE                 :84-99
 *   LineProducingSpecies::set_LTE_specific_levels :105-125
 *   LineProducingSpecies::check_for_convergence   :127-163
 *   LineProducingSpecies::check_for_convergence_trial :165-201
 *   LineProducingSpecies::update_using_acceleration        :266-314
 *   LineProducingSpecies::update_using_acceleration_trial  :316-369
 *   LineProducingSpecies::update_using_statistical_equilibrium_sparse :589-683
 *   LineProducingSpecies::correct_negative_populations     :711-802
 *   LineProducingSpecies::get_emissivity / get_opacity     :23-44
 *   CollisionPartner::interpolate_collision_coefficients   collisionPartner.tpp:5-23
 *   CollisionPartner::adjust_abundance_for_ortho_or_para   collisionPartner.tpp:29-42
 *   search (tools/interpolation.hpp:10-33)
 *   Model::compute_Jeff_sparse                    model.cpp:549-570, with J_lin summed as
 *     Solver::solve_feautrier_order_2_sparse does (solver.tpp:298-363)
 *
 * The QR solve restates Eigen's ColPivHouseholderQR (Businger-Golub column
 * pivoting with the LAPACK xGEQP3 norm downdating, Householder reflectors
 * as in makeHouseholder / applyHouseholderOnTheLeft, rank-revealing solve);
 * it is Eigen's algorithm, not its blocked code, so results agree with the
 * reference to long-double rounding, not bit for bit.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mgx_fmath.h"
#include "mgx_fmath_ref.h"
#include "mgx_oracle.h"

#if defined(MGX_ORACLE_GLIBC)
#define N_EXPF expf
#elif defined(MGX_ORACLE_CR)
#define N_EXPF mgx_expf
#else
#define N_EXPF mgx_ref_expf
#endif

typedef long double ld;

/* src/tools/constants.hpp */
static const double N_KB = 1.38064852E-23;
static const double N_HH = 6.62607004E-34;
#define N_HH_OVER_FOUR_PI (N_HH / (4.0 * 3.14159265358979323846))

/* ------------------------------------------------------------------------ */
/* column-pivoted Householder QR in long double (Eigen ColPivHouseholderQR)  */
/* ------------------------------------------------------------------------ */

/* A is n x n, column-major in a[j*n + i]; overwritten.  x = A^-1 b in the
   rank-revealing sense (components beyond the numerical rank set to zero). */
int oracle_cpqr_solve_ld(int n, ld *a, const ld *b, ld *x)
{
    if (n <= 0) return 0;
    ld *hc = calloc((size_t)n, sizeof(ld));
    ld *nu = calloc((size_t)n, sizeof(ld)); /* updated column norms */
    ld *nd = calloc((size_t)n, sizeof(ld)); /* direct column norms */
    int *perm = malloc((size_t)n * sizeof(int));
    ld *c = malloc((size_t)n * sizeof(ld));
    if (!hc || !nu || !nd || !perm || !c) {
        free(hc); free(nu); free(nd); free(perm); free(c);
        return -1;
    }
#define A_(i, j) a[(size_t)(j) * n + (i)]
    const ld eps = LDBL_EPSILON;
    for (int j = 0; j < n; j++) {
        ld s = 0.0L;
        for (int i = 0; i < n; i++) s += A_(i, j) * A_(i, j);
        nd[j] = nu[j] = sqrtl(s);
        perm[j] = j;
    }
    ld maxnorm = 0.0L;
    for (int j = 0; j < n; j++) maxnorm = nu[j] > maxnorm ? nu[j] : maxnorm;
    const ld threshold_helper = (maxnorm * eps) * (maxnorm * eps) / (ld)n;
    const ld norm_downdate_threshold = sqrtl(eps);
    int nonzero_pivots = n;
    ld maxpivot = 0.0L;
    for (int k = 0; k < n; k++) {
        int big = k;
        for (int j = k + 1; j < n; j++)
            if (nu[j] > nu[big]) big = j;
        const ld big_sq = nu[big] * nu[big];
        if (nonzero_pivots == n && big_sq < threshold_helper * (ld)(n - k)) nonzero_pivots = k;
        if (big != k) {
            for (int i = 0; i < n; i++) {
                const ld t = A_(i, k);
                A_(i, k) = A_(i, big);
                A_(i, big) = t;
            }
            ld t = nu[k]; nu[k] = nu[big]; nu[big] = t;
            t = nd[k]; nd[k] = nd[big]; nd[big] = t;
            const int pt = perm[k]; perm[k] = perm[big]; perm[big] = pt;
        }
        /* makeHouseholderInPlace on A(k:n, k) */
        ld tail = 0.0L;
        for (int i = k + 1; i < n; i++) tail += A_(i, k) * A_(i, k);
        const ld c0 = A_(k, k);
        ld beta, tau;
        if (tail <= LDBL_MIN) {
            tau = 0.0L;
            beta = c0;
            for (int i = k + 1; i < n; i++) A_(i, k) = 0.0L;
        } else {
            beta = sqrtl(c0 * c0 + tail);
            if (c0 >= 0.0L) beta = -beta;
            for (int i = k + 1; i < n; i++) A_(i, k) = A_(i, k) / (c0 - beta);
            tau = (beta - c0) / beta;
        }
        hc[k] = tau;
        A_(k, k) = beta;
        if (fabsl(beta) > maxpivot) maxpivot = fabsl(beta);
        /* applyHouseholderOnTheLeft to A(k:n, k+1:n) */
        if (tau != 0.0L) {
            for (int j = k + 1; j < n; j++) {
                ld t = A_(k, j);
                for (int i = k + 1; i < n; i++) t += A_(i, k) * A_(i, j);
                A_(k, j) -= tau * t;
                for (int i = k + 1; i < n; i++) A_(i, j) -= tau * A_(i, k) * t;
            }
        }
        /* norm downdating (LAPACK xGEQP3) */
        for (int j = k + 1; j < n; j++) {
            if (nu[j] != 0.0L) {
                ld t = fabsl(A_(k, j)) / nu[j];
                t = (1.0L + t) * (1.0L - t);
                if (t < 0.0L) t = 0.0L;
                const ld r = nu[j] / nd[j];
                const ld t2 = t * r * r;
                if (t2 <= norm_downdate_threshold) {
                    ld s = 0.0L;
                    for (int i = k + 1; i < n; i++) s += A_(i, j) * A_(i, j);
                    nd[j] = nu[j] = sqrtl(s);
                } else {
                    nu[j] *= sqrtl(t);
                }
            }
        }
    }
    /* rank(): |R(i,i)| > maxpivot * eps * n over the nonzero pivots */
    const ld pthr = maxpivot * eps * (ld)n;
    int rank = 0;
    for (int i = 0; i < nonzero_pivots; i++)
        if (fabsl(A_(i, i)) > pthr) rank++;
    if (rank == 0) {
        for (int i = 0; i < n; i++) x[i] = 0.0L;
    } else {
        for (int i = 0; i < n; i++) c[i] = b[i];
        /* c = H_{r-1} ... H_0 c */
        for (int k = 0; k < rank; k++) {
            if (hc[k] == 0.0L) continue;
            ld t = c[k];
            for (int i = k + 1; i < n; i++) t += A_(i, k) * c[i];
            c[k] -= hc[k] * t;
            for (int i = k + 1; i < n; i++) c[i] -= hc[k] * A_(i, k) * t;
        }
        /* R(0:r, 0:r) z = c(0:r) */
        for (int i = rank - 1; i >= 0; i--) {
            ld s = c[i];
            for (int j = i + 1; j < rank; j++) s -= A_(i, j) * c[j];
            c[i] = s / A_(i, i);
        }
        for (int i = 0; i < n; i++) x[perm[i]] = i < rank ? c[i] : 0.0L;
    }
#undef A_
    free(hc); free(nu); free(nd); free(perm); free(c);
    return 0;
}

/* double in/out wrapper (tests) */
int oracle_cpqr_solve(int n, const double *a_colmajor, const double *b, double *x)
{
    ld *a = malloc((size_t)n * n * sizeof(ld));
    ld *bb = calloc(n > 0 ? (size_t)n : 1, sizeof(ld));
    ld *xx = malloc((size_t)n * sizeof(ld));
    if (!a || !bb || !xx) { free(a); free(bb); free(xx); return -1; }
    for (int i = 0; i < n * n; i++) a[i] = a_colmajor[i];
    for (int i = 0; i < n; i++) bb[i] = b[i];
    int e = oracle_cpqr_solve_ld(n, a, bb, xx);
    for (int i = 0; i < n; i++) x[i] = (double)xx[i];
    free(a); free(bb); free(xx);
    return e;
}

/* ------------------------------------------------------------------------ */
/* state                                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t num_col_partner;
    int orth_or_para;
    uint32_t ntmp, ncol;
    uint32_t *icol, *jcol;
    double *tmp, *Cd, *Ce;
} ocol;

typedef struct {
    uint32_t num, nlev, nrad, lid0, ncolpar;
    double *energy, *weight, *A, *Bs, *Ba, *frequency;
    uint32_t *irad, *jrad;
    ocol *cp;
    ld *pop, *prev1, *trial, *trial_prev;
    double *pop_tot;
    double *Jlin, *Jeff;
    ld **pops; /* LineProducingSpecies::populations */
    size_t npops, cap_pops;
    ld **res; /* LineProducingSpecies::residuals */
    size_t nres, cap_res;
    double fnc, rcm, rcmax;
} ospec;

struct oracle_nlte {
    uint32_t P, L, Q, F, nlspecs, nspec, hnrays;
    double *abundance, *temperature, *quad_weights, *ray_weight;
    int adaptive;
    mgx_nlte_params prm;
    ospec *s;
};

static void *dup_arr(const void *src, size_t bytes)
{
    void *d = malloc(bytes ? bytes : 1);
    if (d && bytes) memcpy(d, src, bytes);
    return d;
}

static ld *vec_new(size_t n) { return calloc(n ? n : 1, sizeof(ld)); }

static int push(ld ***list, size_t *n, size_t *cap, ld *v)
{
    if (*n == *cap) {
        size_t nc = *cap ? 2 * *cap : 8;
        ld **nl = realloc(*list, nc * sizeof(ld *));
        if (!nl) return -1;
        *list = nl;
        *cap = nc;
    }
    (*list)[(*n)++] = v;
    return 0;
}

oracle_nlte *oracle_nlte_create(const mgx_model *m, uint32_t nlspecs, const mgx_linedata *ld_,
                                const uint32_t *ncolpar, const mgx_colldata *colpar,
                                uint32_t nspec, const double *abundance,
                                const mgx_nlte_params *prm)
{
    oracle_nlte *h = calloc(1, sizeof(oracle_nlte));
    if (!h) return NULL;
    h->P = m->npoints;
    h->L = m->nlines;
    h->Q = m->nquads;
    h->F = m->nfreqs;
    h->hnrays = m->hnrays;
    h->nlspecs = nlspecs;
    h->nspec = nspec;
    h->prm = *prm;
    h->adaptive = m->adaptive_directions;
    const uint32_t P = h->P;
    h->abundance = dup_arr(abundance, (size_t)P * nspec * sizeof(double));
    h->temperature = dup_arr(m->temperature, (size_t)P * sizeof(double));
    h->quad_weights = dup_arr(m->quad_weights, (size_t)nlspecs * h->Q * sizeof(double));
    h->ray_weight = dup_arr(m->ray_weight, (size_t)(h->adaptive ? P : 1) * m->nrays * sizeof(double));
    h->s = calloc(nlspecs, sizeof(ospec));
    uint32_t lid0 = 0, c0 = 0;
    for (uint32_t l = 0; l < nlspecs; l++) {
        ospec *s = &h->s[l];
        const mgx_linedata *d = &ld_[l];
        s->num = d->num;
        s->nlev = d->nlev;
        s->nrad = d->nrad;
        s->lid0 = lid0;
        lid0 += d->nrad;
        s->energy = dup_arr(d->energy, d->nlev * sizeof(double));
        s->weight = dup_arr(d->weight, d->nlev * sizeof(double));
        s->A = dup_arr(d->A, d->nrad * sizeof(double));
        s->Bs = dup_arr(d->Bs, d->nrad * sizeof(double));
        s->Ba = dup_arr(d->Ba, d->nrad * sizeof(double));
        s->frequency = dup_arr(d->frequency, d->nrad * sizeof(double));
        s->irad = dup_arr(d->irad, d->nrad * sizeof(uint32_t));
        s->jrad = dup_arr(d->jrad, d->nrad * sizeof(uint32_t));
        s->ncolpar = ncolpar ? ncolpar[l] : 0;
        s->cp = calloc(s->ncolpar ? s->ncolpar : 1, sizeof(ocol));
        for (uint32_t c = 0; c < s->ncolpar; c++, c0++) {
            const mgx_colldata *cd = &colpar[c0];
            ocol *o = &s->cp[c];
            o->num_col_partner = cd->num_col_partner;
            o->orth_or_para = cd->orth_or_para_H2;
            o->ntmp = cd->ntmp;
            o->ncol = cd->ncol;
            o->icol = dup_arr(cd->icol, cd->ncol * sizeof(uint32_t));
            o->jcol = dup_arr(cd->jcol, cd->ncol * sizeof(uint32_t));
            o->tmp = dup_arr(cd->tmp, cd->ntmp * sizeof(double));
            o->Cd = dup_arr(cd->Cd, (size_t)cd->ntmp * cd->ncol * sizeof(double));
            o->Ce = dup_arr(cd->Ce, (size_t)cd->ntmp * cd->ncol * sizeof(double));
        }
        const size_t n = (size_t)P * s->nlev;
        s->pop = vec_new(n);
        s->prev1 = vec_new(n);
        s->trial = vec_new(n);
        s->trial_prev = vec_new(n);
        s->pop_tot = calloc(P ? P : 1, sizeof(double));
        s->Jlin = calloc((size_t)P * s->nrad + 1, sizeof(double));
        s->Jeff = calloc((size_t)P * s->nrad + 1, sizeof(double));
    }
    return h;
}

static void clear_history(ospec *s)
{
    for (size_t i = 0; i < s->npops; i++) free(s->pops[i]);
    for (size_t i = 0; i < s->nres; i++) free(s->res[i]);
    s->npops = s->nres = 0;
}

void oracle_nlte_destroy(oracle_nlte *h)
{
    if (!h) return;
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        clear_history(s);
        free(s->pops); free(s->res);
        for (uint32_t c = 0; c < s->ncolpar; c++) {
            ocol *o = &s->cp[c];
            free(o->icol); free(o->jcol); free(o->tmp); free(o->Cd); free(o->Ce);
        }
        free(s->cp);
        free(s->energy); free(s->weight); free(s->A); free(s->Bs); free(s->Ba);
        free(s->frequency); free(s->irad); free(s->jrad);
        free(s->pop); free(s->prev1); free(s->trial); free(s->trial_prev);
        free(s->pop_tot); free(s->Jlin); free(s->Jeff);
    }
    free(h->s);
    free(h->abundance); free(h->temperature); free(h->quad_weights); free(h->ray_weight);
    free(h);
}

static ld *vec_copy(const ld *v, size_t n)
{
    ld *d = vec_new(n);
    if (d) memcpy(d, v, n * sizeof(ld));
    return d;
}

/* ------------------------------------------------------------------------ */
/* populations -> line emissivity and opacity                                */
/* ------------------------------------------------------------------------ */

static double get_emissivity(const oracle_nlte *h, const ospec *s, uint32_t p, uint32_t k)
{
    const ld ni = s->pop[(size_t)p * s->nlev + s->irad[k]];
    const ld e = s->frequency[k] * N_HH_OVER_FOUR_PI * s->A[k] * ni;
    const ld mn = h->prm.min_line_emissivity;
    return (double)(e > mn ? e : mn);
}

static double get_opacity(const oracle_nlte *h, const ospec *s, uint32_t p, uint32_t k)
{
    const ld ni = s->pop[(size_t)p * s->nlev + s->irad[k]];
    const ld nj = s->pop[(size_t)p * s->nlev + s->jrad[k]];
    const ld o = s->frequency[k] * N_HH_OVER_FOUR_PI * (nj * s->Ba[k] - ni * s->Bs[k]);
    const ld mn = h->prm.min_line_opacity;
    return (double)(o > mn ? o : mn);
}

/* Lines::set_emissivity_and_opacity + the interpolation criterion */
static void set_emissivity_and_opacity(const oracle_nlte *h, double *eps, double *kap, double *ic)
{
    for (uint32_t p = 0; p < h->P; p++)
        for (uint32_t l = 0; l < h->nlspecs; l++) {
            const ospec *s = &h->s[l];
            for (uint32_t k = 0; k < s->nrad; k++) {
                const size_t idx = (size_t)p * h->L + s->lid0 + k;
                const double e = get_emissivity(h, s, p, k);
                const double o = get_opacity(h, s, p, k);
                if (eps) eps[idx] = e;
                if (kap) kap[idx] = o;
                if (ic) ic[idx] = o / s->pop_tot[p];
            }
        }
}

/* ------------------------------------------------------------------------ */
/* LTE                                                                       */
/* ------------------------------------------------------------------------ */

static void update_using_LTE(oracle_nlte *h, ospec *s)
{
    for (uint32_t p = 0; p < h->P; p++) {
        s->pop_tot[p] = h->abundance[(size_t)p * h->nspec + s->num];
        double pf = 0.0;
        for (uint32_t i = 0; i < s->nlev; i++) {
            const size_t ind = (size_t)p * s->nlev + i;
            s->pop[ind] = s->weight[i] * exp(-s->energy[i] / (N_KB * h->temperature[p]));
            pf = (double)(pf + s->pop[ind]);
        }
        for (uint32_t i = 0; i < s->nlev; i++) {
            const size_t ind = (size_t)p * s->nlev + i;
            s->pop[ind] *= s->pop_tot[p] / pf;
        }
    }
}

int oracle_nlte_set_lte(oracle_nlte *h, double *eps, double *kap, double *ic)
{
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        update_using_LTE(h, s);
        clear_history(s);
        if (push(&s->pops, &s->npops, &s->cap_pops, vec_copy(s->pop, (size_t)h->P * s->nlev)))
            return -1;
    }
    set_emissivity_and_opacity(h, eps, kap, ic);
    return 0;
}

int oracle_nlte_set_populations(oracle_nlte *h, const double *pop, double *eps, double *kap,
                                double *ic)
{
    size_t off = 0;
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        const size_t n = (size_t)h->P * s->nlev;
        for (uint32_t p = 0; p < h->P; p++)
            s->pop_tot[p] = h->abundance[(size_t)p * h->nspec + s->num];
        for (size_t i = 0; i < n; i++) s->pop[i] = pop[off + i];
        off += n;
        clear_history(s);
        if (push(&s->pops, &s->npops, &s->cap_pops, vec_copy(s->pop, n))) return -1;
    }
    set_emissivity_and_opacity(h, eps, kap, ic);
    return 0;
}

int oracle_nlte_get(const oracle_nlte *h, double *pop, double *trial, double *Jlin, double *Jeff)
{
    size_t off = 0, offk = 0;
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        const ospec *s = &h->s[l];
        const size_t n = (size_t)h->P * s->nlev;
        for (size_t i = 0; i < n; i++) {
            if (pop) pop[off + i] = (double)s->pop[i];
            if (trial) trial[off + i] = (double)s->trial[i];
        }
        off += n;
        (void)offk;
    }
    /* J_lin, J_eff as (P, L) */
    for (uint32_t p = 0; p < h->P; p++)
        for (uint32_t l = 0; l < h->nlspecs; l++) {
            const ospec *s = &h->s[l];
            for (uint32_t k = 0; k < s->nrad; k++) {
                if (Jlin) Jlin[(size_t)p * h->L + s->lid0 + k] = s->Jlin[(size_t)p * s->nrad + k];
                if (Jeff) Jeff[(size_t)p * h->L + s->lid0 + k] = s->Jeff[(size_t)p * s->nrad + k];
            }
        }
    return 0;
}

int oracle_nlte_stats(const oracle_nlte *h, double *conv, uint32_t *n_residuals)
{
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        if (conv) {
            conv[3 * l + 0] = h->s[l].fnc;
            conv[3 * l + 1] = h->s[l].rcm;
            conv[3 * l + 2] = h->s[l].rcmax;
        }
        if (n_residuals) n_residuals[l] = (uint32_t)h->s[l].nres;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* J_lin and J_eff                                                           */
/* ------------------------------------------------------------------------ */

/* u (hnrays, P, F) of the plain solve, lambda_diag (P, L).  J_lin(o, k)
   accumulates w_z * (2 w_rr) * u over rr (outer) and z (inner), the order of
   Solver::solve_feautrier_order_2_sparse; u holds Su[centre] or the boundary
   intensity exactly as that loop adds them. */
int oracle_nlte_compute_Jeff(oracle_nlte *h, const double *u, const double *lambda_diag)
{
    const uint32_t P = h->P, F = h->F, Q = h->Q;
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        for (size_t i = 0; i < (size_t)P * s->nrad; i++) s->Jlin[i] = 0.0;
        for (uint32_t rr = 0; rr < h->hnrays; rr++)
            for (uint32_t o = 0; o < P; o++) {
                const double w = h->adaptive ? h->ray_weight[(size_t)o * 2 * h->hnrays + rr]
                                             : h->ray_weight[rr];
                const double wt = w * 2.0;
                for (uint32_t k = 0; k < s->nrad; k++) {
                    const uint32_t lid = s->lid0 + k;
                    double acc = s->Jlin[(size_t)o * s->nrad + k];
                    for (uint32_t z = 0; z < Q; z++) {
                        const size_t f = (size_t)lid * Q + z;
                        acc += h->quad_weights[(size_t)l * Q + z] * wt
                             * u[((size_t)rr * P + o) * F + f];
                    }
                    s->Jlin[(size_t)o * s->nrad + k] = acc;
                }
            }
        for (uint32_t p = 0; p < P; p++)
            for (uint32_t k = 0; k < s->nrad; k++) {
                const ld ni = s->pop[(size_t)p * s->nlev + s->irad[k]];
                double diff = 0.0;
                diff = (double)(diff + lambda_diag[(size_t)p * h->L + s->lid0 + k] * ni);
                s->Jeff[(size_t)p * s->nrad + k] =
                    s->Jlin[(size_t)p * s->nrad + k] - N_HH_OVER_FOUR_PI * diff;
            }
    }
    return 0;
}

/* J_lin given directly (P, L) (tests of the statistical equilibrium alone). */
int oracle_nlte_set_Jeff(oracle_nlte *h, const double *Jlin, const double *lambda_diag)
{
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        for (uint32_t p = 0; p < h->P; p++)
            for (uint32_t k = 0; k < s->nrad; k++) {
                const size_t idx = (size_t)p * h->L + s->lid0 + k;
                const ld ni = s->pop[(size_t)p * s->nlev + s->irad[k]];
                double diff = 0.0;
                diff = (double)(diff + lambda_diag[idx] * ni);
                s->Jlin[(size_t)p * s->nrad + k] = Jlin[idx];
                s->Jeff[(size_t)p * s->nrad + k] = Jlin[idx] - N_HH_OVER_FOUR_PI * diff;
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* statistical equilibrium                                                   */
/* ------------------------------------------------------------------------ */

static uint32_t search_tmp(const double *x, uint32_t n, double v)
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

static void collision_rates(const ocol *o, double T, double *Cd, double *Ce)
{
    const uint32_t t = search_tmp(o->tmp, o->ntmp, T);
    if (t == 0) {
        for (uint32_t k = 0; k < o->ncol; k++) { Ce[k] = o->Ce[k]; Cd[k] = o->Cd[k]; }
    } else if (t == o->ntmp - 1) {
        for (uint32_t k = 0; k < o->ncol; k++) {
            Ce[k] = o->Ce[(size_t)(o->ntmp - 1) * o->ncol + k];
            Cd[k] = o->Cd[(size_t)(o->ntmp - 1) * o->ncol + k];
        }
    } else {
        const double step = (T - o->tmp[t - 1]) / (o->tmp[t] - o->tmp[t - 1]);
        for (uint32_t k = 0; k < o->ncol; k++) {
            const size_t a = (size_t)(t - 1) * o->ncol + k, b = (size_t)t * o->ncol + k;
            Ce[k] = o->Ce[a] + (o->Ce[b] - o->Ce[a]) * step;
            Cd[k] = o->Cd[a] + (o->Cd[b] - o->Cd[a]) * step;
        }
    }
}

static double ortho_para(int op, double T, double abn)
{
    if (op != 'n') {
        const double frac_H2_para = 1.0 / (1.0 + 9.0 * N_EXPF((float)(-170.5 / T)));
        if (op == 'o') abn *= (1.0 - frac_H2_para);
        if (op == 'p') abn *= frac_H2_para;
    }
    return abn;
}

/* one point's system (update_using_statistical_equilibrium_sparse body up to the solve);
   M is nlev x nlev column-major */
static void assemble_point(const oracle_nlte *h, const ospec *s, const double *lambda_diag,
                           uint32_t p, ld *M, ld *y, double *Cd, double *Ce)
{
    const uint32_t n = s->nlev;
    for (size_t i = 0; i < (size_t)n * n; i++) M[i] = 0.0L;
    for (uint32_t i = 0; i < n; i++) y[i] = 0.0L;
#define M_(r, c) M[(size_t)(c) * n + (r)]
    for (uint32_t k = 0; k < s->nrad; k++) {
        const double Jeff = s->Jeff[(size_t)p * s->nrad + k];
        const ld v_IJ = s->A[k] + s->Bs[k] * Jeff;
        const ld v_JI = s->Ba[k] * Jeff;
        const uint32_t I = s->irad[k], J = s->jrad[k];
        M_(J, I) += v_IJ;
        M_(J, J) -= v_JI;
        M_(I, J) += v_JI;
        M_(I, I) -= v_IJ;
    }
    for (uint32_t k = 0; k < s->nrad; k++) {
        const double freq = s->frequency[k];
        const double Ls = lambda_diag[(size_t)p * h->L + s->lid0 + k];
        const ld v_IJ = -Ls * get_opacity(h, s, p, k) / freq;
        const uint32_t I = s->irad[k], J = s->jrad[k];
        M_(J, I) += v_IJ;
        M_(I, I) -= v_IJ;
    }
    for (uint32_t c = 0; c < s->ncolpar; c++) {
        const ocol *o = &s->cp[c];
        double abn = h->abundance[(size_t)p * h->nspec + o->num_col_partner];
        const double tmp = h->temperature[p];
        abn = ortho_para(o->orth_or_para, tmp, abn);
        collision_rates(o, tmp, Cd, Ce);
        for (uint32_t k = 0; k < o->ncol; k++) {
            const ld v_IJ = Cd[k] * abn;
            const ld v_JI = Ce[k] * abn;
            const uint32_t I = o->icol[k], J = o->jcol[k];
            M_(J, I) += v_IJ;
            M_(J, J) -= v_JI;
            M_(I, J) += v_JI;
            M_(I, I) -= v_IJ;
        }
    }
    for (uint32_t c = 0; c < n; c++) M_(n - 1, c) = 1.0L;
#undef M_
    y[n - 1] = s->pop_tot[p];
}

/* the general Lambda operator: per (point, global line) at most cap (point, value) entries */
typedef struct {
    uint32_t cap;
    const uint32_t *count, *nr;
    const double *val;
} lamview;

/* Model::compute_Jeff (model.cpp:514-545): J_lin = sum_z w_z J(p, line bin z), J_eff =
   J_lin - h/4pi sum over the list of (p, k) of Ls n_upper(emitting point) */
int oracle_nlte_compute_Jeff_general(oracle_nlte *h, const double *J, uint32_t cap,
                                     const uint32_t *lam_count, const uint32_t *lam_nr,
                                     const double *lam_val)
{
    const uint32_t P = h->P, F = h->F, Q = h->Q;
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        for (uint32_t p = 0; p < P; p++)
            for (uint32_t k = 0; k < s->nrad; k++) {
                const uint32_t lid = s->lid0 + k;
                double Jlin = 0.0;
                for (uint32_t z = 0; z < Q; z++)
                    Jlin += h->quad_weights[(size_t)l * Q + z] * J[(size_t)p * F + (size_t)lid * Q + z];
                const size_t pl = (size_t)p * h->L + lid;
                double diff = 0.0;
                for (uint32_t m = 0; m < lam_count[pl]; m++) {
                    const uint32_t q = lam_nr[pl * cap + m];
                    diff = (double)(diff + lam_val[pl * cap + m] * s->pop[(size_t)q * s->nlev + s->irad[k]]);
                }
                s->Jlin[(size_t)p * s->nrad + k] = Jlin;
                s->Jeff[(size_t)p * s->nrad + k] = Jlin - N_HH_OVER_FOUR_PI * diff;
            }
    }
    return 0;
}

/* the global rate equations of update_using_statistical_equilibrium
   (lineProducingSpecies.tpp:377-523): triplets summed into a dense column-major
   (P nlev)^2 long-double matrix in the order they are pushed; rows of a point's last level
   skip every term and hold the normalisation.  An approximated-Lambda element of the list
   of (p, k) at emitting point q couples row (p, lower) to column (q, upper) and adds to the
   diagonal of row (q, upper), as the reference does. */
static void assemble_global(const oracle_nlte *h, const ospec *s, const lamview *lv, ld *M,
                            ld *y, double *Cd, double *Ce)
{
    const uint32_t n = s->nlev;
    const size_t N = (size_t)h->P * n;
    for (size_t i = 0; i < N * N; i++) M[i] = 0.0L;
    for (size_t i = 0; i < N; i++) y[i] = 0.0L;
#define G_(r, c) M[(size_t)(c) * N + (r)]
    for (uint32_t p = 0; p < h->P; p++) {
        const size_t b = (size_t)p * n;
        for (uint32_t k = 0; k < s->nrad; k++) {
            const double Jeff = s->Jeff[(size_t)p * s->nrad + k];
            const ld v_IJ = s->A[k] + s->Bs[k] * Jeff;
            const ld v_JI = s->Ba[k] * Jeff;
            const size_t I = b + s->irad[k], J = b + s->jrad[k];
            if (s->jrad[k] != n - 1) {
                G_(J, I) += v_IJ;
                G_(J, J) -= v_JI;
            }
            if (s->irad[k] != n - 1) {
                G_(I, J) += v_JI;
                G_(I, I) -= v_IJ;
            }
        }
        for (uint32_t k = 0; k < s->nrad; k++) {
            const double freq = s->frequency[k];
            const size_t pl = (size_t)p * h->L + s->lid0 + k;
            for (uint32_t m = 0; m < lv->count[pl]; m++) {
                const uint32_t q = lv->nr[pl * lv->cap + m];
                const ld v_IJ = -lv->val[pl * lv->cap + m] * get_opacity(h, s, p, k) / freq;
                const size_t I = (size_t)q * n + s->irad[k], J = b + s->jrad[k];
                if (s->jrad[k] != n - 1) G_(J, I) += v_IJ;
                if (s->irad[k] != n - 1) G_(I, I) -= v_IJ;
            }
        }
        for (uint32_t c = 0; c < s->ncolpar; c++) {
            const ocol *o = &s->cp[c];
            double abn = h->abundance[(size_t)p * h->nspec + o->num_col_partner];
            const double tmp = h->temperature[p];
            abn = ortho_para(o->orth_or_para, tmp, abn);
            collision_rates(o, tmp, Cd, Ce);
            for (uint32_t k = 0; k < o->ncol; k++) {
                const ld v_IJ = Cd[k] * abn;
                const ld v_JI = Ce[k] * abn;
                const size_t I = b + o->icol[k], J = b + o->jcol[k];
                if (o->jcol[k] != n - 1) {
                    G_(J, I) += v_IJ;
                    G_(J, J) -= v_JI;
                }
                if (o->icol[k] != n - 1) {
                    G_(I, J) += v_JI;
                    G_(I, I) -= v_IJ;
                }
            }
        }
        for (uint32_t i = 0; i < n; i++) G_(b + n - 1, b + i) += 1.0L;
        y[b + n - 1] = s->pop_tot[p];
    }
#undef G_
}

static uint32_t max_ncol(const ospec *s)
{
    uint32_t mc = 1;
    for (uint32_t c = 0; c < s->ncolpar; c++) mc = s->cp[c].ncol > mc ? s->cp[c].ncol : mc;
    return mc;
}

/* the system of one point, for tests: M (nlev x nlev, column-major) and y */
int oracle_nlte_stateq_system(const oracle_nlte *h, const double *lambda_diag, uint32_t l,
                              uint32_t p, double *M_out, double *y_out)
{
    const ospec *s = &h->s[l];
    const uint32_t n = s->nlev, mc = max_ncol(s);
    ld *M = malloc((size_t)n * n * sizeof(ld)), *y = malloc(n * sizeof(ld));
    double *Cd = malloc(mc * sizeof(double)), *Ce = malloc(mc * sizeof(double));
    if (!M || !y || !Cd || !Ce) { free(M); free(y); free(Cd); free(Ce); return -1; }
    assemble_point(h, s, lambda_diag, p, M, y, Cd, Ce);
    for (size_t i = 0; i < (size_t)n * n; i++) M_out[i] = (double)M[i];
    for (uint32_t i = 0; i < n; i++) y_out[i] = (double)y[i];
    free(M); free(y); free(Cd); free(Ce);
    return 0;
}

/* LineProducingSpecies::set_LTE (one transition's two levels) */
static void set_LTE_line(const oracle_nlte *h, ospec *s, uint32_t p, uint32_t k)
{
    const size_t i = (size_t)p * s->nlev + s->irad[k];
    const size_t j = (size_t)p * s->nlev + s->jrad[k];
    const double T = h->temperature[p];
    const double tot = (double)(s->pop[i] + s->pop[j]);
    s->pop[i] = s->weight[s->irad[k]] * exp(-s->energy[s->irad[k]] / (N_KB * T));
    s->pop[j] = s->weight[s->jrad[k]] * exp(-s->energy[s->jrad[k]] / (N_KB * T));
    const double pf = (double)(s->pop[i] + s->pop[j]);
    s->pop[i] *= tot / pf;
    s->pop[j] *= tot / pf;
}

static void set_LTE_levels(const oracle_nlte *h, ospec *s, uint32_t p, const char *mask)
{
    const double T = h->temperature[p];
    double tot = 0.0, pf = 0.0;
    for (uint32_t i = 0; i < s->nlev; i++)
        if (mask[i]) {
            const size_t ind = (size_t)p * s->nlev + i;
            tot = (double)(tot + s->pop[ind]);
            s->pop[ind] = s->weight[i] * exp(-s->energy[i] / (N_KB * T));
            pf = (double)(pf + s->pop[ind]);
        }
    for (uint32_t i = 0; i < s->nlev; i++)
        if (mask[i]) s->pop[(size_t)p * s->nlev + i] *= tot / pf;
}

static void correct_negative_populations(const oracle_nlte *h, ospec *s)
{
    const uint32_t n = s->nlev;
    for (uint32_t p = 0; p < h->P; p++) {
        if (s->pop_tot[p] > 0.0) {
            double tpp = 0.0;
            for (uint32_t i = 0; i < n; i++) {
                ld *v = &s->pop[(size_t)p * n + i];
                if (*v < 0.0L) *v = 0.0L;
                else tpp = (double)(tpp + *v);
            }
            for (uint32_t i = 0; i < n; i++) {
                ld *v = &s->pop[(size_t)p * n + i];
                *v = *v * s->pop_tot[p] / tpp;
            }
        }
    }
    const double f = h->prm.population_inversion_fraction;
    char *mod = malloc(n ? n : 1);
    for (uint32_t p = 0; p < h->P; p++) {
        memset(mod, 0, n);
        uint32_t next = 0, curr;
        do {
            curr = next;
            int inv = 0;
            for (uint32_t kk = s->nrad; kk-- > 0;) {
                const ld ni = s->pop[(size_t)p * n + s->irad[kk]];
                const ld nj = s->pop[(size_t)p * n + s->jrad[kk]];
                if (nj < s->Bs[kk] / s->Ba[kk] * ni * f) {
                    set_LTE_line(h, s, p, kk);
                    mod[s->irad[kk]] = mod[s->jrad[kk]] = 1;
                    inv = 1;
                }
            }
            for (uint32_t kk = 0; kk < s->nrad; kk++) {
                const ld ni = s->pop[(size_t)p * n + s->irad[kk]];
                const ld nj = s->pop[(size_t)p * n + s->jrad[kk]];
                if (nj < s->Bs[kk] / s->Ba[kk] * ni * f) {
                    set_LTE_line(h, s, p, kk);
                    mod[s->irad[kk]] = mod[s->jrad[kk]] = 1;
                    inv = 1;
                }
            }
            if (inv) set_LTE_levels(h, s, p, mod);
            next = 0;
            for (uint32_t i = 0; i < n; i++) next += mod[i] ? 1 : 0;
        } while (next != curr);
    }
    free(mod);
}

static void check_for_convergence(const oracle_nlte *h, ospec *s)
{
    const double weight = 1.0 / (double)(h->P * s->nlev);
    double fnc = 0.0, rcm = 0.0, rcmax = 0.0;
    for (uint32_t p = 0; p < h->P; p++)
        for (uint32_t i = 0; i < s->nlev; i++) {
            const size_t ind = (size_t)p * s->nlev + i;
            if (s->pop[ind] > h->prm.min_rel_pop_for_convergence * s->pop_tot[p]) {
                const double rc =
                    (double)(2.0L * fabsl((s->pop[ind] - s->prev1[ind]) / (s->pop[ind] + s->prev1[ind])));
                if (rc > h->prm.pop_prec) fnc += weight;
                rcm += weight * rc;
                rcmax = rc < rcmax ? rcmax : rc;
            }
        }
    s->fnc = fnc;
    s->rcm = rcm;
    s->rcmax = rcmax;
}

static void check_for_convergence_trial(const oracle_nlte *h, ospec *s)
{
    const double weight = 1.0 / (double)(h->P * s->nlev);
    double fnc = 0.0, rcm = 0.0, rcmax = 0.0;
    for (size_t ind = 0; ind < (size_t)h->P * s->nlev; ind++) {
        const double rc = (double)(2.0L * fabsl((s->trial[ind] - s->trial_prev[ind])
                                                / (s->trial[ind] + s->trial_prev[ind])));
        if (rc > h->prm.pop_prec) fnc += weight;
        rcm += weight * rc;
        rcmax = rc < rcmax ? rcmax : rc;
    }
    s->fnc = fnc;
    s->rcm = rcm;
    s->rcmax = rcmax;
}

/* push the current populations and their residual (start of an update) */
static int push_history(ospec *s, size_t n)
{
    ld *r = vec_new(n);
    if (!r) return -1;
    const ld *back = s->pops[s->npops - 1];
    for (size_t i = 0; i < n; i++) r[i] = s->pop[i] - back[i];
    if (push(&s->res, &s->nres, &s->cap_res, r)) return -1;
    return push(&s->pops, &s->npops, &s->cap_pops, vec_copy(s->pop, n));
}

int oracle_nlte_stateq(oracle_nlte *h, const double *lambda_diag, double *eps, double *kap,
                       double *ic)
{
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        const size_t n = (size_t)h->P * s->nlev;
        if (s->npops == 0) return -2; /* no populations set */
        memcpy(s->prev1, s->pop, n * sizeof(ld));
        if (push_history(s, n)) return -1;
        const uint32_t nl = s->nlev, mc = max_ncol(s);
        ld *M = malloc((size_t)nl * nl * sizeof(ld));
        ld *y = malloc(nl * sizeof(ld)), *x = malloc(nl * sizeof(ld));
        double *Cd = malloc(mc * sizeof(double)), *Ce = malloc(mc * sizeof(double));
        int err = (!M || !y || !x || !Cd || !Ce) ? -1 : 0;
        /* point by point in place, as the reference: the Lambda term of point p
           reads get_opacity of p's populations before p is solved */
        for (uint32_t p = 0; p < h->P && !err; p++) {
            assemble_point(h, s, lambda_diag, p, M, y, Cd, Ce);
            err = oracle_cpqr_solve_ld((int)nl, M, y, x);
            for (uint32_t i = 0; i < nl; i++) s->pop[(size_t)p * nl + i] = x[i];
        }
        free(M); free(y); free(x); free(Cd); free(Ce);
        if (err) return err;
        correct_negative_populations(h, s);
        check_for_convergence(h, s);
    }
    set_emissivity_and_opacity(h, eps, kap, ic);
    return 0;
}

/* Lines::iteration_using_statistical_equilibrium (lines.cpp:124-135) with the global
   update_using_statistical_equilibrium: one long-double solve of all points' rate
   equations (the reference: SparseLU; here the column-pivoted QR of the dense matrix, for
   the small models of the tests), then the negative-population correction and the
   convergence statistics */
int oracle_nlte_stateq_global(oracle_nlte *h, uint32_t cap, const uint32_t *lam_count,
                              const uint32_t *lam_nr, const double *lam_val, double *eps,
                              double *kap, double *ic)
{
    const lamview lv = {cap, lam_count, lam_nr, lam_val};
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        const size_t N = (size_t)h->P * s->nlev;
        if (s->npops == 0) return -2;
        memcpy(s->prev1, s->pop, N * sizeof(ld));
        if (push_history(s, N)) return -1;
        const uint32_t mc = max_ncol(s);
        ld *M = malloc(N * N * sizeof(ld)), *y = malloc(N * sizeof(ld)), *x = malloc(N * sizeof(ld));
        double *Cd = malloc(mc * sizeof(double)), *Ce = malloc(mc * sizeof(double));
        int err = (!M || !y || !x || !Cd || !Ce) ? -1 : 0;
        if (!err) {
            assemble_global(h, s, &lv, M, y, Cd, Ce);
            err = oracle_cpqr_solve_ld((int)N, M, y, x);
            if (!err) memcpy(s->pop, x, N * sizeof(ld));
        }
        free(M); free(y); free(x); free(Cd); free(Ce);
        if (err) return err;
        correct_negative_populations(h, s);
        check_for_convergence(h, s);
    }
    set_emissivity_and_opacity(h, eps, kap, ic);
    return 0;
}

/* the global system of species l for tests: M ((P nlev)^2, column-major) and y */
int oracle_nlte_stateq_global_system(const oracle_nlte *h, uint32_t cap,
                                     const uint32_t *lam_count, const uint32_t *lam_nr,
                                     const double *lam_val, uint32_t l, double *M_out,
                                     double *y_out)
{
    const lamview lv = {cap, lam_count, lam_nr, lam_val};
    const ospec *s = &h->s[l];
    const size_t N = (size_t)h->P * s->nlev;
    const uint32_t mc = max_ncol(s);
    ld *M = malloc(N * N * sizeof(ld)), *y = malloc(N * sizeof(ld));
    double *Cd = malloc(mc * sizeof(double)), *Ce = malloc(mc * sizeof(double));
    int err = (!M || !y || !Cd || !Ce) ? -1 : 0;
    if (!err) {
        assemble_global(h, s, &lv, M, y, Cd, Ce);
        for (size_t i = 0; i < N * N; i++) M_out[i] = (double)M[i];
        for (size_t i = 0; i < N; i++) y_out[i] = (double)y[i];
    }
    free(M); free(y); free(Cd); free(Ce);
    return err;
}

static int ng_coefficients(const ospec *s, size_t n, uint32_t order, ld *coef)
{
    ld *RTR = malloc((size_t)order * order * sizeof(ld));
    ld *ones = calloc(order ? order : 1, sizeof(ld));
    if (!RTR || !ones) { free(RTR); free(ones); return -1; }
    const size_t rs = s->nres;
    for (uint32_t i = 0; i < order; i++)
        for (uint32_t j = 0; j < order; j++) {
            const ld *a = s->res[rs - order + i], *b = s->res[rs - order + j];
            ld d = 0.0L;
            for (size_t t = 0; t < n; t++) d += a[t] * b[t];
            RTR[(size_t)j * order + i] = d;
        }
    for (uint32_t i = 0; i < order; i++) ones[i] = 1.0L;
    int e = oracle_cpqr_solve_ld((int)order, RTR, ones, coef);
    ld sum = 0.0L;
    for (uint32_t i = 0; i < order; i++) sum += coef[i];
    for (uint32_t i = 0; i < order; i++) coef[i] /= sum;
    free(RTR); free(ones);
    return e;
}

int oracle_nlte_ng(oracle_nlte *h, uint32_t order, double *eps, double *kap, double *ic)
{
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        const size_t n = (size_t)h->P * s->nlev;
        memcpy(s->prev1, s->pop, n * sizeof(ld));
        if (push_history(s, n)) return -1;
        if (s->nres < order || order == 0) return -3;
        ld *coef = malloc(order * sizeof(ld));
        if (ng_coefficients(s, n, order, coef)) { free(coef); return -1; }
        const size_t ps = s->npops;
        for (size_t t = 0; t < n; t++) s->pop[t] = 0.0L;
        for (uint32_t i = 0; i < order; i++) {
            const ld *v = s->pops[ps - 1 - i];
            const ld c = coef[order - 1 - i];
            for (size_t t = 0; t < n; t++) s->pop[t] += v[t] * c;
        }
        free(coef);
        /* residuals.clear(); populations.erase(begin, end - 1) */
        for (size_t i = 0; i < s->nres; i++) free(s->res[i]);
        s->nres = 0;
        for (size_t i = 0; i + 1 < s->npops; i++) free(s->pops[i]);
        s->pops[0] = s->pops[s->npops - 1];
        s->npops = 1;
        correct_negative_populations(h, s);
        check_for_convergence(h, s);
    }
    set_emissivity_and_opacity(h, eps, kap, ic);
    return 0;
}

int oracle_nlte_trial_ng(oracle_nlte *h, uint32_t order)
{
    for (uint32_t l = 0; l < h->nlspecs; l++) {
        ospec *s = &h->s[l];
        const size_t n = (size_t)h->P * s->nlev;
        if (push_history(s, n)) return -1;
        memcpy(s->trial_prev, s->trial, n * sizeof(ld));
        if (s->npops < order || order == 0 || s->nres < order) return -3;
        ld *coef = malloc(order * sizeof(ld));
        if (ng_coefficients(s, n, order, coef)) { free(coef); return -1; }
        const size_t ps = s->npops;
        for (size_t t = 0; t < n; t++) s->trial[t] = 0.0L;
        for (uint32_t i = 0; i < order; i++) {
            const ld *v = s->pops[ps - 1 - i];
            const ld c = coef[order - 1 - i];
            for (size_t t = 0; t < n; t++) s->trial[t] += v[t] * c;
        }
        free(coef);
        free(s->res[--s->nres]);
        free(s->pops[--s->npops]);
        check_for_convergence_trial(h, s);
    }
    return 0;
}
