/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of Magritte's plain-method Feautrier formal solver,
 * used by tests/ (parity checks), __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.  Nothing in the product (magritte_amd/) links or calls it.
 *
 * It follows the reference function by function and keeps its operation
 * order; it is built with -ffp-contract=off (no FMA), like the reference's gcc
 * Release build.  The float transcendentals are the correctly rounded
 * double-evaluated ones of mgx_fmath.h (shared with the kernels so both
 * produce the same bits); building with -DMGX_ORACLE_GLIBC swaps in the C
 * library's expf/expm1f/logf/powf/erff, which the reference calls, to measure
 * the effect of that difference.
 *
 * Reference files restated (paths relative to the reference root):
 *   src/solver/solver.tpp             setup :1-13,32-77; get_ray_length :80-103;
 *                                     get_ray_lengths_max :106-126;
 *                                     solve_feautrier_order_2(Model&) :459-540;
 *                                     image_feautrier_order_2_new_imager :577-627;
 *                                     trace_ray_for_line :774-800;
 *                                     trace_ray_imaging_get_start :808-830;
 *                                     trace_ray_imaging :842-897;
 *                                     get_ray_length_new_imager :964-989;
 *                                     set_data :1003-1043; set_data_for_line :1057-1097;
 *                                     gaussian :1106-1110; planck :1119-1121;
 *                                     boundary_intensity :1130-1142;
 *                                     get_eta_and_chi<*> :1152-1351;
 *                                     update_Lambda :1761-1850;
 *                                     compute_S_dtau_line_integrated<*> :1879-2111;
 *                                     compute_source_dtau :2140-2186;
 *                                     solve_feautrier_order_2(o,f) :2194-2356;
 *                                     compute_dtau_single_line :2374-2455;
 *                                     image_feautrier_order_2(o,f) :2463-2574
 *   src/solver/solver_interp_helper.tpp get_n_interp :10-46;
 *                                     get_n_interp_for_line :55-75;
 *                                     interpolate_linear :78-86; interpolate_log :89-99
 *   src/model/geometry/geometry.tpp   get_next_general_geometry<*> :14-106;
 *                                     get_next_spherical_symmetry :118-159;
 *                                     get_distance_origin_to_boundary :167-193;
 *                                     get_shift_general_geometry :204-221;
 *                                     get_boundary_point_closer_to_custom_ray :310-341;
 *                                     get_closest_bdy_point_in_custom_raydir :346-367;
 *                                     get_shift_spherical_symmetry :379-415
 *   src/model/thermodynamics/thermodynamics.tpp profile :8-19, widths :27-53
 *   src/model/image/image.cpp         set_coordinates_projection_surface :221-366;
 *                                     surface_coords_to_3D_coordinates :370-398
 *   src/tools/constants.hpp           physical and numerical constants
 */
#include "mgx_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "mgx_fmath.h"
#include "mgx_fmath_ref.h"

/* The float transcendentals, by build:
 *   MGX_ORACLE_GLIBC  the C library's expf, expm1f, logf, powf, erff: the
 *                     reference's own calls (liboracle_glibc.so);
 *   MGX_ORACLE_CR     the correctly rounded functions of mgx_fmath.h, kept to
 *                     measure how much the results depend on these last bits
 *                     (liboracle_cr.so);
 *   neither           the restatements of the C library's functions in
 *                     mgx_fmath_ref.h that the kernels evaluate (liboracle.so). */
#if defined(MGX_ORACLE_GLIBC)
#define O_EXPF expf
#define O_EXPM1F expm1f
#define O_LOGF logf
#define O_POWF powf
#define O_ERFF erff
#elif defined(MGX_ORACLE_CR)
#define O_EXPF mgx_expf
#define O_EXPM1F mgx_expm1f
#define O_LOGF mgx_logf
#define O_POWF mgx_powf
#define O_ERFF mgx_erff
#else
#define O_EXPF mgx_ref_expf
#define O_EXPM1F mgx_ref_expm1f
#define O_LOGF mgx_ref_logf
#define O_POWF mgx_ref_powf
#define O_ERFF mgx_ref_erff
#endif

/* src/tools/constants.hpp */
static const double C_INVERSE_SQRT_PI = 0.5 * M_2_SQRTPI;
static const double C_KB = 1.38064852E-23;
static const double C_HH = 6.62607004E-34;
static const double C_AMU = 1.66053904E-27;
static const double C_T_CMB = 2.7254800;
static const double C_CC_SQUARED = 8.98755179E+16;
#define C_TWO_KB_OVER_AMU_CC_SQUARED (2.0 * C_KB / (C_AMU * C_CC_SQUARED))
#define C_TWO_HH_OVER_CC_SQUARED (2.0 * C_HH / C_CC_SQUARED)
#define C_HH_OVER_KB (C_HH / C_KB)

/* ------------------------------------------------------------------------- */
/* small vector helpers (Paracabs Vector3D: dot = x*x' + y*y' + z*z')         */

static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline const double *ray_dir(const mgx_model *m, uint32_t o, uint32_t r)
{
    /* Rays::get_direction<adaptive> (rays.cpp:94-118) */
    return m->adaptive_directions ? &m->ray_direction[3 * ((size_t)o * m->nrays + r)]
                                  : &m->ray_direction[3 * (size_t)r];
}

static inline double ray_weight(const mgx_model *m, uint32_t o, uint32_t r)
{
    return m->adaptive_directions ? m->ray_weight[(size_t)o * m->nrays + r] : m->ray_weight[r];
}

static inline int not_on_boundary(const mgx_model *m, uint32_t p)
{
    return m->point2boundary[p] == m->npoints;
}

/* ------------------------------------------------------------------------- */
/* geometry (geometry.tpp)                                                   */

/* get_next_general_geometry<Defaulttracer|Imagetracer> :14-44, :56-106 */
static uint32_t get_next_general(const mgx_model *m, const double *origin, const double *raydir,
                                 uint32_t c, double *Z, double *dZ)
{
    const uint32_t n_nbs = m->n_neighbors[c];
    const uint32_t cum = m->cum_n_neighbors[c];
    double dmin = DBL_MAX;
    uint32_t next = m->npoints;
    for (uint32_t i = 0; i < n_nbs; i++) {
        const uint32_t n = m->neighbors[cum + i];
        double R[3];
        R[0] = m->position[3 * (size_t)n + 0] - origin[0];
        R[1] = m->position[3 * (size_t)n + 1] - origin[1];
        R[2] = m->position[3 * (size_t)n + 2] - origin[2];
        const double Z_new = dot3(R, raydir);
        if (Z_new > *Z) {
            const double d2 = dot3(R, R) - Z_new * Z_new;
            if (d2 < dmin) {
                dmin = d2;
                next = n;
                *dZ = Z_new - *Z;
            }
        }
    }
    *Z += *dZ;
    return next;
}

/* get_next_spherical_symmetry :118-159 */
static uint32_t get_next_spherical(const mgx_model *m, const double *origin, const double *raydir,
                                   uint32_t c, double *Z, double *dZ)
{
    uint32_t next;
    const double Rcos = dot3(origin, raydir);
    const double R2 = dot3(origin, origin);
    const double Rsin2 = R2 - Rcos * Rcos;
    const double Rcos_plus_Z = Rcos + *Z;
    if (*Z < -Rcos) {
        if (c <= 0) return m->npoints;
        const double *pm = &m->position[3 * (size_t)(c - 1)];
        if (dot3(pm, pm) >= Rsin2) {
            next = c - 1;
            const double *pn = &m->position[3 * (size_t)next];
            *dZ = -sqrt(dot3(pn, pn) - Rsin2) - Rcos_plus_Z;
        } else {
            next = c;
            *dZ = -2.0 * Rcos_plus_Z;
        }
    } else {
        if (c >= m->npoints - 1) return m->npoints;
        next = c + 1;
        const double *pn = &m->position[3 * (size_t)next];
        *dZ = +sqrt(dot3(pn, pn) - Rsin2) - Rcos_plus_Z;
    }
    *Z += *dZ;
    return next;
}

static uint32_t get_next_xyz(const mgx_model *m, const double *origin, const double *raydir,
                             uint32_t c, double *Z, double *dZ)
{
    return m->spherical_symmetry ? get_next_spherical(m, origin, raydir, c, Z, dZ)
                                 : get_next_general(m, origin, raydir, c, Z, dZ);
}

/* Geometry::get_next<adaptive>(o, r, crt, Z, dZ) :247-258 */
static uint32_t get_next(const mgx_model *m, uint32_t o, uint32_t r, uint32_t c, double *Z,
                         double *dZ)
{
    return get_next_xyz(m, &m->position[3 * (size_t)o], ray_dir(m, o, r), c, Z, dZ);
}

/* get_shift<CoMoving>(origin, origin_velocity, raydir, c, Z, reverse=false) :204-207, :379-396 */
static double shift_comoving(const mgx_model *m, const double *origin, const double *ovel,
                             const double *raydir, uint32_t c, double Z)
{
    if (m->spherical_symmetry) {
        if (dot3(origin, origin) == 0.0) return 1.0;
        const double *pc = &m->position[3 * (size_t)c];
        if (pc[0] == 0.0) return 1.0 - dot3(ovel, raydir);
        const double Rcos_plus_Z = dot3(origin, raydir) + Z;
        return 1.0 + (m->velocity[3 * (size_t)c] * Rcos_plus_Z / pc[0] - dot3(ovel, raydir));
    }
    double dv[3];
    dv[0] = m->velocity[3 * (size_t)c + 0] - ovel[0];
    dv[1] = m->velocity[3 * (size_t)c + 1] - ovel[1];
    dv[2] = m->velocity[3 * (size_t)c + 2] - ovel[2];
    return 1.0 + dot3(dv, raydir);
}

/* get_shift<Rest>(origin, 0, raydir, c, Z, false) :213-221, :405-415 */
static double shift_rest(const mgx_model *m, const double *origin, const double *raydir,
                         uint32_t c, double Z)
{
    if (m->spherical_symmetry) {
        const double *pc = &m->position[3 * (size_t)c];
        if (pc[0] == 0.0) return 1.0;
        const double Rcos_plus_Z = dot3(origin, raydir) + Z;
        return 1.0 + m->velocity[3 * (size_t)c] * Rcos_plus_Z / pc[0];
    }
    return 1.0 + dot3(&m->velocity[3 * (size_t)c], raydir);
}

/* Geometry::get_shift<CoMoving,adaptive>(o, r, c, Z) :429-438 */
static double get_shift(const mgx_model *m, uint32_t o, uint32_t r, uint32_t c, double Z)
{
    return shift_comoving(m, &m->position[3 * (size_t)o], &m->velocity[3 * (size_t)o],
                          ray_dir(m, o, r), c, Z);
}

/* ------------------------------------------------------------------------- */
/* interpolation helper (solver_interp_helper.tpp)                           */

static inline double crit(const mgx_model *m, uint32_t p, uint32_t l)
{
    return m->interp_crit[(size_t)p * m->nlines + l];
}

/* get_n_interp_for_line :55-75 */
static uint32_t get_n_interp_for_line(const mgx_model *m, const mgx_options *op, uint32_t l,
                                      uint32_t c, uint32_t n)
{
    const double curr_source = crit(m, c, l);
    const double next_source = crit(m, n, l);
    const double source_diff = (double)fabsf(O_LOGF((float)(next_source / curr_source)));
    const uint32_t k =
        (uint32_t)ceil(source_diff / (double)O_LOGF((float)op->max_interpolation_diff));
    return (k > 1) ? k : 1;
}

/* get_n_interp :10-46 (largest over all lines) */
static uint32_t get_n_interp(const mgx_model *m, const mgx_options *op, uint32_t c, uint32_t n)
{
    double mx = 0.0;
    for (uint32_t l = 0; l < m->nlines; l++) {
        const double d = (double)fabsf(O_LOGF((float)(crit(m, n, l) / crit(m, c, l))));
        if (l == 0 || mx < d) mx = d;
    }
    const uint32_t k = (uint32_t)ceil(mx / (double)O_LOGF((float)op->max_interpolation_diff));
    return (k > 1) ? k : 1;
}

/* interpolate_linear :78-86 */
static inline double interp_linear(double a, double b, double t)
{
    if (t == 0.0) return a;
    if (t == 1.0) return b;
    return (b - a) * t + a;
}

/* Finite-difference aid of the adjoint checks: with it set, interpolate_log is
   evaluated as a * pow(b/a, t) in double, so the function is smooth in a and b
   (the float powf of the reference rounds b/a and the result to float, which
   puts steps of ~1e-7 relative into a finite difference).  Off by default. */
static int g_smooth_interp = 0;

/* interpolate_log :89-99 */
static inline double interp_log(double a, double b, double t)
{
    if (t == 0.0) return a;
    if (t == 1.0) return b;
    if (g_smooth_interp) return a * pow(b / a, t);
    return a * (double)O_POWF((float)(b / a), (float)t);
}

/* partial derivatives of interpolate_log(a, b, t) = a^(1-t) b^t (r its value) */
static inline void interp_log_partials(double a, double b, double t, double r, double *da,
                                       double *db)
{
    *da = (t == 1.0) ? 0.0 : (1.0 - t) * r / a;
    *db = (t == 0.0) ? 0.0 : t * r / b;
}

/* ------------------------------------------------------------------------- */
/* per-thread scratch (Solver::setup :32-77)                                  */

typedef struct {
    uint32_t length, centre;
    uint32_t first, last, n_tot;
    uint32_t *nr, *nr_interp;
    double *dZ, *shift, *interp_factor;
    double *Su, *A, *C, *inverse_A, *inverse_C, *FF, *FI, *GG, *GP, *L_diag;
    /* off-diagonal Lambda bands (n_off_diag > 0): GI (length) and L_upper,
       L_lower (n_off_diag x length, row-major) */
    uint32_t noff;
    double *GI, *L_upper, *L_lower;
} Scratch;

static int scratch_init(Scratch *s, uint32_t length)
{
    memset(s, 0, sizeof *s);
    s->length = length;
    s->centre = length / 2;
    s->nr = calloc(length, sizeof(uint32_t));
    s->nr_interp = calloc(length, sizeof(uint32_t));
    double **d[] = {&s->dZ, &s->shift, &s->interp_factor, &s->Su, &s->A, &s->C, &s->inverse_A,
                    &s->inverse_C, &s->FF, &s->FI, &s->GG, &s->GP, &s->L_diag};
    int ok = s->nr && s->nr_interp;
    for (size_t i = 0; i < sizeof d / sizeof d[0]; i++) {
        *d[i] = calloc(length, sizeof(double));
        ok = ok && *d[i];
    }
    return ok ? 0 : -1;
}

/* the band arrays of Solver::setup (:50-56) for n_off_diag > 0 */
static int scratch_bands(Scratch *s, uint32_t noff)
{
    s->noff = noff;
    s->GI = calloc(s->length, sizeof(double));
    s->L_upper = calloc((size_t)noff * s->length, sizeof(double));
    s->L_lower = calloc((size_t)noff * s->length, sizeof(double));
    return (s->GI && s->L_upper && s->L_lower) ? 0 : -1;
}

static void scratch_free(Scratch *s)
{
    free(s->GI);
    free(s->L_upper);
    free(s->L_lower);
    free(s->nr);
    free(s->nr_interp);
    double *d[] = {s->dZ, s->shift, s->interp_factor, s->Su, s->A, s->C, s->inverse_A,
                   s->inverse_C, s->FF, s->FI, s->GG, s->GP, s->L_diag};
    for (size_t i = 0; i < sizeof d / sizeof d[0]; i++) free(d[i]);
}

/* ------------------------------------------------------------------------- */
/* tracing (solver.tpp)                                                      */

/* set_data_for_line :1057-1097 (n_interp given) */
static void set_data_n(Scratch *s, uint32_t n_interp, uint32_t crt, uint32_t nxt, double shift_crt,
                       double shift_nxt, double dZ_loc, int increment, uint32_t *id1,
                       uint32_t *id2)
{
    const double dshift = shift_nxt - shift_crt;
    if (n_interp > 1) {
        const double dZ_interpl = dZ_loc / n_interp;
        const double dshift_interpl = dshift / n_interp;
        for (uint32_t i = 1; i <= n_interp; i++) {
            s->nr[*id1] = nxt;
            s->shift[*id1] = shift_crt + i * dshift_interpl;
            s->nr_interp[*id1] = crt;
            s->interp_factor[*id1] = 1.0 - ((double)i) / n_interp;
            s->dZ[*id2] = dZ_interpl;
            *id1 += increment;
            *id2 += increment;
        }
    } else {
        s->nr[*id1] = nxt;
        s->shift[*id1] = shift_nxt;
        s->nr_interp[*id1] = crt;
        s->interp_factor[*id1] = 0.0;
        s->dZ[*id2] = dZ_loc;
        *id1 += increment;
        *id2 += increment;
    }
}

/* trace_ray_for_line :774-800.  Returns id1; *err set on an invalid next point. */
static uint32_t trace_ray_for_line(const mgx_model *m, const mgx_options *op, Scratch *s,
                                   uint32_t l, uint32_t o, uint32_t r, int increment,
                                   uint32_t id1, uint32_t id2, int *err)
{
    double Z = 0.0, dZ = 0.0;
    uint32_t nxt = get_next(m, o, r, o, &Z, &dZ);
    if (nxt < m->npoints) {
        uint32_t crt = o;
        double shift_crt = get_shift(m, o, r, crt, 0.0);
        double shift_nxt = get_shift(m, o, r, nxt, Z);
        set_data_n(s, get_n_interp_for_line(m, op, l, crt, nxt), crt, nxt, shift_crt, shift_nxt,
                   dZ, increment, &id1, &id2);
        while (not_on_boundary(m, nxt)) {
            crt = nxt;
            shift_crt = shift_nxt;
            nxt = get_next(m, o, r, nxt, &Z, &dZ);
            if (nxt >= m->npoints) { /* reference would read out of bounds */
                *err = MGX_E_NO_NEXT;
                return id1;
            }
            shift_nxt = get_shift(m, o, r, nxt, Z);
            set_data_n(s, get_n_interp_for_line(m, op, l, crt, nxt), crt, nxt, shift_crt,
                       shift_nxt, dZ, increment, &id1, &id2);
        }
    }
    return id1;
}

/* get_ray_length :80-103 (all-lines interpolation count) */
static uint32_t get_ray_length(const mgx_model *m, const mgx_options *op, uint32_t o, uint32_t r,
                               int *err)
{
    uint32_t len = 0;
    double Z = 0.0, dZ = 0.0;
    uint32_t nxt = get_next(m, o, r, o, &Z, &dZ);
    if (nxt < m->npoints) {
        uint32_t crt = o;
        len += get_n_interp(m, op, crt, nxt);
        while (not_on_boundary(m, nxt)) {
            crt = nxt;
            nxt = get_next(m, o, r, nxt, &Z, &dZ);
            if (nxt >= m->npoints) {
                *err = MGX_E_NO_NEXT;
                return len;
            }
            len += get_n_interp(m, op, crt, nxt);
        }
    }
    return len;
}

/* get_ray_lengths_max :106-126 */
static uint32_t get_ray_lengths_max(const mgx_model *m, const mgx_options *op, int *err)
{
    uint32_t mx = 0;
    int e = 0;
    for (uint32_t rr = 0; rr < m->hnrays; rr++) {
        const uint32_t ar = m->antipod[rr];
#pragma omp parallel for schedule(static) reduction(max : mx) reduction(min : e)
        for (uint32_t o = 0; o < m->npoints; o++) {
            int le = 0;
            const uint32_t len = get_ray_length(m, op, o, rr, &le) + get_ray_length(m, op, o, ar, &le);
            if (len > mx) mx = len;
            if (le < e) e = le;
        }
    }
    if (e) *err = e;
    return mx;
}

/* scratch length for the one ray pair (rr, antipod) from origin o: the all-lines
   interpolation counts of its two halves bound every line's, so a scratch centred
   at the longer half holds the pair (the single-ray functions below use it in
   place of the maximum over all rays and origins, which costs a walk of every
   ray of the model) */
static uint32_t pair_length(const mgx_model *m, const mgx_options *op, uint32_t o, uint32_t rr,
                            int *err)
{
    int le = 0;
    const uint32_t a = get_ray_length(m, op, o, rr, &le);
    const uint32_t b = get_ray_length(m, op, o, m->antipod[rr], &le);
    if (le) *err = le;
    return 2 * (a > b ? a : b) + 1;
}

/* ------------------------------------------------------------------------- */
/* line profile, Planck, boundary (solver.tpp :1106-1142, thermodynamics.tpp) */

/* Finite-difference aid of the adjoint checks with respect to the inverse line
   widths: with it set, the line profile and the error-function difference of
   the line-integrated optical depth are evaluated in double (the reference's
   float expf and erff put steps of ~6e-8 relative into a difference quotient).
   Off by default. */
static int g_smooth_profile = 0;

static inline double gaussian(double inverse_width, double diff)
{
    const double sqrt_exp = inverse_width * diff;
    if (g_smooth_profile) return inverse_width * C_INVERSE_SQRT_PI * exp(-sqrt_exp * sqrt_exp);
    return inverse_width * C_INVERSE_SQRT_PI * (double)O_EXPF((float)(-sqrt_exp * sqrt_exp));
}

/* erf(next_pos) - erf(curr_pos) of compute_dtau_single_line: the difference of
   two float erff values, rounded to float, as the reference takes it */
static inline double erf_difference(double next_pos, double curr_pos)
{
    if (g_smooth_profile) {
        /* in a wing both values are near +-1: take the difference of the
           complementary functions there, which keeps its relative accuracy */
        if (next_pos >= 0.0 && curr_pos >= 0.0) return erfc(curr_pos) - erfc(next_pos);
        if (next_pos <= 0.0 && curr_pos <= 0.0) return erfc(-next_pos) - erfc(-curr_pos);
        return erf(next_pos) - erf(curr_pos);
    }
    const float d = O_ERFF((float)next_pos) - O_ERFF((float)curr_pos);
    return (double)d;
}

static inline double planck(double temp, double freq)
{
    return C_TWO_HH_OVER_CC_SQUARED * (freq * freq * freq) /
           (double)O_EXPM1F((float)(C_HH_OVER_KB * freq / temp));
}

static inline double boundary_intensity(const mgx_model *m, uint32_t p, double freq)
{
    const uint32_t b = m->point2boundary[p];
    /* an interior point here is undefined behaviour in the reference; use the default */
    const int32_t bc = (b < m->nboundary) ? m->boundary_condition[b] : MGX_BC_CMB;
    switch (bc) {
    case MGX_BC_ZERO:
        return 0.0;
    case MGX_BC_THERMAL:
        return planck(m->boundary_temperature[b], freq);
    default:
        return planck(C_T_CMB, freq);
    }
}

/* Thermodynamics::profile_width_upper_bound_with_linefreq :42-53 */
static inline double pw_upper(const mgx_model *m, uint32_t p, double freq_line)
{
    return freq_line * sqrt(C_TWO_KB_OVER_AMU_CC_SQUARED * m->max_inverse_mass *
                                m->temperature[p] +
                            m->vturb2[p]);
}

/* Thermodynamics::profile(inverse_mass, p, freq_line, freq) :8-35 */
static inline double thermo_profile(const mgx_model *m, double inverse_mass, uint32_t p,
                                    double freq_line, double freq)
{
    const double width =
        freq_line * sqrt(C_TWO_KB_OVER_AMU_CC_SQUARED * inverse_mass * m->temperature[p] +
                         m->vturb2[p]);
    const double inverse_width = 1.0 / width;
    const double sqrtExponent = inverse_width * (freq - freq_line);
    const double exponent = sqrtExponent * sqrtExponent;
    return inverse_width * C_INVERSE_SQRT_PI * (double)O_EXPF((float)(-exponent));
}

/* std::lower_bound / std::upper_bound on sorted_line */
static inline uint32_t lower_bound_d(const double *a, uint32_t n, double v)
{
    uint32_t lo = 0, len = n;
    while (len > 0) {
        const uint32_t h = len / 2;
        if (a[lo + h] < v) {
            lo = lo + h + 1;
            len = len - h - 1;
        } else {
            len = h;
        }
    }
    return lo;
}

static inline uint32_t upper_bound_d(const double *a, uint32_t n, double v)
{
    uint32_t lo = 0, len = n;
    while (len > 0) {
        const uint32_t h = len / 2;
        if (!(v < a[lo + h])) {
            lo = lo + h + 1;
            len = len - h - 1;
        } else {
            len = h;
        }
    }
    return lo;
}

#define LIDX(m, p, l) ((size_t)(p) * (m)->nlines + (l))

/* get_eta_and_chi<approx>(p, ll, freq) :1152-1166, :1214-1240, :1335-1343 */
static void get_eta_and_chi(const mgx_model *m, const mgx_options *op, uint32_t p, uint32_t ll,
                            double freq, double *eta, double *chi)
{
    if (op->approx == MGX_APPROX_ONE_LINE) {
        const double diff = freq - m->line[ll];
        const double prof = gaussian(m->inverse_width[LIDX(m, p, ll)], diff);
        *eta = prof * m->emissivity[LIDX(m, p, ll)];
        *chi = prof * m->opacity[LIDX(m, p, ll)] + op->min_opacity;
        return;
    }
    *eta = 0.0;
    *chi = op->min_opacity;
    if (op->approx == MGX_APPROX_NONE) {
        for (uint32_t l = 0; l < m->nlines; l++) {
            const double diff = freq - m->line[l];
            const double prof = gaussian(m->inverse_width[LIDX(m, p, l)], diff);
            *eta += prof * m->emissivity[LIDX(m, p, l)];
            *chi += prof * m->opacity[LIDX(m, p, l)];
        }
        return;
    }
    const double W = op->max_distance_opacity_contribution * pw_upper(m, p, freq);
    const uint32_t lo = lower_bound_d(m->sorted_line, m->nlines, freq - W);
    const uint32_t hi = upper_bound_d(m->sorted_line, m->nlines, freq + W);
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = m->sorted_line_map[s];
        const double diff = freq - m->sorted_line[s];
        const double prof = gaussian(m->inverse_width[LIDX(m, p, l)], diff);
        *eta += prof * m->emissivity[LIDX(m, p, l)];
        *chi += prof * m->opacity[LIDX(m, p, l)];
    }
}

/* get_eta_and_chi_interpolated<approx>(p1, p2, t, ll, freq) :1177-1202, :1260-1302, :1345-1351 */
static void get_eta_and_chi_interp(const mgx_model *m, const mgx_options *op, uint32_t p1,
                                   uint32_t p2, double t, uint32_t ll, double freq, double *eta,
                                   double *chi)
{
    if (op->approx == MGX_APPROX_ONE_LINE) {
        const double diff = freq - m->line[ll];
        const double iw = interp_linear(m->inverse_width[LIDX(m, p1, ll)],
                                        m->inverse_width[LIDX(m, p2, ll)], t);
        const double prof = gaussian(iw, diff);
        const double e = interp_log(m->emissivity[LIDX(m, p1, ll)], m->emissivity[LIDX(m, p2, ll)], t);
        const double k = interp_log(m->opacity[LIDX(m, p1, ll)], m->opacity[LIDX(m, p2, ll)], t);
        *eta = prof * e;
        *chi = prof * k + op->min_opacity;
        return;
    }
    *eta = 0.0;
    *chi = op->min_opacity;
    uint32_t lo = 0, hi = m->nlines;
    if (op->approx == MGX_APPROX_CLOSE_LINES) {
        const double W = op->max_distance_opacity_contribution * pw_upper(m, p1, freq);
        lo = lower_bound_d(m->sorted_line, m->nlines, freq - W);
        hi = upper_bound_d(m->sorted_line, m->nlines, freq + W);
    }
    for (uint32_t s = lo; s < hi; s++) {
        uint32_t l;
        double diff;
        if (op->approx == MGX_APPROX_CLOSE_LINES) {
            l = m->sorted_line_map[s];
            diff = freq - m->sorted_line[s];
        } else {
            l = s;
            diff = freq - m->line[l];
        }
        const double iw =
            interp_linear(m->inverse_width[LIDX(m, p1, l)], m->inverse_width[LIDX(m, p2, l)], t);
        const double prof = gaussian(iw, diff);
        const double e = interp_log(m->emissivity[LIDX(m, p1, l)], m->emissivity[LIDX(m, p2, l)], t);
        const double k = interp_log(m->opacity[LIDX(m, p1, l)], m->opacity[LIDX(m, p2, l)], t);
        *eta += prof * e;
        *chi += prof * k;
    }
}

/* compute_dtau_single_line :2374-2455 */
static double compute_dtau_single_line(const mgx_model *m, const mgx_options *op, uint32_t ci,
                                       uint32_t ni, uint32_t cii, uint32_t nii, uint32_t l,
                                       double curr_freq, double next_freq, double curr_interp,
                                       double next_interp, double dz)
{
    const double linefreq = m->line[l];
    const double iw_c =
        interp_linear(m->inverse_width[LIDX(m, ci, l)], m->inverse_width[LIDX(m, cii, l)], curr_interp);
    const double iw_n =
        interp_linear(m->inverse_width[LIDX(m, ni, l)], m->inverse_width[LIDX(m, nii, l)], next_interp);
    const double avg_iw = (iw_c + iw_n) / 2.0;
    const double op_c = interp_log(m->opacity[LIDX(m, ci, l)], m->opacity[LIDX(m, cii, l)], curr_interp);
    const double op_n = interp_log(m->opacity[LIDX(m, ni, l)], m->opacity[LIDX(m, nii, l)], next_interp);
    if (curr_freq == next_freq) {
        const double diff = curr_freq - m->line[l];
        const double prof = gaussian(avg_iw, diff);
        const double avg_op = (op_c + op_n) / 2.0;
        return dz * (prof * avg_op + op->min_opacity);
    }
    const double next_pos = (linefreq - next_freq) * avg_iw;
    const double curr_pos = (linefreq - curr_freq) * avg_iw;
    const double diff_pos = next_pos - curr_pos;
    const double avg_op = (op_n + op_c) / 2.0;
    const double erf_diff = erf_difference(next_pos, curr_pos);
    const double erfterm = avg_op / diff_pos / 2.0 * erf_diff;
    return dz * (avg_iw * erfterm + op->min_opacity);
}

/* line source of line l at (p, p_interp, t): interpolate_log(eps/kap, ...) */
static inline double line_source(const mgx_model *m, uint32_t p, uint32_t pi, uint32_t l, double t)
{
    const double S_p = m->emissivity[LIDX(m, p, l)] / m->opacity[LIDX(m, p, l)];
    const double S_i = m->emissivity[LIDX(m, pi, l)] / m->opacity[LIDX(m, pi, l)];
    return interp_log(S_p, S_i, t);
}

/* compute_S_dtau_line_integrated<approx> :1879-1912, :1941-1983, :2013-2111 */
static void compute_S_dtau_line_integrated(const mgx_model *m, const mgx_options *op, uint32_t cp,
                                           uint32_t np, uint32_t cpi, uint32_t npi, uint32_t lineidx,
                                           double currfreq, double nextfreq, double curr_interp,
                                           double next_interp, double dZ, double *dtau,
                                           double *Scurr, double *Snext)
{
    if (op->approx == MGX_APPROX_ONE_LINE) {
        *dtau = compute_dtau_single_line(m, op, cp, np, cpi, npi, lineidx, currfreq, nextfreq,
                                         curr_interp, next_interp, dZ);
        *Scurr = line_source(m, cp, cpi, lineidx, curr_interp);
        *Snext = line_source(m, np, npi, lineidx, next_interp);
        return;
    }
    double sum_dtau = 0.0, sum_Sc = 0.0, sum_Sn = 0.0;
    uint32_t lo = 0, hi = m->nlines;
    if (op->approx == MGX_APPROX_CLOSE_LINES) {
        double left_freq, right_freq;
        if (currfreq < nextfreq) {
            left_freq = currfreq;
            right_freq = nextfreq;
        } else {
            right_freq = currfreq;
            left_freq = nextfreq;
        }
        const double cb = op->max_distance_opacity_contribution * pw_upper(m, cp, right_freq);
        const double nb = op->max_distance_opacity_contribution * pw_upper(m, np, right_freq);
        const double W = (cb < nb) ? nb : cb; /* std::max(cb, nb) */
        lo = lower_bound_d(m->sorted_line, m->nlines, left_freq - W);
        hi = upper_bound_d(m->sorted_line, m->nlines, right_freq + W);
    }
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = (op->approx == MGX_APPROX_CLOSE_LINES) ? m->sorted_line_map[s] : s;
        const double line_dtau = compute_dtau_single_line(m, op, cp, np, cpi, npi, l, currfreq,
                                                          nextfreq, curr_interp, next_interp, dZ);
        const double Sc = line_source(m, cp, cpi, l, curr_interp);
        const double Sn = line_source(m, np, npi, l, next_interp);
        sum_dtau += line_dtau;
        sum_Sc += line_dtau * Sc;
        sum_Sn += line_dtau * Sn;
    }
    double d = sum_dtau;
    if (op->approx == MGX_APPROX_CLOSE_LINES) {
        const double bound_min_dtau = op->min_opacity * dZ;
        if (-bound_min_dtau < d) d = (bound_min_dtau < d) ? d : bound_min_dtau;
    }
    *dtau = d;
    *Scurr = sum_Sc / d;
    *Snext = sum_Sn / d;
}

/* compute_source_dtau<approx> :2140-2186 */
static void compute_source_dtau(const mgx_model *m, const mgx_options *op, uint32_t cp,
                                uint32_t np, uint32_t cpi, uint32_t npi, uint32_t line,
                                double curr_freq, double next_freq, double curr_shift,
                                double next_shift, double dZ, double curr_interp,
                                double next_interp, int *compute_curr_opacity, double *dtaunext,
                                double *chicurr, double *chinext, double *Scurr, double *Snext)
{
    const double dshift = next_shift - curr_shift;
    const double dshift_abs = fabs(dshift);
    const double a = m->dshift_max[cp], b = m->dshift_max[np];
    const double dshift_max = (b < a) ? b : a; /* std::min(a, b) */
    if (dshift_abs > dshift_max) {
        *compute_curr_opacity = 1;
        compute_S_dtau_line_integrated(m, op, cp, np, cpi, npi, line, curr_freq, next_freq,
                                       curr_interp, next_interp, dZ, dtaunext, Scurr, Snext);
    } else {
        if (*compute_curr_opacity) {
            *compute_curr_opacity = 0;
            double eta_c = 0.0;
            get_eta_and_chi_interp(m, op, cp, cpi, curr_interp, line, curr_freq, &eta_c, chicurr);
            *Scurr = eta_c / *chicurr;
        }
        double eta_n = 0.0;
        get_eta_and_chi_interp(m, op, np, npi, next_interp, line, next_freq, &eta_n, chinext);
        *Snext = eta_n / *chinext;
        *dtaunext = 0.5 * (*chicurr + *chinext) * dZ;
    }
}

/* ------------------------------------------------------------------------- */
/* Feautrier solve for one (o, f) on the traced ray in s (:2194-2356, n_off_diag = 0) */

static void solve_feautrier_order_2_of(const mgx_model *m, const mgx_options *op, Scratch *s,
                                       uint32_t o, uint32_t f)
{
    const double freq = m->nu[(size_t)o * m->nfreqs + f];
    const uint32_t l = m->f_line[f];
    double dtau_c, term_c, dtau_n, term_n, chi_c = 0.0, chi_n = 0.0;
    const uint32_t first = s->first, last = s->last, centre = s->centre;
    const uint32_t *nr = s->nr, *nri = s->nr_interp;
    const double *dZ = s->dZ, *shift = s->shift, *ifac = s->interp_factor;
    double *Su = s->Su, *A = s->A, *C = s->C, *iA = s->inverse_A, *iC = s->inverse_C;
    double *FF = s->FF, *FI = s->FI, *GG = s->GG, *GP = s->GP, *L_diag = s->L_diag;
    int cco = 1;

    compute_source_dtau(m, op, nr[first], nr[first + 1], nri[first], nri[first + 1], l,
                        freq * shift[first], freq * shift[first + 1], shift[first],
                        shift[first + 1], dZ[first], ifac[first], ifac[first + 1], &cco, &dtau_n,
                        &chi_c, &chi_n, &term_c, &term_n);

    const double inverse_dtau_f = 1.0 / dtau_n;
    C[first] = 2.0 * inverse_dtau_f * inverse_dtau_f;
    iC[first] = 1.0 / C[first];
    const double Bf_min_Cf = 1.0 + 2.0 * inverse_dtau_f;
    const double Bf = Bf_min_Cf + C[first];
    const double I_bdy_f = boundary_intensity(m, nr[first], freq * shift[first]);
    Su[first] = term_c + 2.0 * I_bdy_f * inverse_dtau_f;
    Su[first] /= Bf;
    FF[first] = 0.5 * Bf_min_Cf * dtau_n * dtau_n;
    FI[first] = 1.0 / (1.0 + FF[first]);

    for (uint32_t n = first + 1; n < last; n++) {
        term_c = term_n;
        dtau_c = dtau_n;
        chi_c = chi_n;
        compute_source_dtau(m, op, nr[n], nr[n + 1], nri[n], nri[n + 1], l, freq * shift[n],
                            freq * shift[n + 1], shift[n], shift[n + 1], dZ[n], ifac[n],
                            ifac[n + 1], &cco, &dtau_n, &chi_c, &chi_n, &term_c, &term_n);
        const double dtau_avg = 0.5 * (dtau_c + dtau_n);
        iA[n] = dtau_avg * dtau_c;
        iC[n] = dtau_avg * dtau_n;
        A[n] = 1.0 / iA[n];
        C[n] = 1.0 / iC[n];
        Su[n] = term_c;
        FF[n] = (A[n] * FF[n - 1] * FI[n - 1] + 1.0) * iC[n];
        FI[n] = 1.0 / (1.0 + FF[n]);
        Su[n] = (A[n] * Su[n - 1] + Su[n]) * FI[n] * iC[n];
    }

    const double inverse_dtau_l = 1.0 / dtau_n;
    A[last] = 2.0 * inverse_dtau_l * inverse_dtau_l;
    const double Bl_min_Al = 1.0 + 2.0 * inverse_dtau_l;
    const double Bl = Bl_min_Al + A[last];
    const double denominator = 1.0 / (Bl * FF[last - 1] + Bl_min_Al);
    const double I_bdy_l = boundary_intensity(m, nr[last], freq * shift[last]);
    Su[last] = term_n + 2.0 * I_bdy_l * inverse_dtau_l;
    Su[last] = (A[last] * Su[last - 1] + Su[last]) * (1.0 + FF[last - 1]) * denominator;

    if (op->n_off_diag > 0) {
        /* all of L_diag and the bands of the inverse (:2319-2355) */
        const uint32_t M = op->n_off_diag, L0 = s->length;
        double *GI = s->GI, *Lu = s->L_upper, *Ll = s->L_lower;
        GG[last] = 0.5 * Bl_min_Al * dtau_n * dtau_n;
        GI[last] = 1.0 / (1.0 + GG[last]);
        GP[last] = GG[last] * GI[last];
        L_diag[last] = (1.0 + FF[last - 1]) / (Bl_min_Al + Bl * FF[last - 1]);
        for (long n = (long)last - 1; n > (long)first; n--) {
            Su[n] += Su[n + 1] * FI[n];
            GG[n] = (C[n] * GP[n + 1] + 1.0) * iA[n];
            GI[n] = 1.0 / (1.0 + GG[n]);
            GP[n] = GG[n] * GI[n];
            L_diag[n] = iC[n] / (FF[n] + GP[n + 1]);
        }
        Su[first] += Su[first + 1] * FI[first];
        L_diag[first] = (1.0 + GG[first + 1]) / (Bf_min_Cf + Bf * GG[first + 1]);
        for (long n = (long)last - 1; n >= (long)first; n--) {
            Lu[n + 1] = L_diag[n + 1] * FI[n];
            Ll[n] = L_diag[n] * GI[n + 1];
        }
        for (uint32_t mm = 1; mm < M && mm < s->n_tot - 1; mm++) {
            for (long n = (long)last - 1 - mm; n >= (long)first; n--) {
                Lu[(size_t)mm * L0 + n + mm + 1] = Lu[(size_t)(mm - 1) * L0 + n + mm + 1] * FI[n];
                Ll[(size_t)mm * L0 + n] = Ll[(size_t)(mm - 1) * L0 + n] * GI[n + mm + 1];
            }
        }
        return;
    }
    if (centre < last) {
        GG[last] = 0.5 * Bl_min_Al * dtau_n * dtau_n;
        GP[last] = GG[last] / (1.0 + GG[last]);
        for (long n = (long)last - 1; n > (long)centre; n--) {
            Su[n] += Su[n + 1] * FI[n];
            GG[n] = (C[n] * GP[n + 1] + 1.0) * iA[n];
            GP[n] = GG[n] / (1.0 + GG[n]);
        }
        Su[centre] += Su[centre + 1] * FI[centre];
        L_diag[centre] = iC[centre] / (FF[centre] + GP[centre + 1]);
    } else {
        L_diag[centre] = (1.0 + FF[centre - 1]) / (Bl_min_Al + Bl * FF[centre - 1]);
    }
}

/* update_Lambda<approx,adaptive>(rr, f), diagonal part (:1761-1800) */
static void update_lambda(const mgx_model *m, const mgx_options *op, const Scratch *s, uint32_t rr,
                          uint32_t f, double *lambda_diag)
{
    if (!m->f_in_line[f]) return;
    const uint32_t c = s->centre;
    const uint32_t p = s->nr[c];
    const double w_ang = 2.0 * ray_weight(m, p, rr);
    const uint32_t spec = m->f_species[f];
    const uint32_t k = m->f_k[f];
    const uint32_t z = m->f_z[f];
    /* global line index of transition k of this species */
    const uint32_t lg = m->f_line[f];
    const double freq_line = m->line[lg];
    const double invr_mass = m->species_inverse_mass[spec];
    const double constante = m->line_A[lg] * m->quad_weights[(size_t)spec * m->nquads + z] * w_ang;
    double eta, chi;
    get_eta_and_chi(m, op, p, k, freq_line, &eta, &chi);
    const double inverse_chi = 1.0 / chi;
    const double frq = m->nu[(size_t)p * m->nfreqs + f] * s->shift[c];
    const double phi = thermo_profile(m, invr_mass, p, freq_line, frq);
    const double L = constante * frq * phi * s->L_diag[c] * inverse_chi;
    lambda_diag[(size_t)p * m->nlines + lg] += L;
}

/* Lambda (lambda.tpp): per receiving point p and line lg a list of (emitting point, value)
   with capacity cap; add_element adds to an existing entry of the same point, else appends */
typedef struct {
    uint32_t cap;
    uint32_t *count, *nr;
    double *val;
} LamList;

static int lam_add(LamList *ll, size_t pl, uint32_t nr, double L)
{
    const uint32_t n = ll->count[pl];
    for (uint32_t i = 0; i < n; i++) {
        if (ll->nr[pl * ll->cap + i] == nr) {
            ll->val[pl * ll->cap + i] += L;
            return 0;
        }
    }
    if (n >= ll->cap) return MGX_E_ARG;
    ll->nr[pl * ll->cap + n] = nr;
    ll->val[pl * ll->cap + n] = L;
    ll->count[pl] = n + 1;
    return 0;
}

/* update_Lambda<approx,adaptive>(rr, f) with the off-diagonal bands (:1761-1850): the
   diagonal element, then for m = 0 .. n_off_diag-1 the entries m+1 before and after the
   centre, each at its own point nr[n] (an interpolated entry counts for its step's next
   point), its opacity at the line centre with the transition index as line index, and its
   profile at nu(nr[n], f) shift[n] */
static int update_lambda_list(const mgx_model *m, const mgx_options *op, const Scratch *s,
                              uint32_t rr, uint32_t f, LamList *ll)
{
    if (!m->f_in_line[f]) return 0;
    const uint32_t c = s->centre, first = s->first, last = s->last;
    const uint32_t p = s->nr[c];
    const double w_ang = 2.0 * ray_weight(m, p, rr);
    const uint32_t spec = m->f_species[f];
    const uint32_t k = m->f_k[f];
    const uint32_t z = m->f_z[f];
    const uint32_t lg = m->f_line[f];
    const double freq_line = m->line[lg];
    const double invr_mass = m->species_inverse_mass[spec];
    const double constante = m->line_A[lg] * m->quad_weights[(size_t)spec * m->nquads + z] * w_ang;
    const size_t pl = (size_t)p * m->nlines + lg;
    double eta, chi;
    get_eta_and_chi(m, op, p, k, freq_line, &eta, &chi);
    double frq = m->nu[(size_t)p * m->nfreqs + f] * s->shift[c];
    double phi = thermo_profile(m, invr_mass, p, freq_line, frq);
    int err = lam_add(ll, pl, p, constante * frq * phi * s->L_diag[c] * (1.0 / chi));
    for (uint32_t mm = 0; mm < op->n_off_diag && mm + 1 < s->n_tot; mm++) {
        if (c >= first + mm + 1) {
            const uint32_t n = c - mm - 1, q = s->nr[n];
            get_eta_and_chi(m, op, q, k, freq_line, &eta, &chi);
            frq = m->nu[(size_t)q * m->nfreqs + f] * s->shift[n];
            phi = thermo_profile(m, invr_mass, q, freq_line, frq);
            const double L =
                constante * frq * phi * s->L_lower[(size_t)mm * s->length + n] * (1.0 / chi);
            if (!err) err = lam_add(ll, pl, q, L);
        }
        if (c + mm + 1 <= last) {
            const uint32_t n = c + mm + 1, q = s->nr[n];
            get_eta_and_chi(m, op, q, k, freq_line, &eta, &chi);
            frq = m->nu[(size_t)q * m->nfreqs + f] * s->shift[n];
            phi = thermo_profile(m, invr_mass, q, freq_line, frq);
            const double L =
                constante * frq * phi * s->L_upper[(size_t)mm * s->length + n] * (1.0 / chi);
            if (!err) err = lam_add(ll, pl, q, L);
        }
    }
    return err;
}

/* ------------------------------------------------------------------------- */
/* public entry points                                                       */

static int check_model(const mgx_model *m, const mgx_options *op)
{
    if (!m || !op) return MGX_E_ARG;
    if (op->n_off_diag != 0) return MGX_E_UNSUPPORTED;
    if (m->npoints == 0 || m->nfreqs == 0) return MGX_E_ARG;
    return 0;
}

static int default_threads(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

static int solve_common(const mgx_model *m, const mgx_options *op, double *u, double *J,
                        double *lambda_diag, LamList *ll, int nthreads)
{
    int err = 0;
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    /* Solver::setup: scratch length from the all-lines ray length */
    const uint32_t length = 2 * get_ray_lengths_max(m, op, &err) + 1;
    if (err) goto done;

    const size_t PF = (size_t)m->npoints * m->nfreqs;
    memset(J, 0, PF * sizeof(double));
    if (lambda_diag) memset(lambda_diag, 0, (size_t)m->npoints * m->nlines * sizeof(double));
    if (ll) memset(ll->count, 0, (size_t)m->npoints * m->nlines * sizeof(uint32_t));

#pragma omp parallel
    {
        Scratch s;
        int serr = scratch_init(&s, length);
        if (!serr && op->n_off_diag > 0) serr = scratch_bands(&s, op->n_off_diag);
        for (uint32_t rr = 0; rr < m->hnrays; rr++) {
            const uint32_t ar = m->antipod[rr];
#pragma omp for schedule(static)
            for (uint32_t o = 0; o < m->npoints; o++) {
                if (serr) continue;
                s.nr[s.centre] = o;
                s.shift[s.centre] = 1.0;
                for (uint32_t l = 0; l < m->nlines; l++) {
                    int terr = 0;
                    s.first = trace_ray_for_line(m, op, &s, l, o, rr, -1, s.centre - 1,
                                                 s.centre - 1, &terr) + 1;
                    s.last = trace_ray_for_line(m, op, &s, l, o, ar, +1, s.centre + 1, s.centre,
                                                &terr) - 1;
                    s.n_tot = (s.last + 1) - s.first;
                    if (terr) serr = terr;
                    const double w2 = 2.0 * ray_weight(m, o, rr);
                    for (uint32_t q = 0; q < m->nquads; q++) {
                        const uint32_t f = l * m->nquads + q;
                        const size_t of = (size_t)o * m->nfreqs + f;
                        if (s.n_tot > 1) {
                            solve_feautrier_order_2_of(m, op, &s, o, f);
                            const double uc = s.Su[s.centre];
                            if (u) u[(size_t)rr * PF + of] = uc;
                            J[of] += uc * 2.0 * ray_weight(m, o, rr);
                            if (lambda_diag) update_lambda(m, op, &s, rr, f, lambda_diag);
                            if (ll) {
                                const int le = update_lambda_list(m, op, &s, rr, f, ll);
                                if (le) serr = le;
                            }
                        } else {
                            const double ub = boundary_intensity(m, o, m->nu[of]);
                            if (u) u[(size_t)rr * PF + of] = ub;
                            J[of] += w2 * ub;
                        }
                    }
                }
            }
        }
#pragma omp critical
        {
            if (serr && !err) err = serr;
        }
        scratch_free(&s);
    }
done:
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return err;
}

int oracle_solve_feautrier_order_2(const mgx_model *m, const mgx_options *op, double *u, double *J,
                                   double *lambda_diag, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    return solve_common(m, op, u, J, lambda_diag, NULL, nthreads);
}

int oracle_solve_feautrier_order_2_lambda(const mgx_model *m, const mgx_options *op, double *u,
                                          double *J, uint32_t cap, uint32_t *lam_count,
                                          uint32_t *lam_nr, double *lam_val, int nthreads)
{
    if (!m || !op || !J || !lam_count || !lam_nr || !lam_val || cap == 0) return MGX_E_ARG;
    if (m->npoints == 0 || m->nfreqs == 0) return MGX_E_ARG;
    LamList ll = {cap, lam_count, lam_nr, lam_val};
    return solve_common(m, op, u, J, NULL, &ll, nthreads);
}

/* ------------------------------------------------------------------------- */
/* Feautrier with v (solver.tpp:238-295, 2767-2921)                          */

/* solve_feautrier_order_2_uv(o, f): the elimination of solve_feautrier_order_2,
   back substitution down to the centre (and one row past it), and v at the
   centre from the second-order difference of u over the two neighbouring
   optical depths (or the boundary difference when the centre is an end) */
static void solve_feautrier_order_2_uv_of(const mgx_model *m, const mgx_options *op, Scratch *s,
                                          uint32_t o, uint32_t f, double *u_c, double *v_c,
                                          double *v_s)
{
    const double freq = m->nu[(size_t)o * m->nfreqs + f];
    const uint32_t l = m->f_line[f];
    double dtau_c, term_c, dtau_n, term_n, chi_c = 0.0, chi_n = 0.0;
    const uint32_t first = s->first, last = s->last, centre = s->centre;
    const uint32_t *nr = s->nr, *nri = s->nr_interp;
    const double *dZ = s->dZ, *shift = s->shift, *ifac = s->interp_factor;
    double *Su = s->Su, *A = s->A, *C = s->C, *iA = s->inverse_A, *iC = s->inverse_C;
    double *FF = s->FF, *FI = s->FI;
    int cco = 1;

    compute_source_dtau(m, op, nr[first], nr[first + 1], nri[first], nri[first + 1], l,
                        freq * shift[first], freq * shift[first + 1], shift[first],
                        shift[first + 1], dZ[first], ifac[first], ifac[first + 1], &cco, &dtau_n,
                        &chi_c, &chi_n, &term_c, &term_n);
    const double inverse_dtau_f = 1.0 / dtau_n;
    C[first] = 2.0 * inverse_dtau_f * inverse_dtau_f;
    iC[first] = 1.0 / C[first];
    const double Bf_min_Cf = 1.0 + 2.0 * inverse_dtau_f;
    const double Bf = Bf_min_Cf + C[first];
    const double I_bdy_f = boundary_intensity(m, nr[first], freq * shift[first]);
    Su[first] = term_c + 2.0 * I_bdy_f * inverse_dtau_f;
    Su[first] /= Bf;
    FF[first] = 0.5 * Bf_min_Cf * dtau_n * dtau_n;
    FI[first] = 1.0 / (1.0 + FF[first]);
    for (uint32_t n = first + 1; n < last; n++) {
        term_c = term_n;
        dtau_c = dtau_n;
        chi_c = chi_n;
        compute_source_dtau(m, op, nr[n], nr[n + 1], nri[n], nri[n + 1], l, freq * shift[n],
                            freq * shift[n + 1], shift[n], shift[n + 1], dZ[n], ifac[n],
                            ifac[n + 1], &cco, &dtau_n, &chi_c, &chi_n, &term_c, &term_n);
        const double dtau_avg = 0.5 * (dtau_c + dtau_n);
        iA[n] = dtau_avg * dtau_c;
        iC[n] = dtau_avg * dtau_n;
        A[n] = 1.0 / iA[n];
        C[n] = 1.0 / iC[n];
        Su[n] = term_c;
        FF[n] = (A[n] * FF[n - 1] * FI[n - 1] + 1.0) * iC[n];
        FI[n] = 1.0 / (1.0 + FF[n]);
        Su[n] = (A[n] * Su[n - 1] + Su[n]) * FI[n] * iC[n];
    }
    const double inverse_dtau_l = 1.0 / dtau_n;
    A[last] = 2.0 * inverse_dtau_l * inverse_dtau_l;
    const double Bl_min_Al = 1.0 + 2.0 * inverse_dtau_l;
    const double Bl = Bl_min_Al + A[last];
    const double denominator = 1.0 / (Bl * FF[last - 1] + Bl_min_Al);
    const double I_bdy_l = boundary_intensity(m, nr[last], freq * shift[last]);
    Su[last] = term_n + 2.0 * I_bdy_l * inverse_dtau_l;
    Su[last] = (A[last] * Su[last - 1] + Su[last]) * (1.0 + FF[last - 1]) * denominator;
    if (centre < last) {
        for (long n = (long)last - 1; n >= (long)centre; n--) Su[n] += Su[n + 1] * FI[n];
    } else {
        *u_c = Su[centre];
        *v_c = Su[last] - I_bdy_l;
        *v_s = fabs(Su[last]) + fabs(I_bdy_l);
        return;
    }
    if (centre > first) {
        Su[centre - 1] += Su[centre] * FI[centre - 1];
        cco = 1;
        compute_source_dtau(m, op, nr[centre - 1], nr[centre], nri[centre - 1], nri[centre], l,
                            freq * shift[centre - 1], freq * shift[centre], shift[centre - 1],
                            shift[centre], dZ[centre - 1], ifac[centre - 1], ifac[centre], &cco,
                            &dtau_n, &chi_c, &chi_n, &term_c, &term_n);
        const double dtaumin = dtau_n;
        compute_source_dtau(m, op, nr[centre], nr[centre + 1], nri[centre], nri[centre + 1], l,
                            freq * shift[centre], freq * shift[centre + 1], shift[centre],
                            shift[centre + 1], dZ[centre], ifac[centre], ifac[centre + 1], &cco,
                            &dtau_n, &chi_c, &chi_n, &term_c, &term_n);
        const double dtauplus = dtau_n;
        const double coeffmin = -dtauplus / (dtaumin * dtaumin + dtaumin * dtauplus);
        const double coeffplus = dtaumin / (dtaumin * dtauplus + dtauplus * dtauplus);
        const double coeffzero = -coeffmin - coeffplus;
        *u_c = Su[centre];
        *v_c = -(coeffmin * Su[centre - 1] + coeffplus * Su[centre + 1] + coeffzero * Su[centre]);
        *v_s = fabs(coeffmin * Su[centre - 1]) + fabs(coeffplus * Su[centre + 1]) +
               fabs(coeffzero * Su[centre]);
    } else {
        *u_c = Su[centre];
        *v_c = I_bdy_f - Su[first];
        *v_s = fabs(I_bdy_f) + fabs(Su[first]);
    }
}

int oracle_feautrier_order_2_uv(const mgx_model *m, const mgx_options *op, double *u, double *v,
                                double *v_scale, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    const uint32_t length = 2 * get_ray_lengths_max(m, op, &err) + 1;
    if (err) goto done;
    const size_t PF = (size_t)m->npoints * m->nfreqs;
#pragma omp parallel
    {
        Scratch s;
        int serr = scratch_init(&s, length);
        for (uint32_t rr = 0; rr < m->hnrays; rr++) {
            const uint32_t ar = m->antipod[rr];
#pragma omp for schedule(static)
            for (uint32_t o = 0; o < m->npoints; o++) {
                if (serr) continue;
                s.nr[s.centre] = o;
                s.shift[s.centre] = 1.0;
                for (uint32_t l = 0; l < m->nlines; l++) {
                    int terr = 0;
                    s.first = trace_ray_for_line(m, op, &s, l, o, rr, -1, s.centre - 1,
                                                 s.centre - 1, &terr) + 1;
                    s.last = trace_ray_for_line(m, op, &s, l, o, ar, +1, s.centre + 1, s.centre,
                                                &terr) - 1;
                    s.n_tot = (s.last + 1) - s.first;
                    if (terr) serr = terr;
                    for (uint32_t q = 0; q < m->nquads; q++) {
                        const uint32_t f = l * m->nquads + q;
                        const size_t of = (size_t)o * m->nfreqs + f;
                        double uc, vc, vs = 0.0;
                        if (s.n_tot > 1) {
                            solve_feautrier_order_2_uv_of(m, op, &s, o, f, &uc, &vc, &vs);
                        } else {
                            uc = boundary_intensity(m, o, m->nu[of]);
                            vc = 0.0;
                        }
                        u[(size_t)rr * PF + of] = uc;
                        v[(size_t)rr * PF + of] = vc;
                        if (v_scale) v_scale[(size_t)rr * PF + of] = vs;
                    }
                }
            }
        }
#pragma omp critical
        {
            if (serr && !err) err = serr;
        }
        scratch_free(&s);
    }
done:
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return err;
}

/* ------------------------------------------------------------------------- */
/* order-0 short characteristics (solver.tpp:151-235, 1565-1757)             */

/* solve_shortchar_order_0_ray_forward (forward = 1) / _backward: the intensity
   arriving at the centre from the first (last) traced entry, and the
   direction's Lambda element in *L (*have_L = 0 when the direction has no
   segment: the reference then reads an uninitialised dtau) */
static double shortchar_ray(const mgx_model *m, const mgx_options *op, const Scratch *s,
                            uint32_t o, uint32_t rr, uint32_t f, double min_dtau, int forward,
                            double *L, int *have_L)
{
    const double freq = m->nu[(size_t)o * m->nfreqs + f];
    const uint32_t l = m->f_line[f];
    const uint32_t *nr = s->nr, *nri = s->nr_interp;
    const double *dZ = s->dZ, *shift = s->shift, *ifac = s->interp_factor;
    double chi_c = 0.0, chi_n = 0.0, source_c = 0.0, source_n = 0.0, dtau = 0.0;
    int cco = 1, any = 0;
    const uint32_t bdy = forward ? s->first : s->last;
    double I = boundary_intensity(m, nr[bdy], freq * shift[bdy]);
    const long n_seg = forward ? (long)s->centre - (long)s->first : (long)s->last - (long)s->centre;
    for (long q = 0; q < n_seg; q++) {
        /* forward: idx = first + 1 + q, (crt, nxt) = (idx - 1, idx);
           backward: indexp1 = last - q, (crt, nxt) = (indexp1, indexp1 - 1) */
        const uint32_t c = forward ? s->first + (uint32_t)q : s->last - (uint32_t)q;
        const uint32_t n = forward ? c + 1 : c - 1;
        const uint32_t dz = forward ? c : n; /* dZ[idx - 1] / dZ[indexp1 - 1] */
        chi_c = chi_n;
        source_c = source_n;
        compute_source_dtau(m, op, nr[c], nr[n], nri[c], nri[n], l, freq * shift[c],
                            freq * shift[n], shift[c], shift[n], dZ[dz], ifac[c], ifac[n], &cco,
                            &dtau, &chi_c, &chi_n, &source_c, &source_n);
        dtau = (min_dtau < dtau) ? dtau : min_dtau; /* std::max(min_dtau, dtau) */
        I = (double)O_EXPF((float)(-dtau)) * I +
            (double)O_EXPM1F((float)(-dtau)) / dtau * (source_c - source_n * (1.0 + dtau)) +
            source_c - source_n;
        any = 1;
    }
    *have_L = any;
    if (any) {
        const double factor = (double)O_EXPM1F((float)(-dtau)) / dtau;
        const uint32_t spec = m->f_species[f];
        const uint32_t k = m->f_k[f];
        const uint32_t z = m->f_z[f];
        const uint32_t lg = m->f_line[f];
        const double w_ang = ray_weight(m, o, rr);
        const double freq_line = m->line[lg];
        const double invr_mass = m->species_inverse_mass[spec];
        const double constante =
            m->line_A[lg] * m->quad_weights[(size_t)spec * m->nquads + z] * w_ang;
        double eta, chi;
        get_eta_and_chi(m, op, o, k, freq_line, &eta, &chi);
        const double inverse_chi = 1.0 / chi;
        const double phi = thermo_profile(m, invr_mass, o, freq_line, freq);
        *L = constante * freq * phi * (factor + 1.0) * inverse_chi;
    }
    return I;
}

int oracle_shortchar_order_0(const mgx_model *m, const mgx_options *op, double min_dtau,
                             double *I, double *u, double *v, double *J, double *lambda_diag,
                             int nthreads)
{
    /* the reference's short characteristics ignore n_off_diag: their Lambda is
       the diagonal (solver.tpp:151-235) */
    mgx_options op0;
    if (op) {
        op0 = *op;
        op0.n_off_diag = 0;
        op = &op0;
    }
    int err = check_model(m, op);
    if (err) return err;
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    const uint32_t length = 2 * get_ray_lengths_max(m, op, &err) + 1;
    if (err) goto done;
    const size_t PF = (size_t)m->npoints * m->nfreqs;
    memset(J, 0, PF * sizeof(double));
    if (lambda_diag) memset(lambda_diag, 0, (size_t)m->npoints * m->nlines * sizeof(double));
#pragma omp parallel
    {
        Scratch s;
        int serr = scratch_init(&s, length);
        for (uint32_t rr = 0; rr < m->hnrays; rr++) {
            const uint32_t ar = m->antipod[rr];
#pragma omp for schedule(static)
            for (uint32_t o = 0; o < m->npoints; o++) {
                if (serr) continue;
                const double wt = ray_weight(m, o, rr) * 2.0;
                s.nr[s.centre] = o;
                s.shift[s.centre] = 1.0;
                for (uint32_t l = 0; l < m->nlines; l++) {
                    int terr = 0;
                    s.first = trace_ray_for_line(m, op, &s, l, o, rr, -1, s.centre - 1,
                                                 s.centre - 1, &terr) + 1;
                    s.last = trace_ray_for_line(m, op, &s, l, o, ar, +1, s.centre + 1, s.centre,
                                                &terr) - 1;
                    s.n_tot = (s.last + 1) - s.first;
                    if (terr) serr = terr;
                    for (uint32_t q = 0; q < m->nquads; q++) {
                        const uint32_t f = l * m->nquads + q;
                        const size_t of = (size_t)o * m->nfreqs + f;
                        double If, Ib, uu, vv;
                        if (s.n_tot > 1) {
                            double Lf = 0.0, Lb = 0.0;
                            int hf = 0, hb = 0;
                            If = shortchar_ray(m, op, &s, o, rr, f, min_dtau, 1, &Lf, &hf);
                            Ib = shortchar_ray(m, op, &s, o, rr, f, min_dtau, 0, &Lb, &hb);
                            uu = 0.5 * (If + Ib);
                            vv = 0.5 * (If - Ib);
                            if (lambda_diag) {
                                const size_t li = (size_t)o * m->nlines + m->f_line[f];
                                if (hf) lambda_diag[li] += Lf;
                                if (hb) lambda_diag[li] += Lb;
                            }
                        } else {
                            If = Ib = uu = boundary_intensity(m, o, m->nu[of]);
                            vv = 0.0;
                        }
                        if (I) {
                            I[(size_t)rr * PF + of] = If;
                            I[(size_t)ar * PF + of] = Ib;
                        }
                        if (u) u[(size_t)rr * PF + of] = uu;
                        if (v) v[(size_t)rr * PF + of] = vv;
                        J[of] += wt * uu;
                    }
                }
            }
        }
#pragma omp critical
        {
            if (serr && !err) err = serr;
        }
        scratch_free(&s);
    }
done:
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return err;
}

int oracle_trace_counts(const mgx_model *m, const mgx_options *op, uint32_t *n_tot, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    const uint32_t length = 2 * get_ray_lengths_max(m, op, &err) + 1;
    if (err) goto done;
#pragma omp parallel
    {
        Scratch s;
        int serr = scratch_init(&s, length);
        for (uint32_t rr = 0; rr < m->hnrays; rr++) {
            const uint32_t ar = m->antipod[rr];
#pragma omp for schedule(static)
            for (uint32_t o = 0; o < m->npoints; o++) {
                if (serr) continue;
                for (uint32_t l = 0; l < m->nlines; l++) {
                    int terr = 0;
                    const uint32_t first =
                        trace_ray_for_line(m, op, &s, l, o, rr, -1, s.centre - 1, s.centre - 1, &terr) + 1;
                    const uint32_t last =
                        trace_ray_for_line(m, op, &s, l, o, ar, +1, s.centre + 1, s.centre, &terr) - 1;
                    if (terr) serr = terr;
                    n_tot[((size_t)rr * m->npoints + o) * m->nlines + l] = (last + 1) - first;
                }
            }
        }
#pragma omp critical
        {
            if (serr && !err) err = serr;
        }
        scratch_free(&s);
    }
done:
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return err;
}

int oracle_trace_ray(const mgx_model *m, const mgx_options *op, uint32_t rr, uint32_t o,
                     uint32_t l, uint32_t capacity, uint32_t *count, uint32_t *centre_idx,
                     uint32_t *nr, uint32_t *nr_interp, double *shift, double *interp_factor,
                     double *dZ)
{
    int err = check_model(m, op);
    if (err) return err;
    const uint32_t length = pair_length(m, op, o, rr, &err);
    if (err) return err;
    Scratch s;
    if (scratch_init(&s, length)) return MGX_E_ARG;
    s.nr[s.centre] = o;
    s.shift[s.centre] = 1.0;
    const uint32_t ar = m->antipod[rr];
    const uint32_t first = trace_ray_for_line(m, op, &s, l, o, rr, -1, s.centre - 1, s.centre - 1, &err) + 1;
    const uint32_t last = trace_ray_for_line(m, op, &s, l, o, ar, +1, s.centre + 1, s.centre, &err) - 1;
    const uint32_t n = (last + 1) - first;
    *count = n;
    *centre_idx = s.centre - first;
    if (!err && n > capacity) err = MGX_E_RAY_LENGTH;
    if (!err) {
        for (uint32_t j = 0; j < n; j++) {
            nr[j] = s.nr[first + j];
            nr_interp[j] = s.nr_interp[first + j];
            shift[j] = s.shift[first + j];
            interp_factor[j] = s.interp_factor[first + j];
            dZ[j] = (j + 1 < n) ? s.dZ[first + j] : 0.0;
        }
    }
    scratch_free(&s);
    return err;
}

/* ------------------------------------------------------------------------- */
/* the projection-surface imager (image_feautrier_order_2_new_imager :577-627) */

/* set_data :1003-1043 (all-lines interpolation) */
static void set_data_all(const mgx_model *m, const mgx_options *op, Scratch *s, uint32_t crt,
                         uint32_t nxt, double shift_crt, double shift_nxt, double dZ_loc,
                         int increment, uint32_t *id1, uint32_t *id2)
{
    set_data_n(s, get_n_interp(m, op, crt, nxt), crt, nxt, shift_crt, shift_nxt, dZ_loc,
               increment, id1, id2);
}

/* get_boundary_point_closer_to_custom_ray :310-341 */
static uint32_t bdy_closer(const mgx_model *m, const double *origin, const double *raydir,
                           uint32_t crt)
{
    const uint32_t n_nbs = m->n_neighbors[crt];
    const uint32_t cum = m->cum_n_neighbors[crt];
    double Rc[3];
    for (int i = 0; i < 3; i++) Rc[i] = m->position[3 * (size_t)crt + i] - origin[i];
    const double Zc = dot3(Rc, raydir);
    double dmin2 = dot3(Rc, Rc) - Zc * Zc;
    uint32_t next = crt;
    for (uint32_t i = 0; i < n_nbs; i++) {
        const uint32_t n = m->neighbors[cum + i];
        if (!not_on_boundary(m, n)) {
            double R[3];
            for (int j = 0; j < 3; j++) R[j] = m->position[3 * (size_t)n + j] - origin[j];
            const double Zn = dot3(R, raydir);
            const double d2 = dot3(R, R) - Zn * Zn;
            if (d2 < dmin2) {
                dmin2 = d2;
                next = n;
            }
        }
    }
    return next;
}

/* get_distance_origin_to_boundary :167-193 */
static double dist_origin_to_bdy(const mgx_model *m, const double *origin, const double *raydir,
                                 uint32_t bdy)
{
    double R[3];
    for (int i = 0; i < 3; i++) R[i] = m->position[3 * (size_t)bdy + i] - origin[i];
    double Z = dot3(raydir, R);
    if (m->spherical_symmetry) {
        const double Rcos = dot3(origin, raydir);
        const double R2 = dot3(origin, origin);
        const double Rsin2 = R2 - Rcos * Rcos;
        const double *pb = &m->position[3 * (size_t)bdy];
        const double h2 = dot3(pb, pb) - Rsin2;
        Z = (h2 > 0) ? (-sqrt(h2) - Rcos) : -Rcos;
    }
    return Z;
}

/* trace_ray_imaging_get_start :808-830 */
static uint32_t imaging_start(const mgx_model *m, const double *origin, uint32_t start_bdy,
                              const double *raydir, double *Z)
{
    uint32_t ip = start_bdy;
    for (;;) {
        const uint32_t na = bdy_closer(m, origin, raydir, ip);
        if (na == ip) break;
        ip = na;
    }
    *Z = dist_origin_to_bdy(m, origin, raydir, ip);
    return ip;
}

/* get_closest_bdy_point_in_custom_raydir :346-367 */
uint32_t oracle_closest_bdy_point(const mgx_model *m, const double *raydir)
{
    if (m->spherical_symmetry) return m->npoints - 1;
    uint32_t best = m->point2boundary[0];
    double pmin = dot3(raydir, &m->position[3 * (size_t)best]);
    for (uint32_t b = 1; b < m->nboundary; b++) {
        const uint32_t p = m->boundary2point[b];
        const double pd = dot3(raydir, &m->position[3 * (size_t)p]);
        if (pd < pmin) {
            best = p;
            pmin = pd;
        }
    }
    return best;
}

/* trace_ray_imaging<Rest> :842-897 */
static uint32_t trace_ray_imaging(const mgx_model *m, const mgx_options *op, Scratch *s,
                                  const double *origin, uint32_t start_bdy, const double *raydir,
                                  int increment, double *Z, uint32_t id1, uint32_t id2)
{
    double dZ = 0.0;
    uint32_t crt = start_bdy;
    uint32_t nxt = get_next_xyz(m, origin, raydir, start_bdy, Z, &dZ);
    int accessed_non_boundary = 0;
    const uint32_t original_id1 = id1;
    if (nxt < m->npoints) {
        double shift_crt = shift_rest(m, origin, raydir, crt, *Z - dZ);
        double shift_nxt = shift_rest(m, origin, raydir, nxt, *Z);
        set_data_all(m, op, s, crt, nxt, shift_crt, shift_nxt, dZ, increment, &id1, &id2);
        for (;;) {
            double rel[3];
            for (int i = 0; i < 3; i++) rel[i] = m->position[3 * (size_t)nxt + i] - m->center[i];
            if (!(not_on_boundary(m, nxt) || dot3(rel, raydir) < 0)) break;
            crt = nxt;
            shift_crt = shift_nxt;
            if (not_on_boundary(m, nxt)) accessed_non_boundary = 1;
            nxt = get_next_xyz(m, origin, raydir, nxt, Z, &dZ);
            if (nxt >= m->npoints) break;
            shift_nxt = shift_rest(m, origin, raydir, nxt, *Z);
            set_data_all(m, op, s, crt, nxt, shift_crt, shift_nxt, dZ, increment, &id1, &id2);
        }
    }
    if (!accessed_non_boundary && !m->spherical_symmetry) return original_id1;
    return id1;
}

/* get_ray_length_new_imager :964-989 */
static uint32_t ray_length_new_imager(const mgx_model *m, const mgx_options *op,
                                      const double *origin, uint32_t start_bdy,
                                      const double *raydir)
{
    uint32_t len = 0;
    double Z = 0.0, dZ = 0.0;
    const uint32_t ip = imaging_start(m, origin, start_bdy, raydir, &Z);
    uint32_t crt = ip;
    uint32_t nxt = get_next_xyz(m, origin, raydir, crt, &Z, &dZ);
    if (nxt < m->npoints) {
        len += get_n_interp(m, op, crt, nxt);
        for (;;) {
            double rel[3];
            for (int i = 0; i < 3; i++) rel[i] = m->position[3 * (size_t)nxt + i] - m->center[i];
            if (!(not_on_boundary(m, nxt) || dot3(rel, raydir) < 0)) break;
            crt = nxt;
            nxt = get_next_xyz(m, origin, raydir, nxt, &Z, &dZ);
            if (nxt >= m->npoints) break;
            len += get_n_interp(m, op, crt, nxt);
        }
    }
    return len;
}

/* image_feautrier_order_2(o, f) :2463-2574 (forward elimination only); returns Su[last] */
static double image_solve_of(const mgx_model *m, const mgx_options *op, Scratch *s, uint32_t o,
                             uint32_t f)
{
    const double freq = m->nu[(size_t)o * m->nfreqs + f];
    const uint32_t l = m->f_line[f];
    double dtau_c, term_c, dtau_n, term_n, chi_c = 0.0, chi_n = 0.0;
    const uint32_t first = s->first, last = s->last;
    const uint32_t *nr = s->nr, *nri = s->nr_interp;
    const double *dZ = s->dZ, *shift = s->shift, *ifac = s->interp_factor;
    double *Su = s->Su, *A = s->A, *C = s->C, *iA = s->inverse_A, *iC = s->inverse_C;
    double *FF = s->FF, *FI = s->FI;
    int cco = 1;
    compute_source_dtau(m, op, nr[first], nr[first + 1], nri[first], nri[first + 1], l,
                        freq * shift[first], freq * shift[first + 1], shift[first],
                        shift[first + 1], dZ[first], ifac[first], ifac[first + 1], &cco, &dtau_n,
                        &chi_c, &chi_n, &term_c, &term_n);
    const double inverse_dtau_f = 1.0 / dtau_n;
    C[first] = 2.0 * inverse_dtau_f * inverse_dtau_f;
    const double Bf_min_Cf = 1.0 + 2.0 * inverse_dtau_f;
    const double Bf = Bf_min_Cf + C[first];
    const double I_bdy_f = boundary_intensity(m, nr[first], freq * shift[first]);
    Su[first] = term_c + 2.0 * I_bdy_f * inverse_dtau_f;
    Su[first] /= Bf;
    FF[first] = 0.5 * Bf_min_Cf * dtau_n * dtau_n;
    FI[first] = 1.0 / (1.0 + FF[first]);
    for (uint32_t n = first + 1; n < last; n++) {
        term_c = term_n;
        dtau_c = dtau_n;
        chi_c = chi_n;
        compute_source_dtau(m, op, nr[n], nr[n + 1], nri[n], nri[n + 1], l, freq * shift[n],
                            freq * shift[n + 1], shift[n], shift[n + 1], dZ[n], ifac[n],
                            ifac[n + 1], &cco, &dtau_n, &chi_c, &chi_n, &term_c, &term_n);
        const double dtau_avg = 0.5 * (dtau_c + dtau_n);
        iA[n] = dtau_avg * dtau_c;
        iC[n] = dtau_avg * dtau_n;
        A[n] = 1.0 / iA[n];
        C[n] = 1.0 / iC[n];
        Su[n] = term_c;
        FF[n] = (A[n] * FF[n - 1] * FI[n - 1] + 1.0) * iC[n];
        FI[n] = 1.0 / (1.0 + FF[n]);
        Su[n] = (A[n] * Su[n - 1] + Su[n]) * FI[n] * iC[n];
    }
    const double inverse_dtau_l = 1.0 / dtau_n;
    A[last] = 2.0 * inverse_dtau_l * inverse_dtau_l;
    const double Bl_min_Al = 1.0 + 2.0 * inverse_dtau_l;
    const double Bl = Bl_min_Al + A[last];
    const double denominator = 1.0 / (Bl * FF[last - 1] + Bl_min_Al);
    const double I_bdy_l = boundary_intensity(m, nr[last], freq * shift[last]);
    Su[last] = term_n + 2.0 * I_bdy_l * inverse_dtau_l;
    Su[last] = (A[last] * Su[last - 1] + Su[last]) * (1.0 + FF[last - 1]) * denominator;
    return Su[last];
}

/* Image::set_coordinates_projection_surface :221-366 and surface_coords_to_3D :370-398 */
typedef struct {
    double ix, iy, jx, jy, jz, denominator;
    double center_pt[3];
} ImageFrame;

static void image_surface_to_3d(const ImageFrame *fr, const double *raydir, double x, double y,
                                double *out)
{
    (void)raydir;
    if (fr->denominator >= 1.0e-9) {
        out[0] = (x * fr->ix + y * fr->jx) + fr->center_pt[0];
        out[1] = (x * fr->iy + y * fr->jy) + fr->center_pt[1];
        out[2] = (y * fr->jz) + fr->center_pt[2];
    } else {
        out[0] = x + fr->center_pt[0];
        out[1] = y + fr->center_pt[1];
        out[2] = 0.0 + fr->center_pt[2];
    }
}

/* Model::compute_image_new :1008-1035 and the frame of Image::set_coordinates_projection_surface
   :221-366: normalised direction, frame, pixel coordinates, start boundary point */
static int image_prepare(const mgx_model *m, const mgx_options *op, const double *raydir_in,
                         uint32_t nx, uint32_t ny, double *rd, ImageFrame *fr, double *ImX,
                         double *ImY, uint32_t *start_bdy)
{
    if (nx == 0 || ny == 0) return MGX_E_ARG;
    const double sq = dot3(raydir_in, raydir_in);
    if (sq == 0.0) return MGX_E_ARG;
    if (op->approx == MGX_APPROX_ONE_LINE) return MGX_E_UNSUPPORTED;
    const double inv = 1 / sqrt(sq);
    rd[0] = raydir_in[0] * inv;
    rd[1] = raydir_in[1] * inv;
    rd[2] = raydir_in[2] * inv;
    if (m->dimension == 1 && (rd[0] != 0.0 || rd[1] != 1.0 || rd[2] != 0.0)) return MGX_E_ARG;

    memset(fr, 0, sizeof *fr);
    double max_x = -DBL_MAX, max_y = -DBL_MAX, min_x = DBL_MAX, min_y = DBL_MAX;
    const double rx = rd[0], ry = rd[1], rz = rd[2];
    fr->denominator = sqrt(rx * rx + ry * ry);
    const double inverse_denominator = 1.0 / fr->denominator;
    fr->ix = ry * inverse_denominator;
    fr->iy = -rx * inverse_denominator;
    fr->jx = rx * rz * inverse_denominator;
    fr->jy = ry * rz * inverse_denominator;
    fr->jz = -fr->denominator;
    if (m->dimension == 1) {
        for (uint32_t b = 0; b < m->nboundary; b++) {
            const double x = m->position[3 * (size_t)m->boundary2point[b]];
            max_x = (max_x < x) ? x : max_x;
            min_x = (x < min_x) ? x : min_x;
        }
        const double ax = fabs(min_x);
        max_x = (ax < max_x) ? max_x : ax;
        min_x = -max_x;
        min_y = min_x;
        max_y = max_x;
    } else {
        for (uint32_t b = 0; b < m->nboundary; b++) {
            const double *p = &m->position[3 * (size_t)m->boundary2point[b]];
            double X, Y;
            if (fr->denominator >= 1.0e-9) {
                X = fr->ix * p[0] + fr->iy * p[1];
                Y = fr->jx * p[0] + fr->jy * p[1] + fr->jz * p[2];
            } else {
                X = p[0];
                Y = p[1];
            }
            min_x = (X < min_x) ? X : min_x;
            max_x = (max_x < X) ? X : max_x;
            min_y = (Y < min_y) ? Y : min_y;
            max_y = (max_y < Y) ? Y : max_y;
        }
    }
    const uint32_t closest = oracle_closest_bdy_point(m, rd);
    const double *cp = &m->position[3 * (size_t)closest];
    double dist = dot3(cp, rd);
    if (m->spherical_symmetry) dist = -cp[0];
    for (int i = 0; i < 3; i++) fr->center_pt[i] = rd[i] * dist;
    const double deltax = (max_x - min_x) / (nx);
    const double deltay = (max_y - min_y) / (ny);
    for (uint32_t xi = 0; xi < nx; xi++) {
        const double xloc = min_x + deltax / 2.0 + xi * deltax;
        for (uint32_t yi = 0; yi < ny; yi++) {
            const size_t t = yi + (size_t)ny * xi;
            ImX[t] = xloc;
            ImY[t] = min_y + deltay / 2.0 + yi * deltay;
        }
    }
    *start_bdy = closest;
    return 0;
}

/* scratch length of setup_new_imager: from the largest per-pixel all-lines length */
static uint32_t image_scratch_length(const mgx_model *m, const mgx_options *op,
                                     const ImageFrame *fr, const double *rd, const double *ImX,
                                     const double *ImY, size_t npix, uint32_t start_bdy)
{
    uint32_t mx = 0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (size_t pix = 0; pix < npix; pix++) {
        double org[3];
        image_surface_to_3d(fr, rd, ImX[pix], ImY[pix], org);
        const uint32_t len = ray_length_new_imager(m, op, org, start_bdy, rd);
        if (len > mx) mx = len;
    }
    return 2 * mx + 1;
}

/* trace of one pixel ray into s (image_feautrier_order_2_new_imager :595-612); returns the
   start boundary point */
static uint32_t image_trace_pixel(const mgx_model *m, const mgx_options *op, Scratch *s,
                                  const ImageFrame *fr, const double *rd, double x, double y,
                                  uint32_t start_bdy)
{
    double org[3];
    image_surface_to_3d(fr, rd, x, y, org);
    double Z = 0.0;
    const uint32_t cb = imaging_start(m, org, start_bdy, rd, &Z);
    s->nr[s->centre] = cb;
    s->shift[s->centre] = shift_rest(m, org, rd, cb, Z);
    s->first = trace_ray_imaging(m, op, s, org, cb, rd, -1, &Z, s->centre - 1, s->centre - 1) + 1;
    s->last = s->centre;
    s->n_tot = (s->last + 1) - s->first;
    return cb;
}

int oracle_image_new(const mgx_model *m, const mgx_options *op, const double *raydir_in,
                     uint32_t nx, uint32_t ny, double *I, double *ImX, double *ImY, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    double rd[3];
    ImageFrame fr;
    uint32_t start_bdy;
    err = image_prepare(m, op, raydir_in, nx, ny, rd, &fr, ImX, ImY, &start_bdy);
    if (err) return err;

    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    const size_t npix = (size_t)nx * ny;
    const uint32_t length = image_scratch_length(m, op, &fr, rd, ImX, ImY, npix, start_bdy);
#pragma omp parallel
    {
        Scratch s;
        const int serr = scratch_init(&s, length);
#pragma omp for schedule(static)
        for (size_t pix = 0; pix < npix; pix++) {
            if (serr) continue;
            const uint32_t cb = image_trace_pixel(m, op, &s, &fr, rd, ImX[pix], ImY[pix], start_bdy);
            for (uint32_t f = 0; f < m->nfreqs; f++) {
                const double nu_cb = m->nu[(size_t)cb * m->nfreqs + f];
                if (s.n_tot > 1) {
                    const double sl = image_solve_of(m, op, &s, cb, f);
                    I[pix * m->nfreqs + f] = 2.0 * sl - boundary_intensity(m, s.nr[s.last], nu_cb);
                } else {
                    I[pix * m->nfreqs + f] = boundary_intensity(m, cb, nu_cb);
                }
            }
        }
        scratch_free(&s);
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return 0;
}

/* image_optical_depth(model, o, f) :2718-2758: the optical depth of the traced ray */
static double optical_depth_of(const mgx_model *m, const mgx_options *op, Scratch *s, uint32_t o,
                               uint32_t f)
{
    const double freq = m->nu[(size_t)o * m->nfreqs + f];
    const uint32_t l = m->f_line[f];
    double term_c, dtau_n, term_n, chi_c = 0.0, chi_n = 0.0;
    const uint32_t first = s->first, last = s->last;
    const uint32_t *nr = s->nr, *nri = s->nr_interp;
    const double *dZ = s->dZ, *shift = s->shift, *ifac = s->interp_factor;
    int cco = 1;
    double tau = 0.0;
    compute_source_dtau(m, op, nr[first], nr[first + 1], nri[first], nri[first + 1], l,
                        freq * shift[first], freq * shift[first + 1], shift[first],
                        shift[first + 1], dZ[first], ifac[first], ifac[first + 1], &cco, &dtau_n,
                        &chi_c, &chi_n, &term_c, &term_n);
    tau += dtau_n;
    for (uint32_t n = first + 1; n < last; n++) {
        term_c = term_n;
        chi_c = chi_n;
        compute_source_dtau(m, op, nr[n], nr[n + 1], nri[n], nri[n + 1], l, freq * shift[n],
                            freq * shift[n + 1], shift[n], shift[n + 1], dZ[n], ifac[n],
                            ifac[n + 1], &cco, &dtau_n, &chi_c, &chi_n, &term_c, &term_n);
        tau += dtau_n;
    }
    return tau;
}

/* image_optical_depth_new_imager :687-736: the pixel rays of oracle_image_new, tau(pix, f) */
int oracle_image_optical_depth_new(const mgx_model *m, const mgx_options *op,
                                   const double *raydir_in, uint32_t nx, uint32_t ny, double *tau,
                                   double *ImX, double *ImY, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    double rd[3];
    ImageFrame fr;
    uint32_t start_bdy;
    err = image_prepare(m, op, raydir_in, nx, ny, rd, &fr, ImX, ImY, &start_bdy);
    if (err) return err;
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    const size_t npix = (size_t)nx * ny;
    const uint32_t length = image_scratch_length(m, op, &fr, rd, ImX, ImY, npix, start_bdy);
#pragma omp parallel
    {
        Scratch s;
        const int serr = scratch_init(&s, length);
#pragma omp for schedule(static)
        for (size_t pix = 0; pix < npix; pix++) {
            if (serr) continue;
            const uint32_t cb = image_trace_pixel(m, op, &s, &fr, rd, ImX[pix], ImY[pix], start_bdy);
            for (uint32_t f = 0; f < m->nfreqs; f++)
                tau[pix * m->nfreqs + f] = (s.n_tot > 1) ? optical_depth_of(m, op, &s, cb, f) : 0.0;
        }
        scratch_free(&s);
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return 0;
}

/* Geometry::get_shift<Rest, false>(o, r, c, Z) :429-465, reversed (2 - shift) on the rays
   r >= hnrays, as the old imager walks them */
static double get_shift_rest_ray(const mgx_model *m, uint32_t o, uint32_t r, uint32_t c, double Z)
{
    const double sh = shift_rest(m, &m->position[3 * (size_t)o], ray_dir(m, o, r), c, Z);
    return (r >= m->hnrays) ? 2.0 - sh : sh;
}

/* trace_ray<Rest, false> :742-767 with set_data's all-lines interpolation count
   (:1003-1043).  Returns id1; *err set on an invalid next point. */
static uint32_t trace_ray_rest(const mgx_model *m, const mgx_options *op, Scratch *s, uint32_t o,
                               uint32_t r, int increment, uint32_t id1, uint32_t id2, int *err)
{
    double Z = 0.0, dZ = 0.0;
    uint32_t nxt = get_next(m, o, r, o, &Z, &dZ);
    if (nxt < m->npoints) {
        uint32_t crt = o;
        double shift_crt = get_shift_rest_ray(m, o, r, crt, 0.0);
        double shift_nxt = get_shift_rest_ray(m, o, r, nxt, Z);
        set_data_all(m, op, s, crt, nxt, shift_crt, shift_nxt, dZ, increment, &id1, &id2);
        while (not_on_boundary(m, nxt)) {
            crt = nxt;
            shift_crt = shift_nxt;
            nxt = get_next(m, o, r, nxt, &Z, &dZ);
            if (nxt >= m->npoints) {
                *err = MGX_E_NO_NEXT;
                return id1;
            }
            shift_nxt = get_shift_rest_ray(m, o, r, nxt, Z);
            set_data_all(m, op, s, crt, nxt, shift_crt, shift_nxt, dZ, increment, &id1, &id2);
        }
    }
    return id1;
}

/*
 * The old imager along model ray rr (image_feautrier_order_2(model, rr) :543-574 and
 * image_optical_depth(model, rr) :656-686), with the coordinates of
 * Image::set_coordinates_all_model_points (image.cpp:154-212).  optical_depth = 0:
 * out(p, f) = 2 Su[last] - I_bdy(nr[last], nu(p, f)); 1: the ray's optical depth.
 */
int oracle_image_ray(const mgx_model *m, const mgx_options *op, uint32_t rr, int optical_depth,
                     double *out, double *ImX, double *ImY, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    if (rr >= m->nrays) return MGX_E_ARG;
    if (op->approx == MGX_APPROX_ONE_LINE || m->adaptive_directions) return MGX_E_UNSUPPORTED;
    const double *rd = &m->ray_direction[3 * (size_t)rr];
    if (m->dimension == 1 && (rd[0] != 0.0 || rd[1] != 1.0 || rd[2] != 0.0)) return MGX_E_ARG;
    const double rx = rd[0], ry = rd[1], rz = rd[2];
    const double denominator = sqrt(rx * rx + ry * ry);
    const double inverse_denominator = 1.0 / denominator;
    const double ix = ry * inverse_denominator, iy = -rx * inverse_denominator;
    const double jx = rx * rz * inverse_denominator, jy = ry * rz * inverse_denominator;
    const double jz = -denominator;
    for (uint32_t p = 0; p < m->npoints; p++) {
        const double *x = &m->position[3 * (size_t)p];
        if (m->dimension == 1) {
            ImX[p] = x[0];
            ImY[p] = 0.0;
        } else if (m->dimension == 3 && denominator >= 1.0e-9) {
            ImX[p] = ix * x[0] + iy * x[1];
            ImY[p] = jx * x[0] + jy * x[1] + jz * x[2];
        } else if (m->dimension == 3) {
            ImX[p] = x[0];
            ImY[p] = x[1];
        } else {
            ImX[p] = 0.0;
            ImY[p] = 0.0;
        }
    }
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    const uint32_t length = 2 * get_ray_lengths_max(m, op, &err) + 1;
    const uint32_t ar = m->antipod[rr];
    const uint32_t F = m->nfreqs;
    if (err) goto done;
#pragma omp parallel
    {
        Scratch s;
        int serr = scratch_init(&s, length);
#pragma omp for schedule(static)
        for (uint32_t o = 0; o < m->npoints; o++) {
            if (serr) continue;
            int terr = 0;
            s.nr[s.centre] = o;
            s.shift[s.centre] = get_shift_rest_ray(m, o, rr, o, 0.0);
            s.first = trace_ray_rest(m, op, &s, o, rr, -1, s.centre - 1, s.centre - 1, &terr) + 1;
            s.last = trace_ray_rest(m, op, &s, o, ar, +1, s.centre + 1, s.centre, &terr) - 1;
            s.n_tot = (s.last + 1) - s.first;
            if (terr) {
                serr = terr;
                continue;
            }
            for (uint32_t f = 0; f < F; f++) {
                const double nu_o = m->nu[(size_t)o * F + f];
                double val;
                if (s.n_tot > 1)
                    val = optical_depth ? optical_depth_of(m, op, &s, o, f)
                                        : 2.0 * image_solve_of(m, op, &s, o, f) -
                                              boundary_intensity(m, s.nr[s.last], nu_o);
                else
                    val = optical_depth ? 0.0 : boundary_intensity(m, o, nu_o);
                out[(size_t)o * F + f] = val;
            }
        }
#pragma omp critical
        {
            if (serr && !err) err = serr;
        }
        scratch_free(&s);
    }
done:
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return err;
}

/*
 * The tridiagonal system behind one (rr, o, f) solve, for the brute-force
 * checks: dtau[j] of the segment between list entries j and j+1 and the row
 * sources S[j] (j = 0 .. n_tot-1, entry 0 = first), the boundary intensities,
 * plus u(centre) and the Lambda diagonal from the reference elimination.
 */
int oracle_feautrier_system(const mgx_model *m, const mgx_options *op, uint32_t rr, uint32_t o,
                            uint32_t f, uint32_t capacity, uint32_t *n_out, uint32_t *centre_out,
                            double *dtau, double *S, double *I_bdy, double *u_c, double *L_c)
{
    int err = check_model(m, op);
    if (err) return err;
    const uint32_t length = pair_length(m, op, o, rr, &err);
    if (err) return err;
    Scratch s;
    if (scratch_init(&s, length)) return MGX_E_ARG;
    const uint32_t l = m->f_line[f];
    const uint32_t ar = m->antipod[rr];
    s.nr[s.centre] = o;
    s.shift[s.centre] = 1.0;
    s.first = trace_ray_for_line(m, op, &s, l, o, rr, -1, s.centre - 1, s.centre - 1, &err) + 1;
    s.last = trace_ray_for_line(m, op, &s, l, o, ar, +1, s.centre + 1, s.centre, &err) - 1;
    s.n_tot = (s.last + 1) - s.first;
    *n_out = s.n_tot;
    *centre_out = s.centre - s.first;
    if (!err && (s.n_tot < 2 || s.n_tot > capacity)) err = MGX_E_RAY_LENGTH;
    if (!err) {
        const double freq = m->nu[(size_t)o * m->nfreqs + f];
        double dtau_n, chi_c = 0.0, chi_n = 0.0, term_c, term_n;
        int cco = 1;
        const uint32_t *nr = s.nr, *nri = s.nr_interp;
        for (uint32_t n = s.first; n < s.last; n++) {
            if (n > s.first) {
                term_c = term_n;
                chi_c = chi_n;
            }
            compute_source_dtau(m, op, nr[n], nr[n + 1], nri[n], nri[n + 1], l, freq * s.shift[n],
                                freq * s.shift[n + 1], s.shift[n], s.shift[n + 1], s.dZ[n],
                                s.interp_factor[n], s.interp_factor[n + 1], &cco, &dtau_n, &chi_c,
                                &chi_n, &term_c, &term_n);
            dtau[n - s.first] = dtau_n;
            S[n - s.first] = term_c;
        }
        S[s.n_tot - 1] = term_n;
        I_bdy[0] = boundary_intensity(m, nr[s.first], freq * s.shift[s.first]);
        I_bdy[1] = boundary_intensity(m, nr[s.last], freq * s.shift[s.last]);
        solve_feautrier_order_2_of(m, op, &s, o, f);
        *u_c = s.Su[s.centre];
        *L_c = s.L_diag[s.centre];
    }
    scratch_free(&s);
    return err;
}

/* ------------------------------------------------------------------------- */
/* adjoint: vector-Jacobian products of J and of images with respect to the   */
/* line emissivity and opacity (P, L).  No reference counterpart; the trace,  */
/* the erf / trapezoid branch of every segment, the CloseLines windows and    */
/* the interpolation counts are held fixed (they do not depend on the         */
/* emissivity, and depend on the opacity only through the discrete counts).   */

void oracle_set_smooth_interp(int on) { g_smooth_interp = on; }
void oracle_set_smooth_profile(int on) { g_smooth_profile = on; }

typedef struct {
    double *eps, *kap; /* (P, L) accumulators */
    double *iw;        /* (P, L) accumulator of the inverse widths, or NULL */
    uint32_t L;
} Grad;

static inline void grad_add(Grad *g, uint32_t p, uint32_t l, double de, double dk)
{
    g->eps[(size_t)p * g->L + l] += de;
    g->kap[(size_t)p * g->L + l] += dk;
}

/* reverse of interpolate_linear(iw(p), iw(pi), t) (partials 1 - t and t) */
static inline void iw_bwd(Grad *g, uint32_t p, uint32_t pi, uint32_t l, double t, double bar)
{
    if (!g->iw) return;
    g->iw[(size_t)p * g->L + l] += (1.0 - t) * bar;
    if (t != 0.0) g->iw[(size_t)pi * g->L + l] += t * bar;
}

/* d(gaussian(w, diff))/dw = exp(-x^2) (1 - 2 x^2) / sqrt(pi), x = w diff: the
   derivative of the smooth profile, in double */
static inline double gaussian_dw(double w, double diff)
{
    const double x = w * diff;
    return C_INVERSE_SQRT_PI * exp(-x * x) * (1.0 - 2.0 * x * x);
}

/* d(compute_dtau_single_line)/d(avg_iw) at fixed frequencies, of the smooth
   function: dz avg_op d(profile)/dw when the two frequencies are equal, else
   dz avg_op (erf(np) - erf(cp)) / (2 (cf - nf)) with np = (lf - nf) w and
   cp = (lf - cf) w, whose derivative is
   dz avg_op (np exp(-np^2) - cp exp(-cp^2)) / (sqrt(pi) (np - cp)) */
static double dtau_line_dw(double avg_iw, double avg_op, double linefreq, double cf, double nf,
                           double dz)
{
    if (cf == nf) return dz * avg_op * gaussian_dw(avg_iw, cf - linefreq);
    const double np = (linefreq - nf) * avg_iw, cp = (linefreq - cf) * avg_iw;
    return dz * avg_op * C_INVERSE_SQRT_PI * (np * exp(-np * np) - cp * exp(-cp * cp)) / (np - cp);
}

/* reverse of get_eta_and_chi_interp: eta_bar d(eta) + chi_bar d(chi) */
static void eta_chi_interp_bwd(const mgx_model *m, const mgx_options *op, uint32_t p1,
                               uint32_t p2, double t, uint32_t ll, double freq, double eta_bar,
                               double chi_bar, Grad *g)
{
    uint32_t lo = 0, hi = m->nlines;
    if (op->approx == MGX_APPROX_ONE_LINE) {
        lo = ll;
        hi = ll + 1;
    } else if (op->approx == MGX_APPROX_CLOSE_LINES) {
        const double W = op->max_distance_opacity_contribution * pw_upper(m, p1, freq);
        lo = lower_bound_d(m->sorted_line, m->nlines, freq - W);
        hi = upper_bound_d(m->sorted_line, m->nlines, freq + W);
    }
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = (op->approx == MGX_APPROX_CLOSE_LINES) ? m->sorted_line_map[s] : s;
        const double diff = freq - m->line[l];
        const double iw =
            interp_linear(m->inverse_width[LIDX(m, p1, l)], m->inverse_width[LIDX(m, p2, l)], t);
        const double prof = gaussian(iw, diff);
        const double e1 = m->emissivity[LIDX(m, p1, l)], e2 = m->emissivity[LIDX(m, p2, l)];
        const double k1 = m->opacity[LIDX(m, p1, l)], k2 = m->opacity[LIDX(m, p2, l)];
        double de1, de2, dk1, dk2;
        const double ei = interp_log(e1, e2, t), ki = interp_log(k1, k2, t);
        interp_log_partials(e1, e2, t, ei, &de1, &de2);
        interp_log_partials(k1, k2, t, ki, &dk1, &dk2);
        grad_add(g, p1, l, prof * eta_bar * de1, prof * chi_bar * dk1);
        if (t != 0.0) grad_add(g, p2, l, prof * eta_bar * de2, prof * chi_bar * dk2);
        /* eta = prof eps_i, chi = chi_min + prof kap_i */
        iw_bwd(g, p1, p2, l, t, gaussian_dw(iw, diff) * (eta_bar * ei + chi_bar * ki));
    }
}

/* the inverse-width part of the reverse of compute_dtau_single_line: bar is the
   adjoint of the line's optical depth */
static void dtau_line_iw_bwd(const mgx_model *m, uint32_t ci, uint32_t ni, uint32_t cii,
                             uint32_t nii, uint32_t l, double curr_freq, double next_freq,
                             double curr_interp, double next_interp, double dz, double bar,
                             Grad *g)
{
    if (!g->iw || bar == 0.0) return;
    const double iw_c =
        interp_linear(m->inverse_width[LIDX(m, ci, l)], m->inverse_width[LIDX(m, cii, l)], curr_interp);
    const double iw_n =
        interp_linear(m->inverse_width[LIDX(m, ni, l)], m->inverse_width[LIDX(m, nii, l)], next_interp);
    const double avg_iw = (iw_c + iw_n) / 2.0;
    const double op_c = interp_log(m->opacity[LIDX(m, ci, l)], m->opacity[LIDX(m, cii, l)], curr_interp);
    const double op_n = interp_log(m->opacity[LIDX(m, ni, l)], m->opacity[LIDX(m, nii, l)], next_interp);
    const double w_bar =
        0.5 * bar * dtau_line_dw(avg_iw, (op_c + op_n) / 2.0, m->line[l], curr_freq, next_freq, dz);
    iw_bwd(g, ci, cii, l, curr_interp, w_bar);
    iw_bwd(g, ni, nii, l, next_interp, w_bar);
}

/* d(compute_dtau_single_line)/d(avg_op) */
static double dtau_single_line_dop(const mgx_model *m, const mgx_options *op, uint32_t ci,
                                   uint32_t ni, uint32_t cii, uint32_t nii, uint32_t l,
                                   double curr_freq, double next_freq, double curr_interp,
                                   double next_interp, double dz)
{
    (void)op;
    const double linefreq = m->line[l];
    const double iw_c =
        interp_linear(m->inverse_width[LIDX(m, ci, l)], m->inverse_width[LIDX(m, cii, l)], curr_interp);
    const double iw_n =
        interp_linear(m->inverse_width[LIDX(m, ni, l)], m->inverse_width[LIDX(m, nii, l)], next_interp);
    const double avg_iw = (iw_c + iw_n) / 2.0;
    if (curr_freq == next_freq) return dz * gaussian(avg_iw, curr_freq - m->line[l]);
    const double next_pos = (linefreq - next_freq) * avg_iw;
    const double curr_pos = (linefreq - curr_freq) * avg_iw;
    const double diff_pos = next_pos - curr_pos;
    const double erf_diff = erf_difference(next_pos, curr_pos);
    return dz * (avg_iw * (erf_diff / diff_pos / 2.0));
}

/* reverse of interpolate_log(kappa(p), kappa(pi), t) */
static void opacity_bwd(const mgx_model *m, uint32_t p, uint32_t pi, uint32_t l, double t,
                        double bar, Grad *g)
{
    const double a = m->opacity[LIDX(m, p, l)], b = m->opacity[LIDX(m, pi, l)];
    double da, db;
    interp_log_partials(a, b, t, interp_log(a, b, t), &da, &db);
    grad_add(g, p, l, 0.0, bar * da);
    if (t != 0.0) grad_add(g, pi, l, 0.0, bar * db);
}

/* reverse of line_source: S = interpolate_log(eps/kap at p, eps/kap at pi, t) */
static void line_source_bwd(const mgx_model *m, uint32_t p, uint32_t pi, uint32_t l, double t,
                            double bar, Grad *g)
{
    const double kp = m->opacity[LIDX(m, p, l)], ki = m->opacity[LIDX(m, pi, l)];
    const double Rp = m->emissivity[LIDX(m, p, l)] / kp;
    const double Ri = m->emissivity[LIDX(m, pi, l)] / ki;
    double dRp, dRi;
    interp_log_partials(Rp, Ri, t, interp_log(Rp, Ri, t), &dRp, &dRi);
    const double bp = bar * dRp;
    grad_add(g, p, l, bp / kp, -bp * Rp / kp);
    if (t != 0.0) {
        const double bi = bar * dRi;
        grad_add(g, pi, l, bi / ki, -bi * Ri / ki);
    }
}

/* reverse of compute_S_dtau_line_integrated */
static void S_dtau_line_integrated_bwd(const mgx_model *m, const mgx_options *op, uint32_t cp,
                                       uint32_t np, uint32_t cpi, uint32_t npi, uint32_t lineidx,
                                       double currfreq, double nextfreq, double curr_interp,
                                       double next_interp, double dZ, double dtau_bar,
                                       double Sc_bar, double Sn_bar, Grad *g)
{
    if (op->approx == MGX_APPROX_ONE_LINE) {
        const double dop = dtau_single_line_dop(m, op, cp, np, cpi, npi, lineidx, currfreq,
                                                nextfreq, curr_interp, next_interp, dZ);
        opacity_bwd(m, cp, cpi, lineidx, curr_interp, 0.5 * dop * dtau_bar, g);
        opacity_bwd(m, np, npi, lineidx, next_interp, 0.5 * dop * dtau_bar, g);
        line_source_bwd(m, cp, cpi, lineidx, curr_interp, Sc_bar, g);
        line_source_bwd(m, np, npi, lineidx, next_interp, Sn_bar, g);
        dtau_line_iw_bwd(m, cp, np, cpi, npi, lineidx, currfreq, nextfreq, curr_interp,
                         next_interp, dZ, dtau_bar, g);
        return;
    }
    uint32_t lo = 0, hi = m->nlines;
    if (op->approx == MGX_APPROX_CLOSE_LINES) {
        double left_freq, right_freq;
        if (currfreq < nextfreq) {
            left_freq = currfreq;
            right_freq = nextfreq;
        } else {
            right_freq = currfreq;
            left_freq = nextfreq;
        }
        const double cb = op->max_distance_opacity_contribution * pw_upper(m, cp, right_freq);
        const double nb = op->max_distance_opacity_contribution * pw_upper(m, np, right_freq);
        const double W = (cb < nb) ? nb : cb;
        lo = lower_bound_d(m->sorted_line, m->nlines, left_freq - W);
        hi = upper_bound_d(m->sorted_line, m->nlines, right_freq + W);
    }
    double sum_dtau = 0.0, sum_Sc = 0.0, sum_Sn = 0.0;
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = (op->approx == MGX_APPROX_CLOSE_LINES) ? m->sorted_line_map[s] : s;
        const double line_dtau = compute_dtau_single_line(m, op, cp, np, cpi, npi, l, currfreq,
                                                          nextfreq, curr_interp, next_interp, dZ);
        sum_dtau += line_dtau;
        sum_Sc += line_dtau * line_source(m, cp, cpi, l, curr_interp);
        sum_Sn += line_dtau * line_source(m, np, npi, l, next_interp);
    }
    double d = sum_dtau;
    int clamped = 0;
    if (op->approx == MGX_APPROX_CLOSE_LINES) {
        const double bound_min_dtau = op->min_opacity * dZ;
        if (-bound_min_dtau < d && d < bound_min_dtau) {
            d = bound_min_dtau;
            clamped = 1;
        }
    }
    const double Scurr = sum_Sc / d, Snext = sum_Sn / d;
    const double sSc_bar = Sc_bar / d, sSn_bar = Sn_bar / d;
    const double sdt_bar = clamped ? 0.0 : dtau_bar - sSc_bar * Scurr - sSn_bar * Snext;
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t l = (op->approx == MGX_APPROX_CLOSE_LINES) ? m->sorted_line_map[s] : s;
        const double line_dtau = compute_dtau_single_line(m, op, cp, np, cpi, npi, l, currfreq,
                                                          nextfreq, curr_interp, next_interp, dZ);
        const double Sc = line_source(m, cp, cpi, l, curr_interp);
        const double Sn = line_source(m, np, npi, l, next_interp);
        /* adjoint of the line's optical depth: dtau_bar + (S_l - S) S_bar / d per side.
           With one line and no clamp S = S_l exactly, and the two source terms, which
           can exceed dtau_bar by many orders, cancel: take dtau_bar itself. */
        const double ld_bar = (hi == lo + 1 && !clamped) ? dtau_bar
                                                         : sdt_bar + sSc_bar * Sc + sSn_bar * Sn;
        const double dop = dtau_single_line_dop(m, op, cp, np, cpi, npi, l, currfreq, nextfreq,
                                                curr_interp, next_interp, dZ);
        opacity_bwd(m, cp, cpi, l, curr_interp, 0.5 * dop * ld_bar, g);
        opacity_bwd(m, np, npi, l, next_interp, 0.5 * dop * ld_bar, g);
        line_source_bwd(m, cp, cpi, l, curr_interp, sSc_bar * line_dtau, g);
        line_source_bwd(m, np, npi, l, next_interp, sSn_bar * line_dtau, g);
        dtau_line_iw_bwd(m, cp, np, cpi, npi, l, currfreq, nextfreq, curr_interp, next_interp,
                         dZ, ld_bar, g);
    }
}

/* reverse of compute_source_dtau for one segment: dtau_bar, Sc_bar, Sn_bar are the
   adjoints of its optical depth and of the two source values */
static void source_dtau_bwd(const mgx_model *m, const mgx_options *op, uint32_t cp, uint32_t np,
                            uint32_t cpi, uint32_t npi, uint32_t line, double curr_freq,
                            double next_freq, double curr_shift, double next_shift, double dZ,
                            double curr_interp, double next_interp, double dtau_bar,
                            double Sc_bar, double Sn_bar, Grad *g)
{
    const double dshift_abs = fabs(next_shift - curr_shift);
    const double a = m->dshift_max[cp], b = m->dshift_max[np];
    const double dshift_max = (b < a) ? b : a;
    if (dshift_abs > dshift_max) {
        S_dtau_line_integrated_bwd(m, op, cp, np, cpi, npi, line, curr_freq, next_freq,
                                   curr_interp, next_interp, dZ, dtau_bar, Sc_bar, Sn_bar, g);
        return;
    }
    double eta_c, chi_c, eta_n, chi_n;
    get_eta_and_chi_interp(m, op, cp, cpi, curr_interp, line, curr_freq, &eta_c, &chi_c);
    get_eta_and_chi_interp(m, op, np, npi, next_interp, line, next_freq, &eta_n, &chi_n);
    const double chi_c_bar = 0.5 * dZ * dtau_bar - Sc_bar * eta_c / (chi_c * chi_c);
    const double chi_n_bar = 0.5 * dZ * dtau_bar - Sn_bar * eta_n / (chi_n * chi_n);
    eta_chi_interp_bwd(m, op, cp, cpi, curr_interp, line, curr_freq, Sc_bar / chi_c, chi_c_bar, g);
    eta_chi_interp_bwd(m, op, np, npi, next_interp, line, next_freq, Sn_bar / chi_n, chi_n_bar, g);
}

/* per-thread work arrays of the ray adjoint */
typedef struct {
    double *d, *S, *F, *FI, *Su, *Sx, *u, *x, *lam, *dbar;
} AdjWork;

static int adjwork_init(AdjWork *w, uint32_t length)
{
    double **a[] = {&w->d, &w->S, &w->F, &w->FI, &w->Su, &w->Sx, &w->u, &w->x, &w->lam, &w->dbar};
    int ok = 1;
    for (size_t i = 0; i < sizeof a / sizeof a[0]; i++) {
        *a[i] = calloc(length, sizeof(double));
        ok = ok && *a[i];
    }
    return ok ? 0 : -1;
}

static void adjwork_free(AdjWork *w)
{
    double *a[] = {w->d, w->S, w->F, w->FI, w->Su, w->Sx, w->u, w->x, w->lam, w->dbar};
    for (size_t i = 0; i < sizeof a / sizeof a[0]; i++) free(a[i]);
}

/*
 * Vector-Jacobian product of one (o, f) solve on the traced ray in s: ubar is
 * dL/du at list entry c (absolute index).  The solve is T u = r with rows
 *   first: (1 + 2/d0) u0 + (2/d0^2)(u0 - u1) = S0 + 2 I0/d0,
 *   inner: u_n + A_n (u_n - u_{n-1}) + C_n (u_n - u_{n+1}) = S_n,
 *          A_n = 2/((d_{n-1}+d_n) d_{n-1}), C_n = 2/((d_{n-1}+d_n) d_n),
 *   last : the first row mirrored, with d_{N-2} and I_l
 * (solver.tpp:2234-2290).  The adjoint lam = T^-T e_c ubar is computed as
 * D T^-1 D^-1 e_c ubar, D = diag(d0/2, (d_{n-1}+d_n)/2, d_{N-2}/2) (D T is
 * symmetric), with T^-1 e_c from the reference's forward elimination and back
 * substitution.  Then dL/dS_n = lam_n and dL/dd_j = -sum_i lam_i dR_i/dd_j,
 * which the segment reverses carry to the emissivity and opacity.
 */
static void ray_adjoint(const mgx_model *m, const mgx_options *op, const Scratch *s, AdjWork *w,
                        uint32_t o, uint32_t f, uint32_t c_abs, double ubar, Grad *g)
{
    const uint32_t first = s->first, N = s->n_tot, c = c_abs - first;
    const uint32_t *nr = s->nr + first, *nri = s->nr_interp + first;
    const double *dZ = s->dZ + first, *shift = s->shift + first, *ifac = s->interp_factor + first;
    const double freq = m->nu[(size_t)o * m->nfreqs + f];
    const uint32_t l = m->f_line[f];
    double *d = w->d, *S = w->S, *F = w->F, *FI = w->FI, *Su = w->Su, *Sx = w->Sx, *u = w->u,
           *x = w->x, *lam = w->lam, *dbar = w->dbar;
    /* segments (compute_source_dtau with its carried state) */
    {
        double chi_c = 0.0, chi_n = 0.0, term_c = 0.0, term_n = 0.0;
        int cco = 1;
        for (uint32_t j = 0; j + 1 < N; j++) {
            if (j > 0) {
                term_c = term_n;
                chi_c = chi_n;
            }
            compute_source_dtau(m, op, nr[j], nr[j + 1], nri[j], nri[j + 1], l,
                                freq * shift[j], freq * shift[j + 1], shift[j], shift[j + 1],
                                dZ[j], ifac[j], ifac[j + 1], &cco, &d[j], &chi_c, &chi_n,
                                &term_c, &term_n);
            S[j] = term_c;
        }
        S[N - 1] = term_n;
    }
    const double I0 = boundary_intensity(m, nr[0], freq * shift[0]);
    const double Il = boundary_intensity(m, nr[N - 1], freq * shift[N - 1]);
    /* forward elimination for u and for x = T^-1 e_c */
    {
        const double idt = 1.0 / d[0];
        const double Cf = 2.0 * idt * idt;
        const double BmC = 1.0 + 2.0 * idt;
        const double B = BmC + Cf;
        Su[0] = (S[0] + 2.0 * I0 * idt) / B;
        Sx[0] = (c == 0) ? 1.0 / B : 0.0;
        F[0] = 0.5 * BmC * d[0] * d[0];
        FI[0] = 1.0 / (1.0 + F[0]);
    }
    for (uint32_t n = 1; n + 1 < N; n++) {
        const double da = 0.5 * (d[n - 1] + d[n]);
        const double iA = da * d[n - 1], iC = da * d[n];
        const double A = 1.0 / iA;
        F[n] = (A * F[n - 1] * FI[n - 1] + 1.0) * iC;
        FI[n] = 1.0 / (1.0 + F[n]);
        Su[n] = (A * Su[n - 1] + S[n]) * FI[n] * iC;
        Sx[n] = (A * Sx[n - 1] + (n == c ? 1.0 : 0.0)) * FI[n] * iC;
    }
    {
        const double idt = 1.0 / d[N - 2];
        const double Al = 2.0 * idt * idt;
        const double BmA = 1.0 + 2.0 * idt;
        const double Bl = BmA + Al;
        const double den = 1.0 / (Bl * F[N - 2] + BmA);
        u[N - 1] = (Al * Su[N - 2] + S[N - 1] + 2.0 * Il * idt) * (1.0 + F[N - 2]) * den;
        x[N - 1] = (Al * Sx[N - 2] + (c == N - 1 ? 1.0 : 0.0)) * (1.0 + F[N - 2]) * den;
    }
    for (long n = (long)N - 2; n >= 0; n--) {
        u[n] = Su[n] + FI[n] * u[n + 1];
        x[n] = Sx[n] + FI[n] * x[n + 1];
    }
    /* lam = D x ubar / D_c */
    for (uint32_t n = 0; n < N; n++) {
        const double Dn = (n == 0) ? 0.5 * d[0] : (n == N - 1) ? 0.5 * d[N - 2]
                                                              : 0.5 * (d[n - 1] + d[n]);
        lam[n] = Dn * x[n];
    }
    {
        const double Dc = (c == 0) ? 0.5 * d[0] : (c == N - 1) ? 0.5 * d[N - 2]
                                                              : 0.5 * (d[c - 1] + d[c]);
        const double sc = ubar / Dc;
        for (uint32_t n = 0; n < N; n++) lam[n] *= sc;
    }
    /* optical-depth adjoints from the rows.  The differences u_n - u_{n+1} are
       taken from the elimination, Su_n - F_n FI_n u_{n+1}: at optically thin
       steps u is constant to ~13 digits and the difference of two u values
       would keep only the last few. */
    double *du = Sx; /* reused: du[n] = u_n - u_{n+1} */
    for (uint32_t n = 0; n + 1 < N; n++) du[n] = Su[n] - F[n] * FI[n] * u[n + 1];
    for (uint32_t j = 0; j + 1 < N; j++) dbar[j] = 0.0;
    {
        const double dd = d[0];
        const double dR = (2.0 / (dd * dd)) * (I0 - u[0]) - (4.0 / (dd * dd * dd)) * du[0];
        dbar[0] -= lam[0] * dR;
    }
    for (uint32_t n = 1; n + 1 < N; n++) {
        const double dm = d[n - 1], dp = d[n], sm = dm + dp;
        const double A = 2.0 / (sm * dm), C = 2.0 / (sm * dp);
        const double gm = -du[n - 1], gp = du[n];
        const double dRdm = -A * (1.0 / sm + 1.0 / dm) * gm - (C / sm) * gp;
        const double dRdp = -(A / sm) * gm - C * (1.0 / sm + 1.0 / dp) * gp;
        dbar[n - 1] -= lam[n] * dRdm;
        dbar[n] -= lam[n] * dRdp;
    }
    {
        const double dd = d[N - 2];
        const double dR = (2.0 / (dd * dd)) * (Il - u[N - 1]) + (4.0 / (dd * dd * dd)) * du[N - 2];
        dbar[N - 2] -= lam[N - 1] * dR;
    }
    /* segments in reverse: row n takes its source from segment n (the last row from
       segment N-2's far value) */
    for (uint32_t j = 0; j + 1 < N; j++) {
        const double Sn_bar = (j + 2 == N) ? lam[N - 1] : 0.0;
        source_dtau_bwd(m, op, nr[j], nr[j + 1], nri[j], nri[j + 1], l, freq * shift[j],
                        freq * shift[j + 1], shift[j], shift[j + 1], dZ[j], ifac[j], ifac[j + 1],
                        dbar[j], lam[j], Sn_bar, g);
    }
}

static int grad_threads_init(Grad *g, uint32_t P, uint32_t L, int want_iw)
{
    g->L = L;
    g->eps = calloc((size_t)P * L, sizeof(double));
    g->kap = calloc((size_t)P * L, sizeof(double));
    g->iw = want_iw ? calloc((size_t)P * L, sizeof(double)) : NULL;
    return (g->eps && g->kap && (g->iw || !want_iw)) ? 0 : -1;
}

static void grad_reduce(const Grad *g, double *d_eps, double *d_kap, double *d_iw, size_t PL)
{
#pragma omp critical
    {
        for (size_t i = 0; i < PL; i++) {
            d_eps[i] += g->eps[i];
            d_kap[i] += g->kap[i];
            if (d_iw) d_iw[i] += g->iw[i];
        }
    }
}

int oracle_J_adjoint(const mgx_model *m, const mgx_options *op, const double *J_bar,
                     double *d_eps, double *d_kap, double *d_iw, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    const uint32_t length = 2 * get_ray_lengths_max(m, op, &err) + 1;
    const size_t PL = (size_t)m->npoints * m->nlines;
    if (err) goto done;
    memset(d_eps, 0, PL * sizeof(double));
    memset(d_kap, 0, PL * sizeof(double));
    if (d_iw) memset(d_iw, 0, PL * sizeof(double));
#pragma omp parallel
    {
        Scratch s;
        AdjWork w;
        Grad g = {0};
        int serr = scratch_init(&s, length);
        if (!serr) serr = adjwork_init(&w, length);
        if (!serr) serr = grad_threads_init(&g, m->npoints, m->nlines, d_iw != NULL);
        for (uint32_t rr = 0; rr < m->hnrays; rr++) {
            const uint32_t ar = m->antipod[rr];
#pragma omp for schedule(static)
            for (uint32_t o = 0; o < m->npoints; o++) {
                if (serr) continue;
                s.nr[s.centre] = o;
                s.shift[s.centre] = 1.0;
                for (uint32_t l = 0; l < m->nlines; l++) {
                    int terr = 0;
                    s.first = trace_ray_for_line(m, op, &s, l, o, rr, -1, s.centre - 1,
                                                 s.centre - 1, &terr) + 1;
                    s.last = trace_ray_for_line(m, op, &s, l, o, ar, +1, s.centre + 1, s.centre,
                                                &terr) - 1;
                    s.n_tot = (s.last + 1) - s.first;
                    s.interp_factor[s.centre] = 0.0;
                    if (terr) serr = terr;
                    if (s.n_tot <= 1) continue;
                    const double w2 = 2.0 * ray_weight(m, o, rr);
                    for (uint32_t q = 0; q < m->nquads; q++) {
                        const uint32_t f = l * m->nquads + q;
                        const double jb = J_bar[(size_t)o * m->nfreqs + f];
                        if (jb == 0.0) continue;
                        ray_adjoint(m, op, &s, &w, o, f, s.centre, w2 * jb, &g);
                    }
                }
            }
        }
        if (!serr) grad_reduce(&g, d_eps, d_kap, d_iw, PL);
#pragma omp critical
        {
            if (serr && !err) err = serr;
        }
        scratch_free(&s);
        adjwork_free(&w);
        free(g.eps);
        free(g.kap);
        free(g.iw);
    }
done:
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return err;
}

int oracle_image_adjoint(const mgx_model *m, const mgx_options *op, const double *raydir_in,
                         uint32_t nx, uint32_t ny, const double *I_bar, double *d_eps,
                         double *d_kap, double *d_iw, int nthreads)
{
    int err = check_model(m, op);
    if (err) return err;
    const size_t npix = (size_t)nx * ny;
    double *ImX = calloc(npix ? npix : 1, sizeof(double));
    double *ImY = calloc(npix ? npix : 1, sizeof(double));
    if (!ImX || !ImY) {
        free(ImX);
        free(ImY);
        return MGX_E_ARG;
    }
    double rd[3];
    ImageFrame fr;
    uint32_t start_bdy = 0;
    err = image_prepare(m, op, raydir_in, nx, ny, rd, &fr, ImX, ImY, &start_bdy);
    const size_t PL = (size_t)m->npoints * m->nlines;
    const int nt = default_threads(nthreads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nt);
#endif
    if (!err) {
        memset(d_eps, 0, PL * sizeof(double));
        memset(d_kap, 0, PL * sizeof(double));
        if (d_iw) memset(d_iw, 0, PL * sizeof(double));
        const uint32_t length = image_scratch_length(m, op, &fr, rd, ImX, ImY, npix, start_bdy);
#pragma omp parallel
        {
            Scratch s;
            AdjWork w;
            Grad g = {0};
            int serr = scratch_init(&s, length);
            if (!serr) serr = adjwork_init(&w, length);
            if (!serr) serr = grad_threads_init(&g, m->npoints, m->nlines, d_iw != NULL);
#pragma omp for schedule(static)
            for (size_t pix = 0; pix < npix; pix++) {
                if (serr) continue;
                const uint32_t cb =
                    image_trace_pixel(m, op, &s, &fr, rd, ImX[pix], ImY[pix], start_bdy);
                if (s.n_tot <= 1) continue;
                for (uint32_t f = 0; f < m->nfreqs; f++) {
                    const double ib = I_bar[pix * m->nfreqs + f];
                    if (ib != 0.0) ray_adjoint(m, op, &s, &w, cb, f, s.last, 2.0 * ib, &g);
                }
            }
            if (!serr) grad_reduce(&g, d_eps, d_kap, d_iw, PL);
#pragma omp critical
            {
                if (serr && !err) err = serr;
            }
            scratch_free(&s);
            adjwork_free(&w);
            free(g.eps);
            free(g.kap);
            free(g.iw);
        }
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    free(ImX);
    free(ImY);
    return err;
}

/* exported transcendentals, for the tests that pin them */
/* per-chunk digests of (argument, result) over float arguments [first, first +
   nchunks 2^chunk_log2), with this build's float functions (mgx_fmath_ref.h's
   mgx_fm_digest; NaN results canonicalised): the host side of the device
   comparison mgx_fmath_digest */
static uint32_t o_eval_bits(int fn, uint32_t xb, float y)
{
    float x, r;
    memcpy(&x, &xb, 4);
    switch (fn) {
    case MGX_FN_EXPF: r = O_EXPF(x); break;
    case MGX_FN_EXPM1F: r = O_EXPM1F(x); break;
    case MGX_FN_ERFF: r = O_ERFF(x); break;
    case MGX_FN_LOGF: r = O_LOGF(x); break;
    default: r = O_POWF(x, y); break;
    }
    if (r != r) return 0x7fc00000u;
    uint32_t rb;
    memcpy(&rb, &r, 4);
    return rb;
}

int oracle_fmath_digest(int fn, float y, uint64_t first, uint32_t nchunks, uint32_t chunk_log2,
                        uint64_t *digest, int nthreads)
{
    if (fn < MGX_FN_EXPF || fn > MGX_FN_POWF || !digest || chunk_log2 > 32) return MGX_E_ARG;
    if (first + ((uint64_t)nchunks << chunk_log2) > (1ull << 32)) return MGX_E_ARG;
    const int nt = default_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t c = 0; c < (int64_t)nchunks; c++) {
        const uint64_t base = first + ((uint64_t)c << chunk_log2);
        uint64_t acc = 0;
        for (uint64_t i = 0; i < (1ull << chunk_log2); i++) {
            const uint32_t xb = (uint32_t)(base + i);
            acc += mgx_fm_digest(xb, o_eval_bits(fn, xb, y));
        }
        digest[c] = acc;
    }
    return 0;
}

float oracle_expf(float x) { return O_EXPF(x); }
float oracle_expm1f(float x) { return O_EXPM1F(x); }
float oracle_logf(float x) { return O_LOGF(x); }
float oracle_powf(float x, float y) { return O_POWF(x, y); }
float oracle_erff(float x) { return O_ERFF(x); }
