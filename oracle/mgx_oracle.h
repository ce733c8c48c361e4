/*
 * CPU ORACLE -- test infrastructure only (tests/, smoke(), bench.py's
 * cpu_baseline).  A C restatement of Magritte's plain-method formal solver;
 * see mgx_oracle.c for the reference functions it follows.
 */
#ifndef MGX_ORACLE_H
#define MGX_ORACLE_H

#include "mgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Model::compute_radiation_field_feautrier_order_2 (setup + solve). u may be NULL. */
int oracle_solve_feautrier_order_2(const mgx_model *m, const mgx_options *op, double *u,
                                   double *J, double *lambda_diag, int nthreads);
/* The same solve with the general Lambda operator (n_off_diag >= 0): for each
   receiving point p and line l a list of at most cap (emitting point, value)
   entries in the reference's order, lam_count (P, L), lam_nr and lam_val
   (P, L, cap); MGX_E_ARG if a list would exceed cap. */
int oracle_solve_feautrier_order_2_lambda(const mgx_model *m, const mgx_options *op, double *u,
                                          double *J, uint32_t cap, uint32_t *lam_count,
                                          uint32_t *lam_nr, double *lam_val, int nthreads);
/* Model::compute_radiation_field_feautrier_order_2_uv: u and v (hnrays, P, F).
   v_scale (may be NULL) receives the sum of the absolute terms v is formed from
   (|coeffmin u_c-1| + |coeffplus u_c+1| + |coeffzero u_c|, or |I_bdy| + |u| at an
   end): the size the rounding of v is relative to. */
int oracle_feautrier_order_2_uv(const mgx_model *m, const mgx_options *op, double *u, double *v,
                                double *v_scale, int nthreads);
/* Model::compute_radiation_field_shortchar_order_0: I (nrays, P, F), u and v
   (hnrays, P, F), J (P, F), lambda_diag (P, L); I, u, v, lambda_diag may be NULL. */
int oracle_shortchar_order_0(const mgx_model *m, const mgx_options *op, double min_dtau,
                             double *I, double *u, double *v, double *J, double *lambda_diag,
                             int nthreads);
/* n_tot(rr, o, l) for every ray pair, origin and line, (hnrays, P, L). */
int oracle_trace_counts(const mgx_model *m, const mgx_options *op, uint32_t *n_tot, int nthreads);
/* expanded list of one (rr, o, l); same layout as mgx_trace_rays_device. */
int oracle_trace_ray(const mgx_model *m, const mgx_options *op, uint32_t rr, uint32_t o,
                     uint32_t l, uint32_t capacity, uint32_t *count, uint32_t *centre_idx,
                     uint32_t *nr, uint32_t *nr_interp, double *shift, double *interp_factor,
                     double *dZ);
/* Model::compute_image_new(raydir, nx, ny): I (nx*ny, F), ImX, ImY (nx*ny). */
int oracle_image_new(const mgx_model *m, const mgx_options *op, const double *raydir, uint32_t nx,
                     uint32_t ny, double *I, double *ImX, double *ImY, int nthreads);
/* Model::compute_image_optical_depth_new: tau (nx*ny, F) on the pixel rays of
   oracle_image_new, ImX, ImY (nx*ny). */
int oracle_image_optical_depth_new(const mgx_model *m, const mgx_options *op,
                                   const double *raydir, uint32_t nx, uint32_t ny, double *tau,
                                   double *ImX, double *ImY, int nthreads);
/* Model::compute_image(rr) (optical_depth = 0) or compute_image_optical_depth(rr) (1):
   out (P, F) and the points' image coordinates ImX, ImY (P). */
int oracle_image_ray(const mgx_model *m, const mgx_options *op, uint32_t rr, int optical_depth,
                     double *out, double *ImX, double *ImY, int nthreads);
uint32_t oracle_closest_bdy_point(const mgx_model *m, const double *raydir);
/* tridiagonal system of one (rr, o, f) solve: dtau (n-1), S (n), I_bdy (2), u_c, L_c */
int oracle_feautrier_system(const mgx_model *m, const mgx_options *op, uint32_t rr, uint32_t o,
                            uint32_t f, uint32_t capacity, uint32_t *n_out, uint32_t *centre_out,
                            double *dtau, double *S, double *I_bdy, double *u_c, double *L_c);

/* Vector-Jacobian products (adjoints) with respect to emissivity, opacity and
   the inverse line widths (P, L): d_eps = (dJ/d eps)^T J_bar and d_kap, d_iw
   likewise (d_iw may be NULL); the image variant takes I_bar (nx*ny, F).  The
   outputs are overwritten.  The inverse-width derivative is that of the smooth
   profile and error function (the reference evaluates them in float). */
int oracle_J_adjoint(const mgx_model *m, const mgx_options *op, const double *J_bar,
                     double *d_eps, double *d_kap, double *d_iw, int nthreads);
int oracle_image_adjoint(const mgx_model *m, const mgx_options *op, const double *raydir,
                         uint32_t nx, uint32_t ny, const double *I_bar, double *d_eps,
                         double *d_kap, double *d_iw, int nthreads);
/* 1: interpolate_log in double (smooth; for finite-difference checks), 0: as the reference */
void oracle_set_smooth_interp(int on);
/* 1: line profile and error-function difference in double (smooth in the inverse
   widths; for finite-difference checks), 0: as the reference */
void oracle_set_smooth_profile(int on);

/* Sparse level-population update in long double (mgx_oracle_nlte.c).  The
   handle copies every array it is given; m is a host model view (sizes,
   temperature, quad_weights, ray_weight).  Emissivity, opacity and the
   interpolation criterion outputs are (P, L) and may be NULL.  Return codes:
   0, -1 allocation, -2 no populations set, -3 Ng order beyond the history. */
typedef struct oracle_nlte oracle_nlte;
oracle_nlte *oracle_nlte_create(const mgx_model *m, uint32_t nlspecs, const mgx_linedata *ld,
                                const uint32_t *ncolpar, const mgx_colldata *colpar,
                                uint32_t nspec, const double *abundance,
                                const mgx_nlte_params *prm);
void oracle_nlte_destroy(oracle_nlte *h);
int oracle_nlte_set_lte(oracle_nlte *h, double *eps, double *kap, double *ic);
int oracle_nlte_set_populations(oracle_nlte *h, const double *pop, double *eps, double *kap,
                                double *ic);
/* J_lin from the plain solve's u (hnrays, P, F), J_eff with lambda_diag (P, L) */
int oracle_nlte_compute_Jeff(oracle_nlte *h, const double *u, const double *lambda_diag);
/* J_lin given as (P, L) */
int oracle_nlte_set_Jeff(oracle_nlte *h, const double *Jlin, const double *lambda_diag);
int oracle_nlte_stateq(oracle_nlte *h, const double *lambda_diag, double *eps, double *kap,
                       double *ic);
int oracle_nlte_ng(oracle_nlte *h, uint32_t order, double *eps, double *kap, double *ic);
int oracle_nlte_trial_ng(oracle_nlte *h, uint32_t order);
/* populations and trial populations (species blocks of (P, nlev)), J_lin, J_eff (P, L) */
int oracle_nlte_get(const oracle_nlte *h, double *pop, double *trial, double *Jlin, double *Jeff);
/* conv: (nlspecs, 3) fraction_not_converged, relative_change_mean, relative_change_max */
int oracle_nlte_stats(const oracle_nlte *h, double *conv, uint32_t *n_residuals);
/* one point's statistical-equilibrium system (column-major nlev x nlev) and right side */
int oracle_nlte_stateq_system(const oracle_nlte *h, const double *lambda_diag, uint32_t l,
                              uint32_t p, double *M, double *y);
/* Model::compute_Jeff with the general Lambda operator: J (P, F) and the lists of
   oracle_solve_feautrier_order_2_lambda */
int oracle_nlte_compute_Jeff_general(oracle_nlte *h, const double *J, uint32_t cap,
                                     const uint32_t *lam_count, const uint32_t *lam_nr,
                                     const double *lam_val);
/* the global statistical-equilibrium update of all points (dense long-double solve) */
int oracle_nlte_stateq_global(oracle_nlte *h, uint32_t cap, const uint32_t *lam_count,
                              const uint32_t *lam_nr, const double *lam_val, double *eps,
                              double *kap, double *ic);
/* its system for species l: M ((P nlev)^2, column-major) and y */
int oracle_nlte_stateq_global_system(const oracle_nlte *h, uint32_t cap,
                                     const uint32_t *lam_count, const uint32_t *lam_nr,
                                     const double *lam_val, uint32_t l, double *M, double *y);
/* long-double column-pivoted QR solve of an n x n column-major system */
int oracle_cpqr_solve(int n, const double *a_colmajor, const double *b, double *x);

int oracle_fmath_digest(int fn, float y, uint64_t first, uint32_t nchunks, uint32_t chunk_log2,
                        uint64_t *digest, int nthreads);
float oracle_expf(float x);
float oracle_expm1f(float x);
float oracle_logf(float x);
float oracle_powf(float x, float y);
float oracle_erff(float x);

#ifdef __cplusplus
}
#endif

#endif
