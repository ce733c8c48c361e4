/*
 * mgx.h - C ABI of the MI355X (gfx950) Feautrier formal solver.
 *
 * The library replaces the plain-method hot path of Magritte:
 *   Model::compute_radiation_field_feautrier_order_2   src/model/model.cpp:353-390
 *     -> Solver::setup<CoMoving,adaptive>              src/solver/solver.tpp:1-13, 32-126
 *     -> Solver::solve_feautrier_order_2<approx,adapt> src/solver/solver.tpp:459-540
 *        (trace_ray_for_line :774-800, set_data_for_line :1057-1097,
 *         solve_feautrier_order_2(o,f) :2194-2356, update_Lambda :1761-1850)
 *   Model::compute_image_new                           src/model/model.cpp:1008-1035
 *     -> Solver::image_feautrier_order_2_new_imager     src/solver/solver.tpp:577-627
 *
 * Conventions
 *   - Plain C: fixed-width integers, plain pointers, no exceptions cross it.
 *   - Every entry point returns 0 on success and a negative MGX_E* code on
 *     failure; mgx_last_error() then returns a message for the calling thread.
 *   - Arrays are row-major and mirror the reference's buffers
 *     (src/bindings/pybindings.cpp:680-822):  position (P,3), u (hnrays,P,F),
 *     J (P,F), image I (Nx*Ny, F) with pixel index y + Ny*x.
 *   - The Lambda diagonal is returned densely as (P, nlines): the entry
 *     (p, l) with l = nrad_cum[species] + k is the sum the reference collects
 *     in LineProducingSpecies::lambda Ls[p][k] for nr == p
 *     (src/model/lines/lineProducingSpecies/lambda/lambda.tpp:79-89) when
 *     n_off_diag == 0.
 *   - Entry points named mgx_*_device take DEVICE pointers for every array in
 *     the model and the outputs, and an optional hipStream_t (NULL = default
 *     stream).  They do not synchronise the stream.
 */
#ifndef MGX_H
#define MGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MGX_OK = 0,
    MGX_E_ARG = -1,        /* invalid argument (sizes, null pointers, options) */
    MGX_E_HIP = -2,        /* HIP runtime error */
    MGX_E_RAY_LENGTH = -3, /* a traced ray exceeds the kernel's staging capacity */
    MGX_E_NO_NEXT = -4,    /* an interior point has no valid next point on a ray
                              (reference prints "ERROR: no valid neighbor",
                              src/solver/solver.tpp:96-98) */
    MGX_E_UNSUPPORTED = -5 /* option combination not implemented */
};

enum { MGX_APPROX_NONE = 0, MGX_APPROX_ONE_LINE = 1, MGX_APPROX_CLOSE_LINES = 2 };
enum { MGX_BC_ZERO = 0, MGX_BC_THERMAL = 1, MGX_BC_CMB = 2 };
enum { MGX_PREC_FP64 = 0, MGX_PREC_MIXED = 1 };
enum { MGX_MAX_OFF_DIAG = 4 }; /* largest n_off_diag of mgx_feautrier_forward_lambda_device */

/*
 * Model view.  Mirrors the data the reference's Solver reads from Model&
 * (src/model/geometry/, src/model/lines/, src/model/thermodynamics/,
 * src/model/radiation/frequencies/).  All arrays are borrowed.
 */
typedef struct mgx_model {
    uint32_t npoints, nrays, hnrays, nboundary;
    uint32_t nlines, nlspecs, nquads, nfreqs;
    int32_t dimension;           /* 1 or 3 (Parameters::dimension) */
    int32_t spherical_symmetry;  /* Parameters::spherical_symmetry */
    int32_t adaptive_directions; /* Rays::use_adaptive_directions */
    int32_t reserved0;

    /* geometry: Points / Rays / Boundary */
    const double *position;            /* (P,3) [m] */
    const double *velocity;            /* (P,3) [fraction of c] */
    const uint32_t *n_neighbors;       /* (P) */
    const uint32_t *cum_n_neighbors;   /* (P) exclusive scan of n_neighbors */
    const uint32_t *neighbors;         /* (sum n_neighbors) */
    const uint32_t *point2boundary;    /* (P) boundary index, or P for interior */
    const uint32_t *boundary2point;    /* (nboundary) */
    const int32_t *boundary_condition; /* (nboundary) MGX_BC_* */
    const double *boundary_temperature;/* (nboundary) [K] */
    const double *ray_direction;       /* (nrays,3) or (P,nrays,3) if adaptive */
    const double *ray_weight;          /* (nrays) or (P,nrays) if adaptive */
    const uint32_t *antipod;           /* (nrays) */
    double center[3];                  /* Points::center, bounding-box centre */

    /* thermodynamics */
    const double *temperature; /* (P) gas temperature [K] */
    const double *vturb2;      /* (P) (turbulent velocity / c)^2 */

    /* lines (global line index l = nrad_cum[species] + k) */
    const double *line;              /* (L) line centre frequency [Hz] */
    const double *sorted_line;       /* (L) ascending (Lines::sorted_line) */
    const uint32_t *sorted_line_map; /* (L) sorted -> line index */
    const double *emissivity;        /* (P,L) */
    const double *opacity;           /* (P,L) */
    const double *inverse_width;     /* (P,L) [1/Hz] */
    const double *interp_crit;       /* (P,L) opacity / population_tot
                                        (InterpHelper::interpolation_criterion) */
    const double *dshift_max;        /* (P) Model::dshift_max */
    double max_inverse_mass;         /* Lines::max_inverse_mass [1/amu] */
    const uint32_t *line_species;    /* (L) species of line l */
    const uint32_t *line_k;          /* (L) transition index k of line l in its species */
    const double *line_A;            /* (L) Einstein A of line l */
    const double *species_inverse_mass; /* (nlspecs) [1/amu] */
    const double *quad_weights;      /* (nlspecs, nquads) */

    /* frequencies (Frequencies) */
    const double *nu;                /* (P,F) [Hz] */
    const uint32_t *f_line;          /* (F) corresponding_line */
    const uint32_t *f_species;       /* (F) corresponding_l_for_spec */
    const uint32_t *f_k;             /* (F) corresponding_k_for_tran */
    const uint32_t *f_z;             /* (F) corresponding_z_for_line */
    const uint8_t *f_in_line;        /* (F) appears_in_line_integral */
} mgx_model;

/* Numerical knobs of Parameters used on this path (src/model/parameters/parameters.hpp:66-124). */
typedef struct mgx_options {
    int32_t approx;      /* MGX_APPROX_*  (model.cpp:366-389 dispatch) */
    uint32_t n_off_diag; /* must be 0 */
    int32_t precision;   /* MGX_PREC_FP64 (bit-faithful) or MGX_PREC_MIXED */
    int32_t reserved0;
    double min_opacity;                       /* default 1e-26 */
    double max_distance_opacity_contribution; /* default 10 */
    double max_interpolation_diff;            /* default 1.4 */
} mgx_options;

/* Message of the last failed call on this thread ("" if none). */
const char *mgx_last_error(void);

/* Library version string. */
const char *mgx_version(void);

/*
 * Plain-method formal solve on the device: fills u (hnrays,P,F) [may be NULL],
 * J (P,F) and lambda_diag (P,L) [may be NULL].
 * Replaces Solver::solve_feautrier_order_2<approx,adaptive>(Model&)
 * (src/solver/solver.tpp:459-540) including Solver::setup (:1-13).
 */
int mgx_feautrier_forward_device(const mgx_model *model_dev, const mgx_options *opts,
                                 double *u, double *J, double *lambda_diag, void *stream);

/*
 * The memory-sparse formal solve: J_line (P, L) receives, for every point o and
 * line l (transition k of species s, l = nrad_cum[s] + k), the line-integrated
 * mean intensity sum_z w_z J(o, l nquads + z) with the quadrature weights of
 * species s -- lspec.J(o, k) -- and lambda_diag (P, L) [may be NULL] the Lambda
 * diagonal; neither u nor J (P, F) is stored.  The view must carry the line
 * frequency grid.  Replaces Solver::solve_feautrier_order_2_sparse
 * (src/solver/solver.tpp:297-363) behind Model::compute_radiation_field_
 * feautrier_order_2_sparse (src/model/model.cpp:474-515).
 */
int mgx_feautrier_sparse_device(const mgx_model *model_dev, const mgx_options *opts,
                                double *J_line, double *lambda_diag, void *stream);

/*
 * The plain-method solve of mgx_feautrier_forward_device (same kernels, same
 * results bit for bit) without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.
 * Its scratch (status words, launch parameters and the packed point and line
 * tables: 256-byte-aligned pieces of 16 B, the parameter block, 80 B per point
 * and 32 B per (point, line)) lives in a static device array of
 * MGX_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a model
 * needing more fails with MGX_E_ARG (e.g. 64 MiB holds the 32^3 wind with 9
 * lines, 12 MB).  The caller passes the inputs and outputs as for
 * mgx_feautrier_forward_device.  Every call on a GPU shares that GPU's array,
 * which keeps the last call's contents: the calls are serialised by a lock,
 * and work that overlaps on one GPU through other means races.  No kernel
 * timing (it would create events).
 */
#define MGX_NOALLOC_SCRATCH_BYTES (64u << 20)
int mgx_feautrier_forward_noalloc(const mgx_model *model_dev, const mgx_options *opts, double *u,
                                  double *J, double *lambda_diag, void *stream);

/*
 * The plain-method formal solve with the general Lambda operator: u
 * (hnrays,P,F) [may be NULL] and J (P,F) as mgx_feautrier_forward_device, and
 * for each receiving point p and line l the list of approximated-Lambda
 * elements (emitting point, value): lam_count (P,L), lam_nr and lam_val
 * (P,L,cap), entries [0, lam_count) in the reference's order, the diagonal
 * element (p itself) first.  With opts->n_off_diag = m the list holds the
 * diagonal and the elements of the m rows before and after p on every ray
 * (an interpolated row counts for its step's next point); a capacity of
 * 1 + 2 m hnrays always suffices.  Replaces Solver::solve_feautrier_order_2
 * with n_off_diag > 0 (src/solver/solver.tpp:459-540, 2319-2355),
 * Solver::update_Lambda (:1761-1850) and Lambda::add_element
 * (src/model/lines/lineProducingSpecies/lambda/lambda.tpp).  Errors:
 * n_off_diag > MGX_MAX_OFF_DIAG -> MGX_E_UNSUPPORTED; a list longer than cap ->
 * MGX_E_ARG.
 */
int mgx_feautrier_forward_lambda_device(const mgx_model *model_dev, const mgx_options *opts,
                                        double *u, double *J, uint32_t cap, uint32_t *lam_count,
                                        uint32_t *lam_nr, double *lam_val, void *stream);

/*
 * Order-0 short-characteristics formal solve on the device: fills I
 * (nrays, P, F) [may be NULL; I(rr) and I(antipod(rr)) are the intensities
 * arriving at each point from the rr and the antipod side], u = (I_f + I_b)/2
 * and v = (I_f - I_b)/2 (hnrays, P, F) [may be NULL], J (P, F) and
 * lambda_diag (P, L) [may be NULL: the elements of both directions].
 * min_dtau is Parameters::min_dtau (default 1e-15).  Replaces
 * Model::compute_radiation_field_shortchar_order_0 (src/model/model.cpp:313-351)
 * -> Solver::solve_shortchar_order_0<approx, adaptive> (src/solver/solver.tpp:
 * 151-235, 1565-1757).  A direction without a segment adds no Lambda element
 * (the reference reads an uninitialised optical depth there).
 */
int mgx_shortchar_order_0_device(const mgx_model *model_dev, const mgx_options *opts,
                                 double min_dtau, double *I, double *u, double *v, double *J,
                                 double *lambda_diag, void *stream);

/*
 * Second-order Feautrier solve with v on the device: u and v (hnrays, P, F),
 * v = -du/dtau at each point by the second-order difference over its two
 * optical-depth increments (the boundary difference at a ray's end).
 * Replaces Model::compute_radiation_field_feautrier_order_2_uv
 * (src/model/model.cpp:395-433) -> Solver::solve_feautrier_order_2_uv
 * (src/solver/solver.tpp:238-295, 2767-2921); like the reference it computes
 * neither J nor Lambda.
 */
int mgx_feautrier_order_2_uv_device(const mgx_model *model_dev, const mgx_options *opts,
                                    double *u, double *v, void *stream);

/*
 * Traced point counts n_tot(rr, o, l) = last + 1 - first for every ray pair,
 * origin and line, written as (hnrays, P, L) uint32.
 * Same trace as Solver::trace_ray_for_line (src/solver/solver.tpp:774-800,
 * 487-495).
 */
int mgx_trace_counts_device(const mgx_model *model_dev, const mgx_options *opts,
                            uint32_t *n_tot, void *stream);

/*
 * Expanded traced lists for nsel selected (rr, o, l) triples.  Triple i
 * writes its entries first..last (first at index 0) into row i of the
 * (nsel, capacity) arrays; centre_idx[i] = centre - first, count[i] = n_tot.
 * Entries are nr, nr_interp, shift, interp_factor and dZ exactly as the
 * reference's per-thread scratch (src/solver/solver.tpp:1057-1097); dZ[j] is
 * the increment between entries j and j+1 (its last element is unused).
 */
int mgx_trace_rays_device(const mgx_model *model_dev, const mgx_options *opts, uint32_t nsel,
                          const uint32_t *sel_rr, const uint32_t *sel_o, const uint32_t *sel_l,
                          uint32_t capacity, uint32_t *count, uint32_t *centre_idx,
                          uint32_t *nr, uint32_t *nr_interp, double *shift,
                          double *interp_factor, double *dZ, void *stream);


/*
 * Projection-surface image (the "new imager"): one ray per pixel from a plane
 * behind the model along raydir, forward Feautrier elimination per image
 * frequency, I(pix, f) = 2 Su[last] - I_bdy.  Replaces
 * Model::compute_image_new(raydir, Nx, Ny) (src/model/model.cpp:1008-1035) ->
 * Solver::image_feautrier_order_2_new_imager (src/solver/solver.tpp:577-627),
 * Image::set_coordinates_projection_surface (src/model/image/image.cpp:221-366).
 * raydir (host, 3 doubles) need not be normalised (it is, as in the
 * reference); the model's frequencies are the image frequencies (nu (P, F)).
 * Outputs (device): I (nx*ny, F) with pixel index y + ny*x, ImX, ImY (nx*ny).
 * Errors: zero direction -> MGX_E_ARG; OneLine approximation ->
 * MGX_E_UNSUPPORTED; 1D model with raydir != (0,1,0) -> MGX_E_ARG.
 */
int mgx_image_device(const mgx_model *model_dev, const mgx_options *opts, const double *raydir,
                     uint32_t nx, uint32_t ny, double *I, double *ImX, double *ImY, void *stream);

/*
 * Optical-depth image on the projection surface: the same pixel rays as
 * mgx_image_device, tau(pix, f) = the sum of the optical-depth increments
 * along the ray at the start point's frequencies (0 for a ray of one point).
 * Replaces Model::compute_image_optical_depth_new(raydir, Nx, Ny)
 * (src/model/model.cpp:1102-1127) -> Solver::image_optical_depth_new_imager
 * (src/solver/solver.tpp:687-736) -> image_optical_depth(o, f) (:2718-2758).
 * Outputs and errors as mgx_image_device.
 */
int mgx_image_optical_depth_device(const mgx_model *model_dev, const mgx_options *opts,
                                   const double *raydir, uint32_t nx, uint32_t ny, double *tau,
                                   double *ImX, double *ImY, void *stream);

/*
 * The old imager: an image at every model point along model ray rr
 * (0 <= rr < nrays).  For each point the ray pair (rr, antipod) is traced with
 * the Rest-frame shifts (2 - shift on rays >= hnrays) and the all-lines
 * interpolation count, the Feautrier system is eliminated from the rr end to
 * the antipod end, and I(p, f) = 2 Su[last] - I_bdy(nr[last], nu(p, f)), the
 * boundary term at the unshifted frequency (I_bdy(p, nu) for a ray of one
 * point).  Outputs (device): I (P, F), ImX, ImY (P), the points' coordinates
 * on the plane normal to the ray.  Replaces Model::compute_image(ray_nr)
 * (src/model/model.cpp:957-976) -> Solver::image_feautrier_order_2(model, rr)
 * (src/solver/solver.tpp:543-574, 742-767, 2463-2560) and
 * Image::set_coordinates_all_model_points (src/model/image/image.cpp:154-212).
 * Errors: OneLine -> MGX_E_UNSUPPORTED; adaptive directions ->
 * MGX_E_UNSUPPORTED; 1D model with ray rr != (0,1,0) -> MGX_E_ARG.
 */
int mgx_image_ray_device(const mgx_model *model_dev, const mgx_options *opts, uint32_t rr,
                         double *I, double *ImX, double *ImY, void *stream);

/*
 * Optical-depth image of the old imager: tau(p, f), the optical depth of the
 * whole ray pair through p along ray rr (0 for a ray of one point).  Replaces
 * Model::compute_image_optical_depth(ray_nr) (src/model/model.cpp:1061-1076)
 * -> Solver::image_optical_depth(model, rr) (src/solver/solver.tpp:656-686,
 * 2718-2758).  Outputs and errors as mgx_image_ray_device.
 */
int mgx_image_optical_depth_ray_device(const mgx_model *model_dev, const mgx_options *opts,
                                       uint32_t rr, double *tau, double *ImX, double *ImY,
                                       void *stream);

/* ------------------------------------------------------------------------ */
/* Adjoint (reverse mode).  No counterpart in the reference: the gradients   */
/* of the formal solve with respect to the line emissivity and opacity       */
/* (model_dev->emissivity / ->opacity, (P, L)), for differentiable use of    */
/* the two entry points above.  The trace, the erf / trapezoid choice of     */
/* every segment, the CloseLines windows and the interpolation counts are    */
/* held fixed; the float rounding steps of interpolate_log's powf are        */
/* differentiated as the smooth a^(1-t) b^t.                                 */
/* ------------------------------------------------------------------------ */

/*
 * Vector-Jacobian product of mgx_feautrier_forward_device's J:
 * d_eps = (dJ/d emissivity)^T J_bar, d_kap = (dJ/d opacity)^T J_bar.
 * J_bar (P, F), d_eps and d_kap (P, L), all device; the outputs are
 * overwritten.  Uses a device workspace kept between calls (grow-only, see
 * mgx_release_workspace); calls are serialised on it.
 */
int mgx_feautrier_adjoint_device(const mgx_model *model_dev, const mgx_options *opts,
                                 const double *J_bar, double *d_eps, double *d_kap, void *stream);

/*
 * Vector-Jacobian product of mgx_image_device's I: I_bar (nx*ny, F),
 * d_eps and d_kap (P, L), all device; the outputs are overwritten.
 */
int mgx_image_adjoint_device(const mgx_model *model_dev, const mgx_options *opts,
                             const double *raydir, uint32_t nx, uint32_t ny, const double *I_bar,
                             double *d_eps, double *d_kap, void *stream);

/*
 * The two products above with a third output, d_inverse_width (P, L), device:
 * the gradient with respect to model_dev->inverse_width, through which the
 * temperature enters the line profile (Lines::set_inverse_width,
 * src/model/lines/lines.tpp:45-60).  The Gaussian profile and the difference
 * of error functions of compute_dtau_single_line (solver.tpp:2374-2455) are
 * differentiated as the smooth functions (the reference evaluates them with
 * float expf / erff).  d_eps and d_kap are the same as without it.
 */
int mgx_feautrier_adjoint_width_device(const mgx_model *model_dev, const mgx_options *opts,
                                       const double *J_bar, double *d_eps, double *d_kap,
                                       double *d_inverse_width, void *stream);
int mgx_image_adjoint_width_device(const mgx_model *model_dev, const mgx_options *opts,
                                   const double *raydir, uint32_t nx, uint32_t ny,
                                   const double *I_bar, double *d_eps, double *d_kap,
                                   double *d_inverse_width, void *stream);

/* Frees the adjoint's device workspace (it is re-allocated on the next call). */
void mgx_release_workspace(void);

/* ------------------------------------------------------------------------ */
/* Batched forward solve.  No counterpart in the reference (which solves one */
/* model per call): many models of uneven size and ray count in one          */
/* persistent launch, captured in a HIP graph, with an autotuner.            */
/* ------------------------------------------------------------------------ */

typedef struct mgx_batch mgx_batch;

/*
 * Plan of a batch: n device model views (models[i] as for
 * mgx_feautrier_forward_device, all with the same options), the device
 * outputs J[i] (P_i, F_i) and lambda_diag[i] (P_i, L_i) (lambda_diag may be
 * NULL).  Packs every model's tables and builds the work list (one item per
 * origin, or per block of ray pairs).  Results of mgx_batch_run equal those of
 * n calls of mgx_feautrier_forward_device (within 1e-15 relative when ray
 * pairs are split over items: their J contributions are added atomically).
 */
int mgx_batch_create(const mgx_model *models, uint32_t n, const mgx_options *opts,
                     double *const *J, double *const *lambda_diag, mgx_batch **out,
                     void *stream);

/* Sets the tiling by hand: ray pairs per work item (0 = all of an origin),
 * item order (0 = decreasing cost, 1 = model by model), register budget in
 * waves per SIMD (2, 3 or 4), and whether runs replay a captured HIP graph. */
int mgx_batch_configure(mgx_batch *batch, uint32_t rays_per_item, uint32_t order, uint32_t wps,
                        int use_graph, void *stream);

/* Autotuner: times the candidate tilings of the solve (without the images) on
 * every sample_stride-th work item and keeps the fastest; best (3 words, may be
 * NULL) receives it.  Leaves the outputs undefined until the next mgx_batch_run. */
int mgx_batch_tune(mgx_batch *batch, uint32_t sample_stride, uint32_t *best, void *stream);

/*
 * Adds an image of every model to the batch (Model::compute_image_new,
 * model.cpp:1008-1035, for each model; replaces any earlier images): model i is
 * imaged on nx x ny pixels along raydirs[3i..3i+2] at the frequencies of
 * image_views[i] -- the view models[i] with nu, f_* and nfreqs replaced by the
 * image's frequency grid (every geometry and line array must be the same
 * pointer as in models[i]) -- into I[i] (nx*ny, F_i), ImX[i] and ImY[i] (nx*ny).
 * mgx_batch_run then launches one persistent kernel over the pixels of all the
 * images after the solve, in the same graph.  Not with the one-line
 * approximation (as the reference).
 */
int mgx_batch_set_images(mgx_batch *batch, const mgx_model *image_views, const double *raydirs,
                         uint32_t nx, uint32_t ny, double *const *I, double *const *ImX,
                         double *const *ImY, void *stream);

/* Runs the batch (line tables, the persistent kernel, the images if set;
 * synchronises). */
int mgx_batch_run(mgx_batch *batch, void *stream);

/* Duration (ms) of the last run when kernel timing is enabled. */
double mgx_batch_last_ms(const mgx_batch *batch);

void mgx_batch_destroy(mgx_batch *batch);

/* ------------------------------------------------------------------------ */
/* Host-side producers of the path's inputs (plain host pointers).           */
/* ------------------------------------------------------------------------ */

/* Line data of one line-producing species
 * (src/model/lines/lineProducingSpecies/linedata/linedata.hpp). */
typedef struct mgx_linedata {
    uint32_t num;            /* species index in the abundance table */
    uint32_t nlev, nrad;
    uint32_t reserved0;
    double inverse_mass;     /* [1/amu] */
    const double *energy;    /* (nlev) [J] */
    const double *weight;    /* (nlev) statistical weights */
    const uint32_t *irad;    /* (nrad) upper level */
    const uint32_t *jrad;    /* (nrad) lower level */
    const double *A;         /* (nrad) [1/s] */
    const double *Bs;        /* (nrad) */
    const double *Ba;        /* (nrad) */
    const double *frequency; /* (nrad) [Hz] */
} mgx_linedata;

/* Lines::set_inverse_width (src/model/lines/lines.tpp:45-60): (P, L). */
int mgx_compute_inverse_widths(uint32_t npoints, uint32_t nlspecs, const mgx_linedata *ld,
                               const double *temperature, const double *vturb2,
                               double *inverse_width);

/* LineProducingSpecies::update_using_LTE (lineProducingSpecies.tpp:52-75) and
 * Lines::set_emissivity_and_opacity (lines.tpp:29-41), populations in long
 * double as the reference (VectorXld).  population: per species s a block of
 * (P, nlev_s) doubles, blocks concatenated in species order; population_tot
 * (nlspecs, P); emissivity, opacity (P, L).  abundance is (P, nspec). */
int mgx_compute_lte(uint32_t npoints, uint32_t nspec, const double *abundance,
                    const double *temperature, uint32_t nlspecs, const mgx_linedata *ld,
                    double min_line_emissivity, double min_line_opacity,
                    double *population, double *population_tot, double *emissivity,
                    double *opacity);

/* Model::compute_spectral_discretisation() (src/model/model.cpp:70-161):
 * nu (P, F) with F = nlines * nquads; roots (nlspecs, nquads); per-frequency
 * maps (F) as in Frequencies. */
int mgx_compute_spectral_discretisation(uint32_t npoints, uint32_t nlspecs,
                                        const mgx_linedata *ld, uint32_t nquads,
                                        const double *roots, const double *temperature,
                                        const double *vturb2, double *nu, uint32_t *f_line,
                                        uint32_t *f_species, uint32_t *f_k, uint32_t *f_z,
                                        uint8_t *f_in_line);

/* Model::compute_spectral_discretisation(nu_min, nu_max, n) (model.cpp:257-298):
 * the same n image frequencies at every point; maps set to "not in a line". */
int mgx_compute_image_discretisation(uint32_t npoints, double nu_min, double nu_max, uint32_t n,
                                     double *nu, uint32_t *f_line, uint32_t *f_species,
                                     uint32_t *f_k, uint32_t *f_z, uint8_t *f_in_line);

/* Model::set_dshift_max (model.cpp:1158-1178): (P). */
int mgx_compute_dshift_max(uint32_t npoints, uint32_t nlspecs, const mgx_linedata *ld,
                           const double *temperature, const double *vturb2,
                           double max_width_fraction, double *dshift_max);

/* InterpHelper constructor (solver_interp_helper.hpp:15-30): opacity / population_tot, (P, L). */
int mgx_compute_interp_crit(uint32_t npoints, uint32_t nlspecs, const mgx_linedata *ld,
                            const double *opacity, const double *population_tot,
                            double *interp_crit);

/* Lines::read line sorting (lines.cpp:44-64, heapsort tools/heapsort.hpp:64-88). */
int mgx_sort_lines(uint32_t nlines, const double *line, double *sorted_line,
                   uint32_t *sorted_line_map);

/* ------------------------------------------------------------------------ */
/* Level populations (non-LTE), the sparse variant: line-integrated J and    */
/* the Lambda diagonal of a formal solve -> J_eff -> per-point statistical   */
/* equilibrium -> negative-population correction -> convergence statistics  */
/* -> line emissivity / opacity / interpolation criterion, plus general-     */
/* order Ng acceleration.  Populations live on the device in FP64 (the       */
/* reference keeps them in long double).                                     */
/* ------------------------------------------------------------------------ */

/* One collision partner of a line-producing species
 * (src/model/lines/lineProducingSpecies/linedata/collisionPartner/collisionPartner.hpp:7-33). */
typedef struct mgx_colldata {
    uint32_t num_col_partner; /* abundance column of the partner */
    int32_t orth_or_para_H2;  /* 'n', 'o' or 'p' */
    uint32_t ntmp, ncol;
    const uint32_t *icol;     /* (ncol) upper level */
    const uint32_t *jcol;     /* (ncol) lower level */
    const double *tmp;        /* (ntmp) ascending [K] */
    const double *Cd;         /* (ntmp, ncol) de-excitation coefficients [m^3/s] */
    const double *Ce;         /* (ntmp, ncol) excitation coefficients [m^3/s] */
} mgx_colldata;

/* Knobs of Parameters read by the level-population loop
 * (src/model/parameters/parameters.hpp:76-104), reference defaults in brackets. */
typedef struct mgx_nlte_params {
    double pop_prec;                      /* [1e-6] convergence threshold */
    double min_rel_pop_for_convergence;   /* [1e-10] */
    double population_inversion_fraction; /* [-1] (inversion -> LTE test off) */
    double min_line_emissivity;           /* [1e-50] */
    double min_line_opacity;              /* [1e-13] */
    double convergence_fraction;          /* [0.995] */
    uint32_t use_adaptive_Ng_acceleration;              /* [1] */
    uint32_t Ng_acceleration_mem_limit;                 /* [9] */
    uint32_t adaptive_Ng_acceleration_min_order;        /* [2] */
    uint32_t adaptive_Ng_acceleration_use_max_criterion;/* [1] */
    uint32_t Ng_acceleration_remove_N_its;              /* [0] */
    uint32_t reserved0;
} mgx_nlte_params;

/* Fills p with the reference's defaults. */
void mgx_nlte_default_params(mgx_nlte_params *p);

/* Convergence statistics of one species (LineProducingSpecies::fraction_not_converged,
 * relative_change_mean, relative_change_max; lineProducingSpecies.tpp:127-163). */
typedef struct mgx_nlte_conv {
    double fraction_not_converged, relative_change_mean, relative_change_max;
} mgx_nlte_conv;

typedef struct mgx_nlte mgx_nlte;

/*
 * Level-population state of every line-producing species of a model:
 * packs the line and collision data (host pointers ld[nlspecs],
 * ncolpar[nlspecs], colpar[sum ncolpar] in species order) into device
 * tables and allocates the populations and the Ng history
 * (Ng_acceleration_mem_limit + Ng_acceleration_remove_N_its + 2 vectors of
 * P * sum(nlev) doubles).  model_dev is the device view the formal solve
 * uses (its temperature, quad_weights and sizes are read; the handle keeps
 * the pointers); abundance_dev is (P, nspec) on the device.  Populations
 * start at zero: call mgx_nlte_set_lte or mgx_nlte_set_populations.
 * Levels per species: at most 64.
 */
int mgx_nlte_create(const mgx_model *model_dev, uint32_t nlspecs, const mgx_linedata *ld,
                    const uint32_t *ncolpar, const mgx_colldata *colpar, uint32_t nspec,
                    const double *abundance_dev, const mgx_nlte_params *params,
                    mgx_nlte **out, void *stream);
void mgx_nlte_destroy(mgx_nlte *h);

/*
 * LineProducingSpecies::update_using_LTE for every species
 * (lineProducingSpecies.tpp:52-75) and Lines::set_emissivity_and_opacity
 * (lines.tpp:29-41) plus the interpolation criterion opacity / population_tot
 * (solver_interp_helper.hpp:15-30) into emissivity, opacity, interp_crit
 * (P, L, device; the arrays of the model view).  The Ng history restarts
 * with the LTE populations.
 */
int mgx_nlte_set_lte(mgx_nlte *h, double *emissivity, double *opacity, double *interp_crit,
                     void *stream);

/* Replaces the populations (device, per species a (P, nlev_s) block, blocks
 * concatenated) and restarts the Ng history with them; emissivity, opacity
 * and interp_crit are recomputed as in mgx_nlte_set_lte. */
int mgx_nlte_set_populations(mgx_nlte *h, const double *population_dev, double *emissivity,
                             double *opacity, double *interp_crit, void *stream);

/* Copies the current populations (layout of mgx_nlte_set_populations) to a
 * device array; population_tot (nlspecs, P) may be NULL. */
int mgx_nlte_get_populations(const mgx_nlte *h, double *population_dev,
                             double *population_tot_dev, void *stream);

/*
 * One statistical-equilibrium iteration of the sparse loop from a formal
 * solve's J (P, F) and lambda_diag (P, L) (device, as written by
 * mgx_feautrier_forward_device):
 *   Model::compute_Jeff_sparse                      model.cpp:549-570
 *     (J_lin(p, k) = sum_z w_z J(p, l*nquads + z), the per-line J of
 *      Solver::solve_feautrier_order_2_sparse, solver.tpp:298-363)
 *   Lines::iteration_using_statistical_equilibrium_sparse  lines.cpp:139-150
 *     update_using_statistical_equilibrium_sparse   lineProducingSpecies.tpp:589-683
 *     correct_negative_populations                  :711-802
 *     check_for_convergence                         :127-163
 *     set_emissivity_and_opacity                    lines.tpp:29-41
 * Each point's nlev x nlev system is solved in FP64 by Gaussian elimination
 * with partial pivoting and two steps of iterative refinement with
 * double-double residuals (the reference: long-double column-pivoted QR).
 * conv (host, nlspecs) receives the statistics; the call synchronises the
 * stream.  J_lin, J_eff (P, L, device) may be NULL; when given they receive
 * the values used.
 */
int mgx_nlte_iteration_stateq(mgx_nlte *h, const double *J, const double *lambda_diag,
                              double *emissivity, double *opacity, double *interp_crit,
                              mgx_nlte_conv *conv, double *J_lin, double *J_eff, void *stream);

/*
 * Lines::iteration_using_Ng_acceleration(abundance, T, pop_prec, order)
 * (lines.cpp:106-117): LineProducingSpecies::update_using_acceleration(order)
 * (lineProducingSpecies.tpp:266-314; the order x order residual Gram matrix
 * on the device, its long-double column-pivoted QR solve on the host), then
 * the negative-population correction, the convergence check and the line
 * emissivity / opacity.  MGX_E_ARG when fewer than order residuals are held.
 */
int mgx_nlte_iteration_ng(mgx_nlte *h, uint32_t order, double *emissivity, double *opacity,
                          double *interp_crit, mgx_nlte_conv *conv, void *stream);

/*
 * Lines::iteration_using_statistical_equilibrium (lines.cpp:124-135) with
 * Model::compute_Jeff (model.cpp:514-545): J_eff from J (P,F) and the general
 * Lambda lists of mgx_feautrier_forward_lambda_device (capacity cap), then the
 * global rate equations of all points
 * (LineProducingSpecies::update_using_statistical_equilibrium,
 * lineProducingSpecies.tpp:377-581), the negative-population correction and the
 * convergence statistics.  The reference factorises the whole sparse system
 * (SparseLU in long double); here block-Jacobi sweeps solve every point's block
 * (elimination with compensated refinement) with the other points' populations
 * on the right-hand side until the largest change is at the rounding level;
 * sweeps (may be NULL) receives their number.  With n_off_diag = 0 the system
 * is block diagonal and one sweep is exact.  Fails with MGX_E_UNSUPPORTED if
 * the sweeps do not converge within 100.
 */
int mgx_nlte_iteration_stateq_global(mgx_nlte *h, const double *J, uint32_t cap,
                                     const uint32_t *lam_count, const uint32_t *lam_nr,
                                     const double *lam_val, double *emissivity, double *opacity,
                                     double *interp_crit, mgx_nlte_conv *conv, double *Jlin,
                                     double *Jeff, uint32_t *sweeps, void *stream);

/* Lines::trial_iteration_using_adaptive_Ng_acceleration (lines.cpp:119-124):
 * the extrapolation into the trial populations and check_for_convergence_trial
 * (lineProducingSpecies.tpp:167-201); the populations are unchanged. */
int mgx_nlte_trial_ng(mgx_nlte *h, uint32_t order, mgx_nlte_conv *conv, void *stream);

/* Number of residuals the Ng history holds (LineProducingSpecies::residuals.size()). */
int mgx_nlte_history(const mgx_nlte *h, uint32_t *n_residuals);

/*
 * The Lambda of the last formal solve an iteration loop ran (what the
 * reference's lspec.lambda holds after compute_level_populations*,
 * lineProducingSpecies.hpp:35): after mgx_compute_level_populations the
 * lists -- *cap entries per (point, line), with *count, *nr and *val set and
 * *diag NULL -- and after the sparse or short-characteristics loop the
 * diagonal (P, L) in *diag with *cap = 0 and the list pointers NULL.  Device
 * memory owned by the handle, valid until the next loop call or
 * mgx_nlte_destroy.  MGX_E_ARG if no loop has run a formal solve yet.
 */
int mgx_nlte_last_lambda(const mgx_nlte *h, uint32_t *cap, const uint32_t **count,
                         const uint32_t **nr, const double **val, const double **diag);

/*
 * Model::compute_level_populations_sparse(use_Ng_acceleration, max_niterations)
 * (model.cpp:756-850) with the Default / Adaptive Ng criteria
 * (model.cpp:580-645): per iteration a formal solve
 * (mgx_feautrier_forward_device with u = NULL into the handle's J and Lambda
 * buffers) and mgx_nlte_iteration_stateq, or an Ng step.  emissivity, opacity
 * and interp_crit must be the arrays model_dev points to.  iterations
 * receives the number of statistical-equilibrium iterations; error_mean and
 * error_max (host, capacity max_niterations * nlspecs, may be NULL) the
 * per-iteration relative_change_mean / _max in the reference's order
 * (Model::error_mean / error_max).
 */
int mgx_compute_level_populations_sparse(const mgx_model *model_dev, const mgx_options *opts,
                                         mgx_nlte *h, int use_Ng_acceleration,
                                         int64_t max_niterations, double *emissivity,
                                         double *opacity, double *interp_crit,
                                         int32_t *iterations, double *error_mean,
                                         double *error_max, void *stream);

/*
 * Model::compute_level_populations(use_Ng_acceleration, max_niterations)
 * (model.cpp:653-744): as mgx_compute_level_populations_sparse, with the
 * formal solve of mgx_feautrier_forward_lambda_device (opts->n_off_diag bands,
 * up to MGX_MAX_OFF_DIAG) and mgx_nlte_iteration_stateq_global.
 */
int mgx_compute_level_populations(const mgx_model *model_dev, const mgx_options *opts,
                                  mgx_nlte *h, int use_ng, int64_t max_niterations,
                                  double *emissivity, double *opacity, double *interp_crit,
                                  int32_t *iterations, double *error_mean, double *error_max,
                                  void *stream);

/*
 * Model::compute_level_populations_shortchar(use_Ng_acceleration, max_niterations)
 * (model.cpp:859-935): as mgx_compute_level_populations_sparse with the formal
 * solve of mgx_shortchar_order_0_device (Parameters::min_dtau).  Its Lambda is
 * diagonal, so the reference's compute_Jeff and global statistical equilibrium
 * reduce to the per-point update of mgx_nlte_iteration_stateq.
 */
int mgx_compute_level_populations_shortchar(const mgx_model *model_dev, const mgx_options *opts,
                                            mgx_nlte *h, int use_ng, int64_t max_niterations,
                                            double min_dtau, double *emissivity, double *opacity,
                                            double *interp_crit, int32_t *iterations,
                                            double *error_mean, double *error_max, void *stream);

/*
 * Workload statistics of the trace, for the algorithmic-byte model of the
 * solve: stats[0] = walk steps S (both halves, all ray pairs and origins),
 * stats[1] = neighbours examined (sum of n_neighbors over the points stepped
 * from), stats[2] = traced entries T (sum of n_tot over ray pairs, origins and
 * lines), stats[3] = interpolated entries (interp_factor != 0).  No
 * counterpart in the reference (measurement aid).
 */
int mgx_trace_stats_device(const mgx_model *model_dev, const mgx_options *opts,
                           uint64_t *stats, void *stream);

/*
 * Kernel timing: when enabled, the device entry points record HIP events
 * around their main kernel on the caller's stream; mgx_last_kernel_ms()
 * returns the duration of the last call's main kernel (ms), summed over the
 * launches that call made (a launch whose walk outgrew its buffer is repeated
 * with a larger one).  No counterpart in the reference (measurement aid).
 */
void mgx_set_kernel_timing(int enable);

/*
 * Verification aid (no counterpart in the reference): evaluates the device
 * build of one of the reference's float functions as this library restates
 * them -- fn 0 expf, 1 expm1f, 2 erff, 3 logf, 4 powf(x, y) -- on every float
 * argument with bits in [first, first + nchunks * 2^chunk_log2) and writes
 * per chunk the wrapping sum of a 64-bit digest of (argument bits, result
 * bits) into digest (device, nchunks entries), NaN results canonicalised.
 * Equal digests to the C library's (computed on the host the same way) show
 * the GPU gives the C library's bits on the whole range.  Synchronises the
 * stream.  MGX_E_ARG for an unknown fn or a range beyond 2^32.
 */
int mgx_fmath_digest(int fn, float y, uint64_t first, uint32_t nchunks, uint32_t chunk_log2,
                     uint64_t *digest, void *stream);
double mgx_last_kernel_ms(void);

#ifdef __cplusplus
}
#endif

#endif /* MGX_H */
