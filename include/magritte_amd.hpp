/*
 * C++ face of the formal-solver path with the reference's class and member
 * names (magritte::Model, magritte::Solver, Parameters, Radiation, Image), for
 * a maintainer who replaces the reference's implementation of
 *   Model::compute_radiation_field_feautrier_order_2()   src/model/model.hpp:59
 *   Model::compute_radiation_field_shortchar_order_0(),
 *          compute_radiation_field_feautrier_order_2_uv() src/model/model.hpp:60, 111
 *   Model::compute_image(...), compute_image_new(...), compute_image_optical_depth(...),
 *          compute_image_optical_depth_new(...)          src/model/model.hpp:72-92
 *   Solver::setup / solve_* / image_*                    src/solver/solver.hpp:81-205
 * The model's arrays are the ones the path reads (the mgx_model view of
 * include/mgx.h), held on the host; the class keeps device mirrors and
 * calls the C ABI.  Errors are thrown as std::runtime_error with the
 * library's message, as the reference throws them.
 */
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "mgx.h"

namespace magritte {

using Real = double;
using Size = uint32_t;
using Vector3D = std::array<Real, 3>;

/* The knobs of src/model/parameters/parameters.hpp that the path reads, plus the
   precision mode of this implementation. */
struct Parameters {
    Size npoints = 0, nrays = 0, hnrays = 0, nboundary = 0;
    Size nlines = 0, nlspecs = 0, nquads = 0, nfreqs = 0;
    Size dimension = 3;
    bool spherical_symmetry = false;
    bool adaptive_ray = false;
    bool one_line_approximation = false;
    bool sum_opacity_emissivity_over_all_lines = false;
    Size n_off_diag = 0;
    Real min_opacity = 1.0e-26;
    Real min_dtau = 1.0e-15; /* short characteristics */
    Real max_distance_opacity_contribution = 10.0;
    Real max_interpolation_diff = 1.4;
    bool mixed_precision = false; /* MGX_PREC_MIXED instead of MGX_PREC_FP64 */
    /* the level-population loop's knobs (parameters.hpp:76-104; mgx_nlte_params) */
    Real pop_prec = 1.0e-6;
    Real min_rel_pop_for_convergence = 1.0e-10;
    Real population_inversion_fraction = -1.0;
    Real min_line_emissivity = 1.0e-50;
    Real min_line_opacity = 1.0e-13;
    Real convergence_fraction = 0.995;
    bool use_adaptive_Ng_acceleration = true;
    Size Ng_acceleration_mem_limit = 9;
    Size adaptive_Ng_acceleration_min_order = 2;
    bool adaptive_Ng_acceleration_use_max_criterion = true;
    Size Ng_acceleration_remove_N_its = 0;
};

/* CollisionPartner (src/model/lines/lineProducingSpecies/linedata/collisionPartner/
   collisionPartner.hpp:7-33): Cd, Ce are (ntmp, ncol) row-major */
struct CollisionPartner {
    Size num_col_partner = 0;  /* abundance column of the partner */
    char orth_or_para_H2 = 'n'; /* 'n', 'o' or 'p' */
    std::vector<Size> icol, jcol;
    std::vector<Real> tmp, Cd, Ce;
};

/* Linedata (linedata.hpp) of one line-producing species */
struct Linedata {
    Size num = 0; /* species column in the abundance table */
    Size nlev = 0, nrad = 0;
    Real inverse_mass = 0.0;
    std::vector<Real> energy, weight;              /* (nlev) */
    std::vector<Size> irad, jrad;                  /* (nrad) upper / lower level */
    std::vector<Real> A, Bs, Ba, frequency;        /* (nrad) */
    std::vector<CollisionPartner> colpar;
};

/* Lambda (src/model/lines/lineProducingSpecies/lambda/lambda.hpp): for every point p
   and transition k of the species the approximated-Lambda elements Ls[p][k][i] and
   their emitting points nr[p][k][i], the diagonal element first */
struct LambdaNested {
    std::vector<std::vector<std::vector<Real>>> Ls;
    std::vector<std::vector<std::vector<Size>>> nr;
};

/* LineProducingSpecies (lineProducingSpecies.hpp): the line data, the level
   populations, (P, nlev) row-major, with their totals (P), and the Lambda operator of
   the last formal solve (filled by compute_radiation_field_feautrier_order_2) */
struct LineProducingSpecies {
    Linedata linedata;
    std::vector<Real> population, population_tot;
    LambdaNested lambda;
};

/* src/model/radiation/radiation.hpp: I (nrays, P, F; short characteristics), u and v
   (hnrays, P, F), J (P, F), row-major; empty when the last solve did not fill them */
struct Radiation {
    std::vector<Real> I, u, v, J;
    Size nrays = 0, hnrays = 0, npoints = 0, nfreqs = 0; /* shapes of the last solve */
};

/* LineProducingSpecies::lambda (lambda.hpp) of every line with n_off_diag > 0: for each
   receiving point p and line l, count(p, l) entries (emitting point, value) at
   nr / val[(p L + l) cap + i], the diagonal element first */
struct LambdaLists {
    Size cap = 0;
    std::vector<uint32_t> count, nr;
    std::vector<Real> val;
};

/* src/model/image/image.hpp: I (npix, F) of the intensity or the optical depth, with the
   image-plane coordinates ImX, ImY (npix).  Projection-surface images (the new imager) have
   npix = nx * ny pixels (index y + ny*x); images at all model points (the old imager) have
   npix = P and nx = P, ny = 1. */
enum ImageType { Intensity = 0, OpticalDepth = 1 };
struct Image {
    ImageType type = Intensity;
    int64_t ray_nr = -1; /* the model ray of an image along one, else -1 */
    Size nx = 0, ny = 0, npix = 0, nfreqs = 0;
    Vector3D direction{0.0, 0.0, 0.0};
    std::vector<Real> I, ImX, ImY, freqs;
};

class Model {
  public:
    Parameters parameters;
    Radiation radiation;
    std::vector<Real> lambda_diag; /* (P, L): the diagonal Lambda entries of every line */
    LambdaLists lambda;            /* the general operator (n_off_diag > 0) */
    std::vector<Image> images;
    /* lines.lineProducingSpecies, chemistry.species.abundance (P, nspec) and the loop's
       per-iteration relative changes (Model::error_mean / error_max, species innermost) */
    std::vector<LineProducingSpecies> lspecs;
    std::vector<Real> abundance;
    Size nspec = 0;
    std::vector<Real> error_mean, error_max;

    Model() = default;
    Model(const Model &) = delete;
    Model &operator=(const Model &) = delete;
    ~Model();

    /* Host arrays by the field names of mgx_model (position, velocity, n_neighbors,
       ..., nu, f_line, ...).  Setting an array marks the device mirror stale. */
    void set_array(const std::string &name, const void *data, size_t bytes);
    const std::vector<unsigned char> &get_array(const std::string &name) const;
    void set_center(const Vector3D &c) { center = c; }
    void set_max_inverse_mass(Real v) { max_inverse_mass = v; }

    /* src/model/model.cpp:353-390; with n_off_diag > 0 also the Lambda lists */
    int compute_radiation_field_feautrier_order_2();
    /* src/model/model.cpp:313-351 */
    int compute_radiation_field_shortchar_order_0();
    /* src/model/model.cpp:395-433 */
    int compute_radiation_field_feautrier_order_2_uv();
    /* the old imager at all model points along ray ray_nr, model.cpp:957-976, 1061-1076 */
    int compute_image(Size ray_nr);
    int compute_image_optical_depth(Size ray_nr);
    /* optical depth on the projection surface, model.cpp:1081-1127 */
    int compute_image_optical_depth_new(Size ray_nr);
    int compute_image_optical_depth_new(Size ray_nr, Size Nxpix, Size Nypix);
    int compute_image_optical_depth_new(Real rx, Real ry, Real rz, Size Nxpix, Size Nypix);
    int compute_image_optical_depth_new(const Vector3D raydir, Size Nxpix, Size Nypix);
    /* src/model/model.cpp:987-1035 */
    int compute_image_new(Size ray_nr);
    int compute_image_new(Size ray_nr, Size Nxpix, Size Nypix);
    int compute_image_new(Real rx, Real ry, Real rz, Size Nxpix, Size Nypix);
    int compute_image_new(const Vector3D raydir, Size Nxpix, Size Nypix);
    /* Model::compute_spectral_discretisation(nu_min, nu_max, n) (model.cpp:257-298):
       a uniform image grid; the line grid is restored by restore_line_grid(). */
    int compute_spectral_discretisation(Real nu_min, Real nu_max, Size n);
    void restore_line_grid();

    /* Model::compute_LTE_level_populations (model.cpp:302-308): LTE populations of every
       species and the line emissivity, opacity and interpolation criterion */
    int compute_LTE_level_populations();
    /* Model::compute_level_populations(use_Ng_acceleration, max_niterations)
       (model.cpp:654-744; the general Lambda operator with n_off_diag bands and the
       global statistical equilibrium) and Model::compute_level_populations_sparse
       (model.cpp:756-850): from the current populations (LTE when none are set);
       returns the number of statistical-equilibrium iterations */
    int compute_level_populations(bool use_Ng_acceleration, long max_niterations);
    int compute_level_populations_sparse(bool use_Ng_acceleration, long max_niterations);
    /* Model::compute_level_populations_shortchar (model.cpp:859-935): the loop with the
       order-0 short-characteristics formal solve (parameters.min_dtau) */
    int compute_level_populations_shortchar(bool use_Ng_acceleration, long max_niterations);

    mgx_options options() const;
    /* the device view (uploads stale arrays first) */
    const mgx_model &device_view();

  private:
    int level_population_loop(int kind, bool use_Ng_acceleration, long max_niterations);
    std::map<std::string, std::vector<unsigned char>> host, saved_line_grid;
    std::map<std::string, void *> dev;
    std::map<std::string, bool> stale;
    Vector3D center{0.0, 0.0, 0.0};
    Real max_inverse_mass = 0.0;
    mgx_model view{};
};

/* src/solver/solver.hpp: the solver is a per-call temporary */
class Solver {
  public:
    void setup(Model &model);
    void solve_feautrier_order_2(Model &model);
    void solve_shortchar_order_0(Model &model);
    void solve_feautrier_order_2_uv(Model &model);
    void image_feautrier_order_2_new_imager(Model &model, const Vector3D raydir, Size Nxpix,
                                            Size Nypix);
    void image_optical_depth_new_imager(Model &model, const Vector3D raydir, Size Nxpix,
                                        Size Nypix);
    void image_feautrier_order_2(Model &model, Size rr);
    void image_optical_depth(Model &model, Size rr);
};

} // namespace magritte
