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

    mgx_options options() const;
    /* the device view (uploads stale arrays first) */
    const mgx_model &device_view();

  private:
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
