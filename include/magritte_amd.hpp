/*
 * C++ face of the formal-solver path with the reference's class and member
 * names (magritte::Model, magritte::Solver, Parameters, Radiation, Image), for
 * a maintainer who replaces the reference's implementation of
 *   Model::compute_radiation_field_feautrier_order_2()   src/model/model.hpp:59
 *   Model::compute_image_new(...)                        src/model/model.hpp:74-80
 *   Solver::setup / solve_feautrier_order_2 / image_feautrier_order_2_new_imager
 *                                                        src/solver/solver.hpp:81-205
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
    Real max_distance_opacity_contribution = 10.0;
    Real max_interpolation_diff = 1.4;
    bool mixed_precision = false; /* MGX_PREC_MIXED instead of MGX_PREC_FP64 */
};

/* src/model/radiation/radiation.hpp: u (hnrays, P, F) and J (P, F), row-major */
struct Radiation {
    std::vector<Real> u, J;
    Size hnrays = 0, npoints = 0, nfreqs = 0; /* shapes of the last solve */
};

/* src/model/image/image.hpp (intensity images of the new imager): I (nx*ny, F)
   with pixel index y + ny*x, pixel coordinates ImX, ImY */
struct Image {
    Size nx = 0, ny = 0, nfreqs = 0;
    Vector3D direction{0.0, 0.0, 0.0};
    std::vector<Real> I, ImX, ImY, freqs;
};

class Model {
  public:
    Parameters parameters;
    Radiation radiation;
    std::vector<Real> lambda_diag; /* (P, L): the diagonal Lambda entries of every line */
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

    /* src/model/model.cpp:353-390 */
    int compute_radiation_field_feautrier_order_2();
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
    void image_feautrier_order_2_new_imager(Model &model, const Vector3D raydir, Size Nxpix,
                                            Size Nypix);
};

} // namespace magritte
