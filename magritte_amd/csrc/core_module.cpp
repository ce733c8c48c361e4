/*
 * pybind11 module magritte_amd.core: the reference's Python surface
 * (src/bindings/pybindings.cpp) for the formal-solver path over magritte::Model:
 *   Model.parameters / geometry / chemistry / thermodynamics / lines / radiation /
 *        images and the compute_* methods                    (:123-271)
 *   Parameters set_* / getters and knobs                      (:274-371)
 *   Geometry, Points, Rays, Boundary, BoundaryCondition       (:372-442)
 *   Thermodynamics, Temperature, Turbulence, Chemistry, Species (:443-469)
 *   Lines, LineProducingSpecies, Lambda, Quadrature, Linedata,
 *        CollisionPartner                                     (:470-640)
 *   Radiation, Frequencies                                    (:641-665)
 *   VReal, VSize, V3DReal vectors, MReal matrices with set() and the buffer
 *        protocol                                             (:666-800)
 * Errors surface as RuntimeError with the library's message.
 */
#include <cstring>

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>

#include "magritte_amd.hpp"

/* species and collision partners are bound as lists that alias the model's own
   (model.lines.lineProducingSpecies[0].linedata.A = ... changes the model, as the
   reference's do) */
PYBIND11_MAKE_OPAQUE(std::vector<magritte::LineProducingSpecies>)
PYBIND11_MAKE_OPAQUE(std::vector<magritte::CollisionPartner>)

namespace py = pybind11;
using namespace magritte;

namespace {

template <typename T>
py::array view_of(std::vector<T> &v, std::vector<py::ssize_t> shape, py::handle base)
{
    if (v.empty()) return py::array(py::dtype::of<T>(), std::vector<py::ssize_t>{0});
    return py::array(py::dtype::of<T>(), shape, v.data(), base);
}

template <typename T>
void bind_vector(py::module_ &m, const char *name)
{
    py::class_<Vector<T>>(m, name, py::buffer_protocol())
        .def(py::init<>())
        .def_buffer([](Vector<T> &v) -> py::buffer_info {
            return py::buffer_info(v.vec.data(), sizeof(T), py::format_descriptor<T>::format(), 1,
                                   {(py::ssize_t)v.vec.size()}, {(py::ssize_t)sizeof(T)});
        })
        .def("set",
             [](Vector<T> &v, py::array_t<T, py::array::c_style | py::array::forcecast> a) {
                 if (a.ndim() != 1) throw std::runtime_error("set: expected a 1-D array");
                 v.vec.assign(a.data(), a.data() + a.size());
             })
        .def("__len__", [](const Vector<T> &v) { return v.vec.size(); });
}

} // namespace

PYBIND11_MODULE(core, m)
{
    m.doc() = "Formal-solver path of Magritte on MI355X (gfx950): the reference's Model object model";
    m.attr("CC") = 2.99792458e+8;
    m.attr("HH") = 6.62607004e-34;
    m.attr("KB") = 1.38064852e-23;
    m.attr("AMU") = 1.66053904e-27;
    m.attr("T_CMB") = 2.72548;

    bind_vector<Real>(m, "VReal");
    bind_vector<Size>(m, "VSize");

    /* Vector<Vector3D>: an (n, 3) buffer; set() takes an (n, 3) array or a list of triples */
    py::class_<Vector<Vector3D>>(m, "VVector3D", py::buffer_protocol())
        .def(py::init<>())
        .def_buffer([](Vector<Vector3D> &v) -> py::buffer_info {
            return py::buffer_info(v.vec.data(), sizeof(Real), py::format_descriptor<Real>::format(),
                                   2, {(py::ssize_t)v.vec.size(), (py::ssize_t)3},
                                   {(py::ssize_t)(3 * sizeof(Real)), (py::ssize_t)sizeof(Real)});
        })
        .def("set",
             [](Vector<Vector3D> &v, py::array_t<Real, py::array::c_style | py::array::forcecast> a) {
                 if (a.ndim() != 2 || a.shape(1) != 3)
                     throw std::runtime_error("set: expected an (n, 3) array");
                 v.vec.resize((size_t)a.shape(0));
                 std::memcpy(v.vec.data(), a.data(), (size_t)a.size() * sizeof(Real));
             })
        .def("__len__", [](const Vector<Vector3D> &v) { return v.vec.size(); });

    /* Matrix<Real>: an (nrows, ncols) buffer */
    py::class_<Matrix<Real>>(m, "MReal", py::buffer_protocol())
        .def(py::init<>())
        .def_buffer([](Matrix<Real> &v) -> py::buffer_info {
            return py::buffer_info(v.vec.data(), sizeof(Real), py::format_descriptor<Real>::format(),
                                   2, {(py::ssize_t)v.nrows, (py::ssize_t)v.ncols},
                                   {(py::ssize_t)(v.ncols * sizeof(Real)), (py::ssize_t)sizeof(Real)});
        })
        .def_readonly("nrows", &Matrix<Real>::nrows)
        .def_readonly("ncols", &Matrix<Real>::ncols)
        .def("set", [](Matrix<Real> &v, py::array_t<Real, py::array::c_style | py::array::forcecast> a) {
            if (a.ndim() != 2) throw std::runtime_error("set: expected a 2-D array");
            v.nrows = (size_t)a.shape(0);
            v.ncols = (size_t)a.shape(1);
            v.vec.assign(a.data(), a.data() + a.size());
        });

    py::class_<Parameters>(m, "Parameters")
        .def(py::init<>())
#define MGX_PARAM(x) .def("set_" #x, &Parameters::set_##x).def(#x, &Parameters::x)
        MGX_PARAM(model_name) MGX_PARAM(dimension) MGX_PARAM(npoints) MGX_PARAM(nrays)
        MGX_PARAM(hnrays) MGX_PARAM(nboundary) MGX_PARAM(nfreqs) MGX_PARAM(nspecs)
        MGX_PARAM(nlspecs) MGX_PARAM(nlines) MGX_PARAM(nquads) MGX_PARAM(use_scattering)
        MGX_PARAM(spherical_symmetry) MGX_PARAM(adaptive_ray_tracing)
#undef MGX_PARAM
        .def("version", &Parameters::version)
        .def_readwrite("n_off_diag", &Parameters::n_off_diag)
        .def_readwrite("max_width_fraction", &Parameters::max_width_fraction)
        .def_readwrite("convergence_fraction", &Parameters::convergence_fraction)
        .def_readwrite("min_rel_pop_for_convergence", &Parameters::min_rel_pop_for_convergence)
        .def_readwrite("pop_prec", &Parameters::pop_prec)
        .def_readwrite("min_opacity", &Parameters::min_opacity)
        .def_readwrite("min_line_opacity", &Parameters::min_line_opacity)
        .def_readwrite("min_line_emissivity", &Parameters::min_line_emissivity)
        .def_readwrite("max_interpolation_diff", &Parameters::max_interpolation_diff)
        .def_readwrite("min_dtau", &Parameters::min_dtau)
        .def_readwrite("population_inversion_fraction", &Parameters::population_inversion_fraction)
        .def_readwrite("store_intensities", &Parameters::store_intensities)
        .def_readwrite("one_line_approximation", &Parameters::one_line_approximation)
        .def_readwrite("sum_opacity_emissivity_over_all_lines",
                       &Parameters::sum_opacity_emissivity_over_all_lines)
        .def_readwrite("max_distance_opacity_contribution",
                       &Parameters::max_distance_opacity_contribution)
        .def_readwrite("use_adaptive_Ng_acceleration", &Parameters::use_adaptive_Ng_acceleration)
        .def_readwrite("Ng_acceleration_mem_limit", &Parameters::Ng_acceleration_mem_limit)
        .def_readwrite("adaptive_Ng_acceleration_use_max_criterion",
                       &Parameters::adaptive_Ng_acceleration_use_max_criterion)
        .def_readwrite("Ng_acceleration_remove_N_its", &Parameters::Ng_acceleration_remove_N_its)
        .def_readwrite("adaptive_Ng_acceleration_min_order",
                       &Parameters::adaptive_Ng_acceleration_min_order)
        .def_readwrite("mixed_precision", &Parameters::mixed_precision,
                       "Single-precision profile and transcendentals with the FP64 elimination "
                       "(this implementation's mixed mode).");

    py::class_<Points>(m, "Points")
        .def_readwrite("position", &Points::position)
        .def_readwrite("velocity", &Points::velocity)
        .def_readwrite("neighbors", &Points::neighbors)
        .def_readwrite("n_neighbors", &Points::n_neighbors)
        .def_readonly("cum_n_neighbors", &Points::cum_n_neighbors)
        .def_property_readonly("center", [](const Points &p) {
            return py::make_tuple(p.center[0], p.center[1], p.center[2]);
        });
    py::class_<Rays>(m, "Rays")
        .def_readwrite("direction", &Rays::direction)
        .def_readwrite("weight", &Rays::weight)
        .def_readwrite("antipod", &Rays::antipod)
        .def_readwrite("use_adaptive_directions", &Rays::use_adaptive_directions)
        .def("get_antipod_index", &Rays::get_antipod_index);
    py::enum_<BoundaryCondition>(m, "BoundaryCondition")
        .value("Zero", Zero)
        .value("Thermal", Thermal)
        .value("CMB", CMB)
        .export_values();
    py::class_<Boundary>(m, "Boundary")
        .def_readwrite("boundary2point", &Boundary::boundary2point)
        .def_readonly("point2boundary", &Boundary::point2boundary)
        .def_readwrite("boundary_temperature", &Boundary::boundary_temperature)
        .def("set_boundary_condition", &Boundary::set_boundary_condition)
        .def("get_boundary_condition", &Boundary::get_boundary_condition);
    py::class_<Geometry>(m, "Geometry")
        .def_readonly("points", &Geometry::points)
        .def_readonly("rays", &Geometry::rays)
        .def_readonly("boundary", &Geometry::boundary);

    py::class_<Species>(m, "Species")
        .def_readwrite("symbol", &Species::symbol)
        .def_readwrite("abundance", &Species::abundance);
    py::class_<Chemistry>(m, "Chemistry").def_readonly("species", &Chemistry::species);
    py::class_<Temperature>(m, "Temperature").def_readwrite("gas", &Temperature::gas);
    py::class_<Turbulence>(m, "Turbulence").def_readwrite("vturb2", &Turbulence::vturb2);
    py::class_<Thermodynamics>(m, "Thermodynamics")
        .def_readonly("temperature", &Thermodynamics::temperature)
        .def_readonly("turbulence", &Thermodynamics::turbulence);

    py::class_<CollisionPartner>(m, "CollisionPartner")
        .def(py::init<>())
        .def_readwrite("num_col_partner", &CollisionPartner::num_col_partner)
        .def_readwrite("orth_or_para_H2", &CollisionPartner::orth_or_para_H2)
        .def_property_readonly("ntmp", [](const CollisionPartner &c) { return c.tmp.size(); })
        .def_property_readonly("ncol", [](const CollisionPartner &c) { return c.icol.size(); })
        .def_readwrite("icol", &CollisionPartner::icol)
        .def_readwrite("jcol", &CollisionPartner::jcol)
        .def_readwrite("tmp", &CollisionPartner::tmp)
        .def_readwrite("Cd", &CollisionPartner::Cd)
        .def_readwrite("Ce", &CollisionPartner::Ce);
    py::bind_vector<std::vector<CollisionPartner>>(m, "VectorCollisionPartner");

    py::class_<Linedata>(m, "Linedata")
        .def(py::init<>())
        .def_readwrite("num", &Linedata::num)
        .def_readwrite("sym", &Linedata::sym)
        .def_readwrite("nlev", &Linedata::nlev)
        .def_readwrite("nrad", &Linedata::nrad)
        .def_readwrite("inverse_mass", &Linedata::inverse_mass)
        .def_readwrite("energy", &Linedata::energy)
        .def_readwrite("weight", &Linedata::weight)
        .def_readwrite("irad", &Linedata::irad)
        .def_readwrite("jrad", &Linedata::jrad)
        .def_readwrite("A", &Linedata::A)
        .def_readwrite("Bs", &Linedata::Bs)
        .def_readwrite("Ba", &Linedata::Ba)
        .def_readwrite("frequency", &Linedata::frequency)
        .def_property_readonly("ncolpar", [](const Linedata &d) { return d.colpar.size(); })
        .def_readwrite("colpar", &Linedata::colpar);

    py::class_<Quadrature>(m, "Quadrature")
        .def_readwrite("roots", &Quadrature::roots)
        .def_readwrite("weights", &Quadrature::weights);

    /* lspec.Lambda.Ls / .nr as nested lists [point][transition][element] (pybindings.cpp:570-580) */
    py::class_<LambdaNested>(m, "Lambda")
        .def_readonly("Ls", &LambdaNested::Ls)
        .def_readonly("nr", &LambdaNested::nr);

    py::class_<LineProducingSpecies>(m, "LineProducingSpecies")
        .def(py::init<>())
        .def_readwrite("linedata", &LineProducingSpecies::linedata)
        .def_readonly("quadrature", &LineProducingSpecies::quadrature)
        .def_readonly("Lambda", &LineProducingSpecies::lambda)
        .def_readonly("J", &LineProducingSpecies::J, "Line-integrated mean intensity (P, nrad) of the sparse solve.")
        .def("index", &LineProducingSpecies::index)
        .def_property(
            "population",
            [](py::object self) {
                LineProducingSpecies &s = self.cast<LineProducingSpecies &>();
                return view_of(s.population, {(py::ssize_t)s.population.size()}, self);
            },
            [](LineProducingSpecies &s, py::array_t<double, py::array::c_style | py::array::forcecast> a) {
                s.population.assign(a.data(), a.data() + a.size());
            })
        .def_property(
            "population_tot",
            [](py::object self) {
                LineProducingSpecies &s = self.cast<LineProducingSpecies &>();
                return view_of(s.population_tot, {(py::ssize_t)s.population_tot.size()}, self);
            },
            [](LineProducingSpecies &s, py::array_t<double, py::array::c_style | py::array::forcecast> a) {
                s.population_tot.assign(a.data(), a.data() + a.size());
            });
    py::bind_vector<std::vector<LineProducingSpecies>>(m, "VectorLineProducingSpecies");

    py::class_<Lines>(m, "Lines")
        .def_readonly("lineProducingSpecies", &Lines::lineProducingSpecies)
        .def_readonly("line", &Lines::line)
        .def_readonly("nrad_cum", &Lines::nrad_cum)
        .def_readonly("emissivity", &Lines::emissivity)
        .def_readonly("opacity", &Lines::opacity)
        .def_readonly("inverse_width", &Lines::inverse_width)
        .def("resize_LineProducingSpecies", &Lines::resize_LineProducingSpecies)
        .def("line_index", &Lines::line_index);

    py::class_<Frequencies>(m, "Frequencies").def_readonly("nu", &Frequencies::nu);

    py::class_<Radiation>(m, "Radiation")
        .def_readonly("frequencies", &Radiation::frequencies)
        .def_property_readonly("u",
                               [](py::object self) {
                                   Radiation &r = self.cast<Radiation &>();
                                   return view_of(r.u, {r.hnrays, r.npoints, r.nfreqs}, self);
                               })
        .def_property_readonly("J",
                               [](py::object self) {
                                   Radiation &r = self.cast<Radiation &>();
                                   return view_of(r.J, {r.npoints, r.nfreqs}, self);
                               })
        .def_property_readonly("I",
                               [](py::object self) {
                                   Radiation &r = self.cast<Radiation &>();
                                   return view_of(r.I, {r.nrays, r.npoints, r.nfreqs}, self);
                               })
        .def_property_readonly("v", [](py::object self) {
            Radiation &r = self.cast<Radiation &>();
            return view_of(r.v, {r.hnrays, r.npoints, r.nfreqs}, self);
        });

    py::class_<Image>(m, "Image")
        .def_readonly("nfreqs", &Image::nfreqs)
        .def_readonly("nx", &Image::nx)
        .def_readonly("ny", &Image::ny)
        .def_readonly("npix", &Image::npix)
        .def_readonly("ray_nr", &Image::ray_nr)
        .def_property_readonly("imageType",
                               [](const Image &im) {
                                   return std::string(im.type == OpticalDepth ? "OpticalDepth"
                                                                              : "Intensity");
                               })
        .def_property_readonly("I",
                               [](py::object self) {
                                   Image &im = self.cast<Image &>();
                                   return view_of(im.I, {(py::ssize_t)im.npix, im.nfreqs}, self);
                               })
        .def_property_readonly("ImX",
                               [](py::object self) {
                                   Image &im = self.cast<Image &>();
                                   return view_of(im.ImX, {(py::ssize_t)im.npix}, self);
                               })
        .def_property_readonly("ImY",
                               [](py::object self) {
                                   Image &im = self.cast<Image &>();
                                   return view_of(im.ImY, {(py::ssize_t)im.npix}, self);
                               })
        .def_property_readonly("freqs",
                               [](py::object self) {
                                   Image &im = self.cast<Image &>();
                                   return view_of(im.freqs, {(py::ssize_t)im.nfreqs}, self);
                               })
        .def_property_readonly("image_direction", [](const Image &im) {
            return py::make_tuple(im.direction[0], im.direction[1], im.direction[2]);
        });

    using RI = py::return_value_policy;
    py::class_<Model>(m, "Model")
        .def(py::init<>())
        .def_readonly("parameters", &Model::parameters, RI::reference_internal)
        .def_readonly("geometry", &Model::geometry, RI::reference_internal)
        .def_readonly("chemistry", &Model::chemistry, RI::reference_internal)
        .def_readonly("thermodynamics", &Model::thermodynamics, RI::reference_internal)
        .def_readonly("lines", &Model::lines, RI::reference_internal)
        .def_readonly("radiation", &Model::radiation, RI::reference_internal)
        .def_readonly("images", &Model::images, RI::reference_internal)
        .def_readonly("error_mean", &Model::error_mean)
        .def_readonly("error_max", &Model::error_max)
        .def_property_readonly("lambda_diag",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda_diag,
                                                  {mo.parameters.npoints(), mo.parameters.nlines()},
                                                  self);
                               })
        .def_property_readonly("lambda_count",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda.count,
                                                  {mo.parameters.npoints(), mo.parameters.nlines()},
                                                  self);
                               })
        .def_property_readonly("lambda_nr",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda.nr,
                                                  {mo.parameters.npoints(), mo.parameters.nlines(),
                                                   mo.lambda.cap},
                                                  self);
                               })
        .def_property_readonly("lambda_val",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda.val,
                                                  {mo.parameters.npoints(), mo.parameters.nlines(),
                                                   mo.lambda.cap},
                                                  self);
                               })
        .def(
            "get_array",
            [](Model &mo, const std::string &name) {
                const auto &v = mo.get_array(name);
                return py::bytes(reinterpret_cast<const char *>(v.data()), v.size());
            },
            "The packed bytes of one array of the device view, by its mgx_model field name.")
        .def("compute_inverse_line_widths", &Model::compute_inverse_line_widths,
             "Compute the inverse line widths (recompute after changing the temperature or "
             "turbulence).")
        .def("compute_spectral_discretisation",
             (int (Model::*)()) & Model::compute_spectral_discretisation,
             "The Gauss-Hermite frequency grid of every line.")
        .def("compute_spectral_discretisation",
             (int (Model::*)(Real, Real)) & Model::compute_spectral_discretisation,
             "A uniform image grid nu_min..nu_max of parameters.nfreqs() bins.")
        .def("compute_spectral_discretisation",
             (int (Model::*)(Real, Real, Size)) & Model::compute_spectral_discretisation,
             "A uniform image grid nu_min..nu_max of n bins.")
        .def("compute_LTE_level_populations", &Model::compute_LTE_level_populations,
             "LTE level populations and the line emissivity and opacity.")
        .def("compute_radiation_field_feautrier_order_2",
             &Model::compute_radiation_field_feautrier_order_2,
             "The radiation field (u, J, Lambda) with the second-order Feautrier solver.")
        .def("compute_radiation_field_feautrier_order_2_sparse",
             &Model::compute_radiation_field_feautrier_order_2_sparse,
             "The line-integrated mean intensity lspec.J and Lambda, without u or J (P, F).")
        .def("compute_radiation_field_shortchar_order_0",
             &Model::compute_radiation_field_shortchar_order_0,
             "The radiation field (I both ways, u, v, J, Lambda diagonal) with the order-0 "
             "short-characteristics solver.")
        .def("compute_radiation_field_feautrier_order_2_uv",
             &Model::compute_radiation_field_feautrier_order_2_uv,
             "u and v with the second-order Feautrier solver.")
        .def("compute_image", &Model::compute_image,
             "Old imager: the intensity at every model point along ray ray_nr.")
        .def("compute_image_optical_depth", &Model::compute_image_optical_depth,
             "Old imager: the optical depth through every model point along ray ray_nr.")
        .def("compute_image_optical_depth_new",
             (int (Model::*)(Size)) & Model::compute_image_optical_depth_new,
             "Optical-depth image along ray ray_nr, 256 x 256 pixels.")
        .def("compute_image_optical_depth_new",
             (int (Model::*)(Size, Size, Size)) & Model::compute_image_optical_depth_new,
             "Optical-depth image along ray ray_nr with Nxpix x Nypix pixels.")
        .def("compute_image_optical_depth_new",
             (int (Model::*)(Real, Real, Real, Size, Size)) & Model::compute_image_optical_depth_new,
             "Optical-depth image along (rx, ry, rz) with Nxpix x Nypix pixels.")
        .def("compute_image_new", (int (Model::*)(Size)) & Model::compute_image_new,
             "New imager along ray ray_nr, 256 x 256 pixels.")
        .def("compute_image_new", (int (Model::*)(Size, Size, Size)) & Model::compute_image_new,
             "New imager along ray ray_nr with Nxpix x Nypix pixels.")
        .def("compute_image_new",
             (int (Model::*)(Real, Real, Real, Size, Size)) & Model::compute_image_new,
             "New imager along (rx, ry, rz) with Nxpix x Nypix pixels.")
        .def("compute_level_populations", &Model::compute_level_populations,
             py::arg("use_Ng_acceleration"), py::arg("max_niterations"),
             "Non-LTE level populations: formal solve with the general Lambda operator "
             "(parameters.n_off_diag bands) and the global statistical equilibrium per "
             "iteration, with Ng acceleration; returns the number of iterations.")
        .def("compute_level_populations_sparse", &Model::compute_level_populations_sparse,
             py::arg("use_Ng_acceleration"), py::arg("max_niterations"),
             "Non-LTE level populations with the Lambda diagonal and the per-point "
             "statistical equilibrium; returns the number of iterations.")
        .def("compute_level_populations_shortchar", &Model::compute_level_populations_shortchar,
             py::arg("use_Ng_acceleration"), py::arg("max_niterations"),
             "Non-LTE level populations with the order-0 short-characteristics formal solve; "
             "returns the number of iterations.");
}
