/*
 * pybind11 module magritte_amd.core: the hot-path subset of the reference's
 * Python surface (src/bindings/pybindings.cpp) over magritte::Model:
 *   Model.compute_radiation_field_feautrier_order_2()          (:185-188)
 *   Model.compute_radiation_field_shortchar_order_0()           (:202-204)
 *   Model.compute_radiation_field_feautrier_order_2_uv()        (:195-197)
 *   Model.compute_image(ray_nr), compute_image_optical_depth(ray_nr) (:225, :241)
 *   Model.compute_image_new / compute_image_optical_depth_new(ray_nr) / (ray_nr, Nx, Ny) /
 *        (rx, ry, rz, Nx, Ny)                                   (:227-258)
 *   Model.parameters knobs of the path                          (:277-327)
 *   model.radiation.u (hnrays, P, F), .J (P, F) and model.images[i].I,
 *   .ImX, .ImY as numpy arrays through the buffer protocol      (:109-118, :735-800)
 * Errors surface as RuntimeError with the library's message.
 */
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>

#include "magritte_amd.hpp"

/* species and collision partners are bound as lists that alias the model's own
   (model.lspecs[0].linedata.A = ... changes the model, as lineProducingSpecies does) */
PYBIND11_MAKE_OPAQUE(std::vector<magritte::LineProducingSpecies>)
PYBIND11_MAKE_OPAQUE(std::vector<magritte::CollisionPartner>)

namespace py = pybind11;
using magritte::CollisionPartner;
using magritte::Image;
using magritte::LambdaNested;
using magritte::Linedata;
using magritte::LineProducingSpecies;
using magritte::Model;
using magritte::Parameters;
using magritte::Radiation;
using magritte::Real;
using magritte::Size;

namespace {

template <typename T>
py::array view_of(std::vector<T> &v, std::vector<py::ssize_t> shape, py::handle base)
{
    if (v.empty()) return py::array(py::dtype::of<T>(), std::vector<py::ssize_t>{0});
    return py::array(py::dtype::of<T>(), shape, v.data(), base);
}

} // namespace

PYBIND11_MODULE(core, m)
{
    m.doc() = "Formal-solver path of Magritte on MI355X (gfx950): Model, Parameters, Radiation, Image";

    py::class_<Parameters>(m, "Parameters")
        .def(py::init<>())
        .def_readwrite("npoints", &Parameters::npoints)
        .def_readwrite("nrays", &Parameters::nrays)
        .def_readwrite("hnrays", &Parameters::hnrays)
        .def_readwrite("nboundary", &Parameters::nboundary)
        .def_readwrite("nlines", &Parameters::nlines)
        .def_readwrite("nlspecs", &Parameters::nlspecs)
        .def_readwrite("nquads", &Parameters::nquads)
        .def_readwrite("nfreqs", &Parameters::nfreqs)
        .def_readwrite("dimension", &Parameters::dimension)
        .def_readwrite("spherical_symmetry", &Parameters::spherical_symmetry)
        .def_readwrite("adaptive_ray", &Parameters::adaptive_ray)
        .def_readwrite("one_line_approximation", &Parameters::one_line_approximation)
        .def_readwrite("sum_opacity_emissivity_over_all_lines",
                       &Parameters::sum_opacity_emissivity_over_all_lines)
        .def_readwrite("n_off_diag", &Parameters::n_off_diag)
        .def_readwrite("min_opacity", &Parameters::min_opacity)
        .def_readwrite("min_dtau", &Parameters::min_dtau)
        .def_readwrite("max_distance_opacity_contribution",
                       &Parameters::max_distance_opacity_contribution)
        .def_readwrite("max_interpolation_diff", &Parameters::max_interpolation_diff)
        .def_readwrite("mixed_precision", &Parameters::mixed_precision)
        .def_readwrite("pop_prec", &Parameters::pop_prec)
        .def_readwrite("min_rel_pop_for_convergence", &Parameters::min_rel_pop_for_convergence)
        .def_readwrite("population_inversion_fraction", &Parameters::population_inversion_fraction)
        .def_readwrite("min_line_emissivity", &Parameters::min_line_emissivity)
        .def_readwrite("min_line_opacity", &Parameters::min_line_opacity)
        .def_readwrite("convergence_fraction", &Parameters::convergence_fraction)
        .def_readwrite("use_adaptive_Ng_acceleration", &Parameters::use_adaptive_Ng_acceleration)
        .def_readwrite("Ng_acceleration_mem_limit", &Parameters::Ng_acceleration_mem_limit)
        .def_readwrite("adaptive_Ng_acceleration_min_order",
                       &Parameters::adaptive_Ng_acceleration_min_order)
        .def_readwrite("adaptive_Ng_acceleration_use_max_criterion",
                       &Parameters::adaptive_Ng_acceleration_use_max_criterion)
        .def_readwrite("Ng_acceleration_remove_N_its", &Parameters::Ng_acceleration_remove_N_its);

    py::class_<CollisionPartner>(m, "CollisionPartner")
        .def(py::init<>())
        .def_readwrite("num_col_partner", &CollisionPartner::num_col_partner)
        .def_readwrite("orth_or_para_H2", &CollisionPartner::orth_or_para_H2)
        .def_readwrite("icol", &CollisionPartner::icol)
        .def_readwrite("jcol", &CollisionPartner::jcol)
        .def_readwrite("tmp", &CollisionPartner::tmp)
        .def_readwrite("Cd", &CollisionPartner::Cd)
        .def_readwrite("Ce", &CollisionPartner::Ce);
    py::bind_vector<std::vector<CollisionPartner>>(m, "VectorCollisionPartner");

    py::class_<Linedata>(m, "Linedata")
        .def(py::init<>())
        .def_readwrite("num", &Linedata::num)
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
        .def_readwrite("colpar", &Linedata::colpar);

    /* lspec.Lambda.Ls / .nr as nested lists [point][transition][element] (pybindings.cpp:570-580) */
    py::class_<LambdaNested>(m, "Lambda")
        .def_readonly("Ls", &LambdaNested::Ls)
        .def_readonly("nr", &LambdaNested::nr);

    py::class_<LineProducingSpecies>(m, "LineProducingSpecies")
        .def(py::init<>())
        .def_readwrite("linedata", &LineProducingSpecies::linedata)
        .def_readonly("Lambda", &LineProducingSpecies::lambda)
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

    py::class_<Radiation>(m, "Radiation")
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
                                   return std::string(im.type == magritte::OpticalDepth
                                                          ? "OpticalDepth"
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

    py::class_<Model>(m, "Model")
        .def(py::init<>())
        .def_readonly("parameters", &Model::parameters, py::return_value_policy::reference_internal)
        .def_readonly("radiation", &Model::radiation, py::return_value_policy::reference_internal)
        .def_readonly("images", &Model::images, py::return_value_policy::reference_internal)
        .def_property_readonly("lambda_diag",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda_diag,
                                                  {mo.parameters.npoints, mo.parameters.nlines},
                                                  self);
                               })
        .def_property_readonly("lambda_count",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda.count,
                                                  {mo.parameters.npoints, mo.parameters.nlines},
                                                  self);
                               })
        .def_property_readonly("lambda_nr",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda.nr,
                                                  {mo.parameters.npoints, mo.parameters.nlines,
                                                   mo.lambda.cap},
                                                  self);
                               })
        .def_property_readonly("lambda_val",
                               [](py::object self) {
                                   Model &mo = self.cast<Model &>();
                                   return view_of(mo.lambda.val,
                                                  {mo.parameters.npoints, mo.parameters.nlines,
                                                   mo.lambda.cap},
                                                  self);
                               })
        .def("set_array",
             [](Model &mo, const std::string &name, py::buffer b) {
                 py::buffer_info info = b.request();
                 mo.set_array(name, info.ptr, (size_t)info.size * info.itemsize);
             })
        .def("set_center",
             [](Model &mo, Real x, Real y, Real z) { mo.set_center({x, y, z}); })
        .def("set_max_inverse_mass", &Model::set_max_inverse_mass)
        .def("compute_radiation_field_feautrier_order_2",
             &Model::compute_radiation_field_feautrier_order_2,
             "Compute the radiation field (u, J, Lambda diagonal) with the second-order "
             "Feautrier solver.")
        .def("compute_radiation_field_shortchar_order_0",
             &Model::compute_radiation_field_shortchar_order_0,
             "Compute the radiation field (I both ways, u, v, J, Lambda diagonal) with the "
             "order-0 short-characteristics solver.")
        .def("compute_radiation_field_feautrier_order_2_uv",
             &Model::compute_radiation_field_feautrier_order_2_uv,
             "Compute u and v with the second-order Feautrier solver.")
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
        .def("compute_spectral_discretisation",
             (int (Model::*)(Real, Real, Size)) & Model::compute_spectral_discretisation,
             "Uniform image frequency grid nu_min..nu_max with n points.")
        .def("restore_line_grid", &Model::restore_line_grid)
        .def_readwrite("lspecs", &Model::lspecs,
                       "lines.lineProducingSpecies: line data, populations and Lambda")
        .def_property(
            "abundance",
            [](py::object self) {
                Model &mo = self.cast<Model &>();
                return view_of(mo.abundance, {mo.parameters.npoints, mo.nspec}, self);
            },
            [](Model &mo, py::array_t<double, py::array::c_style | py::array::forcecast> a) {
                if (a.ndim() != 2) throw std::runtime_error("abundance must be (npoints, nspec)");
                mo.nspec = (Size)a.shape(1);
                mo.abundance.assign(a.data(), a.data() + a.size());
            },
            "chemistry.species.abundance (npoints, nspec)")
        .def(
            "get_array",
            [](const Model &mo, const std::string &name) {
                const auto &v = mo.get_array(name);
                return py::bytes(reinterpret_cast<const char *>(v.data()), v.size());
            },
            "The bytes of a model array (as set, or as the level-population loop left it).")
        .def_readonly("error_mean", &Model::error_mean)
        .def_readonly("error_max", &Model::error_max)
        .def("compute_LTE_level_populations", &Model::compute_LTE_level_populations,
             "LTE level populations and the line emissivity, opacity, interpolation criterion.")
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
