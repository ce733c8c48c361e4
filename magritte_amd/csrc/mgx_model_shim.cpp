/*
 * magritte::Model / magritte::Solver (include/magritte_amd.hpp): the reference's
 * C++ entry points of the formal-solver path on top of the C ABI.
 *   Model::compute_radiation_field_feautrier_order_2   src/model/model.cpp:353-390
 *   Model::compute_radiation_field_shortchar_order_0   src/model/model.cpp:313-351
 *   Model::compute_radiation_field_feautrier_order_2_uv src/model/model.cpp:395-433
 *   Model::compute_image, compute_image_optical_depth  src/model/model.cpp:957-976, 1061-1076
 *   Model::compute_image_new, compute_image_optical_depth_new (all overloads)
 *                                                      src/model/model.cpp:987-1035, 1081-1127
 *   Model::compute_spectral_discretisation(min,max,n)  src/model/model.cpp:257-298
 */
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>

#include "magritte_amd.hpp"

namespace magritte {

namespace {

/* (field name, element size) of every array of mgx_model */
struct Field {
    const char *name;
    size_t elem;
};
const Field kFields[] = {
    {"position", 8},         {"velocity", 8},          {"n_neighbors", 4},
    {"cum_n_neighbors", 4},  {"neighbors", 4},         {"point2boundary", 4},
    {"boundary2point", 4},   {"boundary_condition", 4}, {"boundary_temperature", 8},
    {"ray_direction", 8},    {"ray_weight", 8},        {"antipod", 4},
    {"temperature", 8},      {"vturb2", 8},            {"line", 8},
    {"sorted_line", 8},      {"sorted_line_map", 4},   {"emissivity", 8},
    {"opacity", 8},          {"inverse_width", 8},     {"interp_crit", 8},
    {"dshift_max", 8},       {"line_species", 4},      {"line_k", 4},
    {"line_A", 8},           {"species_inverse_mass", 8}, {"quad_weights", 8},
    {"nu", 8},               {"f_line", 4},            {"f_species", 4},
    {"f_k", 4},              {"f_z", 4},               {"f_in_line", 1},
};

const char *kFreqFields[] = {"nu", "f_line", "f_species", "f_k", "f_z", "f_in_line"};

void check(int code)
{
    if (code != MGX_OK) throw std::runtime_error(mgx_last_error());
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
const T *field_ptr(const std::map<std::string, void *> &dev, const char *name)
{
    auto it = dev.find(name);
    return it == dev.end() ? nullptr : static_cast<const T *>(it->second);
}

/* a device buffer the solver writes, freed at scope end */
struct DevOut {
    double *p = nullptr;
    explicit DevOut(size_t n)
    {
        if (n) hip_check(hipMalloc(&p, n * sizeof(double)), "hipMalloc");
    }
    ~DevOut()
    {
        if (p) (void)hipFree(p);
    }
    void copy_to(std::vector<double> &v, size_t n) const
    {
        v.resize(n);
        if (n) hip_check(hipMemcpy(v.data(), p, n * sizeof(double), hipMemcpyDeviceToHost), "copy");
    }
};

template <class T>
void copy_from_device(std::vector<T> &v, const T *src, size_t n)
{
    v.resize(n);
    if (n) hip_check(hipMemcpy(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost), "copy");
}

} // namespace

Model::~Model()
{
    for (auto &kv : dev)
        if (kv.second) (void)hipFree(kv.second);
}

void Model::set_array(const std::string &name, const void *data, size_t bytes)
{
    bool known = false;
    for (const Field &f : kFields) known = known || name == f.name;
    if (!known) throw std::runtime_error("unknown model array '" + name + "'");
    auto &v = host[name];
    v.assign(static_cast<const unsigned char *>(data), static_cast<const unsigned char *>(data) + bytes);
    stale[name] = true;
}

const std::vector<unsigned char> &Model::get_array(const std::string &name) const
{
    auto it = host.find(name);
    if (it == host.end()) throw std::runtime_error("model array '" + name + "' is not set");
    return it->second;
}

const mgx_model &Model::device_view()
{
    for (auto &kv : host) {
        if (!stale[kv.first]) continue;
        void *&d = dev[kv.first];
        if (d) (void)hipFree(d);
        d = nullptr;
        if (!kv.second.empty()) {
            hip_check(hipMalloc(&d, kv.second.size()), "hipMalloc");
            hip_check(hipMemcpy(d, kv.second.data(), kv.second.size(), hipMemcpyHostToDevice),
                      "hipMemcpy");
        }
        stale[kv.first] = false;
    }
    const Parameters &p = parameters;
    mgx_model &v = view;
    std::memset(&v, 0, sizeof v);
    v.npoints = p.npoints;
    v.nrays = p.nrays;
    v.hnrays = p.hnrays;
    v.nboundary = p.nboundary;
    v.nlines = p.nlines;
    v.nlspecs = p.nlspecs;
    v.nquads = p.nquads;
    v.nfreqs = p.nfreqs;
    v.dimension = (int32_t)p.dimension;
    v.spherical_symmetry = p.spherical_symmetry;
    v.adaptive_directions = p.adaptive_ray;
    v.position = field_ptr<double>(dev, "position");
    v.velocity = field_ptr<double>(dev, "velocity");
    v.n_neighbors = field_ptr<uint32_t>(dev, "n_neighbors");
    v.cum_n_neighbors = field_ptr<uint32_t>(dev, "cum_n_neighbors");
    v.neighbors = field_ptr<uint32_t>(dev, "neighbors");
    v.point2boundary = field_ptr<uint32_t>(dev, "point2boundary");
    v.boundary2point = field_ptr<uint32_t>(dev, "boundary2point");
    v.boundary_condition = field_ptr<int32_t>(dev, "boundary_condition");
    v.boundary_temperature = field_ptr<double>(dev, "boundary_temperature");
    v.ray_direction = field_ptr<double>(dev, "ray_direction");
    v.ray_weight = field_ptr<double>(dev, "ray_weight");
    v.antipod = field_ptr<uint32_t>(dev, "antipod");
    for (int i = 0; i < 3; i++) v.center[i] = center[i];
    v.temperature = field_ptr<double>(dev, "temperature");
    v.vturb2 = field_ptr<double>(dev, "vturb2");
    v.line = field_ptr<double>(dev, "line");
    v.sorted_line = field_ptr<double>(dev, "sorted_line");
    v.sorted_line_map = field_ptr<uint32_t>(dev, "sorted_line_map");
    v.emissivity = field_ptr<double>(dev, "emissivity");
    v.opacity = field_ptr<double>(dev, "opacity");
    v.inverse_width = field_ptr<double>(dev, "inverse_width");
    v.interp_crit = field_ptr<double>(dev, "interp_crit");
    v.dshift_max = field_ptr<double>(dev, "dshift_max");
    v.max_inverse_mass = max_inverse_mass;
    v.line_species = field_ptr<uint32_t>(dev, "line_species");
    v.line_k = field_ptr<uint32_t>(dev, "line_k");
    v.line_A = field_ptr<double>(dev, "line_A");
    v.species_inverse_mass = field_ptr<double>(dev, "species_inverse_mass");
    v.quad_weights = field_ptr<double>(dev, "quad_weights");
    v.nu = field_ptr<double>(dev, "nu");
    v.f_line = field_ptr<uint32_t>(dev, "f_line");
    v.f_species = field_ptr<uint32_t>(dev, "f_species");
    v.f_k = field_ptr<uint32_t>(dev, "f_k");
    v.f_z = field_ptr<uint32_t>(dev, "f_z");
    v.f_in_line = field_ptr<uint8_t>(dev, "f_in_line");
    return view;
}

mgx_options Model::options() const
{
    mgx_options o;
    std::memset(&o, 0, sizeof o);
    /* the approximation dispatch of model.cpp:366-389 */
    o.approx = parameters.one_line_approximation ? MGX_APPROX_ONE_LINE
               : parameters.sum_opacity_emissivity_over_all_lines ? MGX_APPROX_NONE
                                                                  : MGX_APPROX_CLOSE_LINES;
    o.n_off_diag = parameters.n_off_diag;
    o.precision = parameters.mixed_precision ? MGX_PREC_MIXED : MGX_PREC_FP64;
    o.min_opacity = parameters.min_opacity;
    o.max_distance_opacity_contribution = parameters.max_distance_opacity_contribution;
    o.max_interpolation_diff = parameters.max_interpolation_diff;
    return o;
}

int Model::compute_radiation_field_feautrier_order_2()
{
    Solver solver;
    solver.setup(*this);
    solver.solve_feautrier_order_2(*this);
    return 0;
}

int Model::compute_radiation_field_shortchar_order_0()
{
    Solver solver;
    solver.setup(*this);
    solver.solve_shortchar_order_0(*this);
    return 0;
}

int Model::compute_radiation_field_feautrier_order_2_uv()
{
    Solver solver;
    solver.setup(*this);
    solver.solve_feautrier_order_2_uv(*this);
    return 0;
}

int Model::compute_image(Size ray_nr)
{
    if (parameters.one_line_approximation)
        throw std::runtime_error("One line approximation is currently not supported for imaging.");
    Solver solver;
    solver.image_feautrier_order_2(*this, ray_nr);
    return 0;
}

int Model::compute_image_optical_depth(Size ray_nr)
{
    if (parameters.one_line_approximation)
        throw std::runtime_error("One line approximation is currently not supported for imaging.");
    Solver solver;
    solver.image_optical_depth(*this, ray_nr);
    return 0;
}

namespace {

/* geometry.rays.get_direction<false>(0, ray_nr) */
Vector3D ray_direction_of(const Model &model, Size ray_nr)
{
    const auto &d = model.get_array("ray_direction");
    if ((size_t)(ray_nr + 1) * 3 * sizeof(double) > d.size())
        throw std::runtime_error("ray_nr out of range");
    Vector3D r;
    std::memcpy(r.data(), d.data() + (size_t)ray_nr * 3 * sizeof(double), sizeof r);
    return r;
}

void check_image_direction(const Model &model, const Vector3D &raydir)
{
    /* model.cpp:1009-1024: the checks come before any work */
    if (raydir[0] * raydir[0] + raydir[1] * raydir[1] + raydir[2] * raydir[2] == 0.0)
        throw std::runtime_error(
            "The given ray direction vector does not point in a direction. Please use a "
            "non-zero (normed) direction vector to generate an image.");
    if (model.parameters.one_line_approximation)
        throw std::runtime_error("One line approximation is not supported for imaging.");
}

} // namespace

int Model::compute_image_optical_depth_new(Size ray_nr)
{
    return compute_image_optical_depth_new(ray_nr, 256, 256);
}

int Model::compute_image_optical_depth_new(Size ray_nr, Size Nxpix, Size Nypix)
{
    return compute_image_optical_depth_new(ray_direction_of(*this, ray_nr), Nxpix, Nypix);
}

int Model::compute_image_optical_depth_new(Real rx, Real ry, Real rz, Size Nxpix, Size Nypix)
{
    return compute_image_optical_depth_new(Vector3D{rx, ry, rz}, Nxpix, Nypix);
}

int Model::compute_image_optical_depth_new(const Vector3D raydir, Size Nxpix, Size Nypix)
{
    check_image_direction(*this, raydir);
    Solver solver;
    solver.image_optical_depth_new_imager(*this, raydir, Nxpix, Nypix);
    return 0;
}

int Model::compute_image_new(Size ray_nr) { return compute_image_new(ray_nr, 256, 256); }

int Model::compute_image_new(Size ray_nr, Size Nxpix, Size Nypix)
{
    return compute_image_new(ray_direction_of(*this, ray_nr), Nxpix, Nypix);
}

int Model::compute_image_new(Real rx, Real ry, Real rz, Size Nxpix, Size Nypix)
{
    return compute_image_new(Vector3D{rx, ry, rz}, Nxpix, Nypix);
}

int Model::compute_image_new(const Vector3D raydir, Size Nxpix, Size Nypix)
{
    check_image_direction(*this, raydir);
    Solver solver;
    solver.image_feautrier_order_2_new_imager(*this, raydir, Nxpix, Nypix);
    return 0;
}

int Model::compute_spectral_discretisation(Real nu_min, Real nu_max, Size n)
{
    const Size P = parameters.npoints;
    if (saved_line_grid.empty())
        for (const char *f : kFreqFields) saved_line_grid[f] = host[f];
    std::vector<double> nu((size_t)P * n);
    std::vector<uint32_t> fl(n), fs(n), fk(n), fz(n);
    std::vector<uint8_t> fin(n);
    check(mgx_compute_image_discretisation(P, nu_min, nu_max, n, nu.data(), fl.data(), fs.data(),
                                           fk.data(), fz.data(), fin.data()));
    set_array("nu", nu.data(), nu.size() * sizeof(double));
    set_array("f_line", fl.data(), n * 4);
    set_array("f_species", fs.data(), n * 4);
    set_array("f_k", fk.data(), n * 4);
    set_array("f_z", fz.data(), n * 4);
    set_array("f_in_line", fin.data(), n);
    parameters.nfreqs = n;
    return 0;
}

void Model::restore_line_grid()
{
    if (saved_line_grid.empty()) return;
    for (auto &kv : saved_line_grid) set_array(kv.first, kv.second.data(), kv.second.size());
    parameters.nfreqs = (Size)(get_array("f_line").size() / 4);
    saved_line_grid.clear();
}

void Solver::setup(Model &model)
{
    /* the traced-length pass of the reference's setup happens on the device
       (walk capacities grow on demand); only the device mirror is refreshed */
    (void)model.device_view();
}

namespace {

/* lspec.lambda of every species from the model-wide lists (or the diagonal): the
   reference's Lambda::add_element order, line l of the model being transition
   line_k[l] of species line_species[l] */
void fill_species_lambda(Model &model)
{
    const Parameters &p = model.parameters;
    if (model.lspecs.size() != p.nlspecs) return;
    const auto &ls_b = model.get_array("line_species");
    const auto &lk_b = model.get_array("line_k");
    const uint32_t *lsp = reinterpret_cast<const uint32_t *>(ls_b.data());
    const uint32_t *lk = reinterpret_cast<const uint32_t *>(lk_b.data());
    const size_t P = p.npoints, L = p.nlines;
    for (auto &s : model.lspecs) {
        s.lambda.Ls.assign(P, std::vector<std::vector<Real>>(s.linedata.nrad));
        s.lambda.nr.assign(P, std::vector<std::vector<Size>>(s.linedata.nrad));
    }
    const LambdaLists &G = model.lambda;
    for (size_t q = 0; q < P; q++)
        for (size_t l = 0; l < L; l++) {
            LineProducingSpecies &s = model.lspecs.at(lsp[l]);
            if (lk[l] >= s.linedata.nrad) throw std::runtime_error("line_k out of range");
            auto &Ls = s.lambda.Ls[q][lk[l]];
            auto &nr = s.lambda.nr[q][lk[l]];
            if (G.cap == 0) {
                Ls.assign(1, model.lambda_diag[q * L + l]);
                nr.assign(1, (Size)q);
            } else {
                const size_t i = q * L + l, n = G.count[i];
                Ls.assign(G.val.begin() + i * G.cap, G.val.begin() + i * G.cap + n);
                nr.assign(G.nr.begin() + i * G.cap, G.nr.begin() + i * G.cap + n);
            }
        }
}

void set_shapes(Radiation &r, const mgx_model &v)
{
    r.nrays = v.nrays;
    r.hnrays = v.hnrays;
    r.npoints = v.npoints;
    r.nfreqs = v.nfreqs;
}

/* a device buffer of 32-bit integers, freed at scope end */
struct DevOutU32 {
    uint32_t *p = nullptr;
    explicit DevOutU32(size_t n)
    {
        if (n) hip_check(hipMalloc(&p, n * sizeof(uint32_t)), "hipMalloc");
    }
    ~DevOutU32()
    {
        if (p) (void)hipFree(p);
    }
    void copy_to(std::vector<uint32_t> &v, size_t n) const
    {
        v.resize(n);
        if (n) hip_check(hipMemcpy(v.data(), p, n * sizeof(uint32_t), hipMemcpyDeviceToHost), "copy");
    }
};

} // namespace

void Solver::solve_feautrier_order_2(Model &model)
{
    const mgx_model &v = model.device_view();
    const mgx_options o = model.options();
    const size_t PF = (size_t)v.npoints * v.nfreqs;
    const size_t U = (size_t)v.hnrays * PF;
    const size_t PL = (size_t)v.npoints * v.nlines;
    DevOut u(U), J(PF);
    model.radiation.I.clear();
    model.radiation.v.clear();
    if (o.n_off_diag == 0) {
        DevOut lam(PL);
        check(mgx_feautrier_forward_device(&v, &o, u.p, J.p, lam.p, nullptr));
        lam.copy_to(model.lambda_diag, PL);
        model.lambda = LambdaLists{};
    } else {
        /* the general operator: lists of the diagonal and the n_off_diag bands */
        const Size cap = 1 + 2 * o.n_off_diag * v.hnrays;
        DevOutU32 cnt(PL), nr(PL * cap);
        DevOut val(PL * cap);
        check(mgx_feautrier_forward_lambda_device(&v, &o, u.p, J.p, cap, cnt.p, nr.p, val.p,
                                                  nullptr));
        LambdaLists &L = model.lambda;
        L.cap = cap;
        cnt.copy_to(L.count, PL);
        nr.copy_to(L.nr, PL * cap);
        val.copy_to(L.val, PL * cap);
        model.lambda_diag.assign(PL, 0.0);
        for (size_t i = 0; i < PL; i++)
            if (L.count[i]) model.lambda_diag[i] = L.val[i * cap];
    }
    u.copy_to(model.radiation.u, U);
    J.copy_to(model.radiation.J, PF);
    set_shapes(model.radiation, v);
    fill_species_lambda(model);
}

void Solver::solve_shortchar_order_0(Model &model)
{
    const mgx_model &v = model.device_view();
    const mgx_options o = model.options();
    const size_t PF = (size_t)v.npoints * v.nfreqs;
    const size_t PL = (size_t)v.npoints * v.nlines;
    DevOut I((size_t)v.nrays * PF), u((size_t)v.hnrays * PF), vv((size_t)v.hnrays * PF), J(PF),
        lam(PL);
    check(mgx_shortchar_order_0_device(&v, &o, model.parameters.min_dtau, I.p, u.p, vv.p, J.p,
                                       lam.p, nullptr));
    I.copy_to(model.radiation.I, (size_t)v.nrays * PF);
    u.copy_to(model.radiation.u, (size_t)v.hnrays * PF);
    vv.copy_to(model.radiation.v, (size_t)v.hnrays * PF);
    J.copy_to(model.radiation.J, PF);
    lam.copy_to(model.lambda_diag, PL);
    set_shapes(model.radiation, v);
    /* its Lambda is diagonal whatever n_off_diag says (solver.tpp:164-166, 1651, 1754) */
    model.lambda = LambdaLists{};
    fill_species_lambda(model);
}

void Solver::solve_feautrier_order_2_uv(Model &model)
{
    const mgx_model &v = model.device_view();
    const mgx_options o = model.options();
    const size_t U = (size_t)v.hnrays * v.npoints * v.nfreqs;
    DevOut u(U), vv(U);
    check(mgx_feautrier_order_2_uv_device(&v, &o, u.p, vv.p, nullptr));
    u.copy_to(model.radiation.u, U);
    vv.copy_to(model.radiation.v, U);
    model.radiation.I.clear();
    model.radiation.J.clear();
    set_shapes(model.radiation, v);
}

namespace {

/* the image frequencies: the grid at the first point (uniform over points) */
void set_image_freqs(const Model &model, Image &img, Size F)
{
    const auto &nu = model.get_array("nu");
    img.freqs.resize(F);
    std::memcpy(img.freqs.data(), nu.data(), F * sizeof(double));
}

} // namespace

void Solver::image_feautrier_order_2(Model &model, Size rr)
{
    const mgx_model &v = model.device_view();
    const mgx_options o = model.options();
    const size_t P = v.npoints;
    DevOut I(P * v.nfreqs), X(P), Y(P);
    check(mgx_image_ray_device(&v, &o, rr, I.p, X.p, Y.p, nullptr));
    Image img;
    img.type = Intensity;
    img.ray_nr = rr;
    img.nx = v.npoints;
    img.ny = 1;
    img.npix = v.npoints;
    img.nfreqs = v.nfreqs;
    img.direction = ray_direction_of(model, rr);
    I.copy_to(img.I, P * v.nfreqs);
    X.copy_to(img.ImX, P);
    Y.copy_to(img.ImY, P);
    set_image_freqs(model, img, v.nfreqs);
    model.images.push_back(std::move(img));
}

void Solver::image_optical_depth(Model &model, Size rr)
{
    const mgx_model &v = model.device_view();
    const mgx_options o = model.options();
    const size_t P = v.npoints;
    DevOut T(P * v.nfreqs), X(P), Y(P);
    check(mgx_image_optical_depth_ray_device(&v, &o, rr, T.p, X.p, Y.p, nullptr));
    Image img;
    img.type = OpticalDepth;
    img.ray_nr = rr;
    img.nx = v.npoints;
    img.ny = 1;
    img.npix = v.npoints;
    img.nfreqs = v.nfreqs;
    img.direction = ray_direction_of(model, rr);
    T.copy_to(img.I, P * v.nfreqs);
    X.copy_to(img.ImX, P);
    Y.copy_to(img.ImY, P);
    set_image_freqs(model, img, v.nfreqs);
    model.images.push_back(std::move(img));
}

void Solver::image_optical_depth_new_imager(Model &model, const Vector3D raydir, Size Nxpix,
                                            Size Nypix)
{
    const mgx_model &v = model.device_view();
    const mgx_options o = model.options();
    const size_t npix = (size_t)Nxpix * Nypix;
    DevOut T(npix * v.nfreqs), X(npix), Y(npix);
    check(mgx_image_optical_depth_device(&v, &o, raydir.data(), Nxpix, Nypix, T.p, X.p, Y.p,
                                         nullptr));
    Image img;
    img.type = OpticalDepth;
    img.nx = Nxpix;
    img.ny = Nypix;
    img.npix = (Size)npix;
    img.nfreqs = v.nfreqs;
    const Real n = std::sqrt(raydir[0] * raydir[0] + raydir[1] * raydir[1] + raydir[2] * raydir[2]);
    for (int i = 0; i < 3; i++) img.direction[i] = raydir[i] / n;
    T.copy_to(img.I, npix * v.nfreqs);
    X.copy_to(img.ImX, npix);
    Y.copy_to(img.ImY, npix);
    set_image_freqs(model, img, v.nfreqs);
    model.images.push_back(std::move(img));
}

void Solver::image_feautrier_order_2_new_imager(Model &model, const Vector3D raydir, Size Nxpix,
                                                Size Nypix)
{
    const mgx_model &v = model.device_view();
    const mgx_options o = model.options();
    const size_t npix = (size_t)Nxpix * Nypix;
    DevOut I(npix * v.nfreqs), X(npix), Y(npix);
    check(mgx_image_device(&v, &o, raydir.data(), Nxpix, Nypix, I.p, X.p, Y.p, nullptr));
    Image img;
    img.type = Intensity;
    img.nx = Nxpix;
    img.ny = Nypix;
    img.npix = (Size)npix;
    img.nfreqs = v.nfreqs;
    const Real n = std::sqrt(raydir[0] * raydir[0] + raydir[1] * raydir[1] + raydir[2] * raydir[2]);
    for (int i = 0; i < 3; i++) img.direction[i] = raydir[i] / n;
    I.copy_to(img.I, npix * v.nfreqs);
    X.copy_to(img.ImX, npix);
    Y.copy_to(img.ImY, npix);
    set_image_freqs(model, img, v.nfreqs);
    model.images.push_back(std::move(img));
}

/* ------------------------------------------------------------------------- */
/* level populations                                                         */

namespace {

/* the C ABI's view of the species' line and collision data (pointers into lspecs) */
struct NlteTables {
    std::vector<mgx_linedata> ld;
    std::vector<uint32_t> ncolpar;
    std::vector<mgx_colldata> colpar;
    size_t npop = 0; /* sum over species of P * nlev */
};

NlteTables nlte_tables(const Model &model)
{
    const Parameters &p = model.parameters;
    if (model.lspecs.size() != p.nlspecs)
        throw std::runtime_error("lspecs must hold parameters.nlspecs line-producing species");
    if (model.abundance.size() != (size_t)p.npoints * model.nspec)
        throw std::runtime_error("abundance must be (npoints, nspec)");
    NlteTables t;
    for (const LineProducingSpecies &s : model.lspecs) {
        const Linedata &d = s.linedata;
        if (d.energy.size() != d.nlev || d.weight.size() != d.nlev || d.irad.size() != d.nrad ||
            d.jrad.size() != d.nrad || d.A.size() != d.nrad || d.Bs.size() != d.nrad ||
            d.Ba.size() != d.nrad || d.frequency.size() != d.nrad)
            throw std::runtime_error("line data arrays do not match nlev / nrad");
        mgx_linedata x;
        std::memset(&x, 0, sizeof x);
        x.num = d.num;
        x.nlev = d.nlev;
        x.nrad = d.nrad;
        x.inverse_mass = d.inverse_mass;
        x.energy = d.energy.data();
        x.weight = d.weight.data();
        x.irad = d.irad.data();
        x.jrad = d.jrad.data();
        x.A = d.A.data();
        x.Bs = d.Bs.data();
        x.Ba = d.Ba.data();
        x.frequency = d.frequency.data();
        t.ld.push_back(x);
        t.ncolpar.push_back((uint32_t)d.colpar.size());
        for (const CollisionPartner &c : d.colpar) {
            const size_t ntmp = c.tmp.size(), ncol = c.icol.size();
            if (c.jcol.size() != ncol || c.Cd.size() != ntmp * ncol || c.Ce.size() != ntmp * ncol)
                throw std::runtime_error("collision data arrays do not match ntmp / ncol");
            mgx_colldata y;
            std::memset(&y, 0, sizeof y);
            y.num_col_partner = c.num_col_partner;
            y.orth_or_para_H2 = c.orth_or_para_H2;
            y.ntmp = (uint32_t)ntmp;
            y.ncol = (uint32_t)ncol;
            y.icol = c.icol.data();
            y.jcol = c.jcol.data();
            y.tmp = c.tmp.data();
            y.Cd = c.Cd.data();
            y.Ce = c.Ce.data();
            t.colpar.push_back(y);
        }
        t.npop += (size_t)p.npoints * d.nlev;
    }
    return t;
}

mgx_nlte_params nlte_params(const Parameters &p)
{
    mgx_nlte_params q;
    mgx_nlte_default_params(&q);
    q.pop_prec = p.pop_prec;
    q.min_rel_pop_for_convergence = p.min_rel_pop_for_convergence;
    q.population_inversion_fraction = p.population_inversion_fraction;
    q.min_line_emissivity = p.min_line_emissivity;
    q.min_line_opacity = p.min_line_opacity;
    q.convergence_fraction = p.convergence_fraction;
    q.use_adaptive_Ng_acceleration = p.use_adaptive_Ng_acceleration ? 1 : 0;
    q.Ng_acceleration_mem_limit = p.Ng_acceleration_mem_limit;
    q.adaptive_Ng_acceleration_min_order = p.adaptive_Ng_acceleration_min_order;
    q.adaptive_Ng_acceleration_use_max_criterion = p.adaptive_Ng_acceleration_use_max_criterion ? 1 : 0;
    q.Ng_acceleration_remove_N_its = p.Ng_acceleration_remove_N_its;
    return q;
}

/* the populations of every species, concatenated in species order */
std::vector<double> joined_populations(const Model &model, size_t npop)
{
    std::vector<double> pop;
    pop.reserve(npop);
    for (const LineProducingSpecies &s : model.lspecs)
        pop.insert(pop.end(), s.population.begin(), s.population.end());
    return pop;
}

void split_populations(Model &model, const std::vector<double> &pop,
                       const std::vector<double> &ptot)
{
    const size_t P = model.parameters.npoints;
    size_t off = 0;
    for (size_t s = 0; s < model.lspecs.size(); s++) {
        LineProducingSpecies &ls = model.lspecs[s];
        const size_t n = P * ls.linedata.nlev;
        ls.population.assign(pop.begin() + off, pop.begin() + off + n);
        ls.population_tot.assign(ptot.begin() + s * P, ptot.begin() + (s + 1) * P);
        off += n;
    }
}

struct NlteHandle {
    mgx_nlte *h = nullptr;
    ~NlteHandle()
    {
        if (h) mgx_nlte_destroy(h);
    }
};

} // namespace

int Model::compute_LTE_level_populations()
{
    const NlteTables t = nlte_tables(*this);
    const Parameters &p = parameters;
    const size_t P = p.npoints, PL = (size_t)p.npoints * p.nlines;
    const auto &T = get_array("temperature");
    if (T.size() != P * sizeof(double)) throw std::runtime_error("temperature must be (npoints)");
    std::vector<double> pop(t.npop), ptot((size_t)p.nlspecs * P), em(PL), op(PL), ic(PL);
    check(mgx_compute_lte(p.npoints, nspec, abundance.data(),
                          reinterpret_cast<const double *>(T.data()), p.nlspecs, t.ld.data(),
                          p.min_line_emissivity, p.min_line_opacity, pop.data(), ptot.data(),
                          em.data(), op.data()));
    check(mgx_compute_interp_crit(p.npoints, p.nlspecs, t.ld.data(), op.data(), ptot.data(),
                                  ic.data()));
    set_array("emissivity", em.data(), PL * sizeof(double));
    set_array("opacity", op.data(), PL * sizeof(double));
    set_array("interp_crit", ic.data(), PL * sizeof(double));
    split_populations(*this, pop, ptot);
    return 0;
}

/* kind: 0 the general loop, 1 the sparse loop, 2 the short-characteristics loop */
int Model::level_population_loop(int kind, bool use_Ng_acceleration, long max_niterations)
{
    const Parameters &p = parameters;
    NlteTables t = nlte_tables(*this);
    bool have = true;
    for (const LineProducingSpecies &s : lspecs)
        have = have && s.population.size() == (size_t)p.npoints * s.linedata.nlev;
    if (!have) compute_LTE_level_populations();
    const mgx_model &v = device_view();
    const mgx_options o = options();
    const mgx_nlte_params q = nlte_params(p);
    const size_t P = p.npoints, PL = (size_t)p.npoints * p.nlines;
    DevOut ab(abundance.size()), pd(t.npop), pt((size_t)p.nlspecs * P);
    hip_check(hipMemcpy(ab.p, abundance.data(), abundance.size() * sizeof(double),
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    const std::vector<double> pop0 = joined_populations(*this, t.npop);
    hip_check(hipMemcpy(pd.p, pop0.data(), t.npop * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy");
    NlteHandle h;
    check(mgx_nlte_create(&v, p.nlspecs, t.ld.data(), t.ncolpar.data(),
                          t.colpar.empty() ? nullptr : t.colpar.data(), nspec, ab.p, &q, &h.h,
                          nullptr));
    /* the line data the loop rewrites are the device mirrors the view points to */
    double *em = const_cast<double *>(v.emissivity);
    double *op = const_cast<double *>(v.opacity);
    double *ic = const_cast<double *>(v.interp_crit);
    check(mgx_nlte_set_populations(h.h, pd.p, em, op, ic, nullptr));
    const size_t cap = (size_t)(max_niterations > 0 ? max_niterations : 1) * p.nlspecs;
    std::vector<double> emean(cap), emax(cap);
    int32_t it = 0;
    if (kind == 1)
        check(mgx_compute_level_populations_sparse(&v, &o, h.h, use_Ng_acceleration ? 1 : 0,
                                                   max_niterations, em, op, ic, &it,
                                                   emean.data(), emax.data(), nullptr));
    else if (kind == 2)
        check(mgx_compute_level_populations_shortchar(&v, &o, h.h, use_Ng_acceleration ? 1 : 0,
                                                      max_niterations, p.min_dtau, em, op, ic,
                                                      &it, emean.data(), emax.data(), nullptr));
    else
        check(mgx_compute_level_populations(&v, &o, h.h, use_Ng_acceleration ? 1 : 0,
                                            max_niterations, em, op, ic, &it, emean.data(),
                                            emax.data(), nullptr));
    check(mgx_nlte_get_populations(h.h, pd.p, pt.p, nullptr));
    if (it > 0) {
        /* lspec.lambda: the Lambda of the loop's last formal solve */
        uint32_t lcap = 0;
        const uint32_t *lcount = nullptr, *lnr = nullptr;
        const double *lval = nullptr, *ldiag = nullptr;
        check(mgx_nlte_last_lambda(h.h, &lcap, &lcount, &lnr, &lval, &ldiag));
        if (lcap == 0) {
            lambda = LambdaLists{};
            copy_from_device(lambda_diag, ldiag, PL);
        } else {
            lambda = LambdaLists{};
            lambda.cap = lcap;
            copy_from_device(lambda.count, lcount, PL);
            copy_from_device(lambda.nr, lnr, PL * lcap);
            copy_from_device(lambda.val, lval, PL * lcap);
            lambda_diag.assign(PL, 0.0);
            for (size_t i = 0; i < PL; i++)
                if (lambda.count[i]) lambda_diag[i] = lambda.val[i * lcap];
        }
        fill_species_lambda(*this);
    }
    std::vector<double> pop, ptot, line;
    pd.copy_to(pop, t.npop);
    pt.copy_to(ptot, (size_t)p.nlspecs * P);
    split_populations(*this, pop, ptot);
    /* the host arrays take the loop's line data (re-uploaded on the next solve) */
    const char *names[] = {"emissivity", "opacity", "interp_crit"};
    double *arrays[] = {em, op, ic};
    for (int i = 0; i < 3; i++) {
        line.resize(PL);
        hip_check(hipMemcpy(line.data(), arrays[i], PL * sizeof(double), hipMemcpyDeviceToHost),
                  "copy");
        set_array(names[i], line.data(), PL * sizeof(double));
    }
    const size_t n = (size_t)it * p.nlspecs;
    error_mean.assign(emean.begin(), emean.begin() + (n < cap ? n : cap));
    error_max.assign(emax.begin(), emax.begin() + (n < cap ? n : cap));
    return it;
}

int Model::compute_level_populations(bool use_Ng_acceleration, long max_niterations)
{
    return level_population_loop(0, use_Ng_acceleration, max_niterations);
}

int Model::compute_level_populations_sparse(bool use_Ng_acceleration, long max_niterations)
{
    return level_population_loop(1, use_Ng_acceleration, max_niterations);
}

int Model::compute_level_populations_shortchar(bool use_Ng_acceleration, long max_niterations)
{
    return level_population_loop(2, use_Ng_acceleration, max_niterations);
}

} // namespace magritte
