/*
 * magritte::Model / magritte::Solver (include/magritte_amd.hpp): the reference's
 * C++ entry points of the formal-solver path on top of the C ABI.
 *   Model::compute_inverse_line_widths                 src/model/model.cpp:60-66
 *   Model::compute_spectral_discretisation (all)       src/model/model.cpp:70-161, 257-298
 *   Model::compute_LTE_level_populations               src/model/model.cpp:302-308
 *   Model::compute_radiation_field_feautrier_order_2   src/model/model.cpp:353-390
 *   Model::compute_radiation_field_feautrier_order_2_sparse  src/model/model.cpp:474-515
 *   Model::compute_radiation_field_shortchar_order_0   src/model/model.cpp:313-351
 *   Model::compute_radiation_field_feautrier_order_2_uv src/model/model.cpp:395-433
 *   Model::compute_image, compute_image_optical_depth  src/model/model.cpp:957-976, 1061-1076
 *   Model::compute_image_new, compute_image_optical_depth_new (all overloads)
 *                                                      src/model/model.cpp:987-1035, 1081-1127
 *   Model::compute_level_populations*                  src/model/model.cpp:654-935
 * The helper arrays the reference derives when it reads a model are derived by
 * Model::derive (Points::read points.cpp:29-57, Rays::read rays.cpp:22-61,
 * Boundary::read boundary.cpp:9-44, Lines::read lines.cpp:22-68).
 */
#include <hip/hip_runtime_api.h>

#include <cfloat>
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

void check(int code)
{
    if (code != MGX_OK) throw std::runtime_error(mgx_last_error());
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void require(bool ok, const std::string &what)
{
    if (!ok) throw std::runtime_error(what);
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

/* the C ABI's view of the species' line and collision data (pointers into the model) */
struct NlteTables {
    std::vector<mgx_linedata> ld;
    std::vector<uint32_t> ncolpar;
    std::vector<mgx_colldata> colpar;
    size_t npop = 0; /* sum over species of P * nlev */
};

NlteTables nlte_tables(const Model &model)
{
    const auto &species = model.lines.lineProducingSpecies;
    require(!species.empty(), "the model has no line-producing species");
    NlteTables t;
    for (const LineProducingSpecies &s : species) {
        const Linedata &d = s.linedata;
        require(d.energy.size() == d.nlev && d.weight.size() == d.nlev && d.irad.size() == d.nrad &&
                    d.jrad.size() == d.nrad && d.A.size() == d.nrad && d.Bs.size() == d.nrad &&
                    d.Ba.size() == d.nrad && d.frequency.size() == d.nrad,
                "line data arrays do not match nlev / nrad");
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
            require(c.jcol.size() == ncol && c.Cd.size() == ntmp * ncol && c.Ce.size() == ntmp * ncol,
                    "collision data arrays do not match ntmp / ncol");
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
        t.npop += (size_t)model.parameters.npoints() * d.nlev;
    }
    return t;
}

/* chemistry.species.abundance (P rows of nspecs) as one row-major block */
std::vector<double> flat_abundance(const Model &model, Size &nspec)
{
    const auto &ab = model.chemistry.species.abundance;
    const size_t P = model.parameters.npoints();
    require(ab.size() == P, "chemistry.species.abundance must have a row per point");
    nspec = P ? (Size)ab[0].size() : 0;
    std::vector<double> flat;
    flat.reserve(P * nspec);
    for (const auto &row : ab) {
        require(row.size() == nspec, "chemistry.species.abundance rows differ in length");
        flat.insert(flat.end(), row.begin(), row.end());
    }
    return flat;
}

} // namespace

Model::~Model()
{
    for (auto &kv : dev)
        if (kv.second) (void)hipFree(kv.second);
}

void Model::derive()
{
    Parameters &p = parameters;
    Points &pts = geometry.points;
    Rays &rays = geometry.rays;
    Boundary &bdy = geometry.boundary;
    /* Points::read: npoints, the bounding-box centre (with the reference's
       initialisation of the maxima at the smallest positive double), cum_n_neighbors */
    const Size P = (Size)pts.position.size();
    require(P > 0, "geometry.points.position is empty");
    p.set_npoints(P);
    require(pts.velocity.size() == P, "geometry.points.velocity must have npoints entries");
    require(pts.n_neighbors.size() == P, "geometry.points.n_neighbors must have npoints entries");
    Vector3D lo{DBL_MAX, DBL_MAX, DBL_MAX}, hi{DBL_MIN, DBL_MIN, DBL_MIN};
    for (const Vector3D &x : pts.position.vec)
        for (int i = 0; i < 3; i++) {
            if (x[i] < lo[i]) lo[i] = x[i];
            if (x[i] > hi[i]) hi[i] = x[i];
        }
    for (int i = 0; i < 3; i++) pts.center[i] = (lo[i] + hi[i]) / 2;
    pts.cum_n_neighbors.resize(P);
    uint64_t tot = 0;
    for (Size q = 0; q < P; q++) {
        pts.cum_n_neighbors[q] = (Size)tot;
        tot += pts.n_neighbors[q];
    }
    require(tot == pts.neighbors.size(), "geometry.points.neighbors must hold sum(n_neighbors) entries");
    for (Size n : pts.neighbors.vec) require(n < P, "a neighbour index is out of range");
    /* Rays::read: nrays, hnrays, adaptive directions, antipodes (the last r2 with
       |d1 + d2|^2 < 1e-9, from the first point's directions) */
    const size_t nd = rays.direction.size();
    require(nd > 0, "geometry.rays.direction is empty");
    Size nrays = p.nrays_is_set() ? p.nrays() : (Size)nd;
    if (nd == (size_t)nrays) {
        rays.use_adaptive_directions = false;
    } else if (nd == (size_t)nrays * P) {
        rays.use_adaptive_directions = true;
    } else {
        p.set_nrays((Size)nd); /* throws with both sizes, as the reference does */
    }
    p.set_nrays(nrays);
    p.set_hnrays(nrays / 2);
    require(rays.weight.size() == nd, "geometry.rays.weight must match the directions");
    if (rays.antipod.size() != nrays) {
        rays.antipod.vec.assign(nrays, 0);
        for (Size r1 = 0; r1 < nrays; r1++)
            for (Size r2 = 0; r2 < nrays; r2++) {
                const Vector3D &a = rays.direction[r1], &b = rays.direction[r2];
                const Real s0 = a[0] + b[0], s1 = a[1] + b[1], s2 = a[2] + b[2];
                if (s0 * s0 + s1 * s1 + s2 * s2 < 1.0e-9) rays.antipod[r1] = r2;
            }
    }
    /* Boundary::read: nboundary, point2boundary, per-boundary-point conditions */
    const Size nb = (Size)bdy.boundary2point.size();
    require(nb > 0, "geometry.boundary.boundary2point is empty");
    p.set_nboundary(nb);
    bdy.point2boundary.vec.assign(P, P);
    for (Size b = 0; b < nb; b++) {
        require(bdy.boundary2point[b] < P, "a boundary point index is out of range");
        bdy.point2boundary[bdy.boundary2point[b]] = b;
    }
    if (bdy.boundary_condition.size() < nb) bdy.boundary_condition.resize(nb, Zero);
    if (bdy.boundary_temperature.size() < nb) bdy.boundary_temperature.vec.resize(nb, 0.0);
    /* Lines::read: nlspecs, nrad_cum, nlines, the line list and max_inverse_mass */
    auto &species = lines.lineProducingSpecies;
    require(!species.empty(), "lines.lineProducingSpecies is empty");
    p.set_nlspecs((Size)species.size());
    lines.nrad_cum.vec.assign(species.size(), 0);
    lines.line.vec.clear();
    max_inverse_mass = 0.0;
    for (size_t s = 0; s < species.size(); s++) {
        const Linedata &d = species[s].linedata;
        require(d.frequency.size() == d.nrad, "linedata.frequency must have nrad entries");
        if (s > 0) lines.nrad_cum[s] = lines.nrad_cum[s - 1] + species[s - 1].linedata.nrad;
        lines.line.vec.insert(lines.line.vec.end(), d.frequency.begin(), d.frequency.end());
        max_inverse_mass = std::max(max_inverse_mass, d.inverse_mass);
    }
    p.set_nlines((Size)lines.line.size());
    if (!chemistry.species.abundance.empty())
        p.set_nspecs((Size)chemistry.species.abundance[0].size());
    require(thermodynamics.temperature.gas.size() == P,
            "thermodynamics.temperature.gas must have npoints entries");
    require(thermodynamics.turbulence.vturb2.size() == P,
            "thermodynamics.turbulence.vturb2 must have npoints entries");
}

void Model::put(const std::string &name, const void *data, size_t bytes)
{
    auto &v = host[name];
    if (v.size() == bytes && (bytes == 0 || std::memcmp(v.data(), data, bytes) == 0)) return;
    v.assign(static_cast<const unsigned char *>(data), static_cast<const unsigned char *>(data) + bytes);
    stale[name] = true;
}

/* the arrays of the device view, from the model's members */
void Model::pack()
{
    derive();
    const Parameters &p = parameters;
    const size_t P = p.npoints(), L = p.nlines(), S = p.nlspecs();
    const Geometry &g = geometry;
    auto put_vec = [&](const char *name, const auto &v) {
        put(name, v.data(), v.size() * sizeof(v[0]));
    };
    put_vec("position", g.points.position.vec);
    put_vec("velocity", g.points.velocity.vec);
    put_vec("n_neighbors", g.points.n_neighbors.vec);
    put_vec("cum_n_neighbors", g.points.cum_n_neighbors.vec);
    put_vec("neighbors", g.points.neighbors.vec);
    put_vec("point2boundary", g.boundary.point2boundary.vec);
    put_vec("boundary2point", g.boundary.boundary2point.vec);
    std::vector<int32_t> bc(g.boundary.boundary_condition.begin(), g.boundary.boundary_condition.end());
    put_vec("boundary_condition", bc);
    put_vec("boundary_temperature", g.boundary.boundary_temperature.vec);
    put_vec("ray_direction", g.rays.direction.vec);
    put_vec("ray_weight", g.rays.weight.vec);
    put_vec("antipod", g.rays.antipod.vec);
    const auto &T = thermodynamics.temperature.gas.vec;
    const auto &vt2 = thermodynamics.turbulence.vturb2.vec;
    put_vec("temperature", T);
    put_vec("vturb2", vt2);
    /* the line list, its sorted copy and the per-line species tables */
    put_vec("line", lines.line.vec);
    std::vector<double> sl(L);
    std::vector<uint32_t> slm(L), lsp, lk;
    std::vector<double> lA, sim;
    check(mgx_sort_lines((uint32_t)L, lines.line.vec.data(), sl.data(), slm.data()));
    put_vec("sorted_line", sl);
    put_vec("sorted_line_map", slm);
    const auto &species = lines.lineProducingSpecies;
    std::vector<double> qw;
    const Size Q = p.nquads_is_set() ? p.nquads() : 0;
    for (size_t s = 0; s < S; s++) {
        const Linedata &d = species[s].linedata;
        for (Size k = 0; k < d.nrad; k++) {
            lsp.push_back((uint32_t)s);
            lk.push_back(k);
            lA.push_back(d.A.at(k));
        }
        sim.push_back(d.inverse_mass);
        const auto &w = species[s].quadrature.weights.vec;
        require(Q == 0 || w.size() == Q, "lineProducingSpecies[s].quadrature.weights must have nquads entries");
        qw.insert(qw.end(), w.begin(), w.end());
    }
    put_vec("line_species", lsp);
    put_vec("line_k", lk);
    put_vec("line_A", lA);
    put_vec("species_inverse_mass", sim);
    put_vec("quad_weights", qw);
    /* line data of the current populations */
    const NlteTables t = nlte_tables(*this);
    if (lines.emissivity.size() == P * L && lines.opacity.size() == P * L) {
        put_vec("emissivity", lines.emissivity.vec);
        put_vec("opacity", lines.opacity.vec);
        std::vector<double> ptot;
        for (const LineProducingSpecies &s : species) {
            require(s.population_tot.size() == P, "population_tot must have npoints entries");
            ptot.insert(ptot.end(), s.population_tot.begin(), s.population_tot.end());
        }
        std::vector<double> ic(P * L);
        check(mgx_compute_interp_crit((uint32_t)P, (uint32_t)S, t.ld.data(), lines.opacity.vec.data(),
                                      ptot.data(), ic.data()));
        put_vec("interp_crit", ic);
    }
    if (lines.inverse_width.size() == P * L) put_vec("inverse_width", lines.inverse_width.vec);
    /* Model::set_dshift_max, as Solver::setup calls it (solver.tpp:9, 26) */
    std::vector<double> dsm(P);
    check(mgx_compute_dshift_max((uint32_t)P, (uint32_t)S, t.ld.data(), T.data(), vt2.data(),
                                 p.max_width_fraction, dsm.data()));
    put_vec("dshift_max", dsm);
    /* the frequency grid */
    const Frequencies &fr = radiation.frequencies;
    put_vec("nu", fr.nu.vec);
    put_vec("f_line", fr.corresponding_line);
    put_vec("f_species", fr.corresponding_l_for_spec);
    put_vec("f_k", fr.corresponding_k_for_tran);
    put_vec("f_z", fr.corresponding_z_for_line);
    put_vec("f_in_line", fr.appears_in_line_integral);
}

const std::vector<unsigned char> &Model::get_array(const std::string &name)
{
    bool known = false;
    for (const Field &f : kFields) known = known || name == f.name;
    require(known, "unknown model array '" + name + "'");
    pack();
    auto it = host.find(name);
    require(it != host.end(), "model array '" + name + "' is not set");
    return it->second;
}

const mgx_model &Model::device_view()
{
    pack();
    const Parameters &p = parameters;
    require(p.nquads_is_set(), "parameters.nquads is not set");
    require(lines.inverse_width.size() == (size_t)p.npoints() * p.nlines(),
            "no inverse line widths: call compute_inverse_line_widths()");
    require(lines.emissivity.size() == (size_t)p.npoints() * p.nlines(),
            "no line emissivity and opacity: call compute_LTE_level_populations()");
    require(radiation.frequencies.nu.size() == (size_t)p.npoints() * p.nfreqs() && p.nfreqs() > 0,
            "no frequency grid: call compute_spectral_discretisation()");
    for (auto &kv : host) {
        if (!stale[kv.first]) continue;
        void *&d = dev[kv.first];
        size_t &cap = dev_bytes[kv.first];
        if (d && cap < kv.second.size()) {
            (void)hipFree(d);
            d = nullptr;
            cap = 0;
        }
        if (!d && !kv.second.empty()) {
            hip_check(hipMalloc(&d, kv.second.size()), "hipMalloc");
            cap = kv.second.size();
        }
        if (!kv.second.empty())
            hip_check(hipMemcpy(d, kv.second.data(), kv.second.size(), hipMemcpyHostToDevice),
                      "hipMemcpy");
        stale[kv.first] = false;
    }
    mgx_model &v = view;
    std::memset(&v, 0, sizeof v);
    v.npoints = p.npoints();
    v.nrays = p.nrays();
    v.hnrays = p.hnrays();
    v.nboundary = p.nboundary();
    v.nlines = p.nlines();
    v.nlspecs = p.nlspecs();
    v.nquads = p.nquads();
    v.nfreqs = p.nfreqs();
    v.dimension = p.dimension_is_set() ? (int32_t)p.dimension() : 3;
    v.spherical_symmetry = p.spherical_symmetry_is_set() && p.spherical_symmetry();
    v.adaptive_directions = geometry.rays.use_adaptive_directions;
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
    for (int i = 0; i < 3; i++) v.center[i] = geometry.points.center[i];
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

/* ------------------------------------------------------------------------- */
/* producers                                                                 */

int Model::compute_inverse_line_widths()
{
    derive();
    const NlteTables t = nlte_tables(*this);
    const Size P = parameters.npoints(), L = parameters.nlines();
    lines.inverse_width.resize(P, L);
    check(mgx_compute_inverse_widths(P, parameters.nlspecs(), t.ld.data(),
                                     thermodynamics.temperature.gas.vec.data(),
                                     thermodynamics.turbulence.vturb2.vec.data(),
                                     lines.inverse_width.vec.data()));
    return 0;
}

int Model::compute_spectral_discretisation()
{
    derive();
    const Parameters &p = parameters;
    require(p.nquads_is_set(), "parameters.nquads is not set");
    const Size P = p.npoints(), Q = p.nquads(), F = p.nlines() * Q;
    const auto &species = lines.lineProducingSpecies;
    const auto &roots = species[0].quadrature.roots.vec;
    require(roots.size() == Q, "lineProducingSpecies[0].quadrature.roots must have nquads entries");
    for (const auto &s : species)
        require(s.quadrature.roots.vec == roots,
                "every species must use the same quadrature roots");
    const NlteTables t = nlte_tables(*this);
    Frequencies &fr = radiation.frequencies;
    fr.nu.resize(P, F);
    fr.corresponding_line.assign(F, 0);
    fr.corresponding_l_for_spec.assign(F, 0);
    fr.corresponding_k_for_tran.assign(F, 0);
    fr.corresponding_z_for_line.assign(F, 0);
    fr.appears_in_line_integral.assign(F, 0);
    check(mgx_compute_spectral_discretisation(
        P, p.nlspecs(), t.ld.data(), Q, roots.data(), thermodynamics.temperature.gas.vec.data(),
        thermodynamics.turbulence.vturb2.vec.data(), fr.nu.vec.data(), fr.corresponding_line.data(),
        fr.corresponding_l_for_spec.data(), fr.corresponding_k_for_tran.data(),
        fr.corresponding_z_for_line.data(), fr.appears_in_line_integral.data()));
    parameters.set_nfreqs(F);
    return 0;
}

int Model::compute_spectral_discretisation(Real nu_min, Real nu_max)
{
    return compute_spectral_discretisation(nu_min, nu_max, parameters.nfreqs());
}

int Model::compute_spectral_discretisation(Real nu_min, Real nu_max, Size n)
{
    derive();
    const Size P = parameters.npoints();
    require(n > 0, "the image grid needs at least one frequency");
    Frequencies &fr = radiation.frequencies;
    fr.nu.resize(P, n);
    fr.corresponding_line.assign(n, 0);
    fr.corresponding_l_for_spec.assign(n, 0);
    fr.corresponding_k_for_tran.assign(n, 0);
    fr.corresponding_z_for_line.assign(n, 0);
    fr.appears_in_line_integral.assign(n, 0);
    check(mgx_compute_image_discretisation(P, nu_min, nu_max, n, fr.nu.vec.data(),
                                           fr.corresponding_line.data(),
                                           fr.corresponding_l_for_spec.data(),
                                           fr.corresponding_k_for_tran.data(),
                                           fr.corresponding_z_for_line.data(),
                                           fr.appears_in_line_integral.data()));
    parameters.set_nfreqs(n);
    return 0;
}

int Model::compute_radiation_field_feautrier_order_2()
{
    Solver solver;
    solver.setup(*this);
    solver.solve_feautrier_order_2(*this);
    return 0;
}

int Model::compute_radiation_field_feautrier_order_2_sparse()
{
    Solver solver;
    solver.setup(*this);
    solver.solve_feautrier_order_2_sparse(*this);
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
    const auto &d = model.geometry.rays.direction.vec;
    if (ray_nr >= d.size()) throw std::runtime_error("ray_nr out of range");
    return d[ray_nr];
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
    auto &species = model.lines.lineProducingSpecies;
    const auto &ls_b = model.get_array("line_species");
    const auto &lk_b = model.get_array("line_k");
    const uint32_t *lsp = reinterpret_cast<const uint32_t *>(ls_b.data());
    const uint32_t *lk = reinterpret_cast<const uint32_t *>(lk_b.data());
    const size_t P = p.npoints(), L = p.nlines();
    for (auto &s : species) {
        s.lambda.Ls.assign(P, std::vector<std::vector<Real>>(s.linedata.nrad));
        s.lambda.nr.assign(P, std::vector<std::vector<Size>>(s.linedata.nrad));
    }
    const LambdaLists &G = model.lambda;
    for (size_t q = 0; q < P; q++)
        for (size_t l = 0; l < L; l++) {
            LineProducingSpecies &s = species.at(lsp[l]);
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

void Solver::solve_feautrier_order_2_sparse(Model &model)
{
    const mgx_model &v = model.device_view();
    mgx_options o = model.options();
    /* the sparse solve keeps the Lambda diagonal only (solver.tpp:297-363 calls
       update_Lambda, whose bands the statistical equilibrium of the sparse loop
       does not use) */
    o.n_off_diag = 0;
    const size_t P = v.npoints, L = v.nlines;
    DevOut Jl(P * L), lam(P * L);
    check(mgx_feautrier_sparse_device(&v, &o, Jl.p, lam.p, nullptr));
    std::vector<double> jl;
    Jl.copy_to(jl, P * L);
    lam.copy_to(model.lambda_diag, P * L);
    model.lambda = LambdaLists{};
    auto &species = model.lines.lineProducingSpecies;
    for (size_t s = 0; s < species.size(); s++) {
        LineProducingSpecies &ls = species[s];
        const Size nrad = ls.linedata.nrad, l0 = model.lines.nrad_cum[s];
        ls.J.resize(P, nrad);
        for (size_t q = 0; q < P; q++)
            for (Size k = 0; k < nrad; k++) ls.J(q, k) = jl[q * L + l0 + k];
    }
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
    const auto &nu = model.radiation.frequencies.nu.vec;
    img.freqs.assign(nu.begin(), nu.begin() + F);
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
    for (const LineProducingSpecies &s : model.lines.lineProducingSpecies)
        pop.insert(pop.end(), s.population.begin(), s.population.end());
    return pop;
}

void split_populations(Model &model, const std::vector<double> &pop,
                       const std::vector<double> &ptot)
{
    const size_t P = model.parameters.npoints();
    size_t off = 0;
    auto &species = model.lines.lineProducingSpecies;
    for (size_t s = 0; s < species.size(); s++) {
        LineProducingSpecies &ls = species[s];
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
    derive();
    const NlteTables t = nlte_tables(*this);
    const Parameters &p = parameters;
    const size_t P = p.npoints(), L = p.nlines();
    Size nspec = 0;
    const std::vector<double> ab = flat_abundance(*this, nspec);
    std::vector<double> pop(t.npop), ptot((size_t)p.nlspecs() * P);
    lines.emissivity.resize(P, L);
    lines.opacity.resize(P, L);
    check(mgx_compute_lte(p.npoints(), nspec, ab.data(), thermodynamics.temperature.gas.vec.data(),
                          p.nlspecs(), t.ld.data(), p.min_line_emissivity, p.min_line_opacity,
                          pop.data(), ptot.data(), lines.emissivity.vec.data(),
                          lines.opacity.vec.data()));
    split_populations(*this, pop, ptot);
    return 0;
}

/* kind: 0 the general loop, 1 the sparse loop, 2 the short-characteristics loop */
int Model::level_population_loop(int kind, bool use_Ng_acceleration, long max_niterations)
{
    derive();
    const Parameters &p = parameters;
    NlteTables t = nlte_tables(*this);
    bool have = true;
    for (const LineProducingSpecies &s : lines.lineProducingSpecies)
        have = have && s.population.size() == (size_t)p.npoints() * s.linedata.nlev;
    if (!have) compute_LTE_level_populations();
    Size nspec = 0;
    const std::vector<double> abundance = flat_abundance(*this, nspec);
    const mgx_model &v = device_view();
    const mgx_options o = options();
    const mgx_nlte_params q = nlte_params(p);
    const size_t P = p.npoints(), PL = (size_t)p.npoints() * p.nlines();
    DevOut ab(abundance.size()), pd(t.npop), pt((size_t)p.nlspecs() * P);
    hip_check(hipMemcpy(ab.p, abundance.data(), abundance.size() * sizeof(double),
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    const std::vector<double> pop0 = joined_populations(*this, t.npop);
    hip_check(hipMemcpy(pd.p, pop0.data(), t.npop * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy");
    NlteHandle h;
    check(mgx_nlte_create(&v, p.nlspecs(), t.ld.data(), t.ncolpar.data(),
                          t.colpar.empty() ? nullptr : t.colpar.data(), nspec, ab.p, &q, &h.h,
                          nullptr));
    /* the line data the loop rewrites are the device mirrors the view points to */
    double *em = const_cast<double *>(v.emissivity);
    double *op = const_cast<double *>(v.opacity);
    double *ic = const_cast<double *>(v.interp_crit);
    check(mgx_nlte_set_populations(h.h, pd.p, em, op, ic, nullptr));
    const size_t cap = (size_t)(max_niterations > 0 ? max_niterations : 1) * p.nlspecs();
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
    pt.copy_to(ptot, (size_t)p.nlspecs() * P);
    split_populations(*this, pop, ptot);
    /* lines.emissivity / opacity take the loop's line data (the interpolation
       criterion follows from them and the populations on the next solve) */
    lines.emissivity.resize(P, p.nlines());
    lines.opacity.resize(P, p.nlines());
    hip_check(hipMemcpy(lines.emissivity.vec.data(), em, PL * sizeof(double), hipMemcpyDeviceToHost),
              "copy");
    hip_check(hipMemcpy(lines.opacity.vec.data(), op, PL * sizeof(double), hipMemcpyDeviceToHost),
              "copy");
    (void)line;
    const size_t n = (size_t)it * p.nlspecs();
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
