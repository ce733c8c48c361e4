/*
 * A C++ caller of the formal-solver path through the reference's object model
 * (include/magritte_amd.hpp): the all-constant single-ray model of the
 * reference's benchmark (tests/benchmarks/analytic/all_constant_single_ray.py:
 * 37-70, 86-100) built in memory -- parameters, points, rays, boundary,
 * chemistry, thermodynamics, one two-level line species with one quadrature
 * point -- then the producers, and with the argument "solve" the Feautrier
 * solve and the sparse solve on the GPU.  Prints name value lines.
 */
#include <cmath>
#include <cstdio>
#include <cstring>

#include "magritte_amd.hpp"

using namespace magritte;

int main(int argc, char **argv)
{
    const bool solve = argc > 1 && std::strcmp(argv[1], "solve") == 0;
    const Size npoints = 50;
    const Real dx = 1.0e12, HH = 6.62607004e-34, CC = 2.99792458e8, KB = 1.38064852e-23;
    Model model;
    Parameters &p = model.parameters;
    p.set_spherical_symmetry(false);
    p.set_model_name("all_constant_single_ray");
    p.set_dimension(1);
    p.set_npoints(npoints);
    p.set_nrays(2);
    p.set_nspecs(3);
    p.set_nlspecs(1);
    p.set_nquads(1);
    for (Size i = 0; i < npoints; i++) {
        model.geometry.points.position.vec.push_back({i * dx, 0.0, 0.0});
        model.geometry.points.velocity.vec.push_back({0.0, 0.0, 0.0});
        model.chemistry.species.abundance.push_back({1.0e3, 1.0e12, 0.0});
        model.thermodynamics.temperature.gas.vec.push_back(4.5);
        model.thermodynamics.turbulence.vturb2.vec.push_back(0.0);
    }
    model.chemistry.species.symbol = {"test", "H2", "e-"};
    /* neighbours along the line, the two end points as boundary (CMB), rays -x and +x */
    for (Size i = 0; i < npoints; i++) {
        const Size n = (i == 0 || i == npoints - 1) ? 1 : 2;
        model.geometry.points.n_neighbors.vec.push_back(n);
        if (i > 0) model.geometry.points.neighbors.vec.push_back(i - 1);
        if (i + 1 < npoints) model.geometry.points.neighbors.vec.push_back(i + 1);
    }
    model.geometry.boundary.boundary2point.set({0, npoints - 1});
    model.geometry.boundary.boundary_temperature.set({2.72548, 2.72548});
    for (Size b = 0; b < 2; b++) model.geometry.boundary.set_boundary_condition(b, CMB);
    model.geometry.rays.direction.set({{-1.0, 0.0, 0.0}, {1.0, 0.0, 0.0}});
    model.geometry.rays.weight.set({0.5, 0.5});
    /* the two-level test species: levels at 0 and 6 cm^-1 (g 1, 3), A = 1e-4 s^-1, mass 1 */
    model.lines.resize_LineProducingSpecies(1);
    LineProducingSpecies &ls = model.lines.lineProducingSpecies[0];
    Linedata &ld = ls.linedata;
    ld.sym = "test";
    ld.num = 0;
    ld.nlev = 2;
    ld.nrad = 1;
    ld.inverse_mass = 1.0;
    ld.energy = {0.0, 6.0 * 1.0e2 * HH * CC};
    ld.weight = {1.0, 3.0};
    ld.irad = {1};
    ld.jrad = {0};
    ld.A = {1.0e-4};
    const Real nu = (ld.energy[1] - ld.energy[0]) / HH;
    ld.frequency = {nu};
    ld.Bs = {ld.A[0] * CC * CC / (2.0 * HH * nu * nu * nu)};
    ld.Ba = {ld.weight[1] / ld.weight[0] * ld.Bs[0]};
    CollisionPartner c;
    c.num_col_partner = 1;
    c.icol = {1};
    c.jcol = {0};
    c.tmp = {20.0};
    c.Cd = {2.0e-16};
    c.Ce = {2.0e-16 * 3.0 * std::exp(-(ld.energy[1] - ld.energy[0]) / (KB * 20.0))};
    ld.colpar.push_back(c);
    ls.quadrature.roots.set({0.0});
    ls.quadrature.weights.set({1.0});
    try {
        model.compute_spectral_discretisation();
        model.compute_inverse_line_widths();
        model.compute_LTE_level_populations();
        model.derive();
        std::printf("nlines %u\nnboundary %u\nhnrays %u\nnfreqs %u\n", p.nlines(), p.nboundary(),
                    p.hnrays(), p.nfreqs());
        std::printf("antipod %u %u\n", model.geometry.rays.antipod[0], model.geometry.rays.antipod[1]);
        std::printf("nu0 %.17g\nopacity0 %.17g\n", model.radiation.frequencies.nu(0, 0),
                    model.lines.opacity(0, 0));
        if (solve) {
            model.compute_radiation_field_feautrier_order_2();
            std::printf("u0 %.17g\nuend %.17g\nJmid %.17g\n", model.radiation.u[0],
                        model.radiation.u[npoints - 1], model.radiation.J[npoints / 2]);
            model.compute_radiation_field_feautrier_order_2_sparse();
            std::printf("Jline_mid %.17g\n", ls.J(npoints / 2, 0));
        }
    } catch (const std::exception &e) {
        std::printf("error %s\n", e.what());
        return 1;
    }
    return 0;
}
