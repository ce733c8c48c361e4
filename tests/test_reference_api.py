"""Models built through the reference's object model and setup calls, in memory.

The reference's benchmark scripts build their models with parameters.set_*,
.set() on the geometry and thermodynamics vectors, the abundance and symbol
tables and the setup helpers (tests/benchmarks/analytic/all_constant_single_ray.py:37-70,
tests/benchmarks/numeric/vanZadelhoff_1_1D_sparse.py:45-91), then call
compute_spectral_discretisation, compute_inverse_line_widths and
compute_LTE_level_populations before a solve (:86-100).  The same sequence on
magritte_amd.core.Model with magritte_amd.setup (no write and read back):

* on the CPU it packs exactly the arrays of the model generator's equivalent
  model (so the device sees the same inputs), and rejects what the reference
  rejects;
* on the GPU the all-constant single-ray benchmark's u from the Feautrier and
  short-characteristics solves stays within the benchmark's own bounds of the
  analytic solution (1.7e-4 and 1.7e-7) and within 1e-10 of the oracle; the
  sparse solve's lspec.J is the quadrature sum of the full solve's J; and the
  van Zadelhoff 1a sparse loop meets the reference-held RATRAN populations.
"""
import os

import numpy as np
import pytest

from magritte_amd import models

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LAMDA = os.path.join(ROOT, "tests", "golden", "lamda_test.txt")


def _all_constant():
    """tests/benchmarks/analytic/all_constant_single_ray.py's model, as that script
    builds it (dimension 1, 50 points 1e12 m apart, 2 rays, 3 species, one line
    species with one quadrature point, n_H2 = 1e12, n_test = 1e3 m^-3, 4.5 K, no
    turbulence or velocity)."""
    from magritte_amd import core
    from magritte_amd import setup

    npoints, dx = 50, 1.0e12
    model = core.Model()
    model.parameters.set_spherical_symmetry(False)
    model.parameters.set_model_name("all_constant_single_ray")
    model.parameters.set_dimension(1)
    model.parameters.set_npoints(npoints)
    model.parameters.set_nrays(2)
    model.parameters.set_nspecs(3)
    model.parameters.set_nlspecs(1)
    model.parameters.set_nquads(1)
    model.geometry.points.position.set([[i * dx, 0, 0] for i in range(npoints)])
    model.geometry.points.velocity.set([[0.0, 0, 0] for i in range(npoints)])
    model.chemistry.species.abundance = [[1.0e3, 1.0e12, 0.0] for _ in range(npoints)]
    model.chemistry.species.symbol = ["test", "H2", "e-"]
    model.thermodynamics.temperature.gas.set(4.5 * np.ones(npoints))
    model.thermodynamics.turbulence.vturb2.set(0.0 * np.ones(npoints))
    setup.set_Delaunay_neighbor_lists(model)
    setup.set_Delaunay_boundary(model)
    setup.set_boundary_condition_CMB(model)
    setup.set_uniform_rays(model)
    setup.set_linedata_from_LAMDA_file(model, LAMDA)
    setup.set_quadrature(model)
    model.compute_spectral_discretisation()
    model.compute_inverse_line_widths()
    model.compute_LTE_level_populations()
    return model


def _van_zadelhoff_1a():
    """tests/benchmarks/numeric/vanZadelhoff_1_1D_sparse.py's model 'a', as that
    script builds it."""
    from magritte_amd import core
    from magritte_amd import setup

    npoints, r_in, r_out, nH2_in = 250, 1.0e13, 7.8e16, 2.0e13
    rs = np.logspace(np.log10(r_in), np.log10(r_out), npoints, endpoint=True)
    nH2 = nH2_in * np.power(r_in / rs, 2.0)
    model = core.Model()
    model.parameters.set_spherical_symmetry(True)
    model.parameters.set_model_name("vanZadelhoff_1a_1D_sparse")
    model.parameters.set_dimension(1)
    model.parameters.set_npoints(npoints)
    model.parameters.set_nrays(200)
    model.parameters.set_nspecs(3)
    model.parameters.set_nlspecs(1)
    model.parameters.set_nquads(11)
    model.geometry.points.position.set([[r, 0, 0] for r in rs])
    model.geometry.points.velocity.set(np.zeros((npoints, 3)))
    model.chemistry.species.abundance = [[1.0e-8 * n, n, 0.0] for n in nH2]
    model.chemistry.species.symbol = ["test", "H2", "e-"]
    model.thermodynamics.temperature.gas.set(20.0 * np.ones(npoints))
    model.thermodynamics.turbulence.vturb2.set((150.0 / core.CC) ** 2 * np.ones(npoints))
    setup.set_Delaunay_neighbor_lists(model)
    setup.set_Delaunay_boundary(model)
    setup.set_boundary_condition_CMB(model)
    setup.set_rays_spherical_symmetry(model)
    setup.set_linedata_from_LAMDA_file(model, LAMDA)
    setup.set_quadrature(model)
    model.compute_spectral_discretisation()
    model.compute_inverse_line_widths()
    model.compute_LTE_level_populations()
    return model


def _same_arrays(cm, m):
    for k in m.ARRAYS:
        ref = np.ascontiguousarray(getattr(m, k))
        got = np.frombuffer(bytes(cm.get_array(k)), dtype=ref.dtype)
        np.testing.assert_array_equal(got, ref.ravel(), err_msg=k)
    assert tuple(cm.geometry.points.center) == tuple(m.center)


def test_all_constant_built_like_the_reference_packs_the_generator_model():
    cm = _all_constant()
    m = models.model_constant_slab(linedata=models.read_lamda(LAMDA, num=0))
    _same_arrays(cm, m)
    ld = cm.lines.lineProducingSpecies[0].linedata
    assert ld.sym == "test" and ld.num == 0 and ld.ncolpar == 1
    assert ld.colpar[0].num_col_partner == 1  # H2 in the symbol table
    p = cm.parameters
    assert (p.npoints(), p.nrays(), p.hnrays(), p.nboundary(), p.nlines(), p.nfreqs()) == \
        (50, 2, 1, 2, 1, 1)
    assert list(np.array(cm.geometry.boundary.point2boundary)[:2]) == [0, 50]


def test_van_zadelhoff_built_like_the_reference_packs_the_generator_model():
    cm = _van_zadelhoff_1a()
    _same_arrays(cm, models.model_vanzadelhoff_1d(LAMDA, "a"))


def test_reference_errors():
    from magritte_amd import core
    from magritte_amd import setup

    cm = _all_constant()
    with pytest.raises(RuntimeError, match="SetOnce"):
        cm.parameters.set_npoints(51)
    m2 = core.Model()
    m2.parameters.set_dimension(1)
    m2.parameters.set_nrays(4)
    m2.parameters.set_spherical_symmetry(False)
    with pytest.raises(ValueError, match="nrays should be 2"):
        setup.set_uniform_rays(m2)
    m2.parameters.set_nlspecs(1)
    m2.chemistry.species.symbol = ["H2", "e-"]
    with pytest.raises(KeyError, match="Could not find species 'test'"):
        setup.set_linedata_from_LAMDA_file(m2, LAMDA)
    m3 = core.Model()
    m3.parameters.set_dimension(2)
    with pytest.raises(ValueError, match="Dimension = 2"):
        setup.set_Delaunay_neighbor_lists(m3)
    # a solve needs the producers first
    m4 = _all_constant()
    m4.lines.inverse_width.set(np.zeros((0, 0)))
    with pytest.raises(RuntimeError, match="compute_inverse_line_widths"):
        m4.compute_radiation_field_feautrier_order_2()


def _analytic_u(cm):
    """u of the all-constant slab along the single ray pair: the constant source
    S = eta/chi and the CMB at both ends, I_+(x) = S + (B - S) exp(-chi x),
    I_-(x) = S + (B - S) exp(-chi (x_end - x)), u = (I_+ + I_-)/2 (the
    benchmark's closed form, all_constant_single_ray.py:112-130)."""
    from magritte_amd import core

    ld = cm.lines.lineProducingSpecies[0].linedata
    nu = ld.frequency[0]
    T, n = 4.5, 1.0e3
    g, E = np.array(ld.weight), np.array(ld.energy)
    w = g * np.exp(-E / (core.KB * T))
    pop = n * w / w.sum()
    width = nu * np.sqrt(2.0 * core.KB * T / (core.AMU * core.CC**2) * ld.inverse_mass)
    phi = 1.0 / (np.sqrt(np.pi) * width)
    c = core.HH * nu / (4.0 * np.pi)
    eta = c * ld.A[0] * pop[1] * phi
    chi = c * (pop[0] * ld.Ba[0] - pop[1] * ld.Bs[0]) * phi
    S = eta / chi
    B = 2.0 * core.HH * nu**3 / core.CC**2 / np.expm1(core.HH * nu / (core.KB * core.T_CMB))
    x = np.array(cm.geometry.points.position)[:, 0]
    Ip = S + (B - S) * np.exp(-chi * x)
    Im = S + (B - S) * np.exp(-chi * (x[-1] - x))
    return 0.5 * (Ip + Im)


def _rel(a, b):
    return np.abs(2.0 * (a - b) / (a + b))


@pytest.mark.gpu
def test_all_constant_benchmark_on_the_gpu():
    import torch  # noqa: F401  (one HIP runtime: torch's)

    from oracle import oracle

    cm = _all_constant()
    md = models.ModelData.from_core(cm)
    op = md.options()
    u_ex = _analytic_u(cm)
    print("[reference API] all-constant single ray", flush=True)
    cm.compute_radiation_field_shortchar_order_0()
    u0 = np.array(cm.radiation.u)[0, :, 0]
    cm.compute_radiation_field_feautrier_order_2()
    u2 = np.array(cm.radiation.u)[0, :, 0].copy()
    J2 = np.array(cm.radiation.J).copy()
    e0, e2 = float(_rel(u0, u_ex).max()), float(_rel(u2, u_ex).max())
    print(f"  vs the closed form: shortchar {e0:.2e}, Feautrier {e2:.2e}", flush=True)
    assert e0 < 1.7e-7 and e2 < 1.7e-4
    ur, Jr, _ = oracle.solve(md, op)
    assert float(np.max(np.abs(u2 - ur[0, :, 0]) / np.abs(ur[0, :, 0]))) < 1e-10
    cm.compute_radiation_field_feautrier_order_2_sparse()
    Jl = np.array(cm.lines.lineProducingSpecies[0].J)
    w = np.array(cm.lines.lineProducingSpecies[0].quadrature.weights)
    ref = (J2.reshape(md.npoints, md.nlines, md.nquads) * w[None, None, :]).sum(axis=2)
    assert Jl.shape == (md.npoints, 1)
    np.testing.assert_allclose(Jl, ref, rtol=1e-14, atol=0)


@pytest.mark.gpu
def test_sparse_solve_on_a_wind_built_through_the_object_model(small_wind):
    """lspec.J of the sparse solve equals the quadrature sum of the oracle's J on a
    3D wind (nine lines, 31 bins each), and its Lambda the full solve's diagonal."""
    import torch  # noqa: F401

    from oracle import oracle

    m = small_wind
    cm = m.to_core()
    print("[reference API] sparse solve, 12^3 wind", flush=True)
    cm.compute_radiation_field_feautrier_order_2_sparse()
    _, Jr, lr = oracle.solve(m, m.options(), with_u=False)
    w = m.quad_weights[0]
    ref = (Jr.reshape(m.npoints, m.nlines, m.nquads) * w[None, None, :]).sum(axis=2)
    Jl = np.array(cm.lines.lineProducingSpecies[0].J)
    err = float(np.max(np.abs(Jl - ref) / np.abs(ref)))
    print(f"  lspec.J vs the oracle {err:.2e}", flush=True)
    assert err < 1e-10
    np.testing.assert_allclose(np.array(cm.lambda_diag), lr, rtol=1e-10, atol=0)
    lam = cm.lines.lineProducingSpecies[0].Lambda
    assert list(lam.nr[3][2]) == [3]


@pytest.mark.gpu
def test_van_zadelhoff_1a_sparse_loop_through_the_object_model():
    import torch  # noqa: F401
    from test_oracle_nlte import _ratran_errors

    cm = _van_zadelhoff_1a()
    cm.parameters.min_line_opacity = 1.0e-13
    print("[reference API] van Zadelhoff 1a, compute_level_populations_sparse(True, 100)",
          flush=True)
    it = cm.compute_level_populations_sparse(True, 100)
    md = models.ModelData.from_core(cm)
    md.abundance = np.array(cm.chemistry.species.abundance)
    pops = np.array(cm.lines.lineProducingSpecies[0].population).reshape(md.npoints, 2)
    e0, e1 = _ratran_errors(md, pops)
    print(f"  {it} iterations, max rel err vs RATRAN {e0:.4f} / {e1:.4f}", flush=True)
    assert e0 < 0.12 and e1 < 0.13


def _cpp_example(tmp_path):
    """tests/cpp/all_constant_model.cpp compiled against include/ and libmgx.so."""
    import subprocess

    exe = str(tmp_path / "all_constant_model")
    lib = os.path.join(ROOT, "magritte_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "all_constant_model.cpp"), "-o", exe,
                    "-L" + lib, "-l:libmgx.so", "-Wl,-rpath," + lib], check=True)
    return exe


def _run_cpp(exe, *args):
    import subprocess

    r = subprocess.run([exe, *args], stdout=subprocess.PIPE, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    return dict(l.split(" ", 1) for l in r.stdout.strip().splitlines())


def test_cpp_object_model_builds_the_model(tmp_path):
    """A C++ program builds the all-constant model through magritte::Model's nested
    members and runs the producers: the derived counts, antipodes, line grid and LTE
    opacity equal the Python-built model's."""
    out = _run_cpp(_cpp_example(tmp_path))
    cm = _all_constant()
    assert out["nlines"] == "1" and out["nboundary"] == "2" and out["hnrays"] == "1"
    assert out["antipod"] == "1 0"
    assert float(out["nu0"]) == np.array(cm.radiation.frequencies.nu)[0, 0]
    assert float(out["opacity0"]) == np.array(cm.lines.opacity)[0, 0]


@pytest.mark.gpu
def test_cpp_object_model_solves(tmp_path):
    """The same C++ program solves on the GPU: u, J and the sparse lspec.J equal the
    Python-built model's solve bit for bit (the same library, the same inputs)."""
    import torch  # noqa: F401

    exe = _cpp_example(tmp_path)
    out = _run_cpp(exe, "solve")
    cm = _all_constant()
    cm.compute_radiation_field_feautrier_order_2()
    u = np.array(cm.radiation.u)
    J = np.array(cm.radiation.J)
    assert float(out["u0"]) == u[0, 0, 0] and float(out["uend"]) == u[0, -1, 0]
    assert float(out["Jmid"]) == J[25, 0]
    cm.compute_radiation_field_feautrier_order_2_sparse()
    assert float(out["Jline_mid"]) == np.array(cm.lines.lineProducingSpecies[0].J)[25, 0]
