"""The pybind11 module magritte_amd.core (C++ magritte::Model with the
reference's object model and method names, src/bindings/pybindings.cpp:123-800)
loads, is built through the nested objects, derives the arrays the reference's
readers derive, and reports errors as RuntimeError with the library's message
(no GPU is touched on the CPU)."""
import numpy as np
import pytest


def test_core_model_takes_arrays(line_model):
    from magritte_amd import core

    cm = line_model.to_core()
    assert isinstance(cm, core.Model)
    p = cm.parameters
    assert p.npoints() == line_model.npoints and p.nfreqs() == line_model.nfreqs
    assert not p.one_line_approximation and p.n_off_diag == 0
    with pytest.raises(RuntimeError, match="unknown model array"):
        cm.get_array("no_such_array")
    p.one_line_approximation = True
    assert cm.parameters.one_line_approximation
    # set-once parameters (SetOnce, tools/setOnce.hpp)
    p.set_npoints(line_model.npoints)
    with pytest.raises(RuntimeError, match="SetOnce"):
        p.set_npoints(line_model.npoints + 1)


@pytest.mark.parametrize("which", ["line_model", "small_wind", "vgrad_model"])
def test_core_model_packs_the_generator_arrays(which, request):
    """The nested objects, after the derivations of the reference's readers
    (cum_n_neighbors, centre, antipodes, point2boundary, line list and sort) and the
    producers (line grid, inverse widths, dshift_max, interpolation criterion), pack
    to the very arrays the model generator hands the device solver."""
    m = request.getfixturevalue(which)
    cm = m.to_core()
    for k in m.ARRAYS:
        ref = np.ascontiguousarray(getattr(m, k))
        got = np.frombuffer(bytes(cm.get_array(k)), dtype=ref.dtype)
        np.testing.assert_array_equal(got, ref.ravel(), err_msg=k)
    assert tuple(cm.geometry.points.center) == tuple(m.center)
    p = cm.parameters
    assert (p.nlines(), p.nboundary(), p.hnrays(), p.nlspecs()) == (m.nlines, m.nboundary,
                                                                   m.hnrays, m.nlspecs)
    np.testing.assert_array_equal(np.array(cm.radiation.frequencies.nu), m.nu)
    np.testing.assert_array_equal(np.array(cm.lines.inverse_width), m.inverse_width)


def test_core_image_direction_checks(line_model):
    cm = line_model.to_core()
    with pytest.raises(RuntimeError, match="ray_nr out of range"):
        cm.compute_image_new(1000, 4, 4)
    with pytest.raises(RuntimeError, match="does not point in a direction"):
        cm.compute_image_new(0.0, 0.0, 0.0, 4, 4)


@pytest.mark.gpu
def test_core_model_matches_device_solver(small_wind):
    import torch  # noqa: F401  (one HIP runtime: torch's)

    from magritte_amd import solver

    cm = small_wind.to_core()
    cm.compute_radiation_field_feautrier_order_2()
    out = solver.feautrier_forward(solver.DeviceModel(small_wind), small_wind.options())
    np.testing.assert_array_equal(np.array(cm.radiation.u), out["u"].cpu().numpy())
    np.testing.assert_array_equal(np.array(cm.radiation.J), out["J"].cpu().numpy())
    np.testing.assert_array_equal(np.array(cm.lambda_diag), out["lambda_diag"].cpu().numpy())
    # new imager through compute_spectral_discretisation(nu_min, nu_max, n) + compute_image_new
    nu0 = small_wind.line[0]
    cm.compute_spectral_discretisation(nu0 * (1 - 2e-5), nu0 * (1 + 2e-5), 31)
    cm.compute_image_new(0.3, -0.5, 0.81, 12, 10)
    img = cm.images[-1]
    fr = small_wind.image_frequencies(nu0 * (1 - 2e-5), nu0 * (1 + 2e-5), 31)
    I, X, Y = solver.image_new(solver.DeviceModel(small_wind), small_wind.options(),
                               np.array([0.3, -0.5, 0.81]), 12, 10, fr)
    np.testing.assert_array_equal(np.array(img.I), I.cpu().numpy())
    np.testing.assert_array_equal(np.array(img.ImX), X.cpu().numpy())
    cm.compute_spectral_discretisation()
    assert cm.parameters.nfreqs() == small_wind.nfreqs


@pytest.mark.gpu
def test_core_sibling_solvers_match_device_calls(small_wind):
    """The C++ Model's short characteristics, u/v, old imager, optical-depth images and
    general Lambda lists (n_off_diag > 0) give the device entry points' results bit for bit."""
    import torch  # noqa: F401  (one HIP runtime: torch's)

    from magritte_amd import solver

    m = small_wind
    dm = solver.DeviceModel(m)
    op = m.options()
    cm = m.to_core()
    cm.compute_radiation_field_shortchar_order_0()
    sc = solver.shortchar_forward(dm, op)
    for k in ("I", "u", "v", "J"):
        np.testing.assert_array_equal(np.array(getattr(cm.radiation, k)), sc[k].cpu().numpy())
    np.testing.assert_array_equal(np.array(cm.lambda_diag), sc["lambda_diag"].cpu().numpy())
    cm.compute_radiation_field_feautrier_order_2_uv()
    uv = solver.feautrier_uv(dm, op)
    np.testing.assert_array_equal(np.array(cm.radiation.v), uv["v"].cpu().numpy())
    assert np.array(cm.radiation.J).size == 0
    for od, name in ((False, "compute_image"), (True, "compute_image_optical_depth")):
        getattr(cm, name)(8)
        img = cm.images[-1]
        ref, X, _ = solver.image_ray(dm, op, 8, optical_depth=od)
        assert img.imageType == ("OpticalDepth" if od else "Intensity") and img.ray_nr == 8
        np.testing.assert_array_equal(np.array(img.I), ref.cpu().numpy())
        np.testing.assert_array_equal(np.array(img.ImX), X.cpu().numpy())
    nu0 = m.line[0]
    cm.compute_spectral_discretisation(nu0 * (1 - 2e-5), nu0 * (1 + 2e-5), 31)
    cm.compute_image_optical_depth_new(0.3, -0.5, 0.81, 12, 10)
    fr = m.image_frequencies(nu0 * (1 - 2e-5), nu0 * (1 + 2e-5), 31)
    tau, _, _ = solver.image_optical_depth_new(dm, op, np.array([0.3, -0.5, 0.81]), 12, 10, fr)
    np.testing.assert_array_equal(np.array(cm.images[-1].I), tau.cpu().numpy())
    cm.compute_spectral_discretisation()
    cm.parameters.n_off_diag = 2
    cm.compute_radiation_field_feautrier_order_2()
    op2 = m.options()
    op2.n_off_diag = 2
    lam = solver.feautrier_forward_lambda(dm, op2)
    cnt = np.array(cm.lambda_count)
    np.testing.assert_array_equal(cnt, lam["count"].cpu().numpy().view(np.uint32))
    used = np.arange(np.array(cm.lambda_val).shape[2])[None, None, :] < cnt[:, :, None]
    np.testing.assert_array_equal(np.array(cm.lambda_nr)[used],
                                  lam["nr"].cpu().numpy().view(np.uint32)[used])
    np.testing.assert_array_equal(np.array(cm.lambda_val)[used], lam["val"].cpu().numpy()[used])


def test_core_lte_level_populations(small_wind):
    """Model.compute_LTE_level_populations on the host producers: the species'
    populations and the line data equal the model generator's LTE arrays."""
    m = small_wind
    cm = m.to_core()
    species = cm.lines.lineProducingSpecies
    assert len(species) == m.nlspecs
    assert np.array(cm.chemistry.species.abundance).shape == m.abundance.shape
    species[0].population = np.zeros(0)  # forget the copied populations
    cm.compute_LTE_level_populations()
    n = m.npoints * m.linedata[0]["nlev"]
    np.testing.assert_array_equal(np.array(cm.lines.lineProducingSpecies[0].population), m.population[:n])
    np.testing.assert_array_equal(np.array(cm.lines.lineProducingSpecies[0].population_tot), m.population_tot[0])
    for k in ("emissivity", "opacity", "interp_crit"):
        got = np.frombuffer(bytes(cm.get_array(k)), dtype=np.float64).reshape(m.npoints, m.nlines)
        np.testing.assert_array_equal(got, getattr(m, k))
    cm.lines.lineProducingSpecies[0].linedata.A = [1.0]  # the species list aliases the model's
    with pytest.raises(RuntimeError, match="nlev / nrad"):
        cm.compute_LTE_level_populations()


@pytest.mark.gpu
def test_core_level_populations_match_device_loop():
    """Model.compute_level_populations_sparse / compute_level_populations (n_off_diag 2)
    on the C++ Model give the Python device loop's iterations, errors and populations,
    and lspec.Lambda holds the model-wide Lambda lists in the reference's nesting."""
    import torch  # noqa: F401  (one HIP runtime: torch's)

    from magritte_amd import _abi, models, solver

    for noff in (0, 2):
        m = models.model_line_1d()
        cm = m.to_core()
        cm.parameters.n_off_diag = noff
        prm = _abi.default_nlte_params()
        if noff == 0:
            it_c = cm.compute_level_populations_sparse(True, 60)
        else:
            it_c = cm.compute_level_populations(True, 60)
        dm = solver.DeviceModel(m)
        lp = solver.LevelPopulations(dm, prm)
        lp.set_lte()
        op = m.options()
        op.n_off_diag = noff
        loop = solver.compute_level_populations if noff else solver.compute_level_populations_sparse
        it, em, ex = loop(dm, op, lp, True, 60)
        print(f"[core level populations] n_off_diag={noff} iterations {it_c} / {it}", flush=True)
        assert it_c == it
        np.testing.assert_allclose(np.array(cm.error_max), ex, rtol=1e-12, atol=0)
        pop, tot = lp.populations()
        np.testing.assert_allclose(np.array(cm.lines.lineProducingSpecies[0].population), pop.cpu().numpy(),
                                   rtol=1e-12, atol=1e-14 * float(tot.max()))
        got = np.frombuffer(bytes(cm.get_array("opacity")), dtype=np.float64)
        np.testing.assert_allclose(got, dm.tensors["opacity"].cpu().numpy().ravel(), rtol=1e-12)
        # after the loop, lspec.Lambda holds its last formal solve's Lambda: one diagonal
        # element per (point, line) for the sparse loop, the bands for the general one
        lam = cm.lines.lineProducingSpecies[0].Lambda
        lens = [len(lam.Ls[p][0]) for p in range(m.npoints)]
        assert all(len(lam.nr[p][0]) == lens[p] for p in range(m.npoints))
        if noff == 0:
            assert lens == [1] * m.npoints
            assert [list(lam.nr[p][0]) for p in range(m.npoints)] == [[p] for p in range(m.npoints)]
            np.testing.assert_array_equal([lam.Ls[p][0][0] for p in range(m.npoints)],
                                          np.array(cm.lambda_diag)[:, 0])
        else:
            assert max(lens) == 1 + 2 * noff
            np.testing.assert_array_equal([lam.Ls[p][0][0] for p in range(m.npoints)],
                                          np.array(cm.lambda_val)[:, 0, 0])
    # Lambda of the last formal solve in lspec.Lambda.Ls / nr
    cm.compute_radiation_field_feautrier_order_2()
    lam = cm.lines.lineProducingSpecies[0].Lambda
    cnt, nr, val = np.array(cm.lambda_count), np.array(cm.lambda_nr), np.array(cm.lambda_val)
    for p in (0, 7, 50, 99):
        c = int(cnt[p, 0])
        assert list(lam.nr[p][0]) == list(nr[p, 0, :c])
        np.testing.assert_array_equal(lam.Ls[p][0], val[p, 0, :c])


@pytest.mark.gpu
def test_core_shortchar_lambda_replaces_the_previous_lists(small_wind):
    """After a Feautrier solve with Lambda bands, the short-characteristics solve leaves
    its own diagonal Lambda in lspec.Lambda (solver.tpp:164-166, 1651, 1754), whatever
    n_off_diag says, and n_off_diag does not change its results."""
    import torch  # noqa: F401  (one HIP runtime: torch's)

    m = small_wind
    cm = m.to_core()
    cm.parameters.n_off_diag = 2
    cm.compute_radiation_field_feautrier_order_2()
    assert max(len(x) for x in cm.lines.lineProducingSpecies[0].Lambda.Ls[5]) > 1
    cm.compute_radiation_field_shortchar_order_0()
    J2 = np.array(cm.radiation.J).copy()
    diag = np.array(cm.lambda_diag)
    lam = cm.lines.lineProducingSpecies[0].Lambda
    line_k = np.frombuffer(bytes(cm.get_array("line_k")), dtype=np.uint32)
    for p in (0, 5, m.npoints // 2, m.npoints - 1):
        for l in range(m.nlines):
            k = int(line_k[l])
            assert list(lam.nr[p][k]) == [p]
            assert list(lam.Ls[p][k]) == [diag[p, l]]
    cm.parameters.n_off_diag = 0
    cm.compute_radiation_field_shortchar_order_0()
    np.testing.assert_array_equal(np.array(cm.radiation.J), J2)
    np.testing.assert_array_equal(np.array(cm.lambda_diag), diag)
