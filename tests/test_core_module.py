"""The pybind11 module magritte_amd.core (C++ magritte::Model with the
reference's method names, src/bindings/pybindings.cpp:123-271) loads, takes a
model's arrays, and reports errors as RuntimeError with the library's message
(no GPU is touched on the CPU)."""
import numpy as np
import pytest


def test_core_model_takes_arrays(line_model):
    from magritte_amd import core

    cm = line_model.to_core()
    assert isinstance(cm, core.Model)
    p = cm.parameters
    assert p.npoints == line_model.npoints and p.nfreqs == line_model.nfreqs
    assert not p.one_line_approximation and p.n_off_diag == 0
    with pytest.raises(RuntimeError, match="unknown model array"):
        cm.set_array("no_such_array", np.zeros(3))
    p.one_line_approximation = True
    assert cm.parameters.one_line_approximation


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
    cm.restore_line_grid()
    assert cm.parameters.nfreqs == small_wind.nfreqs
