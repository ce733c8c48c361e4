"""Pins the CPU oracle's adjoint (vector-Jacobian products of J and of images
with respect to the line emissivity eps(p, l) and opacity kappa(p, l)).

The adjoint has no counterpart in the reference, so it is pinned against the
oracle's own forward solve by central finite differences (step 1e-6 relative,
the trace and every discrete choice held fixed), on all three approximations,
both branches of compute_source_dtau and the imager.  The finite differences
run with interpolate_log evaluated in double (oracle.set_smooth_interp): the
reference's float powf rounds b/a and its result to float, which puts steps of
~1e-7 relative into a difference quotient; the last test shows that effect on
a point that has interpolated neighbours.
"""
import numpy as np
import pytest

from magritte_amd import _abi, models
from oracle import oracle

H = 1.0e-6


@pytest.fixture(scope="module")
def wind6():
    return models.model_wind(n_edge=6, seed=models.seed_for(1, 11))


@pytest.fixture
def smooth():
    oracle.set_smooth_interp(True)
    yield
    oracle.set_smooth_interp(False)


def _fd_J(m, op, Jb, arr, p, l):
    x0 = arr[p, l]
    arr[p, l] = x0 * (1 + H)
    _, Jp, _ = oracle.solve(m, op, with_u=False)
    arr[p, l] = x0 * (1 - H)
    _, Jm, _ = oracle.solve(m, op, with_u=False)
    arr[p, l] = x0
    return float(np.sum(Jb * (Jp - Jm)) / (2 * H * x0))


def _pick(g, k, rng, frac=1e-3):
    """k indices of (P, L) whose gradient is not negligible (finite differences of a
    vanishing derivative only measure rounding)."""
    cand = np.flatnonzero(np.abs(g) > frac * np.abs(g).max())
    return [divmod(int(i), g.shape[1]) for i in rng.choice(cand, min(k, cand.size), replace=False)]


def _check(m, op, Jb, de, dk, rng, k=5, tol=1e-6):
    worst = 0.0
    for arr, g in ((m.emissivity, de), (m.opacity, dk)):
        for p, l in _pick(g, k, rng):
            fd = _fd_J(m, op, Jb, arr, p, l)
            worst = max(worst, abs(g[p, l] - fd) / abs(fd))
    assert worst < tol, worst
    return worst


@pytest.mark.parametrize("approx", [_abi.APPROX_CLOSE_LINES, _abi.APPROX_NONE, _abi.APPROX_ONE_LINE])
def test_J_adjoint_matches_finite_differences(wind6, smooth, approx):
    op = wind6.options(approx=approx)
    rng = np.random.default_rng(5 + approx)
    Jb = rng.standard_normal((wind6.npoints, wind6.nfreqs))
    de, dk = oracle.J_adjoint(wind6, op, Jb)
    _check(wind6, op, Jb, de, dk, rng)


@pytest.mark.parametrize("branch", ["trapezoid", "erf"])
def test_J_adjoint_each_branch(wind6, smooth, branch):
    """dshift_max huge -> every segment takes the trapezoid rule; zero -> the erf
    (line-integrated) form wherever the shift changes."""
    saved = wind6.dshift_max.copy()
    try:
        wind6.dshift_max[:] = 1.0e10 if branch == "trapezoid" else 0.0
        op = wind6.options()
        rng = np.random.default_rng(17)
        Jb = rng.standard_normal((wind6.npoints, wind6.nfreqs))
        de, dk = oracle.J_adjoint(wind6, op, Jb)
        _check(wind6, op, Jb, de, dk, rng)
    finally:
        wind6.dshift_max[:] = saved


def test_J_adjoint_1d_chain(line_model, smooth):
    op = line_model.options()
    rng = np.random.default_rng(3)
    Jb = rng.standard_normal((line_model.npoints, line_model.nfreqs))
    de, dk = oracle.J_adjoint(line_model, op, Jb)
    _check(line_model, op, Jb, de, dk, rng)


def test_image_adjoint_matches_finite_differences(wind6, smooth):
    op = wind6.options()
    fr = wind6.image_frequencies(wind6.line[0] * (1 - 2e-5), wind6.line[0] * (1 + 2e-5), 31)
    rd = np.array([0.3, -0.5, 0.81])
    nx = ny = 12
    I0, _, _ = oracle.image_new(wind6, op, rd, nx, ny, fr)
    rng = np.random.default_rng(8)
    Ib = rng.standard_normal(I0.shape)
    de, dk = oracle.image_adjoint(wind6, op, rd, nx, ny, Ib, fr)
    worst = 0.0
    for arr, g in ((wind6.emissivity, de), (wind6.opacity, dk)):
        for p, l in _pick(g, 5, rng):
            x0 = arr[p, l]
            arr[p, l] = x0 * (1 + H)
            Ip, _, _ = oracle.image_new(wind6, op, rd, nx, ny, fr)
            arr[p, l] = x0 * (1 - H)
            Im, _, _ = oracle.image_new(wind6, op, rd, nx, ny, fr)
            arr[p, l] = x0
            fd = float(np.sum(Ib * (Ip - Im)) / (2 * H * x0))
            worst = max(worst, abs(g[p, l] - fd) / abs(fd))
    assert worst < 1e-6, worst


def test_adjoint_is_linear_in_the_seed(wind6):
    """The vector-Jacobian product is linear in J_bar (a wrong sign or a dropped
    term in any one row would break the additivity over disjoint seeds)."""
    op = wind6.options()
    rng = np.random.default_rng(21)
    J1 = rng.standard_normal((wind6.npoints, wind6.nfreqs))
    J2 = rng.standard_normal((wind6.npoints, wind6.nfreqs))
    a = oracle.J_adjoint(wind6, op, J1)
    b = oracle.J_adjoint(wind6, op, J2)
    c = oracle.J_adjoint(wind6, op, 2.0 * J1 - 3.0 * J2)
    for x, y, z in zip(a, b, c):
        np.testing.assert_allclose(z, 2.0 * x - 3.0 * y, rtol=1e-9, atol=1e-12 * np.abs(z).max())


def test_reference_powf_limits_finite_differences(wind6):
    """With the reference's float powf, finite differences through interpolated
    entries carry rounding steps; the smooth evaluation removes them.  Documents
    why the checks above use it."""
    op = wind6.options()
    rng = np.random.default_rng(5)
    Jb = rng.standard_normal((wind6.npoints, wind6.nfreqs))
    oracle.set_smooth_interp(True)
    try:
        de, _ = oracle.J_adjoint(wind6, op, Jb)
        errs_s = []
        pts = _pick(de, 12, rng)
        for p, l in pts:
            errs_s.append(abs(de[p, l] - _fd_J(wind6, op, Jb, wind6.emissivity, p, l)) / abs(de[p, l]))
    finally:
        oracle.set_smooth_interp(False)
    errs_r = [abs(de[p, l] - _fd_J(wind6, op, Jb, wind6.emissivity, p, l)) / abs(de[p, l])
              for p, l in pts]
    assert max(errs_s) < 1e-6
    assert max(errs_r) > 10 * max(errs_s)
