"""Pins the CPU oracle's adjoint with respect to the inverse line widths
iw(p, l) (the third output of oracle.J_adjoint / oracle.image_adjoint), which
carries a gradient with respect to temperature through the Doppler width.

The adjoint has no counterpart in the reference, so it is pinned against the
oracle's own forward solve by central finite differences (step 1e-6 relative,
the trace and every discrete choice held fixed) on all three approximations,
both branches of compute_source_dtau (the Gaussian profile of the trapezoid rule
and the error-function difference of the line-integrated depth) and the imager.
The finite differences run with the profile and the error function evaluated in
double (oracle.set_smooth_profile): the reference's float expf and erff round
their results to ~6e-8 relative, which a difference quotient with step 1e-6
turns into errors of several percent (the last test shows it).
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
    oracle.set_smooth_profile(True)
    yield
    oracle.set_smooth_interp(False)
    oracle.set_smooth_profile(False)


def _fd_J(m, op, Jb, p, l):
    arr = m.inverse_width
    x0 = arr[p, l]
    arr[p, l] = x0 * (1 + H)
    _, Jp, _ = oracle.solve(m, op, with_u=False)
    arr[p, l] = x0 * (1 - H)
    _, Jm, _ = oracle.solve(m, op, with_u=False)
    arr[p, l] = x0
    return float(np.sum(Jb * (Jp - Jm)) / (2 * H * x0))


def _pick(g, k, rng, frac=1e-3):
    cand = np.flatnonzero(np.abs(g) > frac * np.abs(g).max())
    return [divmod(int(i), g.shape[1]) for i in rng.choice(cand, min(k, cand.size), replace=False)]


def _check_J(m, op, seed, k=6, tol=1e-6):
    rng = np.random.default_rng(seed)
    Jb = rng.standard_normal((m.npoints, m.nfreqs))
    de, dk, dw = oracle.J_adjoint(m, op, Jb, want_iw=True)
    assert np.abs(dw).max() > 0
    worst = 0.0
    for p, l in _pick(dw, k, rng):
        fd = _fd_J(m, op, Jb, p, l)
        worst = max(worst, abs(dw[p, l] - fd) / abs(fd))
    assert worst < tol, worst
    # the emissivity and opacity gradients do not change when the third is asked for
    # (up to the order in which the threads' partial sums are added)
    de0, dk0 = oracle.J_adjoint(m, op, Jb)
    np.testing.assert_allclose(de, de0, rtol=1e-12, atol=1e-14 * np.abs(de0).max())
    np.testing.assert_allclose(dk, dk0, rtol=1e-12, atol=1e-14 * np.abs(dk0).max())
    return worst


@pytest.mark.parametrize("approx", [_abi.APPROX_CLOSE_LINES, _abi.APPROX_NONE, _abi.APPROX_ONE_LINE])
def test_J_width_adjoint_matches_finite_differences(wind6, smooth, approx):
    _check_J(wind6, wind6.options(approx=approx), 40 + approx)


@pytest.mark.parametrize("branch", ["trapezoid", "erf"])
def test_J_width_adjoint_each_branch(wind6, smooth, branch):
    """dshift_max huge: every segment takes the trapezoid rule (the profile's
    derivative); zero: the line-integrated form wherever the shift changes (the
    error functions' derivative)."""
    saved = wind6.dshift_max.copy()
    try:
        wind6.dshift_max[:] = 1.0e10 if branch == "trapezoid" else 0.0
        _check_J(wind6, wind6.options(), 50)
    finally:
        wind6.dshift_max[:] = saved


def test_J_width_adjoint_1d_chain(line_model, smooth):
    _check_J(line_model, line_model.options(), 7)


def test_image_width_adjoint_matches_finite_differences(wind6, smooth):
    op = wind6.options()
    fr = wind6.image_frequencies(wind6.line[0] * (1 - 2e-5), wind6.line[0] * (1 + 2e-5), 31)
    rd = np.array([0.3, -0.5, 0.81])
    nx = ny = 12
    I0, _, _ = oracle.image_new(wind6, op, rd, nx, ny, fr)
    rng = np.random.default_rng(9)
    Ib = rng.standard_normal(I0.shape)
    de, dk, dw = oracle.image_adjoint(wind6, op, rd, nx, ny, Ib, fr, want_iw=True)
    worst = 0.0
    arr = wind6.inverse_width
    # gradients within 1e-2 of the largest: below that the image's own rounding
    # (~1e-18 absolute in the quotient at this step) reaches 1e-6 of the value
    for p, l in _pick(dw, 6, rng, frac=1e-2):
        x0 = arr[p, l]
        arr[p, l] = x0 * (1 + H)
        Ip, _, _ = oracle.image_new(wind6, op, rd, nx, ny, fr)
        arr[p, l] = x0 * (1 - H)
        Im, _, _ = oracle.image_new(wind6, op, rd, nx, ny, fr)
        arr[p, l] = x0
        fd = float(np.sum(Ib * (Ip - Im)) / (2 * H * x0))
        worst = max(worst, abs(dw[p, l] - fd) / abs(fd))
    assert worst < 1e-6, worst


def test_width_euler_identity():
    """An identity that ties the whole inverse-width adjoint to the (separately
    pinned) emissivity and opacity adjoints: with no velocity field, the
    trapezoid rule everywhere and one frequency bin at the line centre, every
    eta = prof eps and chi - chi_min = prof kap with prof = iw / sqrt(pi), so
    iw dL/d iw = eps dL/d eps + kap dL/d kap at every point, exactly."""
    m = models.model_line_1d(nquads=1, vmax=0.0)
    m.dshift_max[:] = 1.0e10
    op = m.options(approx=_abi.APPROX_ONE_LINE)
    rng = np.random.default_rng(13)
    Jb = rng.standard_normal((m.npoints, m.nfreqs))
    de, dk, dw = oracle.J_adjoint(m, op, Jb, want_iw=True)
    # one bin at the line centre: x = 0, so iw dL/d iw = eps dL/d eps + kap dL/d kap
    assert np.allclose(m.nu[:, 0], m.line[0])
    lhs = m.inverse_width[:, 0] * dw[:, 0]
    rhs = m.emissivity[:, 0] * de[:, 0] + m.opacity[:, 0] * dk[:, 0]
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-14 * np.abs(rhs).max())


def test_reference_profile_limits_finite_differences(wind6):
    """With the reference's float expf and erff, finite differences in the inverse
    widths carry rounding steps far above the smooth evaluation's.  Documents why
    the checks above use it."""
    op = wind6.options()
    rng = np.random.default_rng(41)
    Jb = rng.standard_normal((wind6.npoints, wind6.nfreqs))
    oracle.set_smooth_interp(True)
    oracle.set_smooth_profile(True)
    try:
        _, _, dw = oracle.J_adjoint(wind6, op, Jb, want_iw=True)
        pts = _pick(dw, 6, rng)
        errs_s = [abs(dw[p, l] - _fd_J(wind6, op, Jb, p, l)) / abs(dw[p, l]) for p, l in pts]
    finally:
        oracle.set_smooth_profile(False)
    try:
        errs_r = [abs(dw[p, l] - _fd_J(wind6, op, Jb, p, l)) / abs(dw[p, l]) for p, l in pts]
    finally:
        oracle.set_smooth_interp(False)
    assert max(errs_s) < 1e-6
    assert max(errs_r) > 100 * max(errs_s)
