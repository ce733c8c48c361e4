"""Known-answer and brute-force checks that pin the CPU oracle.

* constant slab: the reference's all-constant single-ray benchmark closed form
  u = (I0 + I1)/2, I = S + (I_bdy - S) exp(-chi x)
  (tests/benchmarks/analytic/all_constant_single_ray.py:112-131, bound :180);
* the elimination: u(centre) and the Lambda diagonal of the oracle against an
  exact (50-digit) solve of the tridiagonal Feautrier system built from the
  same optical depths and sources (the reference's T u = rhs and
  Lambda = diag(T^-1) idea, tests/test_solver_lambda.cpp:15-88);
* the walk: on a 1D chain the traced list is exactly o-1, ..., 0 and
  o+1, ..., P-1; on any model the ends are boundary points, dZ > 0 and the
  interpolation counts equal ceil(|log(ratio)| / log(1.4)) (exact logs).
"""
import math
import os

import mpmath as mp
import numpy as np
import pytest

from magritte_amd import models
from oracle import oracle

import _exact

HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return 2.0 * (a - b) / (a + b)  # magritte/tools.py relative_error


def test_constant_slab_closed_form():
    ld = models.read_lamda(os.path.join(HERE, "golden", "lamda_test.txt"), num=0)
    m = models.model_constant_slab(linedata=ld)
    op = m.options()
    u, J, lam = oracle.solve(m, op)
    # closed form (magritte/tools.py LTEpop, profile, lineEmissivity, lineOpacity, I_CMB)
    T = 4.5
    w = ld["weight"] * np.exp(-ld["energy"] / (models.KB * T))
    pop = w / w.sum() * 1.0e3
    frq = ld["frequency"][0]
    width = frq * math.sqrt(2.0 * models.KB * T / (models.AMU * models.CC_SQUARED) * ld["inverse_mass"])
    phi = 1.0 / (width * math.sqrt(math.pi))
    hv4pi = models.HH * frq / (4.0 * math.pi)
    eta = hv4pi * ld["A"][0] * pop[1] * phi
    chi = hv4pi * (pop[0] * ld["Ba"][0] - pop[1] * ld["Bs"][0]) * phi
    src = eta / chi
    bdy = 2.0 * models.HH / models.CC_SQUARED * frq**3 / math.expm1(models.HH * frq / (models.KB * models.T_CMB))
    x = m.position[:, 0]
    I0 = src + (bdy - src) * np.exp(-chi * x)
    I1 = src + (bdy - src) * np.exp(-chi * (x[-1] - x))
    ua = 0.5 * (I0 + I1)
    err = np.max(np.abs(rel(ua, u[0, :, 0])))
    assert err < 1.7e-4, err
    # and it is a real test: the slab is neither transparent nor opaque everywhere
    assert 0.1 < chi * x[-1] < 1e3


def _exact_tridiag(sys_):
    """u and diag(T^-1) at the centre of the reference's second-order Feautrier system."""
    a, b, c, r = _exact.system(sys_)
    cidx = sys_["centre"]
    u = _exact.solve((a, b, c), r)[cidx]
    e = [mp.mpf(0)] * len(b)
    e[cidx] = mp.mpf(1)
    lam = _exact.solve((a, b, c), e)[cidx]
    return float(u), float(lam)


@pytest.mark.parametrize("which", ["line_model", "small_wind"])
def test_elimination_matches_exact_solve(which, request):
    m = request.getfixturevalue(which)
    op = m.options()
    rng = np.random.default_rng(11)
    checked = 0
    worst_u = worst_l = 0.0
    for _ in range(40):
        rr = int(rng.integers(0, m.hnrays))
        o = int(rng.integers(0, m.npoints))
        f = int(rng.integers(0, m.nfreqs))
        try:
            s = oracle.feautrier_system(m, op, rr, o, f)
        except RuntimeError:
            continue  # n_tot < 2: boundary branch, no system
        u_ex, l_ex = _exact_tridiag(s)
        worst_u = max(worst_u, abs(s["u_c"] - u_ex) / abs(u_ex))
        worst_l = max(worst_l, abs(s["L_c"] - l_ex) / abs(l_ex))
        checked += 1
    assert checked >= 20
    assert worst_u < 1e-11 and worst_l < 1e-11, (worst_u, worst_l)


def test_walk_1d_chain_brute_force(line_model):
    m = line_model
    op = m.options()
    P = m.npoints
    for o in (0, 1, 37, P - 2, P - 1):
        t = oracle.trace_ray(m, op, 0, o, 0)
        c = t["centre"]
        # ray 0 points along -x: entries first..centre-1 are 0..o-1, then o, then o+1..P-1
        np.testing.assert_array_equal(t["nr"], np.arange(P))
        assert c == o
        assert np.all(t["interp_factor"] == 0.0)
        dx = m.position[1, 0] - m.position[0, 0]
        assert np.allclose(t["dZ"][:-1], dx, rtol=1e-12)


def test_walk_invariants_and_interpolation_counts(small_wind):
    m = small_wind
    op = m.options()
    rng = np.random.default_rng(3)
    log14 = float(np.float32(mp.log(mp.mpf(float(np.float32(1.4))))))
    for _ in range(30):
        rr, o, l = int(rng.integers(0, m.hnrays)), int(rng.integers(0, m.npoints)), int(rng.integers(0, m.nlines))
        t = oracle.trace_ray(m, op, rr, o, l)
        n, c = t["count"], t["centre"]
        nr = t["nr"]
        if n > 1:
            # both ends are boundary points (unless the origin itself is the end)
            assert m.point2boundary[nr[0]] != m.npoints or c == 0
            assert m.point2boundary[nr[-1]] != m.npoints or c == n - 1
            assert np.all(t["dZ"][:-1] > 0.0)
        # interpolation counts: group entries by their (nr, nr_interp) segment
        j = 0
        while j < n:
            if j == c:
                j += 1
                continue
            seg = (nr[j], t["nr_interp"][j])
            k = j
            while k < n and k != c and (nr[k], t["nr_interp"][k]) == seg:
                k += 1
            cnt = k - j
            a, b = m.interp_crit[seg[1], l], m.interp_crit[seg[0], l]
            ratio = float(np.float32(b / a))
            lg = float(np.float32(abs(mp.log(mp.mpf(ratio)))))
            expect = max(1, math.ceil(lg / log14))
            assert cnt == expect, (rr, o, l, j, cnt, expect)
            j = k


def _velocity_gradient_closed_form(m, ld, temp, n_species, dx, dv, P):
    """u(ray, r, nu) of the reference's constant-velocity-gradient benchmark
    (constant_velocity_gradient_1D.py:130-177)."""
    from scipy.special import erf

    k = 0
    frq = ld["frequency"][k]
    w = ld["weight"] * np.exp(-ld["energy"] / (models.KB * temp))
    pop = w / w.sum() * n_species
    hv = models.HH * frq / (4.0 * math.pi)
    chi = hv * (pop[0] * ld["Ba"][0] - pop[1] * ld["Bs"][0])
    eta = hv * ld["A"][0] * pop[1]
    src = eta / chi
    dnu = frq * math.sqrt(2.0 * models.KB * temp / (models.AMU * models.CC**2) * ld["inverse_mass"])
    rs = m.position[:, 0]
    r_in, r_out = rs[0], rs[-1]
    L = dx * P
    vmax = dv / models.CC * P

    def v_(r):
        return dv / models.CC * (1 + (r - r_in) / dx)

    def I_cmb(nu):
        return 2.0 * models.HH * nu**3 / models.CC**2 / math.expm1(models.HH * nu / (models.KB * models.T_CMB))

    def bdy(nu, r, th):
        r_shift = math.cos(th) * v_(r)
        if th < math.asin(r_in / r):
            e = math.sqrt(1 - ((r / r_in) * abs(math.sin(th))) ** 2) * v_(r_in)
        else:
            e = math.sqrt(1 - ((r / r_out) * abs(math.sin(th))) ** 2) * v_(r_out)
        return I_cmb(nu * (1.0 + (e - r_shift)))

    def z_max(r, th):
        if th < math.asin(r_in / r):
            return r * math.cos(th) - math.sqrt(r_in**2 - (r * math.sin(th)) ** 2)
        return r * math.cos(th) + math.sqrt(L**2 - (r * math.sin(th)) ** 2)

    def tau(nu, r, th):
        l = z_max(r, th)
        arg = (nu - frq) / dnu
        fct = vmax * nu / dnu
        return chi * L / (fct * dnu) * 0.5 * (erf(-arg) + erf(fct * l / L + arg))

    def I_(nu, r, th):
        return src + (bdy(nu, r, th) - src) * math.exp(-tau(nu, r, th))

    angles = np.arctan2(m.ray_direction[: m.hnrays, 1], m.ray_direction[: m.hnrays, 0])
    nu = m.nu[0]
    return np.array([[[0.5 * (I_(f, r, a) + I_(f, r, np.pi + a)) for f in nu] for r in rs]
                     for a in angles])


@pytest.mark.parametrize("mwf,bound", [(0.5, 1e-6), (0.35, 5e-7), (0.10, 5.31e-8)])
def test_velocity_gradient_erf_branch_closed_form(mwf, bound):
    """Pins the line-integrated (erf) optical depth and the spherical tracer:
    mean error within the reference's bounds for max_width_fraction 0.5 / 0.35 /
    0.10 (constant_velocity_gradient_1D.py:91-95, 249-256).  At 0.10 every
    segment takes the erf branch."""
    ld = models.read_lamda(os.path.join(HERE, "golden", "lamda_test.txt"), num=0)
    m = models.model_velocity_gradient_1d(ld, max_width_fraction=mwf)
    op = m.options()
    u, _, _ = oracle.solve(m, op)
    us = _velocity_gradient_closed_form(m, ld, 45.0, 1.0e8, 1.0e4, 2.5e2, m.npoints)
    err = np.abs(rel(us, u))
    assert np.mean(err) < bound, np.mean(err)


def test_reference_is_sensitive_to_the_erff_rounding(small_wind):
    """Documents why the kernels evaluate the C library's own float functions:
    the reference forms erf(b) - erf(a) in float, so the same solve with
    correctly rounded functions instead of glibc's moves J by ~1e-4 relative,
    while the traced counts do not move."""
    m = small_wind
    op = m.options()
    _, J1, _ = oracle.solve(m, op, with_u=False, kind="cr")
    _, J2, _ = oracle.solve(m, op, with_u=False, kind="glibc")
    d = np.max(np.abs(J2 - J1) / np.abs(J1))
    assert 1e-6 < d < 1e-2, d
    np.testing.assert_array_equal(oracle.trace_counts(m, op, kind="cr"),
                                  oracle.trace_counts(m, op, kind="glibc"))


@pytest.mark.parametrize("mwf,bound", [(0.5, 1.6e-5), (0.35, 1.2e-5), (0.10, 3.3e-7)])
def test_new_imager_closed_form(mwf, bound):
    """Pins the projection-surface imager (frame, start-point search, imaging
    trace, forward elimination, I = 2 Su - I_bdy) against the reference's
    new-imager benchmark closed form and its bounds
    (constant_velocity_gradient_1D_new_imager.py:18-31, 85-160, 221-227)."""
    from scipy.special import erf

    ld = models.read_lamda(os.path.join(HERE, "golden", "lamda_test.txt"), num=0)
    P, dx, dv, temp, nTT = 50, 1.0e4, 2.5e2, 45.0, 1.0e8
    m = models.model_velocity_gradient_1d(ld, npoints=P, nrays=4, nquads=10,
                                          max_width_fraction=mwf)
    fcen = ld["frequency"][0]
    dd = 1.0e4 / models.CC
    fq = m.image_frequencies(fcen - fcen * dd, fcen + fcen * dd, m.nlines * m.nquads)
    raydir = m.ray_direction[m.hnrays - 1]
    assert tuple(raydir) == (0.0, 1.0, 0.0)
    im, imx, imy = oracle.image_new(m, m.options(), raydir, 50, 50, fq)
    nu = fq["nu"][0]
    rs = np.sqrt(imx**2 + imy**2)
    w = ld["weight"] * np.exp(-ld["energy"] / (models.KB * temp))
    pop = w / w.sum() * nTT
    hv = models.HH * fcen / (4.0 * math.pi)
    chi = hv * (pop[0] * ld["Ba"][0] - pop[1] * ld["Bs"][0])
    src = hv * ld["A"][0] * pop[1] / chi
    dnu = fcen * math.sqrt(2.0 * models.KB * temp / (models.AMU * models.CC**2) * ld["inverse_mass"])
    r_in, r_out = m.position[0, 0], m.position[-1, 0]
    L = dx * P
    vmax = dv / models.CC * P

    def z_max(r, th):
        if r >= r_out:
            return 0.0
        if th < math.asin(r_in / r):
            return r * math.cos(th) - math.sqrt(r_in**2 - (r * math.sin(th)) ** 2)
        return r * math.cos(th) + math.sqrt(L**2 - (r * math.sin(th)) ** 2)

    def tau(f, r, th):
        l = z_max(r, th)
        arg = (f - fcen) / dnu
        fct = vmax * f / dnu
        up = chi * L / (fct * dnu) * 0.5 * (erf(arg) - erf(arg - fct * l / L))
        down = chi * L / (-fct * dnu) * 0.5 * (erf(arg) - erf(arg + fct * l / L))
        return up + down

    def I_im(f, r):
        f = f + vmax / models.CC * fcen  # as written in the reference test (vmax is already /c)
        b = 2.0 * models.HH * f**3 / models.CC**2 / math.expm1(models.HH * f / (models.KB * models.T_CMB))
        return src + (b - src) * math.exp(-tau(f, r, 0.5 * math.pi))

    im_a = np.array([[I_im(f, r) for f in nu] for r in rs])
    err = np.abs(rel(im, im_a))[:-1]
    assert np.max(err) < bound, np.max(err)


@pytest.mark.parametrize("approx", [0, 1, 2])
def test_restated_functions_reproduce_the_c_library_solve(small_wind, approx):
    """The oracle built on the restated float functions (the code the kernels
    run) and the one calling the C library give the same u, J and Lambda bit for
    bit, and the same traced counts, on a wind with every branch in use."""
    m = small_wind
    op = m.options()
    op.approx = approx
    a = oracle.solve(m, op, kind="ref")
    b = oracle.solve(m, op, kind="glibc")
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(oracle.trace_counts(m, op, kind="ref"),
                                  oracle.trace_counts(m, op, kind="glibc"))


def test_restated_functions_reproduce_the_c_library_erf_branch():
    """Same on the velocity-gradient benchmark with max_width_fraction 0.10,
    where every segment takes the erf branch and J depends most on erff's bits."""
    ld = models.read_lamda(os.path.join(HERE, "golden", "lamda_test.txt"), num=0)
    m = models.model_velocity_gradient_1d(ld, max_width_fraction=0.10)
    op = m.options()
    a = oracle.solve(m, op, kind="ref")
    b = oracle.solve(m, op, kind="glibc")
    c = oracle.solve(m, op, kind="cr")
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert np.max(np.abs(a[1] - c[1]) / np.abs(b[1])) > 1e-9
