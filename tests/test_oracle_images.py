"""Pins of the oracle's old imager and optical-depth images
(Solver::image_feautrier_order_2(model, rr), image_optical_depth(model, rr),
image_optical_depth_new_imager; solver.tpp:543-574, 656-736, 2718-2758):

- the reference's old-imager benchmark (constant_velocity_gradient_1D_image.py)
  against its analytic image, at the reference's own bounds;
- the same benchmark's analytic optical depth for both optical-depth images;
- a uniform-velocity identity and the optically thin limit, which fix the
  Rest-frame shifts (reversed on rays >= hnrays) and the boundary term taken at
  the unshifted frequency;
- the image coordinates: an orthonormal frame normal to the ray.
"""
import math
import os

import numpy as np
import pytest
from scipy.special import erf

from magritte_amd import _abi, models
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _bench(P, nquads, mwf):
    """The benchmark model of constant_velocity_gradient_1D_image.py:18-70 with its image
    frequency grid (:97-107) and the analytic optical depth and image (:131-158)."""
    ld = models.read_lamda(os.path.join(HERE, "golden", "lamda_test.txt"), num=0)
    dx, dv, temp, nTT = 1.0e4, 2.5e2, 45.0, 1.0e8
    m = models.model_velocity_gradient_1d(ld, npoints=P, nrays=4, nquads=nquads,
                                          max_width_fraction=mwf)
    fcen = ld["frequency"][0]
    dd = 1.0e4 / models.CC
    fq = m.image_frequencies(fcen - fcen * dd, fcen + fcen * dd, m.nlines * m.nquads)
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

    def tau(f, r):
        f = f + vmax / models.CC * fcen  # as the reference test shifts it (vmax is already /c)
        l = z_max(r, 0.5 * math.pi)
        arg = (f - fcen) / dnu
        fct = vmax * f / dnu
        up = chi * L / (fct * dnu) * 0.5 * (erf(arg) - erf(arg - fct * l / L))
        down = chi * L / (-fct * dnu) * 0.5 * (erf(arg) - erf(arg + fct * l / L))
        return up + down

    def image(f, r):
        fs = f + vmax / models.CC * fcen
        b = 2.0 * models.HH * fs**3 / models.CC**2 / math.expm1(models.HH * fs / (models.KB * models.T_CMB))
        return src + (b - src) * math.exp(-tau(f, r))

    return m, fq, tau, image


@pytest.mark.parametrize("mwf,bound", [(0.5, 1.8e-5), (0.35, 8e-6), (0.10, 3.5e-7)])
def test_old_imager_closed_form(mwf, bound):
    """constant_velocity_gradient_1D_image.py:111-158, 215-219: the image along the last ray
    (0, 1, 0) at every point, against the analytic image, all points but the outermost."""
    m, fq, _, image = _bench(100, 100, mwf)
    rr = m.hnrays - 1
    assert tuple(m.ray_direction[rr]) == (0.0, 1.0, 0.0)
    im, imx, imy = oracle.image_ray(m, m.options(), rr, fq)
    np.testing.assert_array_equal(imx, m.position[:, 0])
    np.testing.assert_array_equal(imy, 0.0)
    nu = fq["nu"][0]
    im_a = np.array([[image(f, r) for f in nu] for r in imx])
    err = np.abs((im - im_a) / im_a)[:-1]
    assert np.max(err) < bound, np.max(err)


def test_old_imager_optical_depth_closed_form():
    """The same benchmark's analytic optical depth (tau(nu, r, pi/2)) for
    image_optical_depth(model, rr), on the line-integrated branch (max_width_fraction
    0.10): within 3e-7 of the largest depth (observed 9e-8)."""
    m, fq, tau, _ = _bench(100, 100, 0.10)
    t, imx, _ = oracle.image_ray(m, m.options(), m.hnrays - 1, fq, optical_depth=True)
    nu = fq["nu"][0]
    ta = np.array([[tau(f, r) for f in nu] for r in imx])
    assert np.max(np.abs(t - ta)[:-1]) < 3e-7 * ta.max(), np.max(np.abs(t - ta)[:-1])


def test_new_optical_depth_image_closed_form():
    """image_optical_depth_new_imager on the projection-surface pixels of the new-imager
    benchmark (constant_velocity_gradient_1D_new_imager.py): within 3e-7 of the largest
    depth (observed 6e-8), and exactly 0 for the pixel rays that miss the model."""
    m, fq, tau, _ = _bench(50, 10, 0.10)
    raydir = m.ray_direction[m.hnrays - 1]
    t, imx, imy = oracle.image_optical_depth_new(m, m.options(), raydir, 50, 50, fq)
    rs = np.sqrt(imx**2 + imy**2)
    nu = fq["nu"][0]
    ta = np.array([[tau(f, r) for f in nu] for r in rs])
    assert np.max(np.abs(t - ta)) < 3e-7 * ta.max(), np.max(np.abs(t - ta))
    outside = ta.max(axis=1) == 0.0
    assert outside.sum() > 0
    np.testing.assert_array_equal(t[outside], 0.0)


def _planck_cmb(nu):
    """boundary_intensity for a CMB boundary: 2 h nu^3 / c^2 / expm1f(h nu / k T_CMB), with
    the float expm1 of the reference (the oracle's correctly rounded one)."""
    f = oracle.lib().oracle_expm1f
    a = (models.HH / models.KB * nu / models.T_CMB).astype(np.float32)
    e = np.array([f(float(x)) for x in a.ravel()]).reshape(nu.shape)
    return 2.0 * models.HH / models.CC**2 * nu**3 / e


def _wind(**kw):
    return models.model_wind(n_edge=8, nrays=12, seed=models.seed_for(1, 7), v_out=0.0, **kw)


@pytest.mark.parametrize("rr", [1, 9])
def test_old_imager_uniform_velocity_identity(rr):
    """With one velocity u everywhere the old imager's Rest-frame shifts are a single
    factor s = 1 + u.d over the whole ray pair, d the direction of the pair's ray below
    hnrays (the shifts of rays >= hnrays are reversed, 2 - shift).  So its u at frequency nu
    is the static model's u at nu s, and I + I_bdy(nu) = I_static(nu s) + I_bdy(nu s), the
    boundary term being taken at the unshifted frequency; likewise tau(nu) = tau_static(nu s).
    The reference's line profile takes a float argument, so the identity holds to its
    rounding (observed 8e-7 for I + I_bdy, 1e-6 for tau)."""
    m = _wind()
    mv = _wind()
    u = np.array([3.0e3, -2.0e3, 1.0e3]) / models.CC
    mv.velocity[:] = u
    lo = min(rr, int(m.antipod[rr]))
    assert (rr >= m.hnrays) == (lo != rr)
    s = 1.0 + u @ m.ray_direction[lo]
    op = m.options()
    fq = dict(nu=m.nu * s)
    Im, _, _ = oracle.image_ray(mv, op, rr)
    Is, _, _ = oracle.image_ray(m, op, rr, fq)
    lhs, rhs = Im + _planck_cmb(m.nu), Is + _planck_cmb(m.nu * s)
    assert np.max(np.abs(lhs - rhs) / np.abs(rhs)) < 3e-6
    tm, _, _ = oracle.image_ray(mv, op, rr, optical_depth=True)
    ts, _, _ = oracle.image_ray(m, op, rr, fq, optical_depth=True)
    assert np.max(np.abs(tm - ts) / np.abs(ts)) < 3e-6


@pytest.mark.parametrize("rr", [1, 9])
def test_old_imager_thin_limit(rr):
    """No line emission and only the floor opacity (tau ~ 1e-12): u is the boundary
    intensity at the shifted frequency on both ends, so the reference's image is
    2 I_bdy(nu s) - I_bdy(nu), not I_bdy(nu s).  The test separates the two by > 1e4."""
    mv = _wind()
    u = np.array([3.0e4, -2.0e4, 1.0e4]) / models.CC
    mv.velocity[:] = u
    mv.emissivity[:] = 1.0e-300
    mv.opacity[:] = 1.0e-30
    s = 1.0 + u @ mv.ray_direction[min(rr, int(mv.antipod[rr]))]
    I, _, _ = oracle.image_ray(mv, mv.options(), rr)
    b, bs = _planck_cmb(mv.nu), _planck_cmb(mv.nu * s)
    assert np.max(np.abs(I - (2.0 * bs - b)) / b) < 1e-8
    assert np.min(np.max(np.abs(I - bs) / b, axis=1)) > 1e-5


def test_old_imager_coordinates_are_an_orthonormal_frame(small_wind):
    """ImX, ImY of every point are its coordinates in an orthonormal frame (i, j) normal to
    the ray (image.cpp:173-199): recovered by least squares from the positions."""
    m = small_wind
    for rr in (0, 5, m.hnrays + 2):
        _, X, Y = oracle.image_ray(m, m.options(), rr, optical_depth=True)
        A = np.c_[m.position, np.ones(m.npoints)]
        i = np.linalg.lstsq(A, X, rcond=None)[0]
        j = np.linalg.lstsq(A, Y, rcond=None)[0]
        d = m.ray_direction[rr]
        assert abs(i[3]) < 1e-3 and abs(j[3]) < 1e-3
        i, j = i[:3], j[:3]
        assert abs(i @ i - 1) < 1e-12 and abs(j @ j - 1) < 1e-12
        assert abs(i @ j) < 1e-12 and abs(i @ d) < 1e-12 and abs(j @ d) < 1e-12


def test_old_imager_arguments(small_wind, vgrad_model):
    m = small_wind
    with pytest.raises(RuntimeError):
        oracle.image_ray(m, m.options(approx=_abi.APPROX_ONE_LINE), 0)
    with pytest.raises(RuntimeError):
        oracle.image_ray(m, m.options(), m.nrays)
    # 1D: only the ray (0, 1, 0)
    v = vgrad_model
    assert tuple(v.ray_direction[0]) != (0.0, 1.0, 0.0)
    with pytest.raises(RuntimeError):
        oracle.image_ray(v, v.options(), 0)
