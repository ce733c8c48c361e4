"""Host-side producers (libmgx mgx_compute_*) against the closed forms of
magritte/tools.py (LTEpop, lineEmissivity, lineOpacity, profile width) and
the spectral grid definition of Model::compute_spectral_discretisation."""
import math

import numpy as np

from magritte_amd import models


def test_lte_emissivity_opacity_closed_form(small_wind):
    m = small_wind
    ld = m.linedata[0]
    P = m.npoints
    for p in (0, P // 3, P - 1):
        T = m.temperature[p]
        w = ld["weight"] * np.exp(-ld["energy"] / (models.KB * T))
        pop = w / w.sum() * m.abundance[p, ld["num"]]
        hv = models.HH * ld["frequency"] / (4.0 * math.pi)
        eta = hv * ld["A"] * pop[ld["irad"]]
        chi = hv * (pop[ld["jrad"]] * ld["Ba"] - pop[ld["irad"]] * ld["Bs"])
        eta = np.maximum(eta, 1e-50)
        chi = np.maximum(chi, 1e-13)
        np.testing.assert_allclose(m.emissivity[p], eta, rtol=1e-12)
        np.testing.assert_allclose(m.opacity[p], chi, rtol=1e-10)
        np.testing.assert_allclose(m.interp_crit[p], m.opacity[p] / m.abundance[p, ld["num"]], rtol=1e-15)


def test_widths_spectral_grid_dshift(small_wind):
    m = small_wind
    ld = m.linedata[0]
    P, Q = m.npoints, m.nquads
    roots, weights = models.quadrature(Q)
    assert abs(weights.sum() - 1.0) < 1e-14
    for p in (0, P - 1):
        delta = math.sqrt(2 * models.KB * m.temperature[p] / (models.AMU * models.CC_SQUARED) * ld["inverse_mass"] + m.vturb2[p])
        np.testing.assert_allclose(m.inverse_width[p], 1.0 / (ld["frequency"] * delta), rtol=1e-14)
        nu = (ld["frequency"][:, None] * (1.0 + delta * roots[None, :])).ravel()
        np.testing.assert_allclose(m.nu[p], nu, rtol=1e-14)
        assert abs(m.dshift_max[p] - 0.35 * delta) <= 1e-15 * delta
    f = np.arange(m.nfreqs)
    np.testing.assert_array_equal(m.f_line, f // Q)
    np.testing.assert_array_equal(m.f_z, f % Q)


def test_sorted_lines_and_rotor_data():
    ld = models.linedata_linear_rotor(57.6e9, 0.11, 10)
    np.testing.assert_allclose(ld["frequency"], 2 * 57.6e9 * np.arange(1, 10), rtol=1e-12)
    assert 6.5e-8 < ld["A"][0] < 7.5e-8  # ~7.2e-8 s^-1 for J=1-0 of a 0.11 D rotor
    m = models.model_line_1d(npoints=8)
    np.testing.assert_array_equal(m.sorted_line, np.sort(m.line))


def test_healpix_directions():
    for nside in (1, 2, 4):
        d = models.healpix_pix2vec(nside)
        assert d.shape == (12 * nside * nside, 3)
        np.testing.assert_allclose((d**2).sum(axis=1), 1.0, rtol=1e-14)
        np.testing.assert_allclose(d.sum(axis=0), 0.0, atol=1e-12)  # antipodal symmetry
    r = models.uniform_rays(12)
    np.testing.assert_allclose(r[0], [1.0, 0.0, 0.0], atol=1e-15)
