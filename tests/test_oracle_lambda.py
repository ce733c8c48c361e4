"""Pins of the oracle's general Lambda operator (update_Lambda with n_off_diag > 0,
solver.tpp:2319-2355, 1761-1850; Lambda::add_element, lambda.tpp):

- with no bands the list of every (point, line) is its one diagonal element,
  bit-equal to the diagonal solve;
- the band elements are entries of the inverse of the Feautrier system: on a 1D
  chain with one quadrature bin, no velocity and a varying line opacity, the
  ratio of an element to the diagonal one, times the ratio of the two points'
  opacities, is T^-1(c, n) / T^-1(c, c) of a 50-digit solve (the varying opacity
  makes the system unsymmetric, which separates the bands before and after the
  centre);
- the bands do not change u, J or the diagonal element.
"""
import numpy as np
import pytest

from magritte_amd import models
from oracle import oracle

import _exact


def _opts(m, noff):
    op = m.options()
    op.n_off_diag = noff
    return op


def test_no_bands_is_the_diagonal(small_wind):
    m = small_wind
    _, J0, lam = oracle.solve(m, m.options(), with_u=False)
    _, J, cnt, nr, val = oracle.solve_lambda(m, _opts(m, 0))
    np.testing.assert_array_equal(J, J0)
    assert (cnt == 1).all()
    np.testing.assert_array_equal(nr[:, :, 0], np.broadcast_to(np.arange(m.npoints)[:, None], cnt.shape))
    np.testing.assert_array_equal(val[:, :, 0], lam)


def test_bands_keep_u_J_and_the_diagonal(small_wind):
    m = small_wind
    u0, J0, lam = oracle.solve(m, m.options())
    u, J, cnt, nr, val = oracle.solve_lambda(m, _opts(m, 2), with_u=True)
    np.testing.assert_array_equal(u, u0)
    np.testing.assert_array_equal(J, J0)
    np.testing.assert_array_equal(nr[:, :, 0], np.broadcast_to(np.arange(m.npoints)[:, None], cnt.shape))
    rel = np.abs(val[:, :, 0] - lam) / np.abs(lam)
    assert rel.max() < 1e-12, rel.max()
    assert cnt.max() > 1
    # every list holds distinct points
    for p in range(0, m.npoints, 97):
        for l in range(m.nlines):
            k = int(cnt[p, l])
            assert len(set(nr[p, l, :k].tolist())) == k


def _chain(P=30, noff=3):
    m = models.model_line_1d(npoints=P, nquads=1, vmax=0.0)
    scale = 1.0 + 0.6 * np.sin(0.7 * np.arange(P))
    m.opacity[:, 0] *= scale
    m.emissivity[:, 0] *= scale
    return m, _opts(m, noff)


@pytest.mark.parametrize("o", [0, 1, 7, 15, 28, 29])
def test_bands_are_entries_of_the_inverse(o):
    m, op = _chain()
    _, _, cnt, nr, val = oracle.solve_lambda(m, op)
    sy = oracle.feautrier_system(m, m.options(), 0, o, 0)
    c = sy["centre"]
    row = _exact.inverse_row(sy, c)
    n_tot = sy["n"]
    # the chain's list entries are its points in order: entry j is point j (first = 0)
    assert n_tot == m.npoints
    k = int(cnt[o, 0])
    expect = {o} | {q for q in range(o - op.n_off_diag, o + op.n_off_diag + 1) if 0 <= q < m.npoints}
    assert set(nr[o, 0, :k].tolist()) == expect
    got = dict(zip(nr[o, 0, :k].tolist(), val[o, 0, :k].tolist()))
    kappa = m.opacity[:, 0]
    worst = 0.0
    for q, v in got.items():
        if q == o:
            continue
        ratio = v * kappa[q] / (got[o] * kappa[o])
        exact = float(row[q] / row[c])
        worst = max(worst, abs(ratio - exact) / abs(exact))
    assert worst < 1e-10, worst


def test_band_capacity_is_reported(small_wind):
    m = small_wind
    with pytest.raises(RuntimeError):
        oracle.solve_lambda(m, _opts(m, 2), cap=2)
