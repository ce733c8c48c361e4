"""Pins of the oracle's global statistical equilibrium (Model::compute_Jeff,
LineProducingSpecies::update_using_statistical_equilibrium and
Model::compute_level_populations; model.cpp:514-545, 653-744,
lineProducingSpecies.tpp:377-581):

- without bands (n_off_diag = 0) every list holds only its point, the global
  system is block diagonal and the update equals the per-point (sparse) update,
  itself pinned by closed forms and detailed balance in test_oracle_nlte;
- with bands the update is the solution of the assembled global system (checked
  against an independent double-precision solve of the exported matrix), and the
  off-point coupling is weak: the block-Jacobi iteration matrix of these models
  has spectral radius below 1e-2, which the device's sweeps rely on;
- the non-sparse loop converges like the sparse one.
"""
import numpy as np
import pytest

from magritte_amd import _abi, models
from oracle import oracle


def _wind():
    return models.model_wind(n_edge=6, nrays=12, seed=models.seed_for(1, 7), nlev=5)


def _opts(m, noff):
    op = m.options()
    op.n_off_diag = noff
    return op


@pytest.mark.parametrize("which", ["line", "wind"])
def test_global_without_bands_equals_sparse(which):
    m = models.model_line_1d() if which == "line" else _wind()
    prm = _abi.default_nlte_params()
    a = oracle.NLTE(m, prm)
    a.set_lte()
    b = oracle.NLTE(m, prm)
    b.set_lte()
    # both formal solves first: an update rewrites the model's line data in place
    u, _, lam = oracle.solve(m, m.options())
    _, J, cnt, nr, val = oracle.solve_lambda(m, _opts(m, 0))
    a.compute_Jeff(u, lam)
    a.stateq(lam)
    b.compute_Jeff_general(J, (cnt, nr, val))
    b.stateq_global((cnt, nr, val))
    pa, pb = a.populations(), b.populations()
    np.testing.assert_allclose(pb["J_eff"], pa["J_eff"], rtol=1e-13, atol=0)
    tot = np.repeat(m.abundance[:, 0], a.nlev[0])
    # one dense long-double QR of all points against one per point: the global pivoting
    # mixes blocks whose rates differ by orders of magnitude (observed 3e-12)
    err = np.abs(pb["population"] - pa["population"]) / (np.abs(pa["population"]) + 1e-15 * tot)
    assert err.max() < 1e-10, err.max()
    np.testing.assert_allclose(b.stats()[0], a.stats()[0], rtol=1e-9, atol=1e-15)


def test_global_update_solves_its_system():
    m = _wind()
    prm = _abi.default_nlte_params()
    nl = oracle.NLTE(m, prm)
    nl.set_lte()
    _, J, cnt, nr, val = oracle.solve_lambda(m, _opts(m, 2))
    assert cnt.max() > 1
    nl.compute_Jeff_general(J, (cnt, nr, val))
    M, y = nl.global_system((cnt, nr, val), 0)
    x = np.linalg.solve(M, y)
    nl.stateq_global((cnt, nr, val))
    pop = nl.populations()["population"]
    assert (x > 0).all()  # no negative-population correction on this model
    tot = np.repeat(m.abundance[:, 0], nl.nlev[0])
    err = np.abs(pop - x) / (np.abs(x) + 1e-14 * tot)
    assert err.max() < 1e-9, err.max()
    # the point blocks alone do not reproduce it: the bands couple the points
    n = nl.nlev[0]
    xb = np.concatenate([np.linalg.solve(M[p * n:(p + 1) * n, p * n:(p + 1) * n], y[p * n:(p + 1) * n])
                         for p in range(m.npoints)])
    assert np.max(np.abs(xb - x) / (np.abs(x) + 1e-14 * tot)) > 1e-6


@pytest.mark.parametrize("noff", [1, 2])
def test_block_jacobi_coupling_is_weak(noff):
    m = _wind()
    nl = oracle.NLTE(m, _abi.default_nlte_params())
    nl.set_lte()
    _, J, cnt, nr, val = oracle.solve_lambda(m, _opts(m, noff))
    nl.compute_Jeff_general(J, (cnt, nr, val))
    M, _ = nl.global_system((cnt, nr, val), 0)
    n = nl.nlev[0]
    D = np.zeros_like(M)
    for p in range(m.npoints):
        D[p * n:(p + 1) * n, p * n:(p + 1) * n] = M[p * n:(p + 1) * n, p * n:(p + 1) * n]
    rho = np.abs(np.linalg.eigvals(np.linalg.solve(D, M - D))).max()
    assert 0 < rho < 1e-2, rho


def test_nonsparse_loop_converges_like_the_sparse_one():
    m = models.model_line_1d()
    prm = _abi.default_nlte_params(pop_prec=1e-6)
    _, it0, _, _ = oracle.compute_level_populations_sparse(m, m.options(), prm, True, 60)
    m = models.model_line_1d()
    nl, it1, em, ex = oracle.compute_level_populations(m, _opts(m, 2), prm, True, 60)
    assert it1 < 60 and it1 <= it0 + 1, (it0, it1)
    assert ex[-1] < 1e-5


def test_reference_band_loop_diverges_on_a_clumpy_wind():
    """A property of the reference's method, kept in the restatement: with bands, the
    global rate equations add each off-point element's loss term to the diagonal of the
    emitting point's row (lineProducingSpecies.tpp:455-458).  On a clumpy rotating wind the
    loop then diverges: by the third iteration J_eff is negative by orders of magnitude and
    the largest relative change stays at its bound of 2 (populations flip sign and are
    corrected).  The device's block-Jacobi sweeps stop with an error there."""
    clumps = models.grf_clumps(5, sigma=0.5, corr_length=1.0 / 20.0)
    m = models.model_wind(n_edge=5, nrays=48, nlev=10, v_rot=5.0e3, clumps=clumps)
    nl = oracle.NLTE(m, _abi.default_nlte_params(pop_prec=1e-4))
    nl.set_lte()
    worst = 0.0
    for it in range(3):
        _, J, cnt, nr, val = oracle.solve_lambda(m, _opts(m, 2))
        nl.compute_Jeff_general(J, (cnt, nr, val))
        pj = nl.populations()
        worst = min(worst, float(np.min(pj["J_eff"] / pj["J_lin"])))
        nl.stateq_global((cnt, nr, val))
    assert worst < -1e3, worst
    assert nl.stats()[0][0, 2] > 1.0
