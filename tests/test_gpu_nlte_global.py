"""GPU global statistical equilibrium (mgx_nlte_iteration_stateq_global,
mgx_compute_level_populations) against the oracle's restatement of
Model::compute_Jeff, update_using_statistical_equilibrium (a dense long-double
solve of all points' rate equations) and Model::compute_level_populations:

* one update from the same J and Lambda lists: J_lin, J_eff, populations and the
  line data within 1e-10 relative (the floors of test_gpu_nlte), without bands
  (block diagonal: the block-Jacobi sweeps stop after the second) and with bands;
* the non-sparse loop with bands: the same iteration count as the oracle's loop
  and populations within 1e-8.
"""
import numpy as np
import pytest

from magritte_amd import _abi, models
from test_gpu_nlte import _check_state, _close, _pair, _say

pytestmark = pytest.mark.gpu


def _wind6():
    return models.model_wind(n_edge=6, nrays=12, seed=models.seed_for(1, 7), nlev=5)


def _opts(m, noff):
    op = m.options()
    op.n_off_diag = noff
    return op


@pytest.mark.parametrize("which,noff", [("line", 2), ("wind6", 0), ("wind6", 2)])
def test_global_stateq_matches_oracle(which, noff):
    import torch

    from oracle import oracle

    m = models.model_line_1d() if which == "line" else _wind6()
    prm = _abi.default_nlte_params()
    _say(f"[stateq_global] model={which} n_off_diag={noff} P={m.npoints}")
    dm, lp, nl = _pair(m, prm)
    _, J, cnt, nr, val = oracle.solve_lambda(m, _opts(m, noff), nthreads=16)
    dev = dm.device
    lam = {"count": torch.from_numpy(cnt.view(np.int32)).to(dev),
           "nr": torch.from_numpy(nr.view(np.int32)).to(dev),
           "val": torch.from_numpy(val).to(dev)}
    conv, sweeps, Jl, Je = lp.stateq_global(torch.from_numpy(J).to(dev), lam, want_J=True)
    nl.compute_Jeff_general(J, (cnt, nr, val))
    nl.stateq_global((cnt, nr, val))
    _say(f"  block-Jacobi sweeps {sweeps}")
    if noff == 0:
        assert sweeps == 2
    else:
        assert 2 <= sweeps <= 20
    ref = nl.populations()
    _close(Jl.cpu().numpy(), ref["J_lin"], 1e-12, "J_lin")
    _close(Je.cpu().numpy(), ref["J_eff"], 1e-10, "J_eff")
    _check_state(m, dm, lp, nl, "update")
    rc = nl.stats()[0]
    for s, c in enumerate(conv):
        np.testing.assert_allclose(c[0], rc[s, 0], rtol=0, atol=2.0 / m.npoints)
        np.testing.assert_allclose(c[1:], rc[s, 1:], rtol=1e-6, atol=1e-15)


def test_nonsparse_loop_matches_oracle():
    from magritte_amd import solver
    from oracle import oracle

    m = models.model_line_1d()
    m_ref = models.model_line_1d()
    prm = _abi.default_nlte_params()
    _say(f"[loop global] P={m.npoints}")
    dm = solver.DeviceModel(m)
    lp = solver.LevelPopulations(dm, prm)
    lp.set_lte()
    it, em, ex = solver.compute_level_populations(dm, _opts(m, 2), lp, True, 60)
    nl, it_r, em_r, ex_r = oracle.compute_level_populations(m_ref, _opts(m_ref, 2), prm, True, 60,
                                                            nthreads=16)
    _say(f"  iterations gpu {it} oracle {it_r}; last max change {ex[-1]:.3e} / {ex_r[-1]:.3e}")
    assert it == it_r
    pop, tot = lp.populations()
    ref = nl.populations()["population"]
    ptot = np.repeat(tot.cpu().numpy()[0], lp.nlev[0])
    _close(pop.cpu().numpy(), ref, 1e-8, "populations", floor=1e-14 * ptot)


def test_global_rejects_too_many_bands():
    from magritte_amd import solver

    m = models.model_line_1d()
    dm = solver.DeviceModel(m)
    lp = solver.LevelPopulations(dm, _abi.default_nlte_params())
    lp.set_lte()
    with pytest.raises(_abi.MgxError, match="n_off_diag"):
        solver.compute_level_populations(dm, _opts(m, _abi.MAX_OFF_DIAG + 1), lp, True, 5)
