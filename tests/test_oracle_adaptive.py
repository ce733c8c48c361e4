"""Adaptive ray directions (Rays::use_adaptive_directions; the adaptive
template of Solver::solve_feautrier_order_2, solver.tpp:459-540), pinned by two
invariants of the oracle that need no restated formula:
- every point given the same directions reproduces the global-ray solve bit for
  bit (the per-point indexing reads the same values);
- a ray from origin o uses o's directions for its whole walk, its shifts and its
  weight, so with per-point rotated directions, J, u and Lambda at o equal those
  of a global-ray model whose rays are rotated by o's rotation -- bit for bit --
  while differing from the unrotated model by tens of percent."""
import copy

import numpy as np
import pytest

from magritte_amd import models
from oracle import oracle


@pytest.fixture(scope="module")
def wind8():
    return models.model_wind(n_edge=8, seed=models.seed_for(1, 7))


def _rays(m):
    return m.ray_direction.reshape(-1, 3)[:m.nrays].copy(), m.ray_weight[:m.nrays].copy()


def test_identical_directions_reproduce_global_rays(wind8):
    m = wind8
    d, w = _rays(m)
    ma = copy.deepcopy(m).set_adaptive_directions(np.broadcast_to(d, (m.npoints,) + d.shape),
                                                  np.broadcast_to(w, (m.npoints,) + w.shape))
    op = m.options()
    for a, b in zip(oracle.solve(m, op), oracle.solve(ma, op)):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(oracle.trace_counts(m, op), oracle.trace_counts(ma, op))


def test_origin_uses_its_own_directions(wind8):
    m = wind8
    d, w = _rays(m)
    P = m.npoints
    D = models.rotated_directions(models.random_rotations(P, 123), d)
    ma = copy.deepcopy(m).set_adaptive_directions(D, np.broadcast_to(w, (P,) + w.shape))
    op = m.options()
    u, J, lam = oracle.solve(ma, op)
    _, J0, _ = oracle.solve(m, op)
    for o in (0, 100, 257, P - 1):
        mo = copy.deepcopy(m)
        mo.ray_direction = np.ascontiguousarray(D[o])
        uo, Jo, lo = oracle.solve(mo, op)
        np.testing.assert_array_equal(J[o], Jo[o])
        np.testing.assert_array_equal(u[:, o], uo[:, o])
        np.testing.assert_array_equal(lam[o], lo[o])
        assert np.max(np.abs(J[o] - J0[o]) / J0[o]) > 1e-2
