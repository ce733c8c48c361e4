"""Adaptive ray directions on the GPU: the 12^3 wind with per-point rotated
HEALPix directions against the oracle (traced counts bit-exact, u, J, Lambda
within 1e-10 in FP64, J within 1e-5 mixed), and the two invariants of
tests/test_oracle_adaptive.py on the device (identical directions reproduce the
global-ray solve; J, u, Lambda at an origin equal those of the global model
rotated by that origin's rotation), bit for bit."""
import copy

import numpy as np
import pytest

from magritte_amd import _abi, models
from oracle import oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _adaptive(m, seed):
    d = m.ray_direction.reshape(-1, 3)[:m.nrays].copy()
    w = m.ray_weight[:m.nrays].copy()
    D = models.rotated_directions(models.random_rotations(m.npoints, seed), d)
    return copy.deepcopy(m).set_adaptive_directions(D, np.broadcast_to(w, (m.npoints,) + w.shape)), D


def _gpu(m, op):
    from magritte_amd import solver

    out = solver.feautrier_forward(solver.DeviceModel(m), op)
    return tuple(out[k].cpu().numpy() for k in ("u", "J", "lambda_diag"))


def test_adaptive_directions_match_oracle(small_wind):
    from magritte_amd import solver

    ma, _ = _adaptive(small_wind, 7)
    print(f"[adaptive] P={ma.npoints} nrays={ma.nrays} per-point directions", flush=True)
    op = ma.options()
    n_gpu = solver.trace_counts(solver.DeviceModel(ma), op)
    np.testing.assert_array_equal(n_gpu, oracle.trace_counts(ma, op))
    u, J, lam = _gpu(ma, op)
    ur, Jr, lr = oracle.solve(ma, op)
    for name, a, b in (("u", u, ur), ("J", J, Jr), ("lambda", lam, lr)):
        e = _rel(a, b)
        print(f"  {name} {e:.2e}", flush=True)
        assert e < 1e-10, name
    _, Jm, _ = _gpu(ma, ma.options(precision=_abi.PREC_MIXED))
    assert _rel(Jm, Jr) < 1e-5


def test_adaptive_invariants_on_the_device(small_wind):
    m = small_wind
    d = m.ray_direction.reshape(-1, 3)[:m.nrays].copy()
    w = m.ray_weight[:m.nrays].copy()
    op = m.options()
    ref = _gpu(m, op)
    same = copy.deepcopy(m).set_adaptive_directions(np.broadcast_to(d, (m.npoints,) + d.shape),
                                                    np.broadcast_to(w, (m.npoints,) + w.shape))
    for a, b in zip(_gpu(same, op), ref):
        np.testing.assert_array_equal(a, b)
    ma, D = _adaptive(m, 11)
    u, J, lam = _gpu(ma, op)
    for o in (3, m.npoints // 2, m.npoints - 1):
        mo = copy.deepcopy(m)
        mo.ray_direction = np.ascontiguousarray(D[o])
        uo, Jo, lo = _gpu(mo, op)
        np.testing.assert_array_equal(J[o], Jo[o])
        np.testing.assert_array_equal(u[:, o], uo[:, o])
        np.testing.assert_array_equal(lam[o], lo[o])
