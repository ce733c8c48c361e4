"""GPU parity on the models that pin the oracle's multi-line and tie-rule paths:
- the blended-line slab, the same slab with a one-width CloseLines window, and
  the blended velocity-gradient benchmark (tests/test_oracle_blended.py), under
  None, OneLine and CloseLines: traced counts bit-exact, u, J and Lambda within
  1e-10 of the oracle in FP64 (mixed: J within 1e-5);
- the lattice whose exact ties decide the walk (tests/test_oracle_walk_lattice.py):
  every traced list of every (ray pair, origin) equal to the integer brute force."""
import numpy as np
import pytest

from magritte_amd import _abi
from oracle import oracle
from test_oracle_blended import blended_slab, blended_velocity_gradient
from test_oracle_walk_lattice import expected_list, lattice_model

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _model(which):
    if which == "slab":
        return blended_slab()[0]
    if which == "window":
        return blended_slab(window=1.0)[0]
    return blended_velocity_gradient()[0]


@pytest.mark.parametrize("approx", [_abi.APPROX_NONE, _abi.APPROX_ONE_LINE, _abi.APPROX_CLOSE_LINES])
@pytest.mark.parametrize("which", ["slab", "window", "vgrad"])
def test_blended_lines_match_oracle(which, approx):
    from magritte_amd import solver

    m = _model(which)
    print(f"[blended] {which} approx={approx} P={m.npoints} L={m.nlines} F={m.nfreqs}", flush=True)
    dm = solver.DeviceModel(m)
    op = m.options(approx=approx)
    np.testing.assert_array_equal(solver.trace_counts(dm, op), oracle.trace_counts(m, op))
    out = solver.feautrier_forward(dm, op)
    ref = oracle.solve(m, op)
    for name, a, b in zip(("u", "J", "lambda"), (out["u"], out["J"], out["lambda_diag"]), ref):
        e = _rel(a.cpu().numpy(), b)
        print(f"  {name} {e:.2e}", flush=True)
        assert e < 1e-10, name
    Jm = solver.feautrier_forward(dm, m.options(approx=approx, precision=_abi.PREC_MIXED))["J"]
    assert _rel(Jm.cpu().numpy(), ref[1]) < 1e-5


def test_lattice_walk_matches_brute_force():
    from magritte_amd import solver

    m, idx, nbs, bdy = lattice_model()
    print(f"[lattice] P={m.npoints} rays={m.nrays}", flush=True)
    sel = np.array([(rr, o, 0) for rr in range(m.hnrays) for o in range(m.npoints)], dtype=np.uint32)
    got = solver.trace_rays(solver.DeviceModel(m), m.options(), sel, capacity=64)
    for (rr, o, _), t in zip(sel, got):
        exp = expected_list(m, idx, nbs, bdy, int(rr), int(o))
        assert list(t["nr"]) == exp, (rr, o)
        assert t["centre"] == exp.index(int(o))
