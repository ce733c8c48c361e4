"""GPU parity of the general Lambda operator (mgx_feautrier_forward_lambda_device)
against the oracle's restatement of update_Lambda with n_off_diag bands and
Lambda::add_element: the same lists (counts and emitting points, in the same
order) bit-exact, values within 1e-10 relative (floor 1e-12 of the list's
diagonal element), J within 1e-10; the diagonal element equals the diagonal
solve's Lambda."""
import numpy as np
import pytest

from magritte_amd import _abi

pytestmark = pytest.mark.gpu


def _opts(m, noff, approx=_abi.APPROX_CLOSE_LINES, precision=_abi.PREC_FP64):
    op = m.options(approx=approx, precision=precision)
    op.n_off_diag = noff
    return op


@pytest.mark.parametrize("which,noff,approx", [("line_model", 3, _abi.APPROX_CLOSE_LINES),
                                               ("vgrad_model", 2, _abi.APPROX_CLOSE_LINES),
                                               ("small_wind", 0, _abi.APPROX_CLOSE_LINES),
                                               ("small_wind", 1, _abi.APPROX_NONE),
                                               ("small_wind", 4, _abi.APPROX_CLOSE_LINES)])
def test_lambda_bands_match_oracle(which, noff, approx, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = _opts(m, noff, approx)
    print(f"[lambda] model={which} n_off_diag={noff} approx={approx}", flush=True)
    out = solver.feautrier_forward_lambda(solver.DeviceModel(m), op)
    cnt = out["count"].cpu().numpy().astype(np.uint32)
    nr = out["nr"].cpu().numpy().astype(np.uint32)
    val = out["val"].cpu().numpy()
    J = out["J"].cpu().numpy()
    _, Jr, cr, nrr, vr = oracle.solve_lambda(m, op, nthreads=16)
    np.testing.assert_array_equal(cnt, cr)
    P, L, cap = vr.shape
    mask = np.arange(cap)[None, None, :] < cr[:, :, None]
    np.testing.assert_array_equal(nr[mask], nrr[mask])
    scale = np.abs(vr[:, :, :1])
    err = np.abs(val - vr) / (np.abs(vr) + 1e-2 * scale + 1e-300)
    worst = float(err[mask].max())
    print(f"  lists equal, max count {int(cr.max())}, value worst {worst:.3e}", flush=True)
    assert worst < 1e-10
    assert float(np.max(np.abs(J - Jr) / np.abs(Jr))) < 1e-10
    if not m.spherical_symmetry:
        # in spherical symmetry a ray revisits shells, and a band row at the origin's own
        # shell merges into the diagonal element (as add_element does in the reference)
        _, _, lam = oracle.solve(m, m.options(approx=approx), with_u=False)
        d = np.abs(val[:, :, 0] - lam) / np.abs(lam)
        assert d.max() < 1e-10


def test_lambda_bands_mixed_precision(small_wind):
    from magritte_amd import solver
    from oracle import oracle

    m = small_wind
    out = solver.feautrier_forward_lambda(solver.DeviceModel(m), _opts(m, 2, precision=_abi.PREC_MIXED))
    _, _, cr, nrr, vr = oracle.solve_lambda(m, _opts(m, 2), nthreads=16)
    np.testing.assert_array_equal(out["count"].cpu().numpy().astype(np.uint32), cr)
    val = out["val"].cpu().numpy()
    mask = np.arange(vr.shape[2])[None, None, :] < cr[:, :, None]
    err = np.abs(val - vr) / (np.abs(vr) + 1e-2 * np.abs(vr[:, :, :1]) + 1e-300)
    print(f"  mixed value worst {float(err[mask].max()):.3e}", flush=True)
    assert float(err[mask].max()) < 1e-5


def test_lambda_bands_arguments(small_wind):
    from magritte_amd import solver

    m = small_wind
    dm = solver.DeviceModel(m)
    with pytest.raises(_abi.MgxError, match="n_off_diag"):
        solver.feautrier_forward_lambda(dm, _opts(m, _abi.MAX_OFF_DIAG + 1))
    with pytest.raises(_abi.MgxError, match="capacity"):
        solver.feautrier_forward_lambda(dm, _opts(m, 2), cap=2)
