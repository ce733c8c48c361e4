"""GPU order-0 short characteristics (mgx_shortchar_order_0_device) against the
CPU oracle's restatement of Solver::solve_shortchar_order_0 on the same models
(None, OneLine and CloseLines; the 1D chain, a 12^3 wind and the
velocity-gradient model whose segments all take the line-integrated branch):
I of both directions, u, J and the Lambda diagonal within 1e-10 relative in
FP64 (v = (I_f - I_b)/2 within 1e-10 of u, as it cancels), and J within 1e-5
(above the reference's own rounding noise) in mixed precision."""
import numpy as np
import pytest

from magritte_amd import _abi

pytestmark = pytest.mark.gpu


def _close(got, ref, rtol, what, floor=None):
    """|got - ref| <= rtol |ref| + floor; floor defaults to 1e-14 of the largest |ref|"""
    got = np.asarray(got)
    ref = np.asarray(ref)
    err = np.abs(got - ref)
    if floor is None:
        floor = 1e-14 * np.abs(ref).max()
    bad = err > rtol * np.abs(ref) + floor
    worst = float(np.max(err / (np.abs(ref) + floor / rtol + 1e-300)))
    print(f"  {what}: worst {worst:.3e}", flush=True)
    assert not bad.any(), f"{what}: {int(bad.sum())} off, worst {worst:.3e}"


@pytest.mark.parametrize("which,approx", [("line_model", _abi.APPROX_CLOSE_LINES),
                                          ("small_wind", _abi.APPROX_CLOSE_LINES),
                                          ("small_wind", _abi.APPROX_NONE),
                                          ("small_wind", _abi.APPROX_ONE_LINE),
                                          ("vgrad_model", _abi.APPROX_CLOSE_LINES)])
def test_shortchar_matches_oracle(which, approx, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options(approx=approx)
    print(f"[shortchar] model={which} approx={approx} P={m.npoints}", flush=True)
    out = solver.shortchar_forward(solver.DeviceModel(m), op)
    I, u, v, J, lam = oracle.shortchar(m, op, nthreads=16)
    _close(out["I"].cpu().numpy(), I, 1e-10, "I")
    _close(out["u"].cpu().numpy(), u, 1e-10, "u")
    # v = (I_f - I_b)/2 cancels where both directions see the same intensity:
    # its error is measured against u, the size of the two terms
    _close(out["v"].cpu().numpy(), v, 1e-10, "v",
           floor=1e-10 * np.abs(u) + 1e-14 * np.abs(u).max())
    _close(out["J"].cpu().numpy(), J, 1e-10, "J")
    _close(out["lambda_diag"].cpu().numpy(), lam, 1e-10, "Lambda")


def test_shortchar_mixed_precision(small_wind):
    """Mixed precision against the FP64 oracle: 1e-5 relative, with a floor of 1e-6
    of the largest J.  The floor is the reference's own noise: its recurrence
    adds expm1f(-dt)/dt (S_c - S_n (1 + dt)) + S_c - S_n, where the float expm1f
    leaves ~6e-8 S per segment; in the optically thin far wings of the
    high-frequency lines S ~ 1e-13 while I ~ 1e-22 (the CMB there), so I and J
    are rounding noise of that size (even negative) in both modes."""
    from magritte_amd import solver
    from oracle import oracle

    out = solver.shortchar_forward(solver.DeviceModel(small_wind),
                                   small_wind.options(precision=_abi.PREC_MIXED))
    _, _, _, J, _ = oracle.shortchar(small_wind, small_wind.options(), nthreads=16)
    _close(out["J"].cpu().numpy(), J, 1e-5, "J (mixed)", floor=1e-6 * np.abs(J).max())


def test_shortchar_rejects_nonpositive_min_dtau(line_model):
    from magritte_amd import solver

    with pytest.raises(_abi.MgxError, match="min_dtau"):
        solver.shortchar_forward(solver.DeviceModel(line_model), line_model.options(), min_dtau=0.0)


@pytest.mark.parametrize("which,approx", [("line_model", _abi.APPROX_CLOSE_LINES),
                                          ("small_wind", _abi.APPROX_CLOSE_LINES),
                                          ("small_wind", _abi.APPROX_ONE_LINE),
                                          ("vgrad_model", _abi.APPROX_CLOSE_LINES)])
def test_feautrier_uv_matches_oracle(which, approx, request):
    """mgx_feautrier_order_2_uv_device against the oracle's solve_feautrier_order_2_uv:
    u within 1e-10; v within 1e-10 of itself plus 1e-12 of the sum of the absolute
    terms it is formed from (coeffmin u_c-1, coeffplus u_c+1, coeffzero u_c).  v is
    a second-order difference of u and cancels: in a uniform stretch of the 1D chain
    the terms exceed |v| by 1e12 and more, so v there is rounding noise in the
    reference as well (the two-sided elimination reaches u one row either side of
    the centre through other roundings than the reference's back substitution).
    On the wind, the depth after the centre reuses the opacity of the row before
    it when both segments take the trapezoid branch, as the reference's does."""
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options(approx=approx)
    print(f"[uv] model={which} approx={approx} P={m.npoints}", flush=True)
    out = solver.feautrier_uv(solver.DeviceModel(m), op)
    u, v, v_scale = oracle.feautrier_uv(m, op, nthreads=16, with_scale=True)
    _close(out["u"].cpu().numpy(), u, 1e-10, "u")
    _close(out["v"].cpu().numpy(), v, 1e-10, "v", floor=1e-12 * v_scale)
