"""GPU parity of the formal solver against the CPU oracle (restatement of the
reference's plain method, oracle/mgx_oracle.c) on the same seeded inputs.

Bars: traced point sequences, counts, shifts, interpolation factors and dZ
bit-exact; u, J and the Lambda diagonal within 1e-10 relative (FP64).
"""
import numpy as np
import pytest

from magritte_amd import _abi

pytestmark = pytest.mark.gpu

RTOL = 1e-10  # FP64 parity bar for u, J and Lambda


def _dev(model):
    from magritte_amd.solver import DeviceModel

    return DeviceModel(model)


def _rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den))


@pytest.mark.parametrize("which", ["line_model", "vgrad_model", "small_wind", "wind32"])
def test_trace_counts_bit_exact(which, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options()
    print(f"[trace_counts] model={which} P={m.npoints}", flush=True)
    got = solver.trace_counts(_dev(m), op)
    ref = oracle.trace_counts(m, op)
    assert got.shape == ref.shape
    bad = np.argwhere(got != ref)
    assert len(bad) == 0, f"{len(bad)} mismatching counts, first {bad[:5]} got {got[tuple(bad[0])]} ref {ref[tuple(bad[0])]}"


@pytest.mark.parametrize("which", ["line_model", "vgrad_model", "small_wind", "wind32"])
def test_trace_rays_bit_exact(which, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options()
    rng = np.random.default_rng(5)
    n = 64
    sel = np.stack([rng.integers(0, m.hnrays, n), rng.integers(0, m.npoints, n),
                    rng.integers(0, m.nlines, n)], axis=1)
    # include boundary origins
    sel[0, 1] = m.boundary2point[0]
    sel[1, 1] = m.boundary2point[-1]
    print(f"[trace_rays] model={which} n={n}", flush=True)
    got = solver.trace_rays(_dev(m), op, sel)
    for i, (rr, o, l) in enumerate(sel):
        ref = oracle.trace_ray(m, op, int(rr), int(o), int(l))
        g = got[i]
        assert g["count"] == ref["count"] and g["centre"] == ref["centre"], (i, rr, o, l)
        for k in ("nr", "nr_interp"):
            np.testing.assert_array_equal(g[k], ref[k], err_msg=f"{k} ray {i}")
        for k in ("shift", "interp_factor", "dZ"):
            np.testing.assert_array_equal(g[k].view(np.uint64), ref[k].view(np.uint64),
                                          err_msg=f"{k} ray {i}")


@pytest.mark.parametrize("which,approx", [
    ("line_model", _abi.APPROX_CLOSE_LINES),
    ("line_model", _abi.APPROX_NONE),
    ("line_model", _abi.APPROX_ONE_LINE),
    ("slab_model", _abi.APPROX_CLOSE_LINES),
    ("vgrad_model", _abi.APPROX_CLOSE_LINES),
    ("vgrad_model", _abi.APPROX_NONE),
    ("small_wind", _abi.APPROX_CLOSE_LINES),
    ("small_wind", _abi.APPROX_NONE),
    ("small_wind", _abi.APPROX_ONE_LINE),
    ("wind32", _abi.APPROX_CLOSE_LINES),
])
def test_forward_fp64_parity(which, approx, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options(approx=approx)
    print(f"[forward] model={which} approx={approx}", flush=True)
    out = solver.feautrier_forward(_dev(m), op)
    u = out["u"].cpu().numpy()
    J = out["J"].cpu().numpy()
    lam = out["lambda_diag"].cpu().numpy()
    ur, Jr, lr = oracle.solve(m, op)
    assert np.isfinite(u).all() and np.isfinite(J).all() and np.isfinite(lam).all()
    eu, eJ, eL = _rel_err(u, ur), _rel_err(J, Jr), _rel_err(lam, lr)
    print(f"  max rel err u={eu:.3e} J={eJ:.3e} lambda={eL:.3e}", flush=True)
    assert eu < RTOL and eJ < RTOL and eL < RTOL


# Mixed precision (single-precision profile, erf and log-interpolation
# transcendentals; FP64 walk, interpolation counts and elimination) against the
# FP64 oracle.  Bars: J and Lambda within 1e-5 relative (the metric's J bar),
# u within 1e-4 relative.
MIXED_J_RTOL = 1e-5
MIXED_U_RTOL = 1e-4


@pytest.mark.parametrize("which", ["line_model", "vgrad_model", "small_wind", "wind32"])
def test_forward_mixed_accuracy(which, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options(precision=_abi.PREC_MIXED)
    print(f"[forward mixed] model={which}", flush=True)
    out = solver.feautrier_forward(_dev(m), op)
    u = out["u"].cpu().numpy()
    J = out["J"].cpu().numpy()
    lam = out["lambda_diag"].cpu().numpy()
    ur, Jr, lr = oracle.solve(m, m.options())
    eu, eJ, eL = _rel_err(u, ur), _rel_err(J, Jr), _rel_err(lam, lr)
    print(f"  mixed max rel err u={eu:.3e} J={eJ:.3e} lambda={eL:.3e}", flush=True)
    assert eJ < MIXED_J_RTOL and eL < MIXED_J_RTOL and eu < MIXED_U_RTOL


def test_mixed_mode_keeps_traces(small_wind):
    from magritte_amd import solver

    m = small_wind
    dm = _dev(m)
    a = solver.trace_counts(dm, m.options(precision=_abi.PREC_FP64))
    b = solver.trace_counts(dm, m.options(precision=_abi.PREC_MIXED))
    np.testing.assert_array_equal(a, b)


def _image_case(m, nfreq=31):
    lo = float(m.line[0]) * (1 - 2e-4)
    hi = float(m.line[0]) * (1 + 2e-4)
    return m.image_frequencies(lo, hi, nfreq)


@pytest.mark.parametrize("which,raydir,nx,ny", [
    ("vgrad_model", (0.0, 1.0, 0.0), 12, 10),
    ("imager_model", (0.0, 1.0, 0.0), 50, 50),
    ("small_wind", (0.3, -0.5, 0.8), 16, 12),
    ("small_wind", (0.0, 0.0, 1.0), 9, 9),
    ("wind32", (0.2, 0.7, -0.4), 24, 24),
])
def test_image_fp64_parity(which, raydir, nx, ny, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options()
    fq = _image_case(m)
    print(f"[image] model={which} dir={raydir} {nx}x{ny}", flush=True)
    I, X, Y = solver.image_new(_dev(m), op, raydir, nx, ny, fq)
    I, X, Y = I.cpu().numpy(), X.cpu().numpy(), Y.cpu().numpy()
    Ir, Xr, Yr = oracle.image_new(m, op, raydir, nx, ny, fq)
    np.testing.assert_array_equal(X.view(np.uint64), Xr.view(np.uint64))
    np.testing.assert_array_equal(Y.view(np.uint64), Yr.view(np.uint64))
    assert np.isfinite(I).all()
    scale = np.max(np.abs(Ir))
    err = float(np.max(np.abs(I - Ir) / np.maximum(np.abs(Ir), 1e-6 * scale)))
    print(f"  image max rel err {err:.3e}", flush=True)
    assert err < RTOL
