"""GPU adjoint with respect to the inverse line widths
(mgx_feautrier_adjoint_width_device, mgx_image_adjoint_width_device) and the
(log n, T) image gradient built on it:

* d_iw against the CPU oracle's (pinned by finite differences and an exact
  identity in test_oracle_adjoint_width.py): images within 1e-9 relative, or 1e-10
  of the largest gradient (atomics add in another order); mixed precision 1e-4.
  J within 1e-6, or 1e-7 of the largest: the inverse widths act on J only through
  the optical depths, and in the far wings of the quadrature (depth dz chi_min,
  ~1e-15) the depth adjoints are ill-conditioned -- the rows' derivatives take
  differences of u that agree to ~15 digits -- so the device's two-sided
  elimination and the oracle's one-sided one round them differently (seeding one
  frequency at a time: 1e-9 at the line centre, up to 1% at the outermost bin,
  on contributions 1e-12 of the centre's); the optically thin velocity-gradient
  benchmark, thin at every frequency, at 2e-2.  The eps and kappa gradients carry
  the same depth terms, but their source-function terms dominate;
* d_eps and d_kap of the width variants equal those of the plain entry points;
* the image gradient with respect to (log n, T) through torch.autograd equals the
  oracle's adjoint pulled back through LineThermo on the CPU (the chain pinned by
  finite differences in test_line_thermo.py), within 1e-9.
"""
import numpy as np
import pytest

from magritte_amd import _abi
from test_gpu_adjoint import _close, _image_case, _seed

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,approx,iw_rtol", [("line_model", _abi.APPROX_CLOSE_LINES, 1e-6),
                                                  ("small_wind", _abi.APPROX_CLOSE_LINES, 1e-6),
                                                  ("small_wind", _abi.APPROX_NONE, 1e-6),
                                                  ("small_wind", _abi.APPROX_ONE_LINE, 1e-6),
                                                  ("vgrad_model", _abi.APPROX_CLOSE_LINES, 2e-2)])
def test_J_width_adjoint_matches_oracle(which, approx, iw_rtol, request):
    """The velocity-gradient benchmark is optically thin throughout (line-centre
    depth 1e-5 along a ray, segments ~1e-7): there every depth adjoint is
    conditioning-limited, to ~eps / dtau^2 relative in either implementation (the
    oracle's own finite differences agree with it to ~1e-5 only), so its inverse-
    width gradients are compared at 2e-2 (observed 6e-3)."""
    import torch

    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options(approx=approx)
    rng = np.random.default_rng(31)
    Jb = _seed(m, rng)
    print(f"[J_adjoint_width] model={which} approx={approx} P={m.npoints}", flush=True)
    dm = solver.DeviceModel(m)
    de, dk, dw = solver.feautrier_adjoint(dm, op, torch.from_numpy(Jb), want_iw=True)
    re, rk, rw = oracle.J_adjoint(m, op, Jb, want_iw=True)
    _close(de.cpu().numpy(), re, 1e-9, "d_eps")
    _close(dk.cpu().numpy(), rk, 1e-9, "d_kap")
    _close(dw.cpu().numpy(), rw, iw_rtol, "d_iw")
    de0, dk0 = solver.feautrier_adjoint(dm, op, torch.from_numpy(Jb))
    _close(de.cpu().numpy(), de0.cpu().numpy(), 1e-12, "d_eps vs the plain entry point")
    _close(dk.cpu().numpy(), dk0.cpu().numpy(), 1e-12, "d_kap vs the plain entry point")


@pytest.mark.parametrize("approx,precision,rtol", [(_abi.APPROX_CLOSE_LINES, _abi.PREC_FP64, 1e-9),
                                                   (_abi.APPROX_NONE, _abi.PREC_FP64, 1e-9),
                                                   (_abi.APPROX_CLOSE_LINES, _abi.PREC_MIXED, 1e-4)])
def test_image_width_adjoint_matches_oracle(small_wind, approx, precision, rtol):
    import torch

    from magritte_amd import solver
    from oracle import oracle

    fr, rd = _image_case(small_wind)
    nx = ny = 16
    rng = np.random.default_rng(32)
    Ib = rng.standard_normal((nx * ny, fr["nu"].shape[1]))
    print(f"[image_adjoint_width] approx={approx} precision={precision}", flush=True)
    dm = solver.DeviceModel(small_wind)
    de, dk, dw = solver.image_adjoint(dm, small_wind.options(approx=approx, precision=precision),
                                      rd, nx, ny, fr, torch.from_numpy(Ib), want_iw=True)
    re, rk, rw = oracle.image_adjoint(small_wind, small_wind.options(approx=approx), rd, nx, ny,
                                      Ib, fr, want_iw=True)
    _close(dw.cpu().numpy(), rw, rtol, "d_iw")
    _close(de.cpu().numpy(), re, rtol, "d_eps")
    _close(dk.cpu().numpy(), rk, rtol, "d_kap")


def test_image_log_density_temperature_gradient(small_wind):
    import torch

    from magritte_amd import autograd, solver
    from oracle import oracle

    m = small_wind
    op = m.options()
    fr, rd = _image_case(m)
    nx = ny = 16
    print(f"[image d/d(log n, T)] P={m.npoints}", flush=True)
    dm = solver.DeviceModel(m)
    th = autograd.LineThermo(m, dm.device)
    ln = th.log_density().clone().requires_grad_(True)
    T = th.temperature().clone().requires_grad_(True)
    I = autograd.image_log_density_temperature(ln, T, th, dm, op, rd, nx, ny, fr)
    rng = np.random.default_rng(33)
    Ib = rng.standard_normal(tuple(I.shape))
    (I * torch.from_numpy(Ib).to(I.device)).sum().backward()
    # the oracle's adjoint at the same line data, pulled back through the map on the CPU
    re, rk, rw = oracle.image_adjoint(m, op, rd, nx, ny, Ib, fr, want_iw=True)
    thc = autograd.LineThermo(m, "cpu")
    lnc = thc.log_density().clone().requires_grad_(True)
    Tc = thc.temperature().clone().requires_grad_(True)
    g_ln, g_T = torch.autograd.grad(thc(lnc, Tc), (lnc, Tc),
                                    (torch.from_numpy(re), torch.from_numpy(rk),
                                     torch.from_numpy(rw)))
    _close(ln.grad.cpu().numpy(), g_ln.numpy(), 1e-9, "d/d log n")
    _close(T.grad.cpu().numpy(), g_T.numpy(), 1e-9, "d/d T")


def test_J_width_autograd(small_wind):
    import torch

    from magritte_amd import autograd, solver
    from oracle import oracle

    m = small_wind
    op = m.options()
    dm = solver.DeviceModel(m)
    dev = dm.device
    eps = torch.from_numpy(m.emissivity).to(dev).requires_grad_(True)
    kap = torch.from_numpy(m.opacity).to(dev).requires_grad_(True)
    iw = torch.from_numpy(m.inverse_width).to(dev).requires_grad_(True)
    J = autograd.formal_solve_J_width(eps, kap, iw, dm, op)
    rng = np.random.default_rng(34)
    Jb = _seed(m, rng)
    (J * torch.from_numpy(Jb).to(dev)).sum().backward()
    re, rk, rw = oracle.J_adjoint(m, op, Jb, want_iw=True)
    _close(iw.grad.cpu().numpy(), rw, 1e-6, "d_iw (autograd)")
    _close(eps.grad.cpu().numpy(), re, 1e-9, "d_eps (autograd)")
