"""GPU adjoint (reverse mode) of the formal solve and of the imager.

* against the CPU oracle's adjoint (oracle/mgx_oracle.c, pinned by finite
  differences in test_oracle_adjoint.py) on the same seeded models: FP64
  gradients within 1e-9 relative, or 1e-10 of the largest gradient for the
  entries that are sums of contributions of both signs (the GPU adds them with
  atomics in another order, through the two-sided elimination's factors);
* against central finite differences of the GPU forward itself (step 1e-6
  relative, 64 random (p, l)) through torch.autograd, at 1e-4 -- on a
  configuration without interpolated entries, where the forward is a smooth
  function of the parameters (the float powf of interpolate_log would
  otherwise put rounding steps into the difference quotients);
* mixed precision (FP32 transcendentals, FP64 elimination) against the FP64
  oracle gradient at 1e-4.
"""
import numpy as np
import pytest

from magritte_amd import _abi

pytestmark = pytest.mark.gpu


def _close(got, ref, rtol, what):
    got = np.asarray(got)
    ref = np.asarray(ref)
    scale = np.abs(ref).max()
    assert scale > 0, what
    err = np.abs(got - ref)
    floor = 0.1 * rtol * scale
    bad = err > rtol * np.abs(ref) + floor
    frac = float(np.max(err / (np.abs(ref) + floor / rtol)))
    print(f"  {what}: worst {frac:.3e} (relative, floor {0.1 * rtol:.0e} of the largest)", flush=True)
    assert not bad.any(), f"{what}: {int(bad.sum())} entries off, worst {frac:.3e}"
    return frac


def _seed(m, rng):
    return rng.standard_normal((m.npoints, m.nfreqs))


@pytest.mark.parametrize("which,approx", [("line_model", _abi.APPROX_CLOSE_LINES),
                                          ("small_wind", _abi.APPROX_CLOSE_LINES),
                                          ("small_wind", _abi.APPROX_NONE),
                                          ("small_wind", _abi.APPROX_ONE_LINE),
                                          ("vgrad_model", _abi.APPROX_CLOSE_LINES)])
def test_J_adjoint_matches_oracle(which, approx, request):
    import torch

    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options(approx=approx)
    rng = np.random.default_rng(11)
    Jb = _seed(m, rng)
    print(f"[J_adjoint] model={which} approx={approx} P={m.npoints}", flush=True)
    dm = solver.DeviceModel(m)
    de, dk = solver.feautrier_adjoint(dm, op, torch.from_numpy(Jb))
    re, rk = oracle.J_adjoint(m, op, Jb)
    _close(de.cpu().numpy(), re, 1e-9, "d_eps")
    _close(dk.cpu().numpy(), rk, 1e-9, "d_kap")


def test_J_adjoint_wind32_matches_oracle(wind32):
    import torch

    from magritte_amd import solver
    from oracle import oracle

    op = wind32.options()
    rng = np.random.default_rng(12)
    Jb = _seed(wind32, rng)
    print(f"[J_adjoint] model=wind32 P={wind32.npoints}", flush=True)
    dm = solver.DeviceModel(wind32)
    de, dk = solver.feautrier_adjoint(dm, op, torch.from_numpy(Jb))
    re, rk = oracle.J_adjoint(wind32, op, Jb, nthreads=16)
    _close(de.cpu().numpy(), re, 1e-9, "d_eps")
    _close(dk.cpu().numpy(), rk, 1e-9, "d_kap")


def test_J_adjoint_mixed_precision(small_wind):
    import torch

    from magritte_amd import solver
    from oracle import oracle

    rng = np.random.default_rng(13)
    Jb = _seed(small_wind, rng)
    dm = solver.DeviceModel(small_wind)
    de, dk = solver.feautrier_adjoint(dm, small_wind.options(precision=_abi.PREC_MIXED),
                                      torch.from_numpy(Jb))
    re, rk = oracle.J_adjoint(small_wind, small_wind.options(), Jb)
    _close(de.cpu().numpy(), re, 1e-4, "d_eps (mixed)")
    _close(dk.cpu().numpy(), rk, 1e-4, "d_kap (mixed)")


def _image_case(m):
    fr = m.image_frequencies(m.line[0] * (1 - 2e-5), m.line[0] * (1 + 2e-5), 31)
    return fr, np.array([0.3, -0.5, 0.81])


@pytest.mark.parametrize("precision,rtol", [(_abi.PREC_FP64, 1e-9), (_abi.PREC_MIXED, 1e-4)])
def test_image_adjoint_matches_oracle(small_wind, precision, rtol):
    import torch

    from magritte_amd import solver
    from oracle import oracle

    fr, rd = _image_case(small_wind)
    nx, ny = 24, 20
    rng = np.random.default_rng(14)
    Ib = rng.standard_normal((nx * ny, 31))
    print(f"[image_adjoint] precision={precision}", flush=True)
    dm = solver.DeviceModel(small_wind)
    de, dk = solver.image_adjoint(dm, small_wind.options(precision=precision), rd, nx, ny, fr,
                                  torch.from_numpy(Ib))
    re, rk = oracle.image_adjoint(small_wind, small_wind.options(), rd, nx, ny, Ib, fr)
    _close(de.cpu().numpy(), re, rtol, "d_eps")
    _close(dk.cpu().numpy(), rk, rtol, "d_kap")


def test_image_autograd_loss_gradient(small_wind):
    """L = |I - I*|^2 through torch.autograd (ImageNew) equals the oracle's
    vector-Jacobian product with I_bar = 2 (I - I*)."""
    import torch

    from magritte_amd import autograd, solver
    from oracle import oracle

    fr, rd = _image_case(small_wind)
    nx = ny = 16
    dm = solver.DeviceModel(small_wind)
    op = small_wind.options()
    eps = dm.tensors["emissivity"].clone().requires_grad_(True)
    kap = dm.tensors["opacity"].clone().requires_grad_(True)
    I = autograd.image(eps, kap, dm, op, rd, nx, ny, fr)
    target = I.detach() * (1.0 + 0.1 * torch.sin(torch.arange(I.numel(), device=I.device,
                                                               dtype=torch.float64)).view_as(I))
    loss = ((I - target) ** 2).sum()
    loss.backward()
    Ib = (2.0 * (I - target)).detach().cpu().numpy()
    re, rk = oracle.image_adjoint(small_wind, op, rd, nx, ny, Ib, fr)
    _close(eps.grad.cpu().numpy(), re, 1e-9, "d_eps")
    _close(kap.grad.cpu().numpy(), rk, 1e-9, "d_kap")


def test_autograd_J_against_central_differences(small_wind):
    """torch.autograd gradient of sum(w * J) against central differences of the GPU
    forward at 64 random (p, l), step 1e-6 relative; interpolation off so the
    forward is smooth in the parameters."""
    import torch

    from magritte_amd import autograd, solver

    m = small_wind
    op = m.options()
    op.max_interpolation_diff = 1.0e30  # every step is a single entry
    dm = solver.DeviceModel(m)
    rng = np.random.default_rng(15)
    W = torch.from_numpy(rng.standard_normal((m.npoints, m.nfreqs))).to(dm.device)
    eps = dm.tensors["emissivity"].clone().requires_grad_(True)
    kap = dm.tensors["opacity"].clone().requires_grad_(True)
    (W * autograd.formal_solve_J(eps, kap, dm, op)).sum().backward()
    ge, gk = eps.grad.cpu().numpy(), kap.grad.cpu().numpy()
    h = 1.0e-6
    worst = 0.0
    for name, g in (("eps", ge), ("kap", gk)):
        cand = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
        for i in rng.choice(cand, 32, replace=False):
            p, l = divmod(int(i), m.nlines)
            vals = []
            for sgn in (1.0, -1.0):
                e2 = eps.detach().clone()
                k2 = kap.detach().clone()
                t = e2 if name == "eps" else k2
                t[p, l] = t[p, l] * (1.0 + sgn * h)
                vals.append(autograd.formal_solve_J(e2, k2, dm, op))
            x0 = (eps if name == "eps" else kap)[p, l].item()
            fd = float((W * (vals[0] - vals[1])).sum().item() / (2 * h * x0))
            worst = max(worst, abs(g[p, l] - fd) / abs(fd))
    assert worst < 1e-4, worst
