"""Rays longer than the kernels' first walk buffer.

The reference sizes its scratch with a tracing pass; the solves here size the
LDS walk buffer from a walk-only pass (the longest half ray of the model) and
the imager and adjoints start at 128 steps and relaunch with twice the buffer
when a walk does not fit, up to what the device's LDS per workgroup holds
(160 KiB on gfx950: about 3360 steps), then fail with MGX_E_RAY_LENGTH.  On a
1D chain of 1000 points the half rays from the ends have 999 steps: the forward
solve (u, J and Lambda within 1e-10 of the oracle) and the adjoint (whose
per-wave tape also grows) must come out as on short rays, and so must the
imager's pixel rays through 600 spherical shells (~1200 steps) and a chain of
2000 points (1999 steps); a chain of 4000 points must fail cleanly with the
library's message.
"""
import numpy as np
import pytest

from magritte_amd import _abi, models

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / (np.abs(b) + 1e-300)))


def test_long_chain_forward_and_adjoint():
    import torch

    from magritte_amd import solver
    from oracle import oracle

    m = models.model_line_1d(npoints=1000, nquads=7)
    op = m.options()
    print(f"[long rays] chain P={m.npoints}", flush=True)
    dm = solver.DeviceModel(m)
    out = solver.feautrier_forward(dm, op)
    u_r, J_r, lam_r = oracle.solve(m, op)
    e_J = _rel(out["J"].cpu().numpy(), J_r)
    e_L = _rel(out["lambda_diag"].cpu().numpy(), lam_r)
    e_u = _rel(out["u"].cpu().numpy(), u_r)
    print(f"  J {e_J:.2e}  Lambda {e_L:.2e}  u {e_u:.2e}", flush=True)
    assert e_J < 1e-10 and e_L < 1e-10 and e_u < 1e-10
    rng = np.random.default_rng(3)
    Jb = rng.standard_normal((m.npoints, m.nfreqs))
    de, dk = solver.feautrier_adjoint(dm, op, torch.from_numpy(Jb))
    re, rk = oracle.J_adjoint(m, op, Jb)
    # the opacity gradient carries the optical-depth adjoints, accurate to ~eps / dtau^2
    # relative in either implementation: this chain's segments are ten times thinner than
    # the 100-point chain's (observed 2.3e-9), so its bar is 1e-8
    for name, g, r, bar in (("d_eps", de, re, 1e-9), ("d_kap", dk, rk, 1e-8)):
        g = g.cpu().numpy()
        err = np.abs(g - r) / (np.abs(r) + 1e-10 * np.abs(r).max())
        print(f"  {name} worst {err.max():.2e}", flush=True)
        assert err.max() < bar


def test_long_image_rays_through_many_shells():
    """The imager on a spherically symmetric model of 600 shells (the reference's 1D
    images look along (0, 1, 0)): pixel rays near the centre cross ~1200 shells."""
    import os

    from magritte_amd import solver
    from oracle import oracle

    ld = models.read_lamda(os.path.join(os.path.dirname(__file__), "golden", "lamda_test.txt"),
                           num=0)
    m = models.model_velocity_gradient_1d(ld, npoints=600, nrays=20, nquads=7)
    op = m.options()
    fr = m.image_frequencies(m.line[0] * (1 - 3e-5), m.line[0] * (1 + 3e-5), 11)
    rd = np.array([0.0, 1.0, 0.0])
    print("[long rays] image through 600 shells", flush=True)
    dm = solver.DeviceModel(m)
    I, X, Y = solver.image_new(dm, op, rd, 16, 1, fr)
    Ir, Xr, Yr = oracle.image_new(m, op, rd, 16, 1, fr)
    np.testing.assert_array_equal(X.cpu().numpy(), Xr)
    e = _rel(I.cpu().numpy(), Ir)
    print(f"  I {e:.2e}", flush=True)
    assert e < 1e-10


def test_chain_of_2000_points():
    """1999-step half rays: beyond a 64 KiB walk buffer, within gfx950's 160 KiB."""
    from magritte_amd import solver
    from oracle import oracle

    m = models.model_line_1d(npoints=2000, nquads=3)
    op = m.options()
    print("[long rays] chain P=2000", flush=True)
    out = solver.feautrier_forward(solver.DeviceModel(m), op, with_u=False)
    _, J_r, lam_r = oracle.solve(m, op, with_u=False)
    e_J, e_L = _rel(out["J"].cpu().numpy(), J_r), _rel(out["lambda_diag"].cpu().numpy(), lam_r)
    print(f"  J {e_J:.2e}  Lambda {e_L:.2e}", flush=True)
    assert e_J < 1e-10 and e_L < 1e-10


def test_too_long_chain_fails_cleanly():
    from magritte_amd import solver

    m = models.model_line_1d(npoints=4000, nquads=3)
    print("[long rays] chain P=4000 (beyond the walk buffer)", flush=True)
    dm = solver.DeviceModel(m)
    with pytest.raises(_abi.MgxError, match="walk steps"):
        solver.feautrier_forward(dm, m.options())
    # the library stays usable afterwards
    ms = models.model_line_1d(npoints=50, nquads=3)
    out = solver.feautrier_forward(solver.DeviceModel(ms), ms.options())
    assert np.isfinite(out["J"].cpu().numpy()).all()


def test_interpolation_count_overflow_is_its_own_error():
    """A segment needing more than 65535 interpolation points (here: the criterion
    doubles at one point while max_interpolation_diff is 1 + 1e-6, so
    ceil(log 2 / log(1 + 1e-6)) = 693147) fails at once with its own message,
    not as a walk-buffer overflow; the library stays usable."""
    from magritte_amd import solver

    m = models.model_line_1d(nquads=3)
    m.interp_crit[50, :] *= 2.0
    op = m.options()
    op.max_interpolation_diff = 1.0 + 1e-6
    print("[interpolation overflow] line P=100, one criterion jump of 2", flush=True)
    dm = solver.DeviceModel(m)
    with pytest.raises(_abi.MgxError, match="65535 interpolation points"):
        solver.feautrier_forward(dm, op)
    out = solver.feautrier_forward(dm, m.options())
    assert np.isfinite(out["J"].cpu().numpy()).all()
