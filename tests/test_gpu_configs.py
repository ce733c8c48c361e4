"""The benchmark configurations at their stated sizes.

C3 (differentiation): the 32^3 wind with per-point emissivity and opacity as the
free parameters, L = |I - I*|^2 for a 256 x 256 x 31 image along (0.3, -0.5,
0.81), I* the image of the same wind with seeded 10% log-normal perturbations
of density and temperature.  The gradient through torch.autograd (the adjoint
kernel) equals the oracle's image adjoint with I_bar = 2 (I - I*) (FP64, 1e-9
relative plus 1e-10 of the largest), central differences of the GPU forward
(step 1e-6 relative, interpolation switched off so the forward is smooth) at 64
random (p, l) agree to 1e-4, and the mixed-precision gradient agrees with the
FP64 one to 1e-4 where it is significant (above 1e-3 of the largest)."""
import numpy as np
import pytest

from magritte_amd import _abi, models
from test_gpu_adjoint import _close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    rng = np.random.Generator(np.random.PCG64(models.seed_for(3)))
    m = models.model_wind(n_edge=32)
    P = m.npoints
    pert_n = np.exp(0.1 * rng.standard_normal(P))
    pert_t = np.exp(0.1 * rng.standard_normal(P))
    mt = models.model_wind(n_edge=32, clumps=pert_n, temp_scale=pert_t)
    nu0 = m.line[0]
    fr = m.image_frequencies(nu0 * (1 - 1e-4), nu0 * (1 + 1e-4), 31)
    return m, mt, fr, np.array([0.3, -0.5, 0.81]), rng


def _grad(dm, op, rd, n, fr, Istar):
    from magritte_amd import autograd

    eps = dm.tensors["emissivity"].clone().requires_grad_(True)
    kap = dm.tensors["opacity"].clone().requires_grad_(True)
    I = autograd.image(eps, kap, dm, op, rd, n, n, fr)
    ((I - Istar) ** 2).sum().backward()
    return I.detach(), eps.grad.cpu().numpy(), kap.grad.cpu().numpy()


def test_c3_image_loss_gradient_full_size(c3):
    from magritte_amd import autograd, solver
    from oracle import oracle

    m, mt, fr, rd, rng = c3
    n = 256
    print(f"[C3] P={m.npoints} image {n}x{n}x{fr['nu'].shape[1]}", flush=True)
    dm = solver.DeviceModel(m)
    Istar, _, _ = solver.image_new(solver.DeviceModel(mt), mt.options(), rd, n, n, fr)
    op = m.options()
    I, ge, gk = _grad(dm, op, rd, n, fr, Istar)
    Ib = (2.0 * (I - Istar)).cpu().numpy()
    re, rk = oracle.image_adjoint(m, op, rd, n, n, Ib, fr, nthreads=16)
    _close(ge, re, 1e-9, "d_eps vs oracle")
    _close(gk, rk, 1e-9, "d_kap vs oracle")
    # mixed-precision forward and adjoint against FP64
    _, me, mk = _grad(dm, m.options(precision=_abi.PREC_MIXED), rd, n, fr, Istar)
    for name, a, b in (("eps", me, ge), ("kap", mk, gk)):
        sig = np.abs(b) > 1e-3 * np.abs(b).max()
        e = float(np.max(np.abs(a[sig] - b[sig]) / np.abs(b[sig])))
        print(f"  mixed d_{name} vs FP64: {e:.3e}", flush=True)
        assert e < 1e-4
    # central differences with interpolation off
    opf = m.options()
    opf.max_interpolation_diff = 1.0e30
    _, ge, gk = _grad(dm, opf, rd, n, fr, Istar)
    base = {"eps": dm.tensors["emissivity"], "kap": dm.tensors["opacity"]}
    h = 1.0e-6
    errs = []
    for which, g in (("eps", ge), ("kap", gk)):
        cand = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
        for i in rng.choice(cand, 32, replace=False):
            p, l = divmod(int(i), m.nlines)
            Is = []
            for sgn in (1.0, -1.0):
                e2, k2 = base["eps"].clone(), base["kap"].clone()
                t = e2 if which == "eps" else k2
                t[p, l] = t[p, l] * (1.0 + sgn * h)
                Is.append(autograd.image(e2, k2, dm, opf, rd, n, n, fr))
            x0 = base[which][p, l].item()
            dL = ((Is[0] - Is[1]) * (Is[0] + Is[1] - 2.0 * Istar)).sum().item()
            fd = dL / (2 * h * x0)
            errs.append(abs(g[p, l] - fd) / abs(fd))
    errs = np.array(errs)
    print(f"  central differences at {len(errs)} (p, l): max {errs.max():.3e}, median "
          f"{np.median(errs):.3e}", flush=True)
    assert errs.max() < 1e-4


# ---------------------------------------------------------------------------
# C4 (large non-LTE): the 100^3 clumpy rotating wind, 192 rays, 41-level rotor,
# checked at its full size through what one can afford to recompute on the CPU
# (sampled rays) and through properties that hold at any size.


@pytest.fixture(scope="module")
def c4():
    import time

    t0 = time.time()
    clumps = models.grf_clumps(100, sigma=0.5, corr_length=1.0 / 20.0)
    m = models.model_wind(n_edge=100, nrays=192, nlev=41, v_rot=5.0e3, clumps=clumps)
    print(f"[C4] model built in {time.time() - t0:.0f}s: P={m.npoints} hnrays={m.hnrays} "
          f"F={m.nfreqs}", flush=True)
    return m


@pytest.fixture(scope="module")
def c4_dev(c4):
    from magritte_amd import solver

    return solver.DeviceModel(c4)


def test_c4_sampled_traces_and_solves(c4, c4_dev):
    """At 64 random (ray, origin, line) triples (two of them boundary origins) the
    device's traced lists equal the oracle's bit for bit, and at 64 random
    (origin, frequency) of two ray pairs the device's u (one launch per pair, every
    origin) equals the oracle's single-ray solve within 1e-10."""
    import torch

    from magritte_amd import solver
    from oracle import oracle

    m, dm = c4, c4_dev
    op = m.options()
    rng = np.random.default_rng(41)
    n = 64
    sel = np.stack([rng.integers(0, m.hnrays, n), rng.integers(0, m.npoints, n),
                    rng.integers(0, m.nlines, n)], axis=1)
    sel[0, 1] = m.boundary2point[0]
    sel[1, 1] = m.boundary2point[-1]
    print(f"[C4 trace_rays] n={n}", flush=True)
    got = solver.trace_rays(dm, op, sel, capacity=4096)
    lengths = []
    for i, (rr, o, l) in enumerate(sel):
        ref = oracle.trace_ray(m, op, int(rr), int(o), int(l))
        g = got[i]
        assert g["count"] == ref["count"] and g["centre"] == ref["centre"], (i, rr, o, l)
        for k in ("nr", "nr_interp"):
            np.testing.assert_array_equal(g[k], ref[k], err_msg=f"{k} ray {i}")
        for k in ("shift", "interp_factor", "dZ"):
            np.testing.assert_array_equal(g[k].view(np.uint64), ref[k].view(np.uint64),
                                          err_msg=f"{k} ray {i}")
        lengths.append(ref["count"])
    print(f"  bit-exact; traced lengths {min(lengths)}..{max(lengths)}", flush=True)
    worst = 0.0
    for rr in (int(rng.integers(0, m.hnrays)), m.hnrays - 1):
        v, keep = solver.ray_pair_view(dm, rr)
        out = {"u": torch.empty((1, m.npoints, m.nfreqs), dtype=torch.float64, device=dm.device),
               "J": torch.empty((m.npoints, m.nfreqs), dtype=torch.float64, device=dm.device),
               "lambda_diag": None}
        solver.feautrier_forward(dm, op, out=out, view=v)
        os_ = rng.integers(0, m.npoints, 32)
        fs = rng.integers(0, m.nfreqs, 32)
        os_[0] = m.boundary2point[0]
        ug = out["u"][0, torch.from_numpy(os_).to(dm.device),
                      torch.from_numpy(fs).to(dm.device)].cpu().numpy()
        del out
        for k in range(len(os_)):
            s = oracle.feautrier_system(m, op, rr, int(os_[k]), int(fs[k]))
            worst = max(worst, abs(ug[k] - s["u_c"]) / abs(s["u_c"]))
        print(f"  ray pair {rr}: u at 32 (o, f) within {worst:.2e}", flush=True)
    assert worst < 1e-10


def test_c4_statistical_equilibrium_in_detailed_balance(c4, c4_dev):
    """With J = B_nu(T) at every bin of every line, the Lambda diagonal zero and the
    collision partner's abundance zero, detailed balance makes the LTE populations
    the solution of every point's rate equations.  One device update from LTE
    returns them at all 10^6 points: within 1e-9 of each point's total, and 1e-6
    relative on every level above 1e-6 of it (the Einstein coefficients of the
    rotor, formed with the constants' rounded literals, balance to 3e-10 per line;
    the long-double oracle on a 6^3 wind gives 3e-11 and 3e-9)."""
    import torch

    from magritte_amd import solver

    m, dm = c4, c4_dev
    ab0 = m.abundance.copy()
    m.abundance[:, 0] = 0.0
    try:
        lp = solver.LevelPopulations(dm, _abi.default_nlte_params())
    finally:
        m.abundance[:] = ab0
    lp.set_lte()
    p0, tot = lp.populations()
    nu = torch.as_tensor(np.asarray(m.linedata[0]["frequency"]), device=dm.device)
    T = torch.as_tensor(m.temperature, device=dm.device)
    B = (2.0 * models.HH * nu[None, :] ** 3 / models.CC_SQUARED /
         torch.expm1(models.HH * nu[None, :] / (models.KB * T[:, None])))
    f_line = torch.as_tensor(m.f_line.astype(np.int64), device=dm.device)
    J = B[:, f_line].contiguous()
    lam = torch.zeros((m.npoints, m.nlines), dtype=torch.float64, device=dm.device)
    print(f"[C4 stateq J=B] P={m.npoints} nlev=41", flush=True)
    _, Jl, _ = lp.stateq(J, lam, want_J=True)
    jl_err = float(((Jl - B).abs() / B).max())
    p1, _ = lp.populations()
    a = p0.reshape(m.npoints, -1)
    b = p1.reshape(m.npoints, -1)
    t = tot[0][:, None]
    abs_err = float(((b - a).abs() / t).max())
    sig = a > 1e-6 * t
    rel_err = float(((b - a).abs() / a)[sig].max())
    print(f"  J_lin vs B {jl_err:.2e}; populations: {abs_err:.2e} of the total, {rel_err:.2e} "
          f"relative above 1e-6 of it", flush=True)
    assert jl_err < 1e-13
    assert abs_err < 1e-9 and rel_err < 1e-6


def test_c4_image_gradient_512(c4, c4_dev):
    """The 512^2 x 31 image around line 3 along a seeded direction, L = |I - I*|^2 with
    I* the image after seeded 10% perturbations of log n and T, differentiated with
    respect to the log-density and temperature of every point (LTE populations) through
    torch.autograd.  Central differences of the device forward at 16 points each
    (chosen among those whose gradient is above 1e-3 of the largest; interpolation
    off, so the forward is smooth in the line data): log n with step 1e-6, agreement
    1e-4; T with relative step 1e-4, agreement 5e-3, since the temperature moves the
    arguments of the single-precision erf of the line profile, whose one-ulp steps
    (6e-8) are resolved by a finite difference to ~1e-3 only."""
    import torch

    from magritte_amd import autograd

    m, dm = c4, c4_dev
    n = 512
    th = autograd.LineThermo(m, dm.device)
    ln0, T0 = th.log_density(), th.temperature()
    rng = np.random.Generator(np.random.PCG64(models.seed_for(4)))
    rd = rng.standard_normal(3)
    rd /= np.linalg.norm(rd)
    nu0 = m.line[3]
    fr = m.image_frequencies(nu0 * (1 - 1e-4), nu0 * (1 + 1e-4), 31)
    op = m.options()
    op.max_interpolation_diff = 1.0e30
    P = m.npoints
    pn = torch.from_numpy(0.1 * rng.standard_normal(P)).to(dm.device)
    pt = torch.from_numpy(np.exp(0.1 * rng.standard_normal(P))).to(dm.device)
    print(f"[C4 image gradient] {n}x{n}x31", flush=True)
    with torch.no_grad():
        Istar = autograd.image_log_density_temperature(ln0 + pn, T0 * pt, th, dm, op, rd, n, n, fr)
    ln = ln0.clone().requires_grad_(True)
    T = T0.clone().requires_grad_(True)
    I = autograd.image_log_density_temperature(ln, T, th, dm, op, rd, n, n, fr)
    ((I - Istar) ** 2).sum().backward()
    gl, gT = ln.grad.cpu().numpy(), T.grad.cpu().numpy()

    def loss_diff(x_plus, x_minus, which):
        Is = []
        for x in (x_plus, x_minus):
            with torch.no_grad():
                args = (x, T0) if which == "ln" else (ln0, x)
                Is.append(autograd.image_log_density_temperature(*args, th, dm, op, rd, n, n, fr))
        return ((Is[0] - Is[1]) * (Is[0] + Is[1] - 2.0 * Istar)).sum().item()

    res = {}
    for which, g, x0, tol in (("ln", gl, ln0, 1e-4), ("T", gT, T0, 5e-3)):
        cand = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
        errs = []
        for p in rng.choice(cand, 16, replace=False):
            h = 1e-6 if which == "ln" else 1e-4 * float(x0[p])
            xp, xm = x0.clone(), x0.clone()
            xp[p] += h
            xm[p] -= h
            fd = loss_diff(xp, xm, which) / (2 * h)
            errs.append(abs(g[p] - fd) / abs(fd))
        res[which] = np.array(errs)
        print(f"  d/d{which}: {len(cand)} significant points; central differences at 16: max "
              f"{res[which].max():.3e}, median {np.median(res[which]):.3e}", flush=True)
    assert res["ln"].max() < 1e-4
    assert res["T"].max() < 5e-3


# ---------------------------------------------------------------------------
# C2 (batched): 64 jittered winds, P log-uniform in [2e3, 1.3e5], 12/48/192 rays,
# each with a 256^2 x 31 image along a seeded direction.


def test_c2_batch_with_images_full_size():
    """The 64-model batch with its images, FP64, tuned tiling: every model's J
    equals its own plain solve within 1e-14 (atomics when ray pairs are split),
    every image equals its own mgx_image_device call bit for bit, and the three
    smallest models' J and images equal the oracle's within 1e-10."""
    import time

    import torch

    from magritte_amd import solver
    from oracle import oracle

    t0 = time.time()
    specs = models.batch_wind_specs(n_models=64)
    ms = models.batch_winds(specs, workers=8)
    print(f"[C2] 64 models built in {time.time() - t0:.0f}s, P total "
          f"{sum(m.npoints for m in ms)}", flush=True)
    dms = [solver.DeviceModel(m) for m in ms]
    op = ms[0].options()
    rng = np.random.Generator(np.random.PCG64(models.seed_for(2, 1000)))
    rds = rng.standard_normal((len(ms), 3))
    rds /= np.linalg.norm(rds, axis=1)[:, None]
    frs = [m.image_frequencies(m.line[0] * (1 - 1e-4), m.line[0] * (1 + 1e-4), 31) for m in ms]
    n = 256
    plan = solver.BatchPlan(dms, op)
    best = plan.tune(sample_stride=16)
    plan.set_images(rds, n, n, frs)
    t1 = time.time()
    Js = plan.run()
    torch.cuda.synchronize()
    print(f"[C2] batch (tiling {best}) with images: {time.time() - t1:.1f}s", flush=True)
    worst_J = 0.0
    for i, (dm, m) in enumerate(zip(dms, ms)):
        R = solver.feautrier_forward(dm, op, with_u=False, with_lambda=False)["J"]
        worst_J = max(worst_J, float(((Js[i] - R).abs() / R.abs()).max()))
        del R
        ref = solver.image_new(dm, op, rds[i], n, n, frs[i])
        for k in range(3):
            assert torch.equal(plan.images[i][k], ref[k]), (i, k)
    print(f"  J vs plain solves {worst_J:.1e}; 64 images bit-exact", flush=True)
    assert worst_J <= 1e-14
    small = sorted(range(len(ms)), key=lambda i: ms[i].npoints)[:3]
    for i in small:
        m = ms[i]
        print(f"[C2 oracle] model {i}: P={m.npoints} rays={m.nrays}", flush=True)
        _, Jr, _ = oracle.solve(m, op, with_u=False, nthreads=16)
        eJ = float(np.max(np.abs(Js[i].cpu().numpy() - Jr) / np.abs(Jr)))
        Ir, Xr, Yr = oracle.image_new(m, op, rds[i], n, n, frs[i], nthreads=16)
        I = plan.images[i][0].cpu().numpy()
        eI = float(np.max(np.abs(I - Ir) / np.maximum(np.abs(Ir), 1e-300)))
        print(f"  J {eJ:.2e}, image {eI:.2e}", flush=True)
        assert eJ < 1e-10 and eI < 1e-10
        np.testing.assert_array_equal(plan.images[i][1].cpu().numpy(), Xr)
        np.testing.assert_array_equal(plan.images[i][2].cpu().numpy(), Yr)
    plan.close()
