"""Batched forward solve (mgx_batch_*): several models of uneven size and ray
count in one persistent launch must give each model's own solve -- bit for bit
when every origin is one work item (the same code in the same order), within
1e-14 relative when an origin's ray pairs are split over items (J summed with
atomics) -- on every replay of the captured graph, with eager work between
the replays, and after the autotuner has chosen a tiling."""
import numpy as np
import pytest

from magritte_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_models():
    from magritte_amd import models

    specs = [(7, 12, 101), (9, 48, 102), (6, 12, 103), (11, 12, 104), (8, 48, 105)]
    return models.batch_winds(specs)


def _per_model(dms, op):
    from magritte_amd import solver

    return [solver.feautrier_forward(dm, op, with_u=False)["J"].cpu().numpy() for dm in dms]


@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_batch_equals_per_model_solves(batch_models, precision):
    import torch

    from magritte_amd import solver

    dms = [solver.DeviceModel(m) for m in batch_models]
    op = batch_models[0].options(precision=precision)
    ref = _per_model(dms, op)
    plan = solver.BatchPlan(dms, op)
    for cfg in [dict(rays_per_item=0, order=0, wps=3, use_graph=True),
                dict(rays_per_item=0, order=1, wps=2, use_graph=False),
                dict(rays_per_item=2, order=0, wps=4, use_graph=True),
                dict(rays_per_item=1, order=0, wps=3, use_graph=True)]:
        print(f"[batch] precision={precision} {cfg}", flush=True)
        plan.configure(**cfg)
        for rep in range(3):
            Js = plan.run()
            for i, (J, R) in enumerate(zip(Js, ref)):
                got = J.cpu().numpy()
                if cfg["rays_per_item"] == 0:
                    assert np.array_equal(got, R), (cfg, rep, i)
                else:
                    err = np.max(np.abs(got - R) / np.abs(R))
                    assert err < 1e-14, (cfg, rep, i, err)
            # eager work on the same stream between replays
            torch.cuda.synchronize()
            _ = (torch.ones(1000, device="cuda") * 2.0).sum().item()
    plan.close()


def test_batch_autotuner_and_lambda(batch_models):
    from magritte_amd import solver

    dms = [solver.DeviceModel(m) for m in batch_models]
    op = batch_models[0].options()
    refs = [solver.feautrier_forward(dm, op, with_u=False) for dm in dms]
    plan = solver.BatchPlan(dms, op, with_lambda=True)
    best = plan.tune(sample_stride=3)
    print(f"[batch] tuned {best}", flush=True)
    assert best["wps"] in (2, 3, 4)
    Js, Ls = plan.run()
    for J, L, r in zip(Js, Ls, refs):
        Rj = r["J"].cpu().numpy()
        Rl = r["lambda_diag"].cpu().numpy()
        tol = 0.0 if best["rays_per_item"] == 0 else 1e-14
        assert np.max(np.abs(J.cpu().numpy() - Rj) / np.abs(Rj)) <= tol
        assert np.max(np.abs(L.cpu().numpy() - Rl) / np.maximum(np.abs(Rl), 1e-300)) <= max(tol, 1e-14)
    plan.close()


def _image_case(ms, rng):
    rds = rng.standard_normal((len(ms), 3))
    rds /= np.linalg.norm(rds, axis=1)[:, None]
    frs = []
    for m in ms:
        nu0 = m.line[int(rng.integers(0, m.nlines))]
        frs.append(m.image_frequencies(nu0 * (1 - 1e-4), nu0 * (1 + 1e-4), 31))
    return rds, frs


@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_batch_images_equal_per_model_images(batch_models, precision):
    """mgx_batch_set_images: every model's image from the batch's persistent image
    kernel equals its own mgx_image_device call bit for bit (the same pixel code),
    on every replay of the captured graph, after the tuner, and with the plain
    launch sequence; the solve's J is unchanged by the images."""
    import torch

    from magritte_amd import solver

    dms = [solver.DeviceModel(m) for m in batch_models]
    op = batch_models[0].options(precision=precision)
    rng = np.random.default_rng(77)
    rds, frs = _image_case(batch_models, rng)
    nx, ny = 24, 20
    refJ = _per_model(dms, op)
    refI = [[t.cpu().numpy() for t in solver.image_new(dm, op, rd, nx, ny, fr)]
            for dm, rd, fr in zip(dms, rds, frs)]
    plan = solver.BatchPlan(dms, op)
    plan.set_images(rds, nx, ny, frs)
    for step, cfg in enumerate([None, "tune", dict(rays_per_item=0, order=1, wps=2, use_graph=False)]):
        if cfg == "tune":
            plan.tune(sample_stride=3)
        elif cfg is not None:
            plan.configure(**cfg)
        print(f"[batch images] precision={precision} step {step} {cfg}", flush=True)
        for rep in range(3):
            Js = plan.run()
            for i in range(len(dms)):
                for k, name in enumerate(("I", "ImX", "ImY")):
                    got = plan.images[i][k].cpu().numpy()
                    assert np.array_equal(got, refI[i][k]), (step, rep, i, name)
                err = np.max(np.abs(Js[i].cpu().numpy() - refJ[i]) / np.abs(refJ[i]))
                assert err <= 1e-14, (step, rep, i, err)
            torch.cuda.synchronize()
            _ = (torch.ones(1000, device="cuda") * 2.0).sum().item()
    plan.close()


def test_batch_images_argument_checks(batch_models):
    from magritte_amd import solver

    dms = [solver.DeviceModel(m) for m in batch_models[:2]]
    op = batch_models[0].options()
    rng = np.random.default_rng(78)
    rds, frs = _image_case(batch_models[:2], rng)
    plan = solver.BatchPlan(dms, op)
    with pytest.raises(_abi.MgxError, match="does not point"):
        plan.set_images(np.zeros((2, 3)), 4, 4, frs)
    with pytest.raises(_abi.MgxError, match="at least one pixel"):
        plan.set_images(rds, 0, 4, frs)
    # an image view that does not share the batch model's arrays
    plan.set_images(rds, 4, 4, frs)
    views = (_abi.MgxModel * 2)(plan._img_views[1], plan._img_views[0])
    with pytest.raises(_abi.MgxError, match="does not describe|must share"):
        _abi.check(_abi.lib().mgx_batch_set_images(plan._h, views, plan._rd.ctypes.data, 4, 4,
                                                   plan._img_ptrs[0], plan._img_ptrs[1],
                                                   plan._img_ptrs[2], dms[0].stream_ptr()))
    plan.close()
    plan1 = solver.BatchPlan(dms, batch_models[0].options(approx=_abi.APPROX_ONE_LINE))
    with pytest.raises(_abi.MgxError, match="One line"):
        plan1.set_images(rds, 4, 4, frs)
    plan1.close()
