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
