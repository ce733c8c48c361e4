"""mgx_feautrier_forward_noalloc: the plain-method solve without device
allocations (its scratch is a static device array) gives the allocating
entry point's u, J and Lambda bit for bit, on two calls in a row whose second
model is smaller than the first, and rejects a model that needs more scratch
than the array holds."""
import ctypes as C

import numpy as np
import pytest

from magritte_amd import _abi, models

pytestmark = pytest.mark.gpu


def _solve(entry, dm, op):
    import torch

    m = dm.model
    u = torch.full((m.hnrays, m.npoints, m.nfreqs), np.nan, dtype=torch.float64, device=dm.device)
    J = torch.full((m.npoints, m.nfreqs), np.nan, dtype=torch.float64, device=dm.device)
    lam = torch.full((m.npoints, m.nlines), np.nan, dtype=torch.float64, device=dm.device)
    rc = getattr(_abi.lib(), entry)(C.byref(dm.view), C.byref(op), _abi.ptr(u), _abi.ptr(J),
                                    _abi.ptr(lam), dm.stream_ptr())
    _abi.check(rc)
    torch.cuda.synchronize()
    return u.cpu().numpy(), J.cpu().numpy(), lam.cpu().numpy()


@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_noalloc_forward_equals_the_allocating_forward(precision):
    from magritte_amd import solver

    # the second call's model is smaller than the first: the scratch keeps the first
    # call's larger tables, and the status words must start from zero again
    for m in (models.model_wind(n_edge=12, seed=models.seed_for(1, 7)), models.model_line_1d()):
        print(f"[noalloc] P={m.npoints} L={m.nlines} F={m.nfreqs} precision={precision}",
              flush=True)
        dm = solver.DeviceModel(m)
        op = m.options(precision=precision)
        ref = _solve("mgx_feautrier_forward_device", dm, op)
        got = _solve("mgx_feautrier_forward_noalloc", dm, op)
        for name, a, b in zip(("u", "J", "lambda"), got, ref):
            assert not np.isnan(a).any(), name
            np.testing.assert_array_equal(a, b, err_msg=name)


def test_noalloc_forward_rejects_models_beyond_its_scratch():
    from magritte_amd import solver

    m = models.model_line_1d()
    dm = solver.DeviceModel(m)
    op = m.options()
    big = _abi.MgxModel.from_buffer_copy(dm.view)
    # 2^21 points (one line) need 235 MB of tables: refused before any device work
    big.npoints = 1 << 21
    import torch

    lib = _abi.lib()
    dummy = torch.zeros(1, dtype=torch.float64, device=dm.device)  # never written
    rc = lib.mgx_feautrier_forward_noalloc(C.byref(big), C.byref(op), None, _abi.ptr(dummy), None,
                                           dm.stream_ptr())
    assert rc == _abi.MGX_E_ARG
    assert b"bytes of scratch" in lib.mgx_last_error()
    # the entry point still works afterwards
    got = _solve("mgx_feautrier_forward_noalloc", dm, op)
    ref = _solve("mgx_feautrier_forward_device", dm, op)
    np.testing.assert_array_equal(got[1], ref[1])
