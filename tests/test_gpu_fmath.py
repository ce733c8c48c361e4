"""The GPU build of the restated float functions against the C library, on every
argument: the device evaluates expf, expm1f and erff on all 2^32 float arguments
and logf on all 2^31 non-negative ones (mgx_fmath_digest) and reduces each chunk
of 2^20 (argument, result) pairs to a digest; the oracle computes the same
digests with the C library's own functions; all must be equal.  powf: every
positive finite base for four interpolation exponents, in chunks of 2^16 bases
(each inside one subinterval of the log table): only chunks of subinterval 3
may differ (the one table entry not reproduced to the last bit)."""
import numpy as np
import pytest

from magritte_amd import _abi
from oracle import oracle

pytestmark = pytest.mark.gpu


def _device_digest(fn, first, nchunks, chunk_log2, y=0.0):
    import torch

    d = torch.zeros(nchunks, dtype=torch.int64, device="cuda:0")
    _abi.check(_abi.lib().mgx_fmath_digest(fn, y, first, nchunks, chunk_log2, _abi.ptr(d),
                                           torch.cuda.current_stream().cuda_stream))
    return d.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("fn,name,count", [(oracle.FN_EXPF, "expf", 1 << 32),
                                           (oracle.FN_EXPM1F, "expm1f", 1 << 32),
                                           (oracle.FN_ERFF, "erff", 1 << 32),
                                           (oracle.FN_LOGF, "logf", 1 << 31)])
def test_device_unary_functions_equal_the_c_library_on_every_argument(fn, name, count):
    cl = 20
    n = count >> cl
    print(f"[fmath] {name}: {count} arguments in {n} chunks", flush=True)
    dev = _device_digest(fn, 0, n, cl)
    host = oracle.fmath_digest(fn, 0, n, cl, nthreads=16, kind="glibc")
    bad = np.flatnonzero(dev != host)
    assert bad.size == 0, f"{name}: {bad.size} chunks differ, first at argument bits {bad[0] << cl:#x}"


@pytest.mark.parametrize("y", [0.25, 1.0 / 3.0, 0.6, 0.9])
def test_device_powf_equals_the_c_library_outside_one_table_entry(y):
    cl = 16
    first, n = 1 << cl, (0x7f800000 >> cl) - 1  # positive finite bases, chunk 0 (subnormals) left out
    yf = float(np.float32(y))
    print(f"[fmath] powf(x, {yf}): {n << cl} bases", flush=True)
    dev = _device_digest(oracle.FN_POWF, first, n, cl, yf)
    host = oracle.fmath_digest(oracle.FN_POWF, first, n, cl, y=yf, nthreads=16, kind="glibc")
    bad = np.flatnonzero(dev != host)
    starts = first + (bad.astype(np.uint64) << np.uint64(cl))
    sub = ((starts.astype(np.int64) - 0x3f330000) >> 19) % 16
    print(f"  {bad.size} of {n} chunks differ, subintervals {sorted(set(sub.tolist()))}", flush=True)
    assert np.all(sub == 3)
    assert bad.size <= n // 16
