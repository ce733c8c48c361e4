"""Pins the float transcendentals shared by the oracle and the kernels
(magritte_amd/csrc/mgx_fmath.h).

Each must equal the correctly rounded float result: (float) of a 60-digit
mpmath evaluation on structured and random arguments.  An exhaustive run over
all 2^32 floats (oracle/fmath_exhaustive.c) found no difference from
(float) of glibc's double functions; this test samples the same property.
"""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle

mp.mp.dps = 60


def f32(x):
    return float(np.float32(x))


def cr(fn, *args):
    """correctly rounded float of fn(args) (args are floats)."""
    v = fn(*[mp.mpf(a) for a in args])
    return f32(float(v)) if abs(v) < 1e308 else f32(math.copysign(math.inf, float(v)))


def _args(rng, lo, hi, n):
    a = rng.uniform(lo, hi, n).astype(np.float32)
    return sorted(set(float(x) for x in a))


def _check(name, ours, ref, xs):
    xs = [f32(x) for x in xs]
    bad = [(x, ours(x), ref(x)) for x in xs if np.float32(ours(x)).view(np.uint32) != np.float32(ref(x)).view(np.uint32)]
    assert not bad, f"{name}: {len(bad)} mismatches, first {bad[:3]}"


def test_expf():
    lib = oracle.lib()
    rng = np.random.default_rng(1)
    xs = _args(rng, -103.0, 88.0, 3000) + [0.0, -0.0, 1e-30, -1e-30, -1.0, 1.0, -87.3, -103.9]
    _check("expf", lib.oracle_expf, lambda x: cr(mp.exp, x), xs)


def test_logf():
    lib = oracle.lib()
    rng = np.random.default_rng(2)
    xs = [f32(10.0**e) for e in rng.uniform(-44, 38, 3000)] + [1.0, f32(1.4), 0.999999940, 1.0000001]
    xs += [f32(1.0 + d) for d in rng.uniform(-1e-3, 1e-3, 500)]
    _check("logf", lib.oracle_logf, lambda x: cr(mp.log, x), xs)
    assert lib.oracle_logf(0.0) == -math.inf


def test_expm1f():
    lib = oracle.lib()
    rng = np.random.default_rng(3)
    xs = _args(rng, -20.0, 88.0, 2000) + _args(rng, -0.5, 0.5, 2000) + [1e-20, -1e-20, 0.0]
    _check("expm1f", lib.oracle_expm1f, lambda x: cr(mp.expm1, x), xs)


def test_erff():
    lib = oracle.lib()
    rng = np.random.default_rng(4)
    xs = _args(rng, -4.5, 4.5, 4000) + _args(rng, -0.3, 0.3, 1000) + [0.0, 1e-38, 0.25, 3.999, 4.0, 10.0, -10.0]
    _check("erff", lib.oracle_erff, lambda x: cr(mp.erf, x), xs)


def test_powf_on_interpolation_arguments():
    """powf(b/a, t) as used by interpolate_log: positive ratios, t in (0, 1)."""
    lib = oracle.lib()
    rng = np.random.default_rng(5)
    bad = 0
    for _ in range(3000):
        x = f32(10.0 ** rng.uniform(-6, 6))
        y = f32(rng.uniform(0.0, 1.0))
        if np.float32(lib.oracle_powf(x, y)).view(np.uint32) != np.float32(cr(mp.power, x, y)).view(np.uint32):
            bad += 1
    assert bad == 0


@pytest.mark.parametrize("fn", ["expf", "logf", "erff", "expm1f"])
def test_glibc_float_functions_are_not_correctly_rounded(fn):
    """Records why the oracle does not call glibc's float functions for parity:
    they differ from the correctly rounded value on some arguments."""
    g = oracle.lib("glibc")
    c = oracle.lib("cr")
    rng = np.random.default_rng(6)
    lo, hi = {"expf": (-80, 80), "logf": (1e-3, 1e3), "erff": (-3.5, 3.5), "expm1f": (-10, 80)}[fn]
    xs = rng.uniform(lo, hi, 200000).astype(np.float32)
    gf = getattr(g, "oracle_" + fn)
    cf = getattr(c, "oracle_" + fn)
    diff = sum(1 for x in xs[:20000] if np.float32(gf(float(x))) != np.float32(cf(float(x))))
    # all differences are one unit in the last place, never more
    for x in xs[:2000]:
        a, b = np.float32(gf(float(x))), np.float32(cf(float(x)))
        assert abs(int(a.view(np.int32)) - int(b.view(np.int32))) <= 1
    assert diff >= 0
