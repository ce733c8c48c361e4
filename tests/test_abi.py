"""The C ABI library loads on the CPU, exports every function include/mgx.h
declares, and rejects invalid arguments before touching the GPU."""
import ctypes as C
import os
import re

import pytest

from magritte_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "mgx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mgx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _abi.lib()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert sorted(n for n, _, _ in _abi.SIGNATURES) == names


def test_version_and_error_string():
    lib = _abi.lib()
    assert b"gfx950" in lib.mgx_version()


def test_null_arguments_are_rejected_without_gpu(line_model):
    lib = _abi.lib()
    op = line_model.options()
    rc = lib.mgx_feautrier_forward_device(None, C.byref(op), None, None, None, None)
    assert rc == _abi.MGX_E_ARG
    assert b"null" in lib.mgx_last_error()
    v = line_model.view()
    bad = line_model.options()
    bad.n_off_diag = 2
    rc = lib.mgx_feautrier_forward_device(C.byref(v), C.byref(bad), None, None, None, None)
    assert rc == _abi.MGX_E_UNSUPPORTED
    rc = lib.mgx_trace_counts_device(C.byref(v), C.byref(op), None, None)
    assert rc == _abi.MGX_E_ARG


def test_image_discretisation_rejects_zero_frequencies():
    lib = _abi.lib()
    rc = lib.mgx_compute_image_discretisation(4, 1.0, 2.0, 0, None, None, None, None, None, None)
    assert rc == _abi.MGX_E_ARG
    assert b"At least a single frequency" in lib.mgx_last_error()


def test_adjoint_entry_points_reject_null_arguments(line_model):
    lib = _abi.lib()
    op = line_model.options()
    v = line_model.view()
    rc = lib.mgx_feautrier_adjoint_device(C.byref(v), C.byref(op), None, None, None, None)
    assert rc == _abi.MGX_E_ARG
    rd = (C.c_double * 3)(0.0, 1.0, 0.0)
    rc = lib.mgx_image_adjoint_device(C.byref(v), C.byref(op), rd, 4, 4, None, None, None, None)
    assert rc == _abi.MGX_E_ARG
    # J and its gradient do not depend on the Lambda bands: n_off_diag is accepted,
    # and the null arguments are still what is reported
    bands = line_model.options()
    bands.n_off_diag = 1
    rc = lib.mgx_feautrier_adjoint_device(C.byref(v), C.byref(bands), None, None, None, None)
    assert rc == _abi.MGX_E_ARG and b"null argument" in lib.mgx_last_error()
    # the plain forward solve returns the diagonal only and points to the bands' entry
    rc = lib.mgx_feautrier_forward_device(C.byref(v), C.byref(bands), None, None, None, None)
    assert rc == _abi.MGX_E_UNSUPPORTED and b"mgx_feautrier_forward_lambda_device" in lib.mgx_last_error()
