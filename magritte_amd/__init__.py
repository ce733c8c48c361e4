"""MI355X (gfx950) implementation of Magritte's plain-method Feautrier formal solver.

The compute path is the C ABI of libmgx.so (include/mgx.h): hand-written HIP
kernels for the ray tracing, the second-order Feautrier elimination, the J and
Lambda accumulation and the projection-surface imager.  Python reaches it
through ctypes (``magritte_amd._abi``); ``magritte_amd.models`` builds model
data and runs the host-side producers.
"""
from . import _abi  # noqa: F401

__all__ = ["_abi"]
