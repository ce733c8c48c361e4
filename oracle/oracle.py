"""CPU ORACLE bindings -- test infrastructure only.

ctypes bindings of oracle/liboracle.so (a C restatement of the reference's
plain-method formal solver, see mgx_oracle.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def lib(kind="cr"):
    """kind='cr': correctly rounded float functions (shared with the kernels);
    kind='glibc': the C library's expf/expm1f/logf/powf/erff."""
    if kind not in _libs:
        name = "liboracle.so" if kind == "cr" else "liboracle_glibc.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        h = C.CDLL(path)
        vp = C.c_void_p
        h.oracle_solve_feautrier_order_2.argtypes = [vp, vp, vp, vp, vp, C.c_int]
        h.oracle_trace_counts.argtypes = [vp, vp, vp, C.c_int]
        h.oracle_trace_ray.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       vp, vp, vp, vp, vp, vp, vp]
        h.oracle_image_new.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp, C.c_int]
        h.oracle_closest_bdy_point.argtypes = [vp, vp]
        h.oracle_closest_bdy_point.restype = C.c_uint32
        h.oracle_feautrier_system.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_uint32, vp, vp, vp, vp, vp, vp, vp]
        for f in ("expf", "expm1f", "logf", "erff"):
            fn = getattr(h, "oracle_" + f)
            fn.argtypes = [C.c_float]
            fn.restype = C.c_float
        h.oracle_J_adjoint.argtypes = [vp, vp, vp, vp, vp, C.c_int]
        h.oracle_image_adjoint.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp,
                                           C.c_int]
        h.oracle_set_smooth_interp.argtypes = [C.c_int]
        h.oracle_set_smooth_interp.restype = None
        h.oracle_powf.argtypes = [C.c_float, C.c_float]
        h.oracle_powf.restype = C.c_float
        _libs[kind] = h
    return _libs[kind]


def _check(code, what):
    if code != 0:
        raise RuntimeError(f"oracle {what} failed with code {code}")


def solve(model, opts, with_u=True, nthreads=0, kind="cr"):
    """Returns (u, J, lambda_diag) of the reference plain method."""
    m = model.view()
    P, F, L, R = model.npoints, model.nfreqs, model.nlines, model.hnrays
    u = np.zeros((R, P, F)) if with_u else None
    J = np.zeros((P, F))
    lam = np.zeros((P, L))
    _check(lib(kind).oracle_solve_feautrier_order_2(
        C.addressof(m), C.addressof(opts), None if u is None else u.ctypes.data,
        J.ctypes.data, lam.ctypes.data, nthreads), "solve")
    return u, J, lam


def trace_counts(model, opts, nthreads=0, kind="cr"):
    m = model.view()
    n = np.zeros((model.hnrays, model.npoints, model.nlines), dtype=np.uint32)
    _check(lib(kind).oracle_trace_counts(C.addressof(m), C.addressof(opts), n.ctypes.data,
                                         nthreads), "trace_counts")
    return n


def trace_ray(model, opts, rr, o, l, capacity=4096, kind="cr"):
    m = model.view()
    cnt = np.zeros(1, dtype=np.uint32)
    cen = np.zeros(1, dtype=np.uint32)
    nr = np.zeros(capacity, dtype=np.uint32)
    nri = np.zeros(capacity, dtype=np.uint32)
    sh = np.zeros(capacity)
    fa = np.zeros(capacity)
    dz = np.zeros(capacity)
    _check(lib(kind).oracle_trace_ray(C.addressof(m), C.addressof(opts), rr, o, l, capacity,
                                      cnt.ctypes.data, cen.ctypes.data, nr.ctypes.data,
                                      nri.ctypes.data, sh.ctypes.data, fa.ctypes.data,
                                      dz.ctypes.data), "trace_ray")
    n = int(cnt[0])
    return dict(count=n, centre=int(cen[0]), nr=nr[:n], nr_interp=nri[:n], shift=sh[:n],
                interp_factor=fa[:n], dZ=dz[:n])


def image_new(model, opts, raydir, nx, ny, freqs=None, nthreads=0, kind="cr"):
    F = model.nfreqs if freqs is None else freqs["nu"].shape[1]
    m = model.view(freqs, nfreqs=F)
    rd = np.ascontiguousarray(raydir, dtype=np.float64)
    I = np.zeros((nx * ny, F))
    ImX = np.zeros(nx * ny)
    ImY = np.zeros(nx * ny)
    _check(lib(kind).oracle_image_new(C.addressof(m), C.addressof(opts), rd.ctypes.data, nx, ny,
                                      I.ctypes.data, ImX.ctypes.data, ImY.ctypes.data, nthreads),
           "image_new")
    return I, ImX, ImY


def feautrier_system(model, opts, rr, o, f, capacity=4096, kind="cr"):
    """dtau (n-1), S (n), I_bdy (2), centre, u_c and L_c of one (rr, o, f) solve."""
    m = model.view()
    n = np.zeros(1, dtype=np.uint32)
    c = np.zeros(1, dtype=np.uint32)
    dtau = np.zeros(capacity)
    S = np.zeros(capacity)
    I = np.zeros(2)
    uc = np.zeros(1)
    lc = np.zeros(1)
    _check(lib(kind).oracle_feautrier_system(C.addressof(m), C.addressof(opts), rr, o, f, capacity,
                                             n.ctypes.data, c.ctypes.data, dtau.ctypes.data,
                                             S.ctypes.data, I.ctypes.data, uc.ctypes.data,
                                             lc.ctypes.data), "feautrier_system")
    k = int(n[0])
    return dict(n=k, centre=int(c[0]), dtau=dtau[:k - 1], S=S[:k], I_bdy=I, u_c=float(uc[0]),
                L_c=float(lc[0]))


def set_smooth_interp(on, kind="cr"):
    """Evaluate interpolate_log in double (a smooth function of its end values),
    for finite-difference checks of the adjoint; off restores the reference's
    float powf."""
    lib(kind).oracle_set_smooth_interp(1 if on else 0)


def J_adjoint(model, opts, J_bar, nthreads=0, kind="cr"):
    """(d_eps, d_kap), each (P, L): the gradient of sum(J_bar * J) with respect to
    the line emissivity and opacity."""
    m = model.view()
    P, L = model.npoints, model.nlines
    jb = np.ascontiguousarray(J_bar, dtype=np.float64)
    assert jb.shape == (P, model.nfreqs)
    de = np.zeros((P, L))
    dk = np.zeros((P, L))
    _check(lib(kind).oracle_J_adjoint(C.addressof(m), C.addressof(opts), jb.ctypes.data,
                                      de.ctypes.data, dk.ctypes.data, nthreads), "J_adjoint")
    return de, dk


def image_adjoint(model, opts, raydir, nx, ny, I_bar, freqs=None, nthreads=0, kind="cr"):
    """(d_eps, d_kap): the gradient of sum(I_bar * I) for the image_new image."""
    F = model.nfreqs if freqs is None else freqs["nu"].shape[1]
    m = model.view(freqs, nfreqs=F)
    rd = np.ascontiguousarray(raydir, dtype=np.float64)
    ib = np.ascontiguousarray(I_bar, dtype=np.float64)
    assert ib.shape == (nx * ny, F)
    de = np.zeros((model.npoints, model.nlines))
    dk = np.zeros((model.npoints, model.nlines))
    _check(lib(kind).oracle_image_adjoint(C.addressof(m), C.addressof(opts), rd.ctypes.data, nx,
                                          ny, ib.ctypes.data, de.ctypes.data, dk.ctypes.data,
                                          nthreads), "image_adjoint")
    return de, dk
