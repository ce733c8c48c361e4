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


# The float transcendentals of each build:
#   'glibc'  the C library's expf/expm1f/logf/powf/erff -- the reference's own
#            calls; the default, so parity is checked against the reference's
#            arithmetic;
#   'ref'    the restatements of those functions that the kernels evaluate
#            (mgx_fmath_ref.h), equal to the C library's on every argument
#            checked (tests/test_fmath.py);
#   'cr'     correctly rounded functions (mgx_fmath.h): how far the results
#            move when only these last bits change.
DEFAULT_KIND = "glibc"
_LIB_NAMES = {"glibc": "liboracle_glibc.so", "ref": "liboracle.so", "cr": "liboracle_cr.so"}


def lib(kind=DEFAULT_KIND):
    """The oracle library with the float functions of `kind` ('glibc', 'ref', 'cr')."""
    if kind not in _libs:
        name = _LIB_NAMES[kind]
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
        h.oracle_J_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int]
        h.oracle_image_adjoint.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp,
                                           C.c_int]
        for f in ("oracle_set_smooth_interp", "oracle_set_smooth_profile"):
            getattr(h, f).argtypes = [C.c_int]
            getattr(h, f).restype = None
        h.oracle_powf.argtypes = [C.c_float, C.c_float]
        h.oracle_powf.restype = C.c_float
        _libs[kind] = h
    return _libs[kind]


def _check(code, what):
    if code != 0:
        raise RuntimeError(f"oracle {what} failed with code {code}")


def solve(model, opts, with_u=True, nthreads=0, kind=DEFAULT_KIND):
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


def shortchar(model, opts, min_dtau=1.0e-15, nthreads=0, kind=DEFAULT_KIND):
    """Model::compute_radiation_field_shortchar_order_0: (I (nrays, P, F), u, v
    (hnrays, P, F), J (P, F), lambda_diag (P, L))."""
    m = model.view()
    P, F, L, R = model.npoints, model.nfreqs, model.nlines, model.hnrays
    I = np.zeros((model.nrays, P, F))
    u = np.zeros((R, P, F))
    v = np.zeros((R, P, F))
    J = np.zeros((P, F))
    lam = np.zeros((P, L))
    h = lib(kind)
    h.oracle_shortchar_order_0.argtypes = [C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 5 + [C.c_int]
    _check(h.oracle_shortchar_order_0(C.addressof(m), C.addressof(opts), min_dtau, I.ctypes.data,
                                      u.ctypes.data, v.ctypes.data, J.ctypes.data, lam.ctypes.data,
                                      nthreads), "shortchar")
    return I, u, v, J, lam


def solve_lambda(model, opts, cap=None, with_u=False, nthreads=0, kind=DEFAULT_KIND):
    """The plain-method solve with the general Lambda operator (opts.n_off_diag >= 0):
    returns (u or None, J, count (P, L), nr (P, L, cap), val (P, L, cap)); the list of
    (p, l) is nr[p, l, :count], val[p, l, :count] in the reference's order
    (Lambda::add_element).  cap defaults to 1 + 2 n_off_diag hnrays."""
    m = model.view()
    P, F, L, R = model.npoints, model.nfreqs, model.nlines, model.hnrays
    cap = 1 + 2 * int(opts.n_off_diag) * R if cap is None else cap
    u = np.zeros((R, P, F)) if with_u else None
    J = np.zeros((P, F))
    cnt = np.zeros((P, L), dtype=np.uint32)
    nr = np.zeros((P, L, cap), dtype=np.uint32)
    val = np.zeros((P, L, cap))
    h = lib(kind)
    h.oracle_solve_feautrier_order_2_lambda.argtypes = [
        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
        C.c_void_p, C.c_int]
    _check(h.oracle_solve_feautrier_order_2_lambda(
        C.addressof(m), C.addressof(opts), u.ctypes.data if with_u else None, J.ctypes.data,
        cap, cnt.ctypes.data, nr.ctypes.data, val.ctypes.data, nthreads), "solve_lambda")
    return u, J, cnt, nr, val


def feautrier_uv(model, opts, nthreads=0, kind=DEFAULT_KIND, with_scale=False):
    """Model::compute_radiation_field_feautrier_order_2_uv: (u, v), each (hnrays, P, F).
    with_scale=True also returns v_scale, the sum of the absolute terms each v is
    formed from (v is a difference of neighbouring u's and cancels)."""
    m = model.view()
    P, F, R = model.npoints, model.nfreqs, model.hnrays
    u = np.zeros((R, P, F))
    v = np.zeros((R, P, F))
    vs = np.zeros((R, P, F))
    h = lib(kind)
    h.oracle_feautrier_order_2_uv.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_int]
    _check(h.oracle_feautrier_order_2_uv(C.addressof(m), C.addressof(opts), u.ctypes.data,
                                         v.ctypes.data, vs.ctypes.data, nthreads),
           "feautrier_uv")
    return (u, v, vs) if with_scale else (u, v)


def trace_counts(model, opts, nthreads=0, kind=DEFAULT_KIND):
    m = model.view()
    n = np.zeros((model.hnrays, model.npoints, model.nlines), dtype=np.uint32)
    _check(lib(kind).oracle_trace_counts(C.addressof(m), C.addressof(opts), n.ctypes.data,
                                         nthreads), "trace_counts")
    return n


def trace_ray(model, opts, rr, o, l, capacity=4096, kind=DEFAULT_KIND):
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


def image_new(model, opts, raydir, nx, ny, freqs=None, nthreads=0, kind=DEFAULT_KIND):
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


def image_optical_depth_new(model, opts, raydir, nx, ny, freqs=None, nthreads=0, kind=DEFAULT_KIND):
    """Model::compute_image_optical_depth_new: (tau (nx*ny, F), ImX, ImY)."""
    F = model.nfreqs if freqs is None else freqs["nu"].shape[1]
    m = model.view(freqs, nfreqs=F)
    rd = np.ascontiguousarray(raydir, dtype=np.float64)
    tau = np.zeros((nx * ny, F))
    ImX = np.zeros(nx * ny)
    ImY = np.zeros(nx * ny)
    h = lib(kind)
    h.oracle_image_optical_depth_new.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                 C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_int]
    _check(h.oracle_image_optical_depth_new(C.addressof(m), C.addressof(opts), rd.ctypes.data, nx,
                                            ny, tau.ctypes.data, ImX.ctypes.data, ImY.ctypes.data,
                                            nthreads), "image_optical_depth_new")
    return tau, ImX, ImY


def image_ray(model, opts, rr, freqs=None, optical_depth=False, nthreads=0, kind=DEFAULT_KIND):
    """Model::compute_image(rr) (or compute_image_optical_depth(rr) with
    optical_depth=True), the old imager at every model point: (out (P, F), ImX, ImY)."""
    F = model.nfreqs if freqs is None else freqs["nu"].shape[1]
    m = model.view(freqs, nfreqs=F)
    P = model.npoints
    out = np.zeros((P, F))
    ImX = np.zeros(P)
    ImY = np.zeros(P)
    h = lib(kind)
    h.oracle_image_ray.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int]
    _check(h.oracle_image_ray(C.addressof(m), C.addressof(opts), rr, int(bool(optical_depth)),
                              out.ctypes.data, ImX.ctypes.data, ImY.ctypes.data, nthreads),
           "image_ray")
    return out, ImX, ImY


def feautrier_system(model, opts, rr, o, f, capacity=4096, kind=DEFAULT_KIND):
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


def set_smooth_interp(on, kind=DEFAULT_KIND):
    """Evaluate interpolate_log in double (a smooth function of its end values),
    for finite-difference checks of the adjoint; off restores the reference's
    float powf."""
    lib(kind).oracle_set_smooth_interp(1 if on else 0)


def set_smooth_profile(on, kind=DEFAULT_KIND):
    """Evaluate the line profile and the error-function difference of the
    line-integrated optical depth in double (smooth in the inverse widths), for
    finite-difference checks of the adjoint; off restores the float expf/erff."""
    lib(kind).oracle_set_smooth_profile(1 if on else 0)


def J_adjoint(model, opts, J_bar, nthreads=0, kind=DEFAULT_KIND, want_iw=False):
    """(d_eps, d_kap), each (P, L): the gradient of sum(J_bar * J) with respect to
    the line emissivity and opacity; with want_iw also d_iw, the gradient with
    respect to the inverse line widths."""
    m = model.view()
    P, L = model.npoints, model.nlines
    jb = np.ascontiguousarray(J_bar, dtype=np.float64)
    assert jb.shape == (P, model.nfreqs)
    de = np.zeros((P, L))
    dk = np.zeros((P, L))
    dw = np.zeros((P, L)) if want_iw else None
    _check(lib(kind).oracle_J_adjoint(C.addressof(m), C.addressof(opts), jb.ctypes.data,
                                      de.ctypes.data, dk.ctypes.data,
                                      dw.ctypes.data if want_iw else None, nthreads), "J_adjoint")
    return (de, dk, dw) if want_iw else (de, dk)


def image_adjoint(model, opts, raydir, nx, ny, I_bar, freqs=None, nthreads=0, kind=DEFAULT_KIND,
                  want_iw=False):
    """(d_eps, d_kap[, d_iw]): the gradient of sum(I_bar * I) for the image_new image."""
    F = model.nfreqs if freqs is None else freqs["nu"].shape[1]
    m = model.view(freqs, nfreqs=F)
    rd = np.ascontiguousarray(raydir, dtype=np.float64)
    ib = np.ascontiguousarray(I_bar, dtype=np.float64)
    assert ib.shape == (nx * ny, F)
    de = np.zeros((model.npoints, model.nlines))
    dk = np.zeros((model.npoints, model.nlines))
    dw = np.zeros((model.npoints, model.nlines)) if want_iw else None
    _check(lib(kind).oracle_image_adjoint(C.addressof(m), C.addressof(opts), rd.ctypes.data, nx,
                                          ny, ib.ctypes.data, de.ctypes.data, dk.ctypes.data,
                                          dw.ctypes.data if want_iw else None, nthreads),
           "image_adjoint")
    return (de, dk, dw) if want_iw else (de, dk)


# ---------------------------------------------------------------------------
# sparse level populations (mgx_oracle_nlte.c)


def _nlte_sigs(h):
    vp = C.c_void_p
    h.oracle_nlte_create.restype = vp
    h.oracle_nlte_create.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp]
    h.oracle_nlte_destroy.restype = None
    h.oracle_nlte_destroy.argtypes = [vp]
    for f, n in (("set_lte", 3), ("set_populations", 4), ("compute_Jeff", 2), ("set_Jeff", 2),
                 ("stateq", 4), ("get", 4), ("stats", 2)):
        fn = getattr(h, "oracle_nlte_" + f)
        fn.argtypes = [vp] * (n + 1)
    h.oracle_nlte_ng.argtypes = [vp, C.c_uint32, vp, vp, vp]
    h.oracle_nlte_trial_ng.argtypes = [vp, C.c_uint32]
    h.oracle_nlte_stateq_system.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp, vp]
    h.oracle_cpqr_solve.argtypes = [C.c_int, vp, vp, vp]
    h.oracle_nlte_compute_Jeff_general.argtypes = [vp, vp, C.c_uint32, vp, vp, vp]
    h.oracle_nlte_stateq_global.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, vp, vp]
    h.oracle_nlte_stateq_global_system.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp]


def cpqr_solve(A, b, kind=DEFAULT_KIND):
    """x = A^-1 b by the long-double column-pivoted Householder QR."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    a = np.ascontiguousarray(A.T)  # column-major
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    h = lib(kind)
    _nlte_sigs(h)
    _check(h.oracle_cpqr_solve(n, a.ctypes.data, bb.ctypes.data, x.ctypes.data), "cpqr_solve")
    return x


class NLTE:
    """Long-double level-population state of a ModelData (the oracle's
    LineProducingSpecies set).  Writes the model's emissivity, opacity and
    interp_crit arrays in place, as Lines::set_emissivity_and_opacity does."""

    def __init__(self, model, params, kind=DEFAULT_KIND):
        from magritte_amd import _abi  # noqa: F401  (ctypes structs)

        self.model = model
        self.params = params
        self.h_lib = lib(kind)
        _nlte_sigs(self.h_lib)
        self._view = model.view()
        ld, self._keep_ld = model._ld_array()
        ncol, cp, self._keep_cp = model._colpar_arrays()
        self._ncol, self._cp = ncol, cp
        self.h = self.h_lib.oracle_nlte_create(C.addressof(self._view), model.nlspecs, ld,
                                               ncol.ctypes.data, cp, model.nspec,
                                               model.abundance.ctypes.data, C.addressof(params))
        if not self.h:
            raise RuntimeError("oracle_nlte_create failed")
        self.nlev = [d["nlev"] for d in model.linedata]

    def close(self):
        if self.h:
            self.h_lib.oracle_nlte_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _eoi(self):
        m = self.model
        return m.emissivity.ctypes.data, m.opacity.ctypes.data, m.interp_crit.ctypes.data

    def set_lte(self):
        _check(self.h_lib.oracle_nlte_set_lte(self.h, *self._eoi()), "nlte_set_lte")

    def set_populations(self, pop):
        pop = np.ascontiguousarray(pop, dtype=np.float64)
        _check(self.h_lib.oracle_nlte_set_populations(self.h, pop.ctypes.data, *self._eoi()),
               "nlte_set_populations")

    def compute_Jeff(self, u, lam):
        u = np.ascontiguousarray(u, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        _check(self.h_lib.oracle_nlte_compute_Jeff(self.h, u.ctypes.data, lam.ctypes.data),
               "nlte_compute_Jeff")

    def set_Jeff(self, Jlin, lam):
        Jlin = np.ascontiguousarray(Jlin, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        _check(self.h_lib.oracle_nlte_set_Jeff(self.h, Jlin.ctypes.data, lam.ctypes.data),
               "nlte_set_Jeff")

    def stateq(self, lam):
        self._lam = np.ascontiguousarray(lam, dtype=np.float64)
        _check(self.h_lib.oracle_nlte_stateq(self.h, self._lam.ctypes.data, *self._eoi()),
               "nlte_stateq")

    def compute_Jeff_general(self, J, lists):
        """Model::compute_Jeff: J (P, F) and the Lambda lists (count, nr, val) of
        solve_lambda."""
        J = np.ascontiguousarray(J, dtype=np.float64)
        cnt, nr, val = (np.ascontiguousarray(a) for a in lists)
        _check(self.h_lib.oracle_nlte_compute_Jeff_general(
            self.h, J.ctypes.data, nr.shape[2], cnt.ctypes.data, nr.ctypes.data, val.ctypes.data),
            "nlte_compute_Jeff_general")

    def stateq_global(self, lists):
        """Lines::iteration_using_statistical_equilibrium: the global update of all points."""
        self._lists = tuple(np.ascontiguousarray(a) for a in lists)
        cnt, nr, val = self._lists
        _check(self.h_lib.oracle_nlte_stateq_global(self.h, nr.shape[2], cnt.ctypes.data,
                                                    nr.ctypes.data, val.ctypes.data, *self._eoi()),
               "nlte_stateq_global")

    def global_system(self, lists, l):
        """The global rate equations of species l: (M (N, N), y), N = P nlev."""
        cnt, nr, val = (np.ascontiguousarray(a) for a in lists)
        N = self.nlev[l] * self.model.npoints
        M = np.zeros(N * N)
        y = np.zeros(N)
        _check(self.h_lib.oracle_nlte_stateq_global_system(
            self.h, nr.shape[2], cnt.ctypes.data, nr.ctypes.data, val.ctypes.data, l,
            M.ctypes.data, y.ctypes.data), "nlte_stateq_global_system")
        return M.reshape(N, N).T.copy(), y

    def ng(self, order):
        _check(self.h_lib.oracle_nlte_ng(self.h, order, *self._eoi()), "nlte_ng")

    def trial_ng(self, order):
        _check(self.h_lib.oracle_nlte_trial_ng(self.h, order), "nlte_trial_ng")

    def populations(self):
        n = sum(self.nlev) * self.model.npoints
        pop = np.zeros(n)
        trial = np.zeros(n)
        P, L = self.model.npoints, self.model.nlines
        Jlin = np.zeros((P, L))
        Jeff = np.zeros((P, L))
        _check(self.h_lib.oracle_nlte_get(self.h, pop.ctypes.data, trial.ctypes.data,
                                          Jlin.ctypes.data, Jeff.ctypes.data), "nlte_get")
        return dict(population=pop, trial=trial, J_lin=Jlin, J_eff=Jeff)

    def stats(self):
        conv = np.zeros((self.model.nlspecs, 3))
        nres = np.zeros(self.model.nlspecs, dtype=np.uint32)
        _check(self.h_lib.oracle_nlte_stats(self.h, conv.ctypes.data, nres.ctypes.data),
               "nlte_stats")
        return conv, nres

    def system(self, lam, l, p):
        n = self.nlev[l]
        M = np.zeros(n * n)
        y = np.zeros(n)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        _check(self.h_lib.oracle_nlte_stateq_system(self.h, lam.ctypes.data, l, p, M.ctypes.data,
                                                    y.ctypes.data), "nlte_stateq_system")
        return M.reshape(n, n).T.copy(), y


def ng_decision(nlte, params, use_ng, prior):
    """Model::ng_acceleration_criterion<Default / Adaptive> (model.cpp:580-645)."""
    skip = params.Ng_acceleration_remove_N_its
    if params.use_adaptive_Ng_acceleration:
        if not use_ng or prior < params.adaptive_Ng_acceleration_min_order + skip:
            return False, 0
        col = 2 if params.adaptive_Ng_acceleration_use_max_criterion else 1
        curr = float(nlte.stats()[0][:, col].sum())
        nlte.trial_ng(prior - skip)
        trial = float(nlte.stats()[0][:, col].sum())
        if trial < curr or prior == params.Ng_acceleration_mem_limit + skip:
            return True, prior - skip
        return False, 0
    if not use_ng:
        return False, 0
    if prior == params.Ng_acceleration_mem_limit + skip:
        return True, params.Ng_acceleration_mem_limit
    return False, 0


def compute_level_populations(model, opts, params, use_ng=True, max_iterations=100, nlte=None,
                              nthreads=0, kind=DEFAULT_KIND):
    """Model::compute_level_populations (model.cpp:653-744) on the CPU: each iteration the
    formal solve with the general Lambda operator (opts.n_off_diag bands), compute_Jeff and
    the global statistical-equilibrium update; Ng steps per the criterion.  Starts from LTE
    unless an NLTE state is given.  Returns (nlte, iterations, error_mean, error_max)."""
    if nlte is None:
        nlte = NLTE(model, params, kind)
        nlte.set_lte()
    iteration = 0
    normal = 0
    err_mean, err_max = [], []
    some_not_converged = True
    while some_not_converged and iteration < max_iterations:
        use_step, order = ng_decision(nlte, params, use_ng, normal)
        if use_step:
            nlte.ng(order)
            normal = 0
            continue
        iteration += 1
        some_not_converged = False
        _, J, cnt, nr, val = solve_lambda(model, opts, nthreads=nthreads, kind=kind)
        nlte.compute_Jeff_general(J, (cnt, nr, val))
        nlte.stateq_global((cnt, nr, val))
        normal += 1
        conv = nlte.stats()[0]
        for l in range(model.nlspecs):
            err_mean.append(float(conv[l, 1]))
            err_max.append(float(conv[l, 2]))
            if conv[l, 0] > 1.0 - params.convergence_fraction:
                some_not_converged = True
    return nlte, iteration, err_mean, err_max


def compute_level_populations_sparse(model, opts, params, use_ng=True, max_iterations=100,
                                     nlte=None, nthreads=0, kind=DEFAULT_KIND):
    """Model::compute_level_populations_sparse (model.cpp:756-850) on the CPU:
    each iteration the plain formal solve (u and the Lambda diagonal), then
    compute_Jeff_sparse and the statistical-equilibrium update; Ng steps per
    the criterion.  Starts from LTE unless an NLTE state is given.  Returns
    (nlte, iterations, error_mean, error_max)."""
    if nlte is None:
        nlte = NLTE(model, params, kind)
        nlte.set_lte()
    iteration = 0
    normal = 0
    err_mean, err_max = [], []
    some_not_converged = True
    while some_not_converged and iteration < max_iterations:
        use_step, order = ng_decision(nlte, params, use_ng, normal)
        if use_step:
            nlte.ng(order)
            normal = 0
            continue
        iteration += 1
        some_not_converged = False
        u, _, lam = solve(model, opts, with_u=True, nthreads=nthreads, kind=kind)
        nlte.compute_Jeff(u, lam)
        nlte.stateq(lam)
        normal += 1
        conv = nlte.stats()[0]
        for l in range(model.nlspecs):
            err_mean.append(float(conv[l, 1]))
            err_max.append(float(conv[l, 2]))
            if conv[l, 0] > 1.0 - params.convergence_fraction:
                some_not_converged = True
    return nlte, iteration, err_mean, err_max


def compute_level_populations_shortchar(model, opts, params, use_ng=True, max_iterations=100,
                                        min_dtau=1.0e-15, nlte=None, nthreads=0,
                                        kind=DEFAULT_KIND):
    """Model::compute_level_populations_shortchar (model.cpp:859-935) on the CPU: each
    iteration the order-0 short-characteristics solve (solver.tpp:151-235), whose Lambda
    is the diagonal whatever opts.n_off_diag says, then Model::compute_Jeff and the
    global statistical-equilibrium update over those one-element lists, as the reference
    does; Ng steps per the criterion.  Starts from LTE unless an NLTE state is given.
    Returns (nlte, iterations, error_mean, error_max)."""
    if nlte is None:
        nlte = NLTE(model, params, kind)
        nlte.set_lte()
    P, L = model.npoints, model.nlines
    cnt = np.ones((P, L), dtype=np.uint32)
    nr = np.repeat(np.arange(P, dtype=np.uint32), L).reshape(P, L, 1)
    iteration = 0
    normal = 0
    err_mean, err_max = [], []
    some_not_converged = True
    while some_not_converged and iteration < max_iterations:
        use_step, order = ng_decision(nlte, params, use_ng, normal)
        if use_step:
            nlte.ng(order)
            normal = 0
            continue
        iteration += 1
        some_not_converged = False
        _, _, _, J, lam = shortchar(model, opts, min_dtau=min_dtau, nthreads=nthreads, kind=kind)
        lists = (cnt, nr, np.ascontiguousarray(lam.reshape(P, L, 1)))
        nlte.compute_Jeff_general(J, lists)
        nlte.stateq_global(lists)
        normal += 1
        conv = nlte.stats()[0]
        for l in range(model.nlspecs):
            err_mean.append(float(conv[l, 1]))
            err_max.append(float(conv[l, 2]))
            if conv[l, 0] > 1.0 - params.convergence_fraction:
                some_not_converged = True
    return nlte, iteration, err_mean, err_max


FN_EXPF, FN_EXPM1F, FN_ERFF, FN_LOGF, FN_POWF = range(5)


def fmath_digest(fn, first, nchunks, chunk_log2, y=0.0, nthreads=0, kind=DEFAULT_KIND):
    """Per-chunk digests of (argument, result) of one float function of this build
    over the float arguments [first, first + nchunks 2^chunk_log2): the host side
    of the device check mgx_fmath_digest."""
    h = lib(kind)
    h.oracle_fmath_digest.argtypes = [C.c_int, C.c_float, C.c_uint64, C.c_uint32, C.c_uint32,
                                      C.c_void_p, C.c_int]
    out = np.zeros(nchunks, dtype=np.uint64)
    _check(h.oracle_fmath_digest(fn, y, first, nchunks, chunk_log2, out.ctypes.data, nthreads),
           "fmath_digest")
    return out
