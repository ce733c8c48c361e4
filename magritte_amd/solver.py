"""Device entry points of the formal solver (ctypes over include/mgx.h).

Inputs are copied to HBM once (``ModelData.to_device``); every call then runs
the HIP kernels on those resident arrays.  torch is used for device memory
and streams only.  There is no CPU fallback: without the GPU library these
functions raise.
"""
import ctypes as C

import numpy as np

from . import _abi


def _require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("magritte_amd: no GPU visible; the formal solver runs only on the device")
    return torch


class DeviceModel:
    """Model arrays resident on the GPU plus the mgx_model view of them."""

    def __init__(self, model, device="cuda"):
        torch = _require_gpu()
        self.model = model
        self.device = torch.device(device)
        self.view, self.tensors = model.to_device(self.device)

    def stream_ptr(self):
        import torch

        return torch.cuda.current_stream(self.device).cuda_stream


def feautrier_forward(dm, opts, with_u=True, with_lambda=True, out=None, view=None):
    """Plain-method formal solve: returns dict(u=(hnrays,P,F), J=(P,F), lambda_diag=(P,L)).
    ``view`` replaces the resident model's view (e.g. with other emissivities)."""
    torch = _require_gpu()
    m = dm.model
    dev = dm.device
    if out is None:
        out = {}
        out["u"] = torch.empty((m.hnrays, m.npoints, m.nfreqs), dtype=torch.float64,
                               device=dev) if with_u else None
        out["J"] = torch.empty((m.npoints, m.nfreqs), dtype=torch.float64, device=dev)
        out["lambda_diag"] = torch.empty((m.npoints, m.nlines), dtype=torch.float64,
                                         device=dev) if with_lambda else None
    _abi.check(_abi.lib().mgx_feautrier_forward_device(
        C.byref(dm.view if view is None else view), C.byref(opts), _abi.ptr(out["u"]), _abi.ptr(out["J"]),
        _abi.ptr(out["lambda_diag"]), dm.stream_ptr()))
    return out


def ray_pair_view(dm, rr):
    """(view, tensors): the resident model's view restricted to ray pair rr -- its two
    directions and weights, antipods 1 and 0.  A formal solve of the view gives that
    pair's u (as ray 0) and its part of J and of the Lambda diagonal, which add up over
    the pairs to the full solve's (Solver::solve_feautrier_order_2 adds the pairs'
    contributions one pair at a time, solver.tpp:475-520); keep ``tensors`` alive while
    the view is used.  Models with per-point (adaptive) directions are not supported."""
    torch = _require_gpu()
    m = dm.model
    if m.adaptive_directions:
        raise ValueError("ray_pair_view: per-point ray directions are not supported")
    if not 0 <= rr < m.hnrays:
        raise ValueError(f"ray_pair_view: ray pair {rr} out of range (hnrays {m.hnrays})")
    d = m.ray_direction.reshape(-1, 3)
    ar = int(m.antipod[rr])
    t = dict(dm.tensors)
    t["ray_direction"] = torch.from_numpy(np.ascontiguousarray([d[rr], d[ar]])).to(dm.device)
    t["ray_weight"] = torch.from_numpy(np.array([m.ray_weight[rr], m.ray_weight[ar]])).to(dm.device)
    t["antipod"] = torch.from_numpy(np.array([1, 0], dtype=np.uint32)).to(dm.device)
    v = m.view(t)
    v.nrays, v.hnrays = 2, 1
    return v, t


def feautrier_forward_lambda(dm, opts, cap=None, with_u=False):
    """The formal solve with the general Lambda operator (opts.n_off_diag bands on either
    side, up to _abi.MAX_OFF_DIAG): dict(u, J, count (P, L), nr (P, L, cap), val (P, L, cap));
    the list of (p, l) is nr[p, l, :count], val[p, l, :count], diagonal element first.
    cap defaults to 1 + 2 n_off_diag hnrays, which always suffices."""
    torch = _require_gpu()
    m = dm.model
    dev = dm.device
    P, L = m.npoints, m.nlines
    cap = 1 + 2 * int(opts.n_off_diag) * m.hnrays if cap is None else int(cap)
    out = {"u": torch.empty((m.hnrays, P, m.nfreqs), dtype=torch.float64, device=dev)
           if with_u else None,
           "J": torch.empty((P, m.nfreqs), dtype=torch.float64, device=dev),
           "count": torch.empty((P, L), dtype=torch.int32, device=dev),
           "nr": torch.zeros((P, L, cap), dtype=torch.int32, device=dev),
           "val": torch.zeros((P, L, cap), dtype=torch.float64, device=dev)}
    _abi.check(_abi.lib().mgx_feautrier_forward_lambda_device(
        C.byref(dm.view), C.byref(opts), _abi.ptr(out["u"]), _abi.ptr(out["J"]), cap,
        _abi.ptr(out["count"]), _abi.ptr(out["nr"]), _abi.ptr(out["val"]), dm.stream_ptr()))
    return out


def shortchar_forward(dm, opts, min_dtau=1.0e-15, with_I=True, with_uv=True, with_lambda=True):
    """Order-0 short characteristics (Model::compute_radiation_field_shortchar_order_0):
    dict(I=(nrays,P,F), u=(hnrays,P,F), v=(hnrays,P,F), J=(P,F), lambda_diag=(P,L))."""
    torch = _require_gpu()
    m = dm.model
    dev = dm.device
    f64 = dict(dtype=torch.float64, device=dev)
    out = {"I": torch.empty((m.nrays, m.npoints, m.nfreqs), **f64) if with_I else None,
           "u": torch.empty((m.hnrays, m.npoints, m.nfreqs), **f64) if with_uv else None,
           "v": torch.empty((m.hnrays, m.npoints, m.nfreqs), **f64) if with_uv else None,
           "J": torch.empty((m.npoints, m.nfreqs), **f64),
           "lambda_diag": torch.empty((m.npoints, m.nlines), **f64) if with_lambda else None}
    _abi.check(_abi.lib().mgx_shortchar_order_0_device(
        C.byref(dm.view), C.byref(opts), float(min_dtau), _abi.ptr(out["I"]), _abi.ptr(out["u"]),
        _abi.ptr(out["v"]), _abi.ptr(out["J"]), _abi.ptr(out["lambda_diag"]), dm.stream_ptr()))
    return out


def feautrier_uv(dm, opts):
    """Model::compute_radiation_field_feautrier_order_2_uv: dict(u, v), each (hnrays, P, F)."""
    torch = _require_gpu()
    m = dm.model
    f64 = dict(dtype=torch.float64, device=dm.device)
    out = {"u": torch.empty((m.hnrays, m.npoints, m.nfreqs), **f64),
           "v": torch.empty((m.hnrays, m.npoints, m.nfreqs), **f64)}
    _abi.check(_abi.lib().mgx_feautrier_order_2_uv_device(
        C.byref(dm.view), C.byref(opts), _abi.ptr(out["u"]), _abi.ptr(out["v"]), dm.stream_ptr()))
    return out


def trace_counts(dm, opts):
    torch = _require_gpu()
    m = dm.model
    n = torch.empty((m.hnrays, m.npoints, m.nlines), dtype=torch.int32, device=dm.device)
    _abi.check(_abi.lib().mgx_trace_counts_device(C.byref(dm.view), C.byref(opts), _abi.ptr(n),
                                                  dm.stream_ptr()))
    return n.cpu().numpy().view(np.uint32)


def trace_rays(dm, opts, sel, capacity=1024):
    """Expanded traced lists of the (rr, o, l) triples in ``sel`` (array (n,3))."""
    torch = _require_gpu()
    sel = np.ascontiguousarray(sel, dtype=np.uint32)
    n = len(sel)
    dev = dm.device
    s = [torch.from_numpy(np.ascontiguousarray(sel[:, i]).view(np.int32)).to(dev) for i in range(3)]
    cnt = torch.empty(n, dtype=torch.int32, device=dev)
    cen = torch.empty(n, dtype=torch.int32, device=dev)
    nr = torch.zeros((n, capacity), dtype=torch.int32, device=dev)
    nri = torch.zeros((n, capacity), dtype=torch.int32, device=dev)
    sh = torch.zeros((n, capacity), dtype=torch.float64, device=dev)
    fa = torch.zeros((n, capacity), dtype=torch.float64, device=dev)
    dz = torch.zeros((n, capacity), dtype=torch.float64, device=dev)
    _abi.check(_abi.lib().mgx_trace_rays_device(
        C.byref(dm.view), C.byref(opts), n, _abi.ptr(s[0]), _abi.ptr(s[1]), _abi.ptr(s[2]),
        capacity, _abi.ptr(cnt), _abi.ptr(cen), _abi.ptr(nr), _abi.ptr(nri), _abi.ptr(sh),
        _abi.ptr(fa), _abi.ptr(dz), dm.stream_ptr()))
    cnt = cnt.cpu().numpy().view(np.uint32)
    cen = cen.cpu().numpy().view(np.uint32)
    nr, nri = nr.cpu().numpy().view(np.uint32), nri.cpu().numpy().view(np.uint32)
    sh, fa, dz = sh.cpu().numpy(), fa.cpu().numpy(), dz.cpu().numpy()
    res = []
    for i in range(n):
        k = int(cnt[i])
        res.append(dict(count=k, centre=int(cen[i]), nr=nr[i, :k], nr_interp=nri[i, :k],
                        shift=sh[i, :k], interp_factor=fa[i, :k], dZ=dz[i, :k]))
    return res


def trace_stats(dm, opts):
    """(walk steps S, neighbours examined, traced entries T, interpolated entries)."""
    _require_gpu()
    st = np.zeros(4, dtype=np.uint64)
    _abi.check(_abi.lib().mgx_trace_stats_device(C.byref(dm.view), C.byref(opts),
                                                 st.ctypes.data, dm.stream_ptr()))
    return tuple(int(x) for x in st)


def feautrier_adjoint(dm, opts, J_bar, view=None, want_iw=False):
    """(d_eps, d_kap), each (P, L): the gradient of sum(J_bar * J) with respect to
    the line emissivity and opacity (mgx_feautrier_adjoint_device); with want_iw
    also d_iw, the gradient with respect to the inverse line widths
    (mgx_feautrier_adjoint_width_device)."""
    torch = _require_gpu()
    m = dm.model
    jb = J_bar.to(device=dm.device, dtype=torch.float64).contiguous()
    if tuple(jb.shape) != (m.npoints, m.nfreqs):
        raise ValueError(f"J_bar must be ({m.npoints}, {m.nfreqs})")
    de = torch.empty((m.npoints, m.nlines), dtype=torch.float64, device=dm.device)
    dk = torch.empty_like(de)
    v = dm.view if view is None else view
    if want_iw:
        dw = torch.empty_like(de)
        _abi.check(_abi.lib().mgx_feautrier_adjoint_width_device(
            C.byref(v), C.byref(opts), _abi.ptr(jb), _abi.ptr(de), _abi.ptr(dk), _abi.ptr(dw),
            dm.stream_ptr()))
        return de, dk, dw
    _abi.check(_abi.lib().mgx_feautrier_adjoint_device(
        C.byref(v), C.byref(opts), _abi.ptr(jb), _abi.ptr(de), _abi.ptr(dk), dm.stream_ptr()))
    return de, dk


def image_view(dm, freqs, arrays=None):
    """mgx_model view of the resident model with the image frequency grid ``freqs``
    (and any array replaced from ``arrays``)."""
    torch = _require_gpu()
    m = dm.model
    t = dict(dm.tensors)
    for k in m.FREQ_ARRAYS:
        t[k] = torch.from_numpy(np.ascontiguousarray(freqs[k])).to(dm.device)
    if arrays:
        t.update(arrays)
    return m.view(t, nfreqs=freqs["nu"].shape[1]), t


def image_adjoint(dm, opts, raydir, nx, ny, freqs, I_bar, view=None, want_iw=False):
    """(d_eps, d_kap[, d_iw]): the gradient of sum(I_bar * I) for image_new's I
    (mgx_image_adjoint_device, or mgx_image_adjoint_width_device with want_iw)."""
    torch = _require_gpu()
    m = dm.model
    F = freqs["nu"].shape[1]
    ib = I_bar.to(device=dm.device, dtype=torch.float64).contiguous()
    if tuple(ib.shape) != (nx * ny, F):
        raise ValueError(f"I_bar must be ({nx * ny}, {F})")
    if view is None:
        view, keep = image_view(dm, freqs)
    de = torch.empty((m.npoints, m.nlines), dtype=torch.float64, device=dm.device)
    dk = torch.empty_like(de)
    rd = np.ascontiguousarray(raydir, dtype=np.float64)
    if want_iw:
        dw = torch.empty_like(de)
        _abi.check(_abi.lib().mgx_image_adjoint_width_device(
            C.byref(view), C.byref(opts), rd.ctypes.data, nx, ny, _abi.ptr(ib), _abi.ptr(de),
            _abi.ptr(dk), _abi.ptr(dw), dm.stream_ptr()))
        return de, dk, dw
    _abi.check(_abi.lib().mgx_image_adjoint_device(
        C.byref(view), C.byref(opts), rd.ctypes.data, nx, ny, _abi.ptr(ib), _abi.ptr(de),
        _abi.ptr(dk), dm.stream_ptr()))
    return de, dk


def image_new(dm, opts, raydir, nx, ny, freqs, view=None):
    """Projection-surface image along raydir (Model::compute_image_new): returns
    (I (nx*ny, F), ImX, ImY) as torch tensors; ``freqs`` is the image frequency
    grid from ModelData.image_frequencies."""
    torch = _require_gpu()
    dev = dm.device
    F = freqs["nu"].shape[1]
    if view is None:
        view, keep = image_view(dm, freqs)
    I = torch.empty((nx * ny, F), dtype=torch.float64, device=dev)
    ImX = torch.empty(nx * ny, dtype=torch.float64, device=dev)
    ImY = torch.empty(nx * ny, dtype=torch.float64, device=dev)
    rd = np.ascontiguousarray(raydir, dtype=np.float64)
    _abi.check(_abi.lib().mgx_image_device(C.byref(view), C.byref(opts), rd.ctypes.data, nx, ny,
                                           _abi.ptr(I), _abi.ptr(ImX), _abi.ptr(ImY),
                                           dm.stream_ptr()))
    return I, ImX, ImY


def image_optical_depth_new(dm, opts, raydir, nx, ny, freqs, view=None):
    """Optical-depth image on the projection surface (Model::compute_image_optical_depth_new):
    (tau (nx*ny, F), ImX, ImY) as torch tensors, the pixel rays of image_new."""
    torch = _require_gpu()
    dev = dm.device
    F = freqs["nu"].shape[1]
    if view is None:
        view, keep = image_view(dm, freqs)
    tau = torch.empty((nx * ny, F), dtype=torch.float64, device=dev)
    ImX = torch.empty(nx * ny, dtype=torch.float64, device=dev)
    ImY = torch.empty(nx * ny, dtype=torch.float64, device=dev)
    rd = np.ascontiguousarray(raydir, dtype=np.float64)
    _abi.check(_abi.lib().mgx_image_optical_depth_device(
        C.byref(view), C.byref(opts), rd.ctypes.data, nx, ny, _abi.ptr(tau), _abi.ptr(ImX),
        _abi.ptr(ImY), dm.stream_ptr()))
    return tau, ImX, ImY


def image_ray(dm, opts, rr, freqs=None, optical_depth=False, view=None):
    """The old imager (Model::compute_image(rr), or compute_image_optical_depth(rr) with
    optical_depth=True): an image at every model point along model ray rr.  Returns
    (I or tau (P, F), ImX, ImY) as torch tensors; ``freqs`` is an image frequency grid
    (default: the model's own frequencies)."""
    torch = _require_gpu()
    m = dm.model
    dev = dm.device
    if view is None:
        if freqs is None:
            view = dm.view
        else:
            view, keep = image_view(dm, freqs)
    F = view.nfreqs
    out = torch.empty((m.npoints, F), dtype=torch.float64, device=dev)
    ImX = torch.empty(m.npoints, dtype=torch.float64, device=dev)
    ImY = torch.empty(m.npoints, dtype=torch.float64, device=dev)
    fn = (_abi.lib().mgx_image_optical_depth_ray_device if optical_depth
          else _abi.lib().mgx_image_ray_device)
    _abi.check(fn(C.byref(view), C.byref(opts), int(rr), _abi.ptr(out), _abi.ptr(ImX),
                  _abi.ptr(ImY), dm.stream_ptr()))
    return out, ImX, ImY


class BatchPlan:
    """Many models solved together in one persistent launch (mgx_batch_*): the
    models' work items go through a single atomic queue, runs replay a captured
    HIP graph, and tune() picks the tiling (ray pairs per work item, item order,
    register budget) by timing a sample of the work.  All DeviceModels must be
    on the same device and use the same options."""

    def __init__(self, dms, opts, with_lambda=False):
        torch = _require_gpu()
        self.dms = list(dms)
        self.opts = opts
        n = len(self.dms)
        dev = self.dms[0].device
        self.views = (_abi.MgxModel * n)(*[dm.view for dm in self.dms])
        self.J = [torch.empty((dm.model.npoints, dm.model.nfreqs), dtype=torch.float64, device=dev)
                  for dm in self.dms]
        self.lam = ([torch.empty((dm.model.npoints, dm.model.nlines), dtype=torch.float64,
                                 device=dev) for dm in self.dms] if with_lambda else None)
        self._Jp = (C.c_void_p * n)(*[t.data_ptr() for t in self.J])
        self._Lp = (C.c_void_p * n)(*[t.data_ptr() for t in self.lam]) if with_lambda else None
        h = C.c_void_p()
        _abi.check(_abi.lib().mgx_batch_create(self.views, n, C.byref(opts), self._Jp, self._Lp,
                                               C.byref(h), self.dms[0].stream_ptr()))
        self._h = h

    def configure(self, rays_per_item=0, order=0, wps=3, use_graph=True):
        _abi.check(_abi.lib().mgx_batch_configure(self._h, rays_per_item, order, wps,
                                                  1 if use_graph else 0,
                                                  self.dms[0].stream_ptr()))

    def tune(self, sample_stride=8):
        best = np.zeros(3, dtype=np.uint32)
        _abi.check(_abi.lib().mgx_batch_tune(self._h, sample_stride, best.ctypes.data,
                                             self.dms[0].stream_ptr()))
        return dict(rays_per_item=int(best[0]), order=int(best[1]), wps=int(best[2]))

    def set_images(self, raydirs, nx, ny, freqs):
        """Adds an nx x ny image of every model (mgx_batch_set_images): model i along
        raydirs[i] at the frequency grid freqs[i] (ModelData.image_frequencies);
        run() then computes them after the solve.  The images are self.images, a
        list of (I (nx*ny, F_i), ImX, ImY) tensors."""
        torch = _require_gpu()
        n = len(self.dms)
        if len(raydirs) != n or len(freqs) != n:
            raise ValueError("one ray direction and one frequency grid per model")
        dev = self.dms[0].device
        views, self._img_keep = [], []
        self.images = []
        for dm, fr in zip(self.dms, freqs):
            v, keep = image_view(dm, fr)
            views.append(v)
            self._img_keep.append(keep)
            F = fr["nu"].shape[1]
            self.images.append((torch.empty((nx * ny, F), dtype=torch.float64, device=dev),
                                torch.empty(nx * ny, dtype=torch.float64, device=dev),
                                torch.empty(nx * ny, dtype=torch.float64, device=dev)))
        self._img_views = (_abi.MgxModel * n)(*views)
        self._rd = np.ascontiguousarray(np.asarray(raydirs, dtype=np.float64).reshape(n, 3))
        ptrs = [(C.c_void_p * n)(*[im[k].data_ptr() for im in self.images]) for k in range(3)]
        self._img_ptrs = ptrs
        _abi.check(_abi.lib().mgx_batch_set_images(self._h, self._img_views, self._rd.ctypes.data,
                                                   nx, ny, ptrs[0], ptrs[1], ptrs[2],
                                                   self.dms[0].stream_ptr()))

    def run(self):
        """Solves every model (and images them, after set_images); returns the list of
        J tensors (and Lambda, if kept)."""
        _abi.check(_abi.lib().mgx_batch_run(self._h, self.dms[0].stream_ptr()))
        return self.J if self.lam is None else (self.J, self.lam)

    def last_ms(self):
        return _abi.lib().mgx_batch_last_ms(self._h)

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().mgx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LevelPopulations:
    """Device level-population state (mgx_nlte) of a DeviceModel, with the
    reference's LineProducingSpecies operations: set_lte (update_using_LTE),
    stateq (compute_Jeff_sparse + iteration_using_statistical_equilibrium_sparse),
    ng / trial_ng (iteration_using_Ng_acceleration, the adaptive trial).  Every
    update rewrites the DeviceModel's emissivity, opacity and interp_crit."""

    def __init__(self, dm, params=None):
        torch = _require_gpu()
        m = dm.model
        self.dm = dm
        self.params = _abi.default_nlte_params() if params is None else params
        ld, self._keep_ld = m._ld_array()
        ncol, cp, self._keep_cp = m._colpar_arrays()
        self.abundance = torch.from_numpy(np.ascontiguousarray(m.abundance)).to(dm.device)
        self.nlev = [d["nlev"] for d in m.linedata]
        h = C.c_void_p()
        _abi.check(_abi.lib().mgx_nlte_create(C.byref(dm.view), m.nlspecs, ld, _abi.ptr(ncol), cp,
                                              m.nspec, _abi.ptr(self.abundance),
                                              C.byref(self.params), C.byref(h), dm.stream_ptr()))
        self.h = h

    def close(self):
        if self.h:
            _abi.lib().mgx_nlte_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _eoi(self):
        t = self.dm.tensors
        return _abi.ptr(t["emissivity"]), _abi.ptr(t["opacity"]), _abi.ptr(t["interp_crit"])

    def _conv(self, c):
        return [(c[s].fraction_not_converged, c[s].relative_change_mean, c[s].relative_change_max)
                for s in range(len(self.nlev))]

    def set_lte(self):
        _abi.check(_abi.lib().mgx_nlte_set_lte(self.h, *self._eoi(), self.dm.stream_ptr()))

    def set_populations(self, pop):
        torch = _require_gpu()
        p = torch.as_tensor(np.ascontiguousarray(pop, dtype=np.float64)).to(self.dm.device)
        _abi.check(_abi.lib().mgx_nlte_set_populations(self.h, _abi.ptr(p), *self._eoi(),
                                                       self.dm.stream_ptr()))

    def stateq(self, J, lambda_diag, want_J=False):
        """One statistical-equilibrium iteration from a formal solve's J (P, F) and
        lambda_diag (P, L); returns the convergence statistics per species (and
        J_lin, J_eff (P, L) when want_J)."""
        torch = _require_gpu()
        m = self.dm.model
        conv = (_abi.MgxNlteConv * len(self.nlev))()
        Jl = Je = None
        if want_J:
            Jl = torch.empty((m.npoints, m.nlines), dtype=torch.float64, device=self.dm.device)
            Je = torch.empty_like(Jl)
        _abi.check(_abi.lib().mgx_nlte_iteration_stateq(self.h, _abi.ptr(J), _abi.ptr(lambda_diag),
                                                        *self._eoi(), conv, _abi.ptr(Jl),
                                                        _abi.ptr(Je), self.dm.stream_ptr()))
        return (self._conv(conv), Jl, Je) if want_J else self._conv(conv)

    def stateq_global(self, J, lam, want_J=False):
        """One global statistical-equilibrium iteration (Lines::iteration_using_statistical_
        equilibrium with Model::compute_Jeff) from J (P, F) and the Lambda lists ``lam`` of
        feautrier_forward_lambda (a dict with count, nr, val); returns the convergence
        statistics per species and the number of block-Jacobi sweeps (and J_lin, J_eff)."""
        torch = _require_gpu()
        m = self.dm.model
        conv = (_abi.MgxNlteConv * len(self.nlev))()
        sweeps = np.zeros(1, dtype=np.uint32)
        Jl = Je = None
        if want_J:
            Jl = torch.empty((m.npoints, m.nlines), dtype=torch.float64, device=self.dm.device)
            Je = torch.empty_like(Jl)
        _abi.check(_abi.lib().mgx_nlte_iteration_stateq_global(
            self.h, _abi.ptr(J), int(lam["nr"].shape[2]), _abi.ptr(lam["count"]),
            _abi.ptr(lam["nr"]), _abi.ptr(lam["val"]), *self._eoi(), conv, _abi.ptr(Jl),
            _abi.ptr(Je), sweeps.ctypes.data, self.dm.stream_ptr()))
        out = (self._conv(conv), int(sweeps[0]))
        return out + (Jl, Je) if want_J else out

    def ng(self, order):
        conv = (_abi.MgxNlteConv * len(self.nlev))()
        _abi.check(_abi.lib().mgx_nlte_iteration_ng(self.h, order, *self._eoi(), conv,
                                                    self.dm.stream_ptr()))
        return self._conv(conv)

    def trial_ng(self, order):
        conv = (_abi.MgxNlteConv * len(self.nlev))()
        _abi.check(_abi.lib().mgx_nlte_trial_ng(self.h, order, conv, self.dm.stream_ptr()))
        return self._conv(conv)

    def populations(self):
        """(populations, population_tot): species blocks of (P, nlev) and (nlspecs, P)."""
        torch = _require_gpu()
        m = self.dm.model
        pop = torch.empty(sum(self.nlev) * m.npoints, dtype=torch.float64, device=self.dm.device)
        tot = torch.empty((m.nlspecs, m.npoints), dtype=torch.float64, device=self.dm.device)
        _abi.check(_abi.lib().mgx_nlte_get_populations(self.h, _abi.ptr(pop), _abi.ptr(tot),
                                                       self.dm.stream_ptr()))
        return pop, tot

    def n_residuals(self):
        n = np.zeros(1, dtype=np.uint32)
        _abi.check(_abi.lib().mgx_nlte_history(self.h, n.ctypes.data))
        return int(n[0])


def compute_level_populations(dm, opts, lp, use_ng=True, max_iterations=100):
    """Model::compute_level_populations on the device: the loop with the general Lambda
    operator (opts.n_off_diag bands) and the global update; returns (iterations,
    error_mean, error_max)."""
    m = dm.model
    err_mean = np.zeros(max(1, max_iterations) * m.nlspecs)
    err_max = np.zeros_like(err_mean)
    it = np.zeros(1, dtype=np.int32)
    t = dm.tensors
    _abi.check(_abi.lib().mgx_compute_level_populations(
        C.byref(dm.view), C.byref(opts), lp.h, 1 if use_ng else 0, int(max_iterations),
        _abi.ptr(t["emissivity"]), _abi.ptr(t["opacity"]), _abi.ptr(t["interp_crit"]),
        it.ctypes.data, err_mean.ctypes.data, err_max.ctypes.data, dm.stream_ptr()))
    n = int(it[0]) * m.nlspecs
    return int(it[0]), err_mean[:n].tolist(), err_max[:n].tolist()


def compute_level_populations_shortchar(dm, opts, lp, use_ng=True, max_iterations=100,
                                        min_dtau=1.0e-15):
    """Model::compute_level_populations_shortchar on the device: the loop with the
    order-0 short-characteristics formal solve; returns (iterations, error_mean,
    error_max)."""
    m = dm.model
    err_mean = np.zeros(max(1, max_iterations) * m.nlspecs)
    err_max = np.zeros_like(err_mean)
    it = np.zeros(1, dtype=np.int32)
    t = dm.tensors
    _abi.check(_abi.lib().mgx_compute_level_populations_shortchar(
        C.byref(dm.view), C.byref(opts), lp.h, 1 if use_ng else 0, int(max_iterations),
        float(min_dtau), _abi.ptr(t["emissivity"]), _abi.ptr(t["opacity"]),
        _abi.ptr(t["interp_crit"]), it.ctypes.data, err_mean.ctypes.data, err_max.ctypes.data,
        dm.stream_ptr()))
    n = int(it[0]) * m.nlspecs
    return int(it[0]), err_mean[:n].tolist(), err_max[:n].tolist()


def compute_level_populations_sparse(dm, opts, lp, use_ng=True, max_iterations=100):
    """Model::compute_level_populations_sparse on the device (the native loop of
    libmgx): returns (iterations, error_mean, error_max)."""
    m = dm.model
    err_mean = np.zeros(max(1, max_iterations) * m.nlspecs)
    err_max = np.zeros_like(err_mean)
    it = np.zeros(1, dtype=np.int32)
    t = dm.tensors
    _abi.check(_abi.lib().mgx_compute_level_populations_sparse(
        C.byref(dm.view), C.byref(opts), lp.h, 1 if use_ng else 0, int(max_iterations),
        _abi.ptr(t["emissivity"]), _abi.ptr(t["opacity"]), _abi.ptr(t["interp_crit"]),
        it.ctypes.data, err_mean.ctypes.data, err_max.ctypes.data, dm.stream_ptr()))
    n = int(it[0]) * m.nlspecs
    return int(it[0]), err_mean[:n].tolist(), err_max[:n].tolist()
