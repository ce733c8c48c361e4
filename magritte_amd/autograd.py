"""Differentiable formal solve: torch.autograd.Function wrappers whose backward
runs the hand-written adjoint kernels (mgx_feautrier_adjoint_device,
mgx_image_adjoint_device and their _width variants).

The free parameters are the line emissivity and opacity, (P, L) float64
tensors on the model's device, and optionally the inverse line widths; every
other model array stays as resident in the DeviceModel.  LineThermo maps the
log-density and temperature of every point to those three, so images can be
differentiated with respect to (log n, T).  The trace, the branch of every
segment, the CloseLines windows and the interpolation counts (from the
opacities the model was built with) are held fixed, so the gradient is that of
the solve at a frozen discrete structure.
"""
import numpy as np
import torch

from . import solver


def _lines(dm, eps, kap):
    """the two parameter arrays as contiguous float64 device tensors"""
    return {"emissivity": eps.detach().to(dm.device, torch.float64).contiguous(),
            "opacity": kap.detach().to(dm.device, torch.float64).contiguous()}


class FormalSolveJ(torch.autograd.Function):
    """J (P, F) of the plain-method formal solve as a function of (eps, kap)."""

    @staticmethod
    def forward(ctx, eps, kap, dm, opts):
        arrays = dict(dm.tensors)
        arrays.update(_lines(dm, eps, kap))
        view = dm.model.view(arrays)
        out = solver.feautrier_forward(dm, opts, with_u=False, with_lambda=False, view=view)
        ctx.dm, ctx.opts, ctx.view, ctx.arrays = dm, opts, view, arrays
        return out["J"]

    @staticmethod
    def backward(ctx, gJ):
        de, dk = solver.feautrier_adjoint(ctx.dm, ctx.opts, gJ.contiguous(), view=ctx.view)
        return de, dk, None, None


class ImageNew(torch.autograd.Function):
    """Projection-surface image I (nx*ny, F) as a function of (eps, kap)."""

    @staticmethod
    def forward(ctx, eps, kap, dm, opts, raydir, nx, ny, freqs):
        view, arrays = solver.image_view(dm, freqs, _lines(dm, eps, kap))
        I, _, _ = solver.image_new(dm, opts, raydir, nx, ny, freqs, view=view)
        ctx.dm, ctx.opts, ctx.view, ctx.arrays = dm, opts, view, arrays
        ctx.raydir, ctx.nx, ctx.ny, ctx.freqs = raydir, nx, ny, freqs
        return I

    @staticmethod
    def backward(ctx, gI):
        de, dk = solver.image_adjoint(ctx.dm, ctx.opts, ctx.raydir, ctx.nx, ctx.ny, ctx.freqs,
                                      gI.contiguous(), view=ctx.view)
        return de, dk, None, None, None, None, None, None


def formal_solve_J(eps, kap, dm, opts):
    return FormalSolveJ.apply(eps, kap, dm, opts)


def image(eps, kap, dm, opts, raydir, nx, ny, freqs):
    return ImageNew.apply(eps, kap, dm, opts, raydir, nx, ny, freqs)


# ---------------------------------------------------------------------------
# line data from the density and temperature of every point

HH_OVER_FOUR_PI = 6.62607004e-34 / (4.0 * 3.14159265358979323846)
KB = 1.38064852e-23
TWO_KB_OVER_AMU_CC_SQUARED = 2.0 * KB / (1.66053904e-27 * 8.98755179e16)


class LineThermo:
    """The line emissivity, opacity and inverse width of every (point, line) as a
    differentiable function of the log-density and the temperature of every point,
    as the reference's producers compute them:

    * inverse width: Lines::set_inverse_width (src/model/lines/lines.tpp:45-60),
      1 / (nu_l sqrt(2 kB T / (amu c^2) / m_s + vturb^2 / c^2));
    * populations, ``populations="lte"``: LineProducingSpecies::update_using_LTE
      (lineProducingSpecies.tpp:52-75), n_i = n_s g_i exp(-E_i / kB T) / Z(T);
      ``populations="fixed"``: n_i = n_s x_i with the level fractions x_i held at
      ``fractions`` (per species (P, nlev), e.g. converged non-LTE populations
      divided by their totals), so that the temperature enters through the widths;
    * eps = nu h/(4 pi) A n_u and kap = nu h/(4 pi) (n_l B_lu - n_u B_ul), floored at
      min_line_emissivity / min_line_opacity (Lines::set_emissivity_and_opacity,
      lines.tpp:29-41).

    The density scales every abundance of a point: n_s(p) = abundance(p, num_s) n(p) / n0(p)
    with n0 the model's column ``density_column`` of the abundance table (the H2
    density of the reference's models).  Evaluated in float64 torch arithmetic on
    ``device`` (the reference forms the populations in long double; the two agree to
    ~1e-15 relative).
    """

    def __init__(self, model, device, populations="lte", fractions=None, density_column=0):
        if populations not in ("lte", "fixed"):
            raise ValueError("populations must be 'lte' or 'fixed'")
        if populations == "fixed" and fractions is None:
            raise ValueError("populations='fixed' needs the level fractions")
        t = lambda x, dt=torch.float64: torch.as_tensor(np.asarray(x), dtype=dt, device=device)
        self.model, self.device, self.populations = model, device, populations
        self.abundance = t(model.abundance)
        self.n0 = self.abundance[:, density_column].clone()
        self.vturb2 = t(model.vturb2)
        self.min_eps = float(model.params["min_line_emissivity"])
        self.min_kap = float(model.params["min_line_opacity"])
        self.species = []
        for s, d in enumerate(model.linedata):
            sp = dict(num=int(d["num"]), inverse_mass=float(d["inverse_mass"]),
                      energy=t(d["energy"]), weight=t(d["weight"]),
                      irad=t(d["irad"], torch.long), jrad=t(d["jrad"], torch.long),
                      A=t(d["A"]), Bs=t(d["Bs"]), Ba=t(d["Ba"]), freq=t(d["frequency"]))
            if populations == "fixed":
                x = t(fractions[s])
                if tuple(x.shape) != (model.npoints, int(d["nlev"])):
                    raise ValueError(f"fractions[{s}] must be ({model.npoints}, {d['nlev']})")
                sp["fractions"] = x
            self.species.append(sp)

    def log_density(self):
        """log n0 of the model (the point at which the map reproduces its abundances)"""
        return torch.log(self.n0)

    def temperature(self):
        return torch.as_tensor(self.model.temperature, dtype=torch.float64, device=self.device)

    def __call__(self, log_n, T):
        """(eps, kap, iw), each (P, L), from log_n (P) and T (P)."""
        scale = torch.exp(log_n) / self.n0
        eps, kap, iw = [], [], []
        for sp in self.species:
            ptot = self.abundance[:, sp["num"]] * scale
            if self.populations == "lte":
                w = sp["weight"][None, :] * torch.exp(-sp["energy"][None, :] / (KB * T[:, None]))
                pop = w * (ptot / w.sum(dim=1))[:, None]
            else:
                pop = sp["fractions"] * ptot[:, None]
            ni, nj = pop[:, sp["irad"]], pop[:, sp["jrad"]]
            c = sp["freq"][None, :] * HH_OVER_FOUR_PI
            eps.append(torch.clamp(c * sp["A"][None, :] * ni, min=self.min_eps))
            kap.append(torch.clamp(c * (nj * sp["Ba"][None, :] - ni * sp["Bs"][None, :]),
                                   min=self.min_kap))
            width = sp["freq"][None, :] * torch.sqrt(
                TWO_KB_OVER_AMU_CC_SQUARED * sp["inverse_mass"] * T[:, None] + self.vturb2[:, None])
            iw.append(1.0 / width)
        return torch.cat(eps, dim=1), torch.cat(kap, dim=1), torch.cat(iw, dim=1)


def _lines3(dm, eps, kap, iw):
    d = _lines(dm, eps, kap)
    d["inverse_width"] = iw.detach().to(dm.device, torch.float64).contiguous()
    return d


class FormalSolveJWidth(torch.autograd.Function):
    """J (P, F) as a function of (eps, kap, iw).  The model's quadrature frequencies
    (placed by the widths of its own temperatures) are held fixed."""

    @staticmethod
    def forward(ctx, eps, kap, iw, dm, opts):
        arrays = dict(dm.tensors)
        arrays.update(_lines3(dm, eps, kap, iw))
        view = dm.model.view(arrays)
        out = solver.feautrier_forward(dm, opts, with_u=False, with_lambda=False, view=view)
        ctx.dm, ctx.opts, ctx.view, ctx.arrays = dm, opts, view, arrays
        return out["J"]

    @staticmethod
    def backward(ctx, gJ):
        de, dk, dw = solver.feautrier_adjoint(ctx.dm, ctx.opts, gJ.contiguous(), view=ctx.view,
                                              want_iw=True)
        return de, dk, dw, None, None


class ImageNewWidth(torch.autograd.Function):
    """Projection-surface image I (nx*ny, F) as a function of (eps, kap, iw)."""

    @staticmethod
    def forward(ctx, eps, kap, iw, dm, opts, raydir, nx, ny, freqs):
        view, arrays = solver.image_view(dm, freqs, _lines3(dm, eps, kap, iw))
        I, _, _ = solver.image_new(dm, opts, raydir, nx, ny, freqs, view=view)
        ctx.dm, ctx.opts, ctx.view, ctx.arrays = dm, opts, view, arrays
        ctx.raydir, ctx.nx, ctx.ny, ctx.freqs = raydir, nx, ny, freqs
        return I

    @staticmethod
    def backward(ctx, gI):
        de, dk, dw = solver.image_adjoint(ctx.dm, ctx.opts, ctx.raydir, ctx.nx, ctx.ny, ctx.freqs,
                                          gI.contiguous(), view=ctx.view, want_iw=True)
        return de, dk, dw, None, None, None, None, None, None


def formal_solve_J_width(eps, kap, iw, dm, opts):
    return FormalSolveJWidth.apply(eps, kap, iw, dm, opts)


def image_width(eps, kap, iw, dm, opts, raydir, nx, ny, freqs):
    return ImageNewWidth.apply(eps, kap, iw, dm, opts, raydir, nx, ny, freqs)


def image_log_density_temperature(log_n, T, thermo, dm, opts, raydir, nx, ny, freqs):
    """The image as a function of the log-density and temperature of every point
    (through LineThermo ``thermo``).  The trace, every segment's branch, the
    CloseLines windows, the interpolation counts and the image frequencies are
    held at the model's own values, so the gradient is that of the image at a
    frozen discrete structure."""
    eps, kap, iw = thermo(log_n, T)
    return ImageNewWidth.apply(eps, kap, iw, dm, opts, raydir, nx, ny, freqs)
