"""Differentiable formal solve: torch.autograd.Function wrappers whose backward
runs the hand-written adjoint kernels (mgx_feautrier_adjoint_device,
mgx_image_adjoint_device).

The free parameters are the line emissivity and opacity, (P, L) float64
tensors on the model's device; every other model array stays as resident in
the DeviceModel.  The trace, the branch of every segment, the CloseLines
windows and the interpolation counts (from the opacities the model was built
with) are held fixed, so the gradient is that of the solve at a frozen
discrete structure.
"""
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
