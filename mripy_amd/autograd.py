"""Differentiable NUFFT for PyTorch.

``nufft_forward(image, coords, plan)`` is y = A(x) image (reference type 2) and
``nufft_adjoint(y, coords, plan)`` is image = A(x)^H y (the exact adjoint, no
1/M, no density weighting).  Their hand-written backward passes are

* image / data gradients: the other operator (A^H for the forward map, A for
  the adjoint map) -- PyTorch's convention for C-linear maps;
* trajectory gradients: the ``gn_traj_grad`` kernel, the analytic derivative of
  the Gaussian interpolation kernel, contracted with the incoming gradient
  (forward map) or with the data (adjoint map).

``coords`` is a (d, M) real tensor in radians; float32 is used on the device.
"""
from __future__ import annotations

import torch

from .plan import Plan


def _bind(plan, coords):
    """Make sure the plan holds exactly these coordinates.  The plan keeps a reference
    to the tensor it binned (so its memory cannot be reused by another tensor while
    bound) and the value of its version counter: another tensor, or an in-place
    update of this one (or of a tensor it views), makes the plan re-bin."""
    if getattr(plan, "_bound_coords", None) is not coords or plan._bound_ver != coords._version:
        c32 = coords.detach().to(torch.float32).contiguous()
        # keep the plan's df and density weights (the exact adjoint ignores the weights)
        plan.set_points([c32[i] for i in range(c32.shape[0])], df=plan.df, dcf=plan._dcf)
        plan._bound_coords = coords
        plan._bound_ver = coords._version


class _Forward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, coords, plan):
        _bind(plan, coords)
        y = plan.type2(image.detach().contiguous())
        ctx.plan = plan
        ctx.coords = coords
        ctx.save_for_backward(image)
        return y

    @staticmethod
    def backward(ctx, gy):
        (image,) = ctx.saved_tensors
        plan = ctx.plan
        _bind(plan, ctx.coords)
        gy = gy.contiguous()
        g_crd = None
        if ctx.needs_input_grad[1]:
            # first: when nothing ran on the plan since the forward, its grid still holds
            # the type-2 grid of image (saved, so alive and unmodified) and is reused
            g_crd = plan.traj_grad(image.detach(), gy, reuse_grid=True)
        g_img = plan.adjoint(gy) if ctx.needs_input_grad[0] else None
        return g_img, g_crd, None


class _Adjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, coords, plan):
        _bind(plan, coords)
        img = plan.adjoint(y.detach().contiguous())
        ctx.plan = plan
        ctx.coords = coords
        ctx.save_for_backward(y)
        return img

    @staticmethod
    def backward(ctx, g_img):
        (y,) = ctx.saved_tensors
        plan = ctx.plan
        _bind(plan, ctx.coords)
        g_img = g_img.contiguous()
        g_y = plan.type2(g_img) if ctx.needs_input_grad[0] else None
        g_crd = None
        if ctx.needs_input_grad[1]:
            # the type 2 just above left g_img's grid in the plan
            g_crd = plan.traj_grad(g_img, y.detach(), reuse_grid=True)
        return g_y, g_crd, None


def nufft_forward(image, coords, plan):
    """Differentiable y = A(coords) image."""
    return _Forward.apply(image, coords, plan)


def nufft_adjoint(y, coords, plan):
    """Differentiable image = A(coords)^H y."""
    return _Adjoint.apply(y, coords, plan)


class NUFFT(torch.nn.Module):
    """Module holding one plan: ``forward(image, coords)`` and ``adjoint(y, coords)``."""

    def __init__(self, modes, ncoils=1, rat=2, nspread=3, eps=None, iflag=1, df=1.0, device=None, **kw):
        super().__init__()
        self.plan = Plan(modes, ncoils=ncoils, rat=None if eps is not None else rat,
                         nspread=None if eps is not None else nspread, eps=eps, iflag=iflag, device=device, **kw)
        self.plan.df = df

    def forward(self, image, coords):
        return nufft_forward(image, coords, self.plan)

    def adjoint(self, y, coords):
        return nufft_adjoint(y, coords, self.plan)
