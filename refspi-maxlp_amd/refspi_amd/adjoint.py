"""Differentiable Poisson log-likelihood of the rotation form (torch.autograd.Function).

    loss = -sum_d log((1/R) sum_r exp LL[d, r]),
    LL[d, r] = sum_photons K log W - sum_pixels W,  W = scale_d |O_r + R_d|^2

with O_r the bilinear rotation of the complex object model (kernels.cu:6-42) and R_d the
reference sphere term of frame d with shift (sx_d, sy_d) and diameter dia_d
(kernels.cu:44-82).  The forward pass is ``rsp_rot_estep`` (split-BF16, FP32 or FP64 cross
term, FP64 log-sum-exp); the backward pass is the hand-written adjoint ``rsp_rot_adjoint``,
which recomputes the rotated slices and reference rows in FP64 instead of storing them and
keeps only P(r|d) from the forward pass.

The reference has no gradient code on its path; its nearest is the never-called
``get_grad_pixel`` of maxLP.py:199-214.  The reference's pattern search over (shift,
diameter) stays the judge of any gradient refinement.
"""
from __future__ import annotations

import math

import torch

__all__ = ['RotLoss', 'rot_loss']


class RotLoss(torch.autograd.Function):
    """Inputs: model (npix, 2) or, for several classes, (C, npix, 2) real (re, im; FP32 or
    FP64), sx, sy, dia (N,) real; the plan (:class:`~refspi_amd.rotem.RotEM`) and per-frame
    scales (N,) or None are constants.  Output: the FP64 scalar loss
    -sum_d log((1/H) sum_h exp LL[d, h]) over the plan's H = C * R hypotheses."""

    @staticmethod
    def forward(ctx, model, sx, sy, dia, plan, scales):
        f64 = torch.float64
        m = torch.view_as_complex(model.detach().to(f64).contiguous())
        a = [v.detach().to(f64).contiguous() for v in (sx, sy, dia)]
        sc = None if scales is None else torch.as_tensor(scales).to(m.device, f64)
        _, prob, logz = plan.estep_device(m, *a, sc, want_ll=False, want_prob=True)
        loss = -(logz - math.log(plan.num_hyp)).sum()
        ctx.plan = plan
        ctx.scales = sc
        ctx.dtypes = (model.dtype, sx.dtype, sy.dtype, dia.dtype)
        ctx.save_for_backward(m, *a, prob)
        return loss

    @staticmethod
    def backward(ctx, grad):
        m, sx, sy, dia, prob = ctx.saved_tensors
        gm, gx, gy, gd = ctx.plan.adjoint_device(m, sx, sy, dia, ctx.scales, prob)
        grad = grad.to(torch.float64)
        out = [torch.view_as_real(gm) * grad, gx * grad, gy * grad, gd * grad]
        out = [o.to(dt) for o, dt in zip(out, ctx.dtypes)]
        return out[0], out[1], out[2], out[3], None, None


def rot_loss(plan, model, sx, sy, dia, scales=None):
    """loss (0-dim FP64 tensor) with gradients to ``model`` (npix, 2), ``sx``, ``sy``,
    ``dia`` through :class:`RotLoss`."""
    return RotLoss.apply(model, sx, sy, dia, plan, scales)
