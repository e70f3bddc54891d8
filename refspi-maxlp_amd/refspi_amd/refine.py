"""Per-frame refinement of the reference parameters (shift, diameter) and the grid search
that judges it.

The reference finds each frame's reference sphere by a joint grid search over
(shift_x, shift_y, diameter) states (emc.py:67-87, 159-216; ``shiftx``/``shifty``/
``sphere_dia`` as ``lo hi n`` triples).  :func:`grid_search` is that search in the
rotation form: for every state of the grid it evaluates each frame's log-evidence
log Z_d = log sum_h exp LL_dh over the plan's hypotheses and keeps the best state per frame.
:func:`refine` moves the parameters continuously instead, with the frame gradients of the
hand-written adjoint (``rsp_rot_adjoint`` / the pass API): a per-frame trust-region ascent
on log Z_d, each step accepted only if it raises that frame's log Z_d.
"""
from __future__ import annotations

import math
import types

import numpy as np

from .rotem import PASS_FRAME_GRAD, RotEM, StreamedRotEM


def _pass(plan, models, sx, sy, dia, scales, grad):
    """(logz, g_sx, g_sy, g_dia) at the given parameters (device tensors) for a resident or
    a streamed plan; gradients of -log Z_d (None unless ``grad``)."""
    if isinstance(plan, StreamedRotEM):
        out = plan.run_pass(models, sx, sy, dia, scales, flags=PASS_FRAME_GRAD if grad else 0,
                            want_best=False)
        return out['logz'], out.get('g_sx'), out.get('g_sy'), out.get('g_dia')
    _, prob, logz = plan.estep_device(models, sx, sy, dia, scales, want_ll=False,
                                      want_prob=grad)
    if not grad:
        return logz, None, None, None
    _, gx, gy, gd = plan.adjoint_device(models, sx, sy, dia, scales, prob)
    return logz, gx, gy, gd


def refine(plan, models, sx, sy, dia, scales=None, steps=4, radius=0.25, grow=1.5,
           shrink=0.3, min_radius=1e-3, on_step=None, grad_precision=None):
    """Trust-region ascent of every frame's log Z_d over (sx, sy, dia).

    Each step moves a frame by ``radius_d`` (px) along its gradient of log Z_d (the three
    parameters in pixels, unscaled), evaluates log Z_d there, and keeps the move only if
    log Z_d rose; the frame's radius then grows by ``grow``, else shrinks by ``shrink``.
    One pass (E-step + frame adjoint) per step.  ``grad_precision``: the frame-gradient
    GEMM's ('fp64' or 'split'; default 'split' for a split-mode plan, else 'fp64') -- only
    the direction of a step depends on it.  Returns (sx, sy, dia, logz, history) with
    device tensors; history[i] = (mean log Z, accepted fraction).
    """
    if grad_precision is None:
        grad_precision = 'split' if plan.mode == 'split' else 'fp64'
    before = plan.frame_grad
    plan.set_frame_grad(grad_precision)
    try:
        return _refine(plan, models, sx, sy, dia, scales, steps, radius, grow, shrink,
                       min_radius, on_step)
    finally:
        plan.set_frame_grad(before)


def _refine(plan, models, sx, sy, dia, scales, steps, radius, grow, shrink, min_radius,
            on_step):
    import torch
    f64 = torch.float64
    dev = 'cuda:%d' % plan.device
    sx, sy, dia = (torch.as_tensor(v, device=dev).to(f64).clone() for v in (sx, sy, dia))
    N = sx.numel()
    sc = None if scales is None else torch.as_tensor(scales, device=dev).to(f64)
    rad = torch.full((N,), float(radius), dtype=f64, device=dev)
    logz, gx, gy, gd = _pass(plan, models, sx, sy, dia, sc, True)
    hist = [(float(logz.mean()), 1.)]
    for it in range(steps):
        # ascent direction of log Z = -gradient of the loss
        g = torch.stack([-gx, -gy, -gd])
        norm = torch.linalg.vector_norm(g, dim=0).clamp_min(1e-300)
        step = g / norm * rad
        nx, ny, nd = sx + step[0], sy + step[1], dia + step[2]
        last = it == steps - 1
        lz, hx, hy, hd = _pass(plan, models, nx, ny, nd, sc, not last)
        ok = (lz > logz) & (rad > min_radius)
        sx, sy, dia = torch.where(ok, nx, sx), torch.where(ok, ny, sy), torch.where(ok, nd, dia)
        logz = torch.where(ok, lz, logz)
        if not last:
            gx, gy, gd = (torch.where(ok, a, b) for a, b in ((hx, gx), (hy, gy), (hd, gd)))
        rad = torch.where(ok, rad * grow, rad * shrink)
        hist.append((float(logz.mean()), float(ok.to(f64).mean())))
        if on_step is not None:
            on_step(it, hist[-1])
    return sx, sy, dia, logz, hist


def _subset_frames(data, idx):
    """Frames ``idx`` (in that order) of a flat sparse dataset as a new flat dataset."""
    ones, multi = np.asarray(data.ones), np.asarray(data.multi)
    oa = np.concatenate([[0], np.cumsum(ones, dtype='i8')])
    ma = np.concatenate([[0], np.cumsum(multi, dtype='i8')])
    po = np.concatenate([np.asarray(data.place_ones)[oa[d]:oa[d + 1]] for d in idx])
    pm = np.concatenate([np.asarray(data.place_multi)[ma[d]:ma[d + 1]] for d in idx])
    cm = np.concatenate([np.asarray(data.count_multi)[ma[d]:ma[d + 1]] for d in idx])
    return types.SimpleNamespace(num_pix=data.num_pix, ones=ones[idx], multi=multi[idx],
                                 place_ones=po.astype('i4'), place_multi=pm.astype('i4'),
                                 count_multi=cm.astype('i4'), bg=None)


def grid_search(data, size, num_rot, rel_scale, models, shiftx=(-4., 4., 17),
                shifty=(-4., 4., 17), sphere_dia=(7., 9., 21), scales=None, classes=1,
                full_turn=False, mode='split', device=0):
    """The reference's (shift_x, shift_y, diameter) grid (emc.py:67-78 'ij' meshgrid order)
    searched per frame for the largest log Z_d in the rotation form.

    The frames are replicated once per (shift_x, shift_y) state into one resident plan; one
    E-step per diameter evaluates all replicas.  Returns (sx, sy, dia, logz, state) as NumPy
    arrays; ``state`` is the first-occurrence argmax in the reference's state order.
    """
    import torch
    gx = np.linspace(*shiftx[:2], int(shiftx[2]))
    gy = np.linspace(*shifty[:2], int(shifty[2]))
    gd = np.linspace(*sphere_dia[:2], int(sphere_dia[2]))
    n = len(data.ones)
    nxy = len(gx) * len(gy)
    rep = _subset_frames(data, np.tile(np.arange(n), nxy))
    plan = RotEM(rep, size, num_rot, rel_scale=rel_scale, device=device, mode=mode,
                 classes=classes, full_turn=full_turn)
    sxr = np.repeat(np.repeat(gx, len(gy)), n)          # replica (ix, iy) major, frame minor
    syr = np.repeat(np.tile(gy, len(gx)), n)
    scr = None if scales is None else np.tile(np.asarray(scales, dtype='f8'), nxy)
    best = np.full(n, -np.inf)
    state = np.zeros(n, dtype='i8')
    for k, dval in enumerate(gd):
        _, _, lz = plan.estep_device(models, sxr, syr, np.full(n * nxy, dval), scr,
                                     want_ll=False, want_prob=False)
        lz = lz.cpu().numpy().reshape(nxy, n)
        # state index (ix * ny + iy) * nd + k, compared in the reference's order
        for q in range(nxy):
            s = q * len(gd) + k
            better = (lz[q] > best) | ((lz[q] == best) & (s < state))
            best = np.where(better, lz[q], best)
            state = np.where(better, s, state)
    del plan
    torch.cuda.empty_cache()
    q, k = state // len(gd), state % len(gd)
    return gx[q // len(gy)], gy[q % len(gy)], gd[k], best, state
