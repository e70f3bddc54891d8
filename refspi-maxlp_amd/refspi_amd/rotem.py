"""Rotation-form E-step and merge on the GPU (per-frame reference parameters).

The reference searches the joint (state x rotation) grid with rotated intensities
(emc.py:159-216).  This form keeps each frame's reference sphere (shift, diameter) as a
frame parameter and searches rotations only, so that the dense part of the Poisson
log-likelihood becomes the cross-term GEMM sum_p O_r(p) conj(R_d(p)) (csrc/xterm.hip) and
the rotation posterior is a log-sum-exp over r (csrc/rotem.hip):

    W_dr(p) = scale_d |O_r(p) + R_d(p)|^2,  LL_dr = sum_photons K log W_dr - sum_p W_dr

over the scored pixels of emc.py:89-98, with O_r the bilinear rotation of kernels.cu:6-42
and R_d the sphere term of kernels.cu:44-82.  ``merge`` is merge_all + slice_merge
(kernels.cu:139-152, 84-114) for every rotation with the weight normalisation.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib

MODES = {'fp32': 0, 'split': 1, 'fp64': 2}


def rot_cos_sin(num_rot):
    """cos/sin of j*pi/num_rot (emc.py:192) from the C library."""
    a = [j * math.pi / num_rot for j in range(num_rot)]
    return np.array([math.cos(v) for v in a]), np.array([math.sin(v) for v in a])


class RotEM:
    """E-step + merge plan for one GPU.

    Args:
        data: a :class:`~refspi_amd.dset.Dataset` (or an object with the flat arrays)
        size: q-grid side
        num_rot: rotations j*pi/num_rot (<= 768)
        rel_scale: amplitude of the reference term
        mode: cross-term precision 'split' (BF16 hi/lo on MFMA), 'fp32' (rocBLAS) or
            'fp64' (rocBLAS, with an FP64 log term)
        chunk: frames per GEMM chunk (0 = automatic)
    """

    def __init__(self, data, size, num_rot, rel_scale=1., device=0, mode='split', chunk=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('RotEM needs a GPU (HIP device)')
        self.size, self.num_rot, self.rel_scale, self.device = size, num_rot, rel_scale, device
        self.num_data = len(data.ones)
        lib = _lib.load()
        rc, rs = rot_cos_sin(num_rot)
        self._rc, self._rs = np.ascontiguousarray(rc), np.ascontiguousarray(rs)
        h = C.c_void_p()
        _lib.check(lib.rsp_rot_create(device, size, num_rot, _lib.host_ptr(self._rc, C.c_double),
                                      _lib.host_ptr(self._rs, C.c_double), float(rel_scale),
                                      C.byref(h)), 'rsp_rot_create')
        self._h = h
        arr = lambda a: np.ascontiguousarray(np.asarray(a, dtype='i4'))
        ones, multi = arr(data.ones), arr(data.multi)
        po, pm, cm = arr(data.place_ones), arr(data.place_multi), arr(data.count_multi)
        _lib.check(lib.rsp_rot_set_frames(h, self.num_data, _lib.host_ptr(ones, C.c_int32),
                                          _lib.host_ptr(multi, C.c_int32),
                                          _lib.host_ptr(po, C.c_int32),
                                          _lib.host_ptr(pm, C.c_int32),
                                          _lib.host_ptr(cm, C.c_int32)), 'rsp_rot_set_frames')
        self.set_mode(mode)
        if chunk:
            _lib.call('rsp_rot_set_option', h, 2, int(chunk))

    def set_mode(self, mode):
        self.mode = mode
        _lib.call('rsp_rot_set_option', self._h, 1, MODES[mode])

    def set_timing(self, on=True):
        _lib.call('rsp_rot_set_option', self._h, 5, 1 if on else 0)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.rsp_rot_destroy(h)
            self._h = None

    def _dev(self, a, dtype):
        import torch
        if isinstance(a, torch.Tensor):
            return a.to(device='cuda:%d' % self.device, dtype=dtype).contiguous()
        return torch.as_tensor(np.ascontiguousarray(a), device='cuda:%d' % self.device).to(dtype)

    def estep_device(self, model, sx, sy, dia, scales=None, want_ll=True, want_prob=True):
        """Device tensors: (ll (N, R) or None, prob (N, R) or None, logz (N,))."""
        import torch
        f64 = torch.float64
        N, R = self.num_data, self.num_rot
        m = self._dev(model, torch.complex128).reshape(-1)
        if m.numel() != self.size ** 2:
            raise ValueError('model must have size*size entries')
        args = [self._dev(v, f64) for v in (sx, sy, dia)]
        sc = self._dev(np.ones(N) if scales is None else scales, f64)
        for v in args + [sc]:
            if v.numel() != N:
                raise ValueError('per-frame arrays must have num_data entries')
        dev = m.device
        ll = torch.empty(N, R, dtype=f64, device=dev) if want_ll else None
        prob = torch.empty(N, R, dtype=f64, device=dev) if want_prob else None
        logz = torch.empty(N, dtype=f64, device=dev)
        p = lambda t: _lib.ptr(t) if t is not None else None
        _lib.call('rsp_rot_estep', self._h, _lib.ptr(torch.view_as_real(m)), *[p(v) for v in args],
                  p(sc), p(ll), p(prob), p(logz), _lib.stream_handle())
        return ll, prob, logz

    def estep(self, model, sx, sy, dia, scales=None):
        """NumPy (ll, prob, logz)."""
        ll, prob, logz = self.estep_device(model, sx, sy, dia, scales)
        return ll.cpu().numpy(), prob.cpu().numpy(), logz.cpu().numpy()

    def merge_device(self, prob, scales=None):
        """Merged intensity model, slice_merge accumulator and weights (device, npix)."""
        import torch
        f64 = torch.float64
        N, npix = self.num_data, self.size ** 2
        pr = self._dev(prob, f64)
        sc = self._dev(np.ones(N) if scales is None else scales, f64)
        out = [torch.empty(npix, dtype=f64, device=pr.device) for _ in range(3)]
        _lib.call('rsp_rot_merge', self._h, _lib.ptr(pr), _lib.ptr(sc),
                  *[_lib.ptr(o) for o in out], _lib.stream_handle())
        return tuple(out)

    def merge(self, prob, scales=None):
        return tuple(o.cpu().numpy() for o in self.merge_device(prob, scales))

    def adjoint_device(self, model, sx, sy, dia, scales, prob):
        """Gradient of L = -sum_d log((1/R) sum_r exp LL[d, r]) (FP64, device tensors).

        ``prob`` is P(r|d) of :meth:`estep_device` at the same inputs.  Returns
        (g_model complex128 (npix,) = dL/dRe + i dL/dIm, g_sx, g_sy, g_dia (N,)).
        """
        import torch
        f64 = torch.float64
        N = self.num_data
        m = self._dev(model, torch.complex128).reshape(-1)
        if m.numel() != self.size ** 2:
            raise ValueError('model must have size*size entries')
        args = [self._dev(v, f64) for v in (sx, sy, dia)]
        sc = self._dev(np.ones(N) if scales is None else scales, f64)
        pr = self._dev(prob, f64)
        if pr.numel() != N * self.num_rot:
            raise ValueError('prob must have num_data * num_rot entries')
        for v in args + [sc]:
            if v.numel() != N:
                raise ValueError('per-frame arrays must have num_data entries')
        gm = torch.empty(self.size ** 2, dtype=torch.complex128, device=m.device)
        g = [torch.empty(N, dtype=f64, device=m.device) for _ in range(3)]
        _lib.call('rsp_rot_adjoint', self._h, _lib.ptr(torch.view_as_real(m)),
                  *[_lib.ptr(v) for v in args], _lib.ptr(sc), _lib.ptr(pr),
                  _lib.ptr(torch.view_as_real(gm)), *[_lib.ptr(v) for v in g],
                  _lib.stream_handle())
        return (gm, *g)

    def loss_grad(self, model, sx, sy, dia, scales=None):
        """NumPy (loss, g_model, g_sx, g_sy, g_dia) at the plan's forward precision."""
        _, prob, logz = self.estep_device(model, sx, sy, dia, scales, want_ll=False)
        loss = -float((logz - math.log(self.num_rot)).sum())
        out = self.adjoint_device(model, sx, sy, dia, scales, prob)
        return (loss,) + tuple(o.cpu().numpy() for o in out)

    def timing(self):
        out = np.zeros(5)
        _lib.call('rsp_rot_timing', self._h, _lib.host_ptr(out, C.c_double))
        return dict(slices_ms=out[0], ref_rows_ms=out[1], gemm_ms=out[2], logterm_lse_ms=out[3],
                    chunks=int(out[4]))

    def info(self):
        out = np.zeros(7, dtype='i8')
        _lib.call('rsp_rot_info', self._h, _lib.host_ptr(out, C.c_int64))
        keys = ('frames', 'scored_pixels', 'k_padded', 'rot_padded', 'chunk', 'photons',
                'radii')
        return dict(zip(keys, (int(v) for v in out)))
