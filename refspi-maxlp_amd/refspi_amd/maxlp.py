"""MaxLP object update on the GPU (mirror of the reference's maxLP.py:12-340 live path).

``MaxLPhaser._run_phaser(model, sx, sy, dia, ang)`` returns the updated complex model as
the reference does (maxLP.py:304-340): phase 1 pattern search over the 0.39<q<0.41 ring
with per-candidate rescale, phase 2 mean ring rescale, phase 3 pattern search over
4/size<q<0.5 with that rescale.  All pixels of a phase run in parallel in
``k_pattern`` (csrc/maxlp.hip).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib


def frame_rot_cos_sin(ang):
    """cos/sin of ``-ang`` per frame (maxLP.py:311 ``_get_rot(-ang)``), C library per value."""
    ang = np.asarray(ang, dtype='f8')
    uniq, inv = np.unique(ang, return_inverse=True)
    c = np.array([math.cos(-a) for a in uniq])
    s = np.array([math.sin(-a) for a in uniq])
    return c[inv], s[inv]


def qvals(size):
    """maxLP.py:79-95: qx, qy, qrad (0 -> 1e-6), mask, rmask (flat, 'ij' order)."""
    ind = (np.arange(size, dtype='f8') - size // 2) / size
    qx, qy = np.meshgrid(ind, ind, indexing='ij')
    qx, qy = qx.ravel(), qy.ravel()
    qrad = np.sqrt(qx ** 2 + qy ** 2)
    qrad[qrad == 0] = 1.e-6
    mask = (qrad > 4 / size) & (qrad < 0.5)
    rmask = (qrad > 0.39) & (qrad < 0.41)
    return qx, qy, qrad, mask, rmask


class MaxLPhaser:
    """maxLP.py:14-35 ``MaxLPhaser(data_fname, size, num_data)``.

    ``data`` may be a photons-file path or a :class:`~refspi_amd.dset.Dataset`.
    """

    def __init__(self, data, size=185, num_data=-1, device=0):
        import torch
        from .dset import Dataset
        if not torch.cuda.is_available():
            raise RuntimeError('MaxLPhaser needs a GPU (HIP device)')
        self.size = size
        self.device = device
        self.dset = data if isinstance(data, Dataset) else Dataset(data, size * size)
        n = self.dset.num_data if num_data < 0 else num_data
        self.num_data = n
        self.qx, self.qy, self.qrad, self.mask, self.rmask = qvals(size)
        ind = np.arange(-8, 8.5, 0.5)
        x, y = np.meshgrid(ind, ind, indexing='ij')
        self.pattern = (x + 1j * y).ravel()
        self.rescale = None
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.rsp_maxlp_create(device, size, C.byref(h)), 'rsp_maxlp_create')
        self._h = h
        d = self.dset
        ones, multi = d.ones[:n], d.multi[:n]
        no, nm = int(ones.sum()), int(multi.sum())
        _lib.check(lib.rsp_maxlp_set_frames(
            h, n, _lib.host_ptr(np.ascontiguousarray(ones), C.c_int32),
            _lib.host_ptr(np.ascontiguousarray(multi), C.c_int32),
            _lib.host_ptr(np.ascontiguousarray(d.place_ones[:no]), C.c_int32),
            _lib.host_ptr(np.ascontiguousarray(d.place_multi[:nm]), C.c_int32),
            _lib.host_ptr(np.ascontiguousarray(d.count_multi[:nm]), C.c_int32)),
            'rsp_maxlp_set_frames')

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.rsp_maxlp_destroy(h)
            self._h = None

    def info(self):
        out = np.zeros(8, dtype='i8')
        _lib.call('rsp_maxlp_info', self._h, _lib.host_ptr(out, C.c_int64))
        return dict(frames=int(out[0]), photons=int(out[1]), model_entries=int(out[2]),
                    ring_pixels=int(out[3]), mask_pixels=int(out[4]),
                    entry_candidate_products=int(out[5]), search_steps=int(out[6]),
                    pattern_ms=out[7] / 1000.)

    def run_device(self, model_dev, sx, sy, dia, ang, stream=None):
        """Update ``model_dev`` (torch complex128 on the GPU, npix) in place."""
        import torch
        dev = model_dev.device
        f64 = dict(dtype=torch.float64, device=dev)
        to = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype='f8')[:self.num_data]
                                       if not torch.is_tensor(a) else a[:self.num_data],
                                       **f64).contiguous()
        ang_h = ang.detach().cpu().numpy() if torch.is_tensor(ang) else np.asarray(ang)
        rc, rs = frame_rot_cos_sin(ang_h[:self.num_data])
        t_sx, t_sy, t_dia = to(sx), to(sy), to(dia)
        t_rc, t_rs = to(rc), to(rs)
        resc = C.c_double(0.)
        _lib.call('rsp_maxlp_run', self._h, _lib.ptr(model_dev), _lib.ptr(t_sx), _lib.ptr(t_sy),
                  _lib.ptr(t_dia), _lib.ptr(t_rc), _lib.ptr(t_rs), C.byref(resc),
                  _lib.stream_handle(stream))
        self.rescale = resc.value
        return model_dev

    def get_photons_pixel(self, t):
        """maxLP.py:245-253 for all frames: K_d(t) (NumPy f8[N]) of model pixel t at the
        rotations of the last run, read from the photon lists that run pulled back."""
        import torch
        out = torch.empty(self.num_data, dtype=torch.float64, device='cuda:%d' % self.device)
        _lib.call('rsp_maxlp_photons_pixel', self._h, int(t), _lib.ptr(out), _lib.stream_handle())
        return out.cpu().numpy()

    def get_logq_pixel(self, cands, t, rescale=None):
        """maxLP.py:181-197: LL of the candidate values (up to 1280) at mask pixel t with the
        frames and per-frame parameters of the last run; rescale None = per candidate."""
        import torch
        dev = 'cuda:%d' % self.device
        c = torch.as_tensor(np.ascontiguousarray(np.asarray(cands, dtype='c16').ravel()),
                            device=dev)
        out = torch.empty(c.numel(), dtype=torch.float64, device=dev)
        _lib.call('rsp_maxlp_logq_pixel', self._h, int(t), _lib.ptr(c), c.numel(),
                  -1. if rescale is None else float(rescale), _lib.ptr(out),
                  _lib.stream_handle())
        return out.cpu().numpy()

    def _run_phaser(self, model, sx_vals, sy_vals, dia_vals, ang_vals):
        """maxLP.py:304-340: returns the updated model as NumPy c16 (flat)."""
        import torch
        m = torch.as_tensor(np.ascontiguousarray(np.asarray(model, dtype='c16').ravel()),
                            device='cuda:%d' % self.device)
        self.run_device(m, sx_vals, sy_vals, dia_vals, ang_vals)
        return m.cpu().numpy()
