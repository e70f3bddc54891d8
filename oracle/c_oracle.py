"""ctypes binding of the oracle's C restatement (``oracle/emc_oracle.c``).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``tests/golden/`` fixture scripts,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this
module, and only as the checker or the timed CPU baseline.  The product
(``refspi-maxlp_amd/``) never imports it.

The C file restates the same reference lines as ``emc_oracle.py`` in the same operation
order; ``tests/test_oracle_c.py`` pins it to ``emc_oracle.py`` function by function on small
inputs.  It is the form that finishes the full 6069 x 360 hypothesis grid at 185^2 and
MaxLP over thousands of pixels in minutes (OpenMP threads).  Python here only prepares
arrays and composes the phases of ``_run_phaser`` (maxLP.py:304-340), whose mean rescale
is taken with NumPy as the reference takes it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import emc_oracle as E

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_build', 'libemc_oracle.so')

_lib = None

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags='C_CONTIGUOUS')
_I64 = C.c_int64
_D = C.c_double


def build():
    """Compile the library with the committed Makefile (gcc, OpenMP)."""
    r = subprocess.run(['make', '-s', '-B', '-C', HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('oracle build failed:\n' + r.stdout + r.stderr)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError('oracle library not built (%s); run `make -C oracle`' % LIB)
    lib = C.CDLL(LIB)
    lib.ora_slice_gen_holo.argtypes = [_d, _D, _D, _D, _D, _D, _I64, _d, _I64, _d]
    lib.ora_slice_gen.argtypes = [_d, _D, _D, _D, _I64, _d, _I64, _d]
    lib.ora_slice_merge.argtypes = [_d, _D, _D, _I64, _d, _d]
    lib.ora_calc_prob_all.argtypes = [_d, _i32, _I64, _i32, _i32, _i64, _i64, _i32, _i32, _i32,
                                      _D, _d, _d]
    lib.ora_merge_all.argtypes = [_d, _I64, _i32, _i32, _i64, _i64, _i32, _i32, _i32, _d]
    lib.ora_estep_top2.argtypes = [_I64, _d, _d, _I64, _d, _d, _d, _d, _I64, _d, _d, _I64,
                                   _i32, _i32, _i64, _i64, _i32, _i32, _i32, _i32, _d, C.c_int,
                                   _i32, _d, _d]
    lib.ora_estep_top2.restype = C.c_int
    lib.ora_photon_index.argtypes = [_I64, _I64, _i32, _d, _d, _d, _d, _i32]
    lib.ora_maxlp_pixels.argtypes = [_I64, _I64, _i64, _i32, _d, _d, _d, _d, _d, _d, _d, _d,
                                     _d, _I64, _i32, _d, _D, C.c_int, C.c_int, _d, _d, _i32, _d]
    lib.ora_maxlp_pixels.restype = C.c_int
    lib.ora_maxlp_logq.argtypes = [_I64, _I64, _i64, _i32, _d, _d, _d, _d, _d, _d, _d, _d, _d,
                                   C.c_int32, _I64, _d, _D, C.c_int, _d, _d, _d, _d]
    lib.ora_maxlp_logq.restype = C.c_int
    _lib = lib
    return lib


def _f8(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c16(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).ravel()).view(np.float64)


def threads():
    """Threads for the OpenMP loops: the cores this process may run on, at most
    OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get('OMP_NUM_THREADS', '')
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


# ----------------------------------------------------------------------------------
# kernels.cu single functions (pinned against emc_oracle in tests/test_oracle_c.py)
# ----------------------------------------------------------------------------------

def slice_gen_holo(model, shiftx, shifty, diameter, rel_scale, scale, size, bg, log_flag):
    out = np.empty(size * size)
    load().ora_slice_gen_holo(_c16(model), shiftx, shifty, diameter, rel_scale, scale, size,
                              _f8(bg), int(log_flag), out)
    return out


def slice_gen(model, angle, scale, size, bg, log_flag, cos_sin=None):
    ac, as_ = cos_sin if cos_sin is not None else (np.cos(angle), np.sin(angle))
    out = np.empty(size * size)
    load().ora_slice_gen(_f8(model), float(ac), float(as_), scale, size, _f8(bg),
                         int(log_flag), out)
    return out


def slice_merge(view, angle, size, model, mweights, cos_sin=None):
    ac, as_ = cos_sin if cos_sin is not None else (np.cos(angle), np.sin(angle))
    assert model.dtype == np.float64 and mweights.dtype == np.float64
    load().ora_slice_merge(_f8(view), float(ac), float(as_), size, model, mweights)


def _frames_args(fr: E.SparseFramesNP):
    return (fr.ones, fr.multi, np.ascontiguousarray(fr.ones_accum, dtype=np.int64),
            np.ascontiguousarray(fr.multi_accum, dtype=np.int64), fr.place_ones,
            fr.place_multi, fr.count_multi)


def calc_prob_all(lview, mask, fr: E.SparseFramesNP, init, scales):
    out = np.empty(fr.num_data)
    load().ora_calc_prob_all(_f8(lview), np.ascontiguousarray(mask, dtype=np.int32),
                             fr.num_data, *_frames_args(fr), float(init), _f8(scales), out)
    return out


def merge_all(prob_r, fr: E.SparseFramesNP, npix, view=None):
    view = np.zeros(npix) if view is None else _f8(view).copy()
    load().ora_merge_all(_f8(prob_r), fr.num_data, *_frames_args(fr), view)
    return view


# ----------------------------------------------------------------------------------
# emc.py:159-216 joint search with top-2 per frame
# ----------------------------------------------------------------------------------

def estep_top2(model, size, states, num_rot, fr: E.SparseFramesNP, scales=None, nthreads=None):
    """The joint (state x rotation) E-step of emc.py:159-216 for every frame.

    msums and vscale (emc.py:172-181) come from ``emc_oracle`` (NumPy's summation order);
    the log-views and per-frame sums over all S*R hypotheses from the C restatement.
    Returns (rmax i4[N], maxll f8[N], second f8[N], msums, vscale).
    """
    sx, sy, dia = (_f8(a) for a in states)
    S = len(sx)
    views, msums, sum_views = E.estep_views(model, size, (sx, sy, dia), fr.bg)
    del views
    vscale = E.estep_vscale(fr.get_powder(), sum_views, size, S)
    init = _f8(-(msums * vscale))
    scales = np.ones(fr.num_data) if scales is None else _f8(scales)
    rc, rs = E.rotation_table(num_rot)
    rmax = np.empty(fr.num_data, dtype=np.int32)
    maxll = np.empty(fr.num_data)
    second = np.empty(fr.num_data)
    load().ora_estep_top2(size, _c16(model), _f8(fr.bg), S, sx, sy, dia, init, num_rot,
                          _f8(rc), _f8(rs), fr.num_data, *_frames_args(fr),
                          np.ascontiguousarray(E.probmask(size), dtype=np.int32), scales,
                          int(nthreads or threads()), rmax, maxll, second)
    return rmax, maxll, second, msums, vscale


# ----------------------------------------------------------------------------------
# MaxLP (maxLP.py)
# ----------------------------------------------------------------------------------

def photon_index(t, size, rc, rs):
    """C form of emc_oracle.photon_index (one model pixel per frame)."""
    qx, qy, _, _, _ = E.qvals(size)
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.int32), np.shape(rc)))
    out = np.empty(len(t), dtype=np.int32)
    load().ora_photon_index(size, len(t), t, _f8(rc), _f8(rs), qx, qy, out)
    return out


class MaxLPC:
    """maxLP.py:304-340 ``_run_phaser`` over the C restatement of the pixel search."""

    def __init__(self, fr: E.SparseFramesNP, size):
        self.size = size
        self.N = fr.num_data
        npix = size * size
        # per frame: distinct detector pixels (ascending) with summed counts
        # (the photons CSR of maxLP.py:60-71 sums duplicates)
        keys = np.concatenate([fr.frame_of_one.astype(np.int64) * npix + fr.place_ones,
                               fr.frame_of_multi.astype(np.int64) * npix + fr.place_multi])
        cnts = np.concatenate([np.ones(len(fr.place_ones)), fr.count_multi.astype('f8')])
        uk, inv = np.unique(keys, return_inverse=True)
        self.fcnt = np.bincount(inv, weights=cnts, minlength=len(uk)).astype('f8')
        self.fpix = (uk % npix).astype(np.int32)
        self.fo = np.searchsorted(uk // npix, np.arange(self.N + 1)).astype(np.int64)
        self.qx, self.qy, self.qrad, self.mask, self.rmask = E.qvals(size)

    def set_frames(self, sx, sy, dia, ang):
        """maxLP.py:306-311 (cos/sin of -ang through the C library, as emc_oracle)."""
        self.sx, self.sy, self.dia = _f8(sx), _f8(sy), _f8(dia)
        rc, rs = E.frame_rot_cos_sin(ang)
        self.rc, self.rs = _f8(rc), _f8(rs)

    def pixels(self, model_vals, pixels, rescale=None, nthreads=None, num_iter=10):
        """run_pixel_pattern (maxLP.py:154-179) for each listed pixel from its value in
        ``model_vals``; rescale None = per-candidate rescale (phase 1).  Returns
        (new values c16, get_rescale_pixel of them, accepted steps, smallest relative top-2
        margin over the steps)."""
        pixels = np.ascontiguousarray(pixels, dtype=np.int32)
        n = len(pixels)
        f_out = np.empty(2 * n)
        resc = np.empty(n)
        steps = np.empty(n, dtype=np.int32)
        margin = np.empty(n)
        load().ora_maxlp_pixels(self.size, self.N, self.fo, self.fpix, self.fcnt, self.sx,
                                self.sy, self.dia, self.rc, self.rs, self.qx, self.qy,
                                self.qrad, n, pixels, _c16(model_vals), -1. if rescale is None
                                else float(rescale), num_iter, int(nthreads or threads()),
                                f_out, resc, steps, margin)
        return f_out.view(np.complex128), resc, steps, margin

    def logq(self, t, cands, rescale=None, nthreads=None):
        """get_logq_pixel (maxLP.py:181-197) of the candidates at pixel t; also returns
        K_d(t) and R_d(t)."""
        cands = np.asarray(cands, dtype=np.complex128).ravel()
        out = np.empty(len(cands))
        k = np.empty(self.N)
        rr = np.empty(self.N)
        ri = np.empty(self.N)
        load().ora_maxlp_logq(self.size, self.N, self.fo, self.fpix, self.fcnt, self.sx,
                              self.sy, self.dia, self.rc, self.rs, self.qx, self.qy, self.qrad,
                              int(t), len(cands), _c16(cands), -1. if rescale is None
                              else float(rescale), int(nthreads or threads()), out, k, rr, ri)
        return out, k, rr + 1j * ri

    def run_phaser(self, model, sx, sy, dia, ang, pixels=None, nthreads=None):
        """maxLP.py:304-340: phase 1 over rmask (per-candidate rescale), phase 2 the mean of
        the ring pixels' rescales (NumPy mean over the rmask pixels in index order, as
        maxLP.py:321-327), phase 3 over mask (or the listed subset of it) with that
        rescale.  Returns (model c16, rescale, info dict with per-pixel steps/margins)."""
        self.set_frames(sx, sy, dia, ang)
        fconv = np.asarray(model, dtype=np.complex128).ravel().copy()
        ring = np.nonzero(self.rmask)[0]
        f1, resc1, st1, mg1 = self.pixels(fconv[ring], ring, None, nthreads)
        fconv[ring] = f1
        rescales = np.zeros(self.size ** 2)
        rescales[ring] = resc1
        rescale = np.mean(rescales[self.rmask])
        todo = np.nonzero(self.mask)[0] if pixels is None else \
            np.asarray([t for t in pixels if self.mask[t]], dtype=np.int64)
        f3, _, st3, mg3 = self.pixels(fconv[todo], todo, rescale, nthreads)
        fconv[todo] = f3
        info = {'ring': ring, 'ring_steps': st1, 'ring_margin': mg1, 'phase3': todo,
                'phase3_steps': st3, 'phase3_margin': mg3}
        return fconv, rescale, info
