"""Rotation-form E-step, merge and adjoint on the GPU (per-frame reference parameters).

The reference searches the joint (state x rotation) grid with rotated intensities
(emc.py:159-216).  This form keeps each frame's reference sphere (shift, diameter) as a
frame parameter and searches hypotheses h = (class c, rotation r), so that the dense part of
the Poisson log-likelihood becomes the cross-term GEMM sum_p O_h(p) conj(R_d(p))
(csrc/xterm.hip) and the posterior is a log-sum-exp over h (csrc/rotem.hip):

    W_dh(p) = scale_d |O_h(p) + R_d(p)|^2,  LL_dh = sum_photons K log W_dh - sum_p W_dh

over the scored pixels of emc.py:89-98, with O_h the bilinear rotation (kernels.cu:6-42) of
class c's model and R_d the sphere term of kernels.cu:44-82.  ``merge`` is merge_all +
slice_merge (kernels.cu:139-152, 84-114) for every hypothesis with the weight normalisation
per class.  :class:`RotEM` keeps the frames resident in HBM; :class:`StreamedRotEM` keeps
them packed in pinned host memory and streams chunks in on a copy stream.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib

MODES = {'fp32': 0, 'split': 1, 'fp64': 2}
PASS_MERGE, PASS_FRAME_GRAD, PASS_MODEL_GRAD = 1, 2, 4


def rot_cos_sin(num_rot, full_turn=False):
    """cos/sin of j*pi/num_rot (emc.py:192; a half turn, enough when the whole hologram
    rotates, as in the reference) or of j*2pi/num_rot (``full_turn``: a detector-frame
    reference, |F_obj(Rq) + F_ref(q)|^2, is not centrosymmetric under R -> -R), from the C
    library."""
    step = (2. if full_turn else 1.) * math.pi / num_rot
    a = [j * step for j in range(num_rot)]
    return np.array([math.cos(v) for v in a]), np.array([math.sin(v) for v in a])


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype='i4'))


class _Plan:
    """Owns an ``rsp_rot`` plan (creation, options, device transfers)."""

    def __init__(self, size, num_rot, rel_scale=1., device=0, mode='split', chunk=0,
                 classes=1, full_turn=False):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('the rotation form needs a GPU (HIP device)')
        self.size, self.num_rot, self.rel_scale, self.device = size, num_rot, rel_scale, device
        self.classes, self.full_turn = int(classes), bool(full_turn)
        lib = _lib.load()
        rc, rs = rot_cos_sin(num_rot, full_turn)
        self._rc, self._rs = np.ascontiguousarray(rc), np.ascontiguousarray(rs)
        h = C.c_void_p()
        _lib.check(lib.rsp_rot_create(device, size, num_rot, _lib.host_ptr(self._rc, C.c_double),
                                      _lib.host_ptr(self._rs, C.c_double), float(rel_scale),
                                      C.byref(h)), 'rsp_rot_create')
        self._h = h
        if self.classes != 1:
            _lib.call('rsp_rot_set_option', h, 3, self.classes)
        self.num_hyp = self.classes * num_rot
        self.frame_grad = 'fp64'
        self.set_mode(mode)
        if chunk:
            _lib.call('rsp_rot_set_option', h, 2, int(chunk))

    def set_mode(self, mode):
        self.mode = mode
        _lib.call('rsp_rot_set_option', self._h, 1, MODES[mode])

    def set_timing(self, on=True):
        _lib.call('rsp_rot_set_option', self._h, 5, 1 if on else 0)

    def set_frame_grad(self, precision):
        """Precision of the frame-gradient GEMM Obar = P O of the adjoint: 'fp64'
        (rocBLAS, the default) or 'split' (BF16 hi/lo on MFMA, about 1e-5 relative in Obar;
        for gradient refinement, where a step is kept only if log Z rose)."""
        if precision not in ('fp64', 'split'):
            raise ValueError("frame-gradient precision must be 'fp64' or 'split'")
        _lib.call('rsp_rot_set_option', self._h, 6, 1 if precision == 'split' else 0)
        self.frame_grad = precision

    def set_ref_rows(self, how):
        """Reference rows R_d of the split cross term: 'generated' (computed inside the GEMM
        kernel from three numbers per frame, never stored; the default up to 513 x 513) or
        'stored' (written to HBM as bf16 hi/lo rows, then read by the GEMM)."""
        if how not in ('generated', 'stored'):
            raise ValueError("reference rows must be 'generated' or 'stored'")
        _lib.call('rsp_rot_set_option', self._h, 8, 1 if how == 'generated' else 0)
        self.ref_rows = how

    def set_model_grad(self, precision):
        """Precision of the model-gradient GEMM GT = Q^T B of the adjoint (Q = P * scale,
        B = the reference rows): 'fp64' (rocBLAS, the default) or 'split' (BF16 hi/lo on
        MFMA with FP32 accumulation over each sub-chunk's frames, added in FP64)."""
        if precision not in ('fp64', 'split'):
            raise ValueError("model-gradient precision must be 'fp64' or 'split'")
        _lib.call('rsp_rot_set_option', self._h, 7, 1 if precision == 'split' else 0)
        self.model_grad = precision

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

    def _models(self, model):
        import torch
        m = self._dev(model, torch.complex128).reshape(-1)
        if m.numel() != self.classes * self.size ** 2:
            raise ValueError('model must have classes * size * size entries')
        return m

    def _shape_model(self, t):
        """(C, npix) results as (npix,) when there is one class."""
        return t.reshape(self.size ** 2) if self.classes == 1 else \
            t.reshape(self.classes, self.size ** 2)

    def timing(self):
        out = np.zeros(5)
        _lib.call('rsp_rot_timing', self._h, _lib.host_ptr(out, C.c_double))
        return dict(slices_ms=out[0], ref_rows_ms=out[1], gemm_ms=out[2], logterm_lse_ms=out[3],
                    chunks=int(out[4]))

    def info(self):
        out = np.zeros(10, dtype='i8')
        _lib.call('rsp_rot_info', self._h, _lib.host_ptr(out, C.c_int64))
        keys = ('frames', 'scored_pixels', 'k_padded', 'hyp_padded', 'chunk', 'photons',
                'radii', 'classes', 'hypotheses', 'hyp_stride')
        return dict(zip(keys, (int(v) for v in out)))


class RotEM(_Plan):
    """E-step, merge and adjoint plan for one GPU with the frames resident in HBM.

    Args:
        data: a :class:`~refspi_amd.dset.Dataset` (or an object with the flat arrays)
        size: q-grid side
        num_rot: rotations j*pi/num_rot
        rel_scale: amplitude of the reference term
        mode: cross-term precision 'split' (BF16 hi/lo on MFMA), 'fp32' (rocBLAS) or
            'fp64' (rocBLAS, with an FP64 log term)
        chunk: frames per GEMM chunk (0 = automatic)
        classes: object classes C (hypotheses h = c * num_rot + r; classes * num_rot <= 1536)
        full_turn: rotations j*2pi/num_rot instead of the reference's j*pi/num_rot
    """

    def __init__(self, data, size, num_rot, rel_scale=1., device=0, mode='split', chunk=0,
                 classes=1, full_turn=False):
        super().__init__(size, num_rot, rel_scale, device, mode, chunk, classes, full_turn)
        self.num_data = len(data.ones)
        ones, multi = _i32(data.ones), _i32(data.multi)
        po, pm, cm = _i32(data.place_ones), _i32(data.place_multi), _i32(data.count_multi)
        _lib.check(_lib.load().rsp_rot_set_frames(
            self._h, self.num_data, _lib.host_ptr(ones, C.c_int32),
            _lib.host_ptr(multi, C.c_int32), _lib.host_ptr(po, C.c_int32),
            _lib.host_ptr(pm, C.c_int32), _lib.host_ptr(cm, C.c_int32)), 'rsp_rot_set_frames')

    def _frame_args(self, sx, sy, dia, scales):
        import torch
        f64 = torch.float64
        N = self.num_data
        args = [self._dev(v, f64) for v in (sx, sy, dia)]
        sc = self._dev(np.ones(N) if scales is None else scales, f64)
        for v in args + [sc]:
            if v.numel() != N:
                raise ValueError('per-frame arrays must have num_data entries')
        return args, sc

    def estep_device(self, model, sx, sy, dia, scales=None, want_ll=True, want_prob=True,
                     want_best=False):
        """Device tensors: (ll (N, H) or None, prob (N, H) or None, logz (N,)), plus best
        (N,) int32 (first-occurrence argmax over h) when ``want_best``."""
        import torch
        f64 = torch.float64
        N, H = self.num_data, self.num_hyp
        m = self._models(model)
        args, sc = self._frame_args(sx, sy, dia, scales)
        dev = m.device
        ll = torch.empty(N, H, dtype=f64, device=dev) if want_ll else None
        prob = torch.empty(N, H, dtype=f64, device=dev) if want_prob else None
        logz = torch.empty(N, dtype=f64, device=dev)
        best = torch.empty(N, dtype=torch.int32, device=dev) if want_best else None
        p = lambda t: _lib.ptr(t) if t is not None else None
        _lib.call('rsp_rot_estep_best', self._h, _lib.ptr(torch.view_as_real(m)),
                  *[p(v) for v in args], p(sc), p(ll), p(prob), p(logz), p(best),
                  _lib.stream_handle())
        return (ll, prob, logz, best) if want_best else (ll, prob, logz)

    def estep(self, model, sx, sy, dia, scales=None):
        """NumPy (ll, prob, logz)."""
        ll, prob, logz = self.estep_device(model, sx, sy, dia, scales)
        return ll.cpu().numpy(), prob.cpu().numpy(), logz.cpu().numpy()

    def merge_device(self, prob, scales=None):
        """Merged intensity models, slice_merge accumulators and weights (device;
        (npix,) for one class, else (C, npix))."""
        import torch
        f64 = torch.float64
        N, n = self.num_data, self.classes * self.size ** 2
        pr = self._dev(prob, f64)
        if pr.numel() != N * self.num_hyp:
            raise ValueError('prob must have num_data * hypotheses entries')
        sc = self._dev(np.ones(N) if scales is None else scales, f64)
        out = [torch.empty(n, dtype=f64, device=pr.device) for _ in range(3)]
        _lib.call('rsp_rot_merge', self._h, _lib.ptr(pr), _lib.ptr(sc),
                  *[_lib.ptr(o) for o in out], _lib.stream_handle())
        return tuple(self._shape_model(o) for o in out)

    def merge(self, prob, scales=None):
        return tuple(o.cpu().numpy() for o in self.merge_device(prob, scales))

    def adjoint_device(self, model, sx, sy, dia, scales, prob):
        """Gradient of L = -sum_d log((1/H) sum_h exp LL[d, h]) (FP64, device tensors).

        ``prob`` is P(h|d) of :meth:`estep_device` at the same inputs.  Returns
        (g_model complex128 ((npix,) or (C, npix)) = dL/dRe + i dL/dIm, g_sx, g_sy, g_dia).
        """
        import torch
        f64 = torch.float64
        N = self.num_data
        m = self._models(model)
        args, sc = self._frame_args(sx, sy, dia, scales)
        pr = self._dev(prob, f64)
        if pr.numel() != N * self.num_hyp:
            raise ValueError('prob must have num_data * hypotheses entries')
        gm = torch.empty(m.numel(), dtype=torch.complex128, device=m.device)
        g = [torch.empty(N, dtype=f64, device=m.device) for _ in range(3)]
        _lib.call('rsp_rot_adjoint', self._h, _lib.ptr(torch.view_as_real(m)),
                  *[_lib.ptr(v) for v in args], _lib.ptr(sc), _lib.ptr(pr),
                  _lib.ptr(torch.view_as_real(gm)), *[_lib.ptr(v) for v in g],
                  _lib.stream_handle())
        return (self._shape_model(gm), *g)

    def loss_grad(self, model, sx, sy, dia, scales=None):
        """NumPy (loss, g_model, g_sx, g_sy, g_dia) at the plan's forward precision."""
        _, prob, logz = self.estep_device(model, sx, sy, dia, scales, want_ll=False)
        loss = -float((logz - math.log(self.num_hyp)).sum())
        out = self.adjoint_device(model, sx, sy, dia, scales, prob)
        return (loss,) + tuple(o.cpu().numpy() for o in out)


class StreamedRotEM(_Plan):
    """The rotation form over frames kept in pinned host memory, in chunks.

    Each chunk's frames are packed once (``rsp_rot_pack``) into the scored CSR of both
    orders and held in pinned host tensors.  A pass copies chunk i+1 to the device on a
    copy stream while chunk i is evaluated on the compute stream (two device buffers,
    events in both directions), so HBM holds two chunks of photons, the plan's operands and
    the outputs the caller asks for.

    Args: as :class:`RotEM`, plus ``chunk_frames`` (frames per streamed chunk).
    """

    def __init__(self, data, size, num_rot, rel_scale=1., device=0, mode='split', chunk=0,
                 classes=1, full_turn=False, chunk_frames=50_000):
        import torch
        super().__init__(size, num_rot, rel_scale, device, mode, chunk, classes, full_turn)
        lib = _lib.load()
        self.num_data = N = len(data.ones)
        self.chunk_frames = cf = int(min(chunk_frames, N))
        ones, multi = _i32(data.ones), _i32(data.multi)
        o_acc = np.concatenate([[0], np.cumsum(ones, dtype='i8')])
        m_acc = np.concatenate([[0], np.cumsum(multi, dtype='i8')])
        po_all, pm_all, cm_all = (_i32(data.place_ones), _i32(data.place_multi),
                                  _i32(data.count_multi))
        K = self.info()['scored_pixels']
        self.bounds = [(a, min(N, a + cf)) for a in range(0, N, cf)]
        self.chunks = []
        for a, b in self.bounds:
            n = b - a
            args = [ones[a:b], multi[a:b], po_all[o_acc[a]:o_acc[b]],
                    pm_all[m_acc[a]:m_acc[b]], cm_all[m_acc[a]:m_acc[b]]]
            args = [np.ascontiguousarray(v) for v in args]
            nnz = C.c_int64()
            cargs = [_lib.host_ptr(v, C.c_int32) for v in args]
            _lib.check(lib.rsp_rot_pack(self._h, n, *cargs, None, None, None, None, None, None,
                                        C.byref(nnz)), 'rsp_rot_pack')
            nz = max(1, nnz.value)
            t = dict(fo=torch.empty(n + 1, dtype=torch.int64, pin_memory=True),
                     fj=torch.empty(nz, dtype=torch.int32, pin_memory=True),
                     fk=torch.empty(nz, dtype=torch.float32, pin_memory=True),
                     po=torch.empty(K + 1, dtype=torch.int64, pin_memory=True),
                     pd=torch.empty(nz, dtype=torch.int32, pin_memory=True),
                     pk=torch.empty(nz, dtype=torch.float32, pin_memory=True))
            ptr = lambda k: C.c_void_p(t[k].data_ptr())
            _lib.check(lib.rsp_rot_pack(self._h, n, *cargs, ptr('fo'), ptr('fj'), ptr('fk'),
                                        ptr('po'), ptr('pd'), ptr('pk'), C.byref(nnz)),
                       'rsp_rot_pack')
            t['n'], t['nnz'] = n, nnz.value
            self.chunks.append(t)
        self.host_bytes = sum(v.numel() * v.element_size() for t in self.chunks
                              for k, v in t.items() if k not in ('n', 'nnz'))
        dev = 'cuda:%d' % device
        maxz = max(c['fj'].numel() for c in self.chunks)
        self._dbuf = [dict(fo=torch.empty(cf + 1, dtype=torch.int64, device=dev),
                           fj=torch.empty(maxz, dtype=torch.int32, device=dev),
                           fk=torch.empty(maxz, dtype=torch.float32, device=dev),
                           po=torch.empty(K + 1, dtype=torch.int64, device=dev),
                           pd=torch.empty(maxz, dtype=torch.int32, device=dev),
                           pk=torch.empty(maxz, dtype=torch.float32, device=dev))
                      for _ in range(2)]
        self._copy_stream = torch.cuda.Stream(device=dev)

    def run_pass(self, model, sx, sy, dia, scales=None, flags=0, want_ll=False,
                 want_prob=False, want_best=True, on_chunk=None):
        """Evaluate ``model`` over all chunks.  ``sx``, ``sy``, ``dia``, ``scales``: device
        or host arrays of all frames (moved to the device once).  ``flags``: PASS_MERGE,
        PASS_FRAME_GRAD, PASS_MODEL_GRAD (or'ed).  Returns a dict of device tensors: logz,
        best (if asked), ll/prob (N, H) (if asked), g_sx/g_sy/g_dia (frame gradients),
        model/acc/weights (merge) and g_model (model gradient).  ``on_chunk(i)`` is called
        after chunk i is queued (e.g. to sample memory use)."""
        import torch
        f64 = torch.float64
        dev = 'cuda:%d' % self.device
        N, H = self.num_data, self.num_hyp
        m = self._models(model)
        fr = [self._dev(v, f64) for v in (sx, sy, dia)]
        sc = self._dev(np.ones(N) if scales is None else scales, f64)
        out = dict(logz=torch.empty(N, dtype=f64, device=dev))
        if want_best:
            out['best'] = torch.empty(N, dtype=torch.int32, device=dev)
        if want_ll:
            out['ll'] = torch.empty(N, H, dtype=f64, device=dev)
        if want_prob:
            out['prob'] = torch.empty(N, H, dtype=f64, device=dev)
        if flags & PASS_FRAME_GRAD:
            for k in ('g_sx', 'g_sy', 'g_dia'):
                out[k] = torch.empty(N, dtype=f64, device=dev)
        comp = torch.cuda.current_stream(dev)
        cst = self._copy_stream
        st = _lib.stream_handle(comp)
        _lib.call('rsp_rot_pass_begin', self._h, _lib.ptr(torch.view_as_real(m)), int(flags),
                  self.chunk_frames, st)
        ready = [torch.cuda.Event(), torch.cuda.Event()]
        free = [torch.cuda.Event(), torch.cuda.Event()]

        def copy_in(i):
            b = i % 2
            with torch.cuda.stream(cst):
                cst.wait_event(free[b])
                src, dst = self.chunks[i], self._dbuf[b]
                for k in ('fo', 'fj', 'fk', 'po', 'pd', 'pk'):
                    n = src[k].numel()
                    dst[k][:n].copy_(src[k], non_blocking=True)
                ready[b].record(cst)

        for e in free:
            e.record(comp)
        copy_in(0)
        p = lambda t: _lib.ptr(t) if t is not None else None
        for i, (a, b_) in enumerate(self.bounds):
            if i + 1 < len(self.bounds):
                copy_in(i + 1)
            b = i % 2
            comp.wait_event(ready[b])
            d = self._dbuf[b]
            sl = lambda t: t[a:b_] if t is not None else None
            _lib.call('rsp_rot_pass_chunk', self._h, b_ - a, p(d['fo']), p(d['fj']), p(d['fk']),
                      p(d['po']), p(d['pd']), p(d['pk']), p(fr[0][a:b_]), p(fr[1][a:b_]),
                      p(fr[2][a:b_]), p(sc[a:b_]), p(sl(out.get('ll'))), p(sl(out.get('prob'))),
                      p(out['logz'][a:b_]), p(sl(out.get('best'))), p(sl(out.get('g_sx'))),
                      p(sl(out.get('g_sy'))), p(sl(out.get('g_dia'))), st)
            free[b].record(comp)
            if on_chunk is not None:
                on_chunk(i)
        n = self.classes * self.size ** 2
        mo = acc = wts = gm = None
        if flags & PASS_MERGE:
            mo, acc, wts = (torch.empty(n, dtype=f64, device=dev) for _ in range(3))
        if flags & PASS_MODEL_GRAD:
            gm = torch.empty(n, dtype=torch.complex128, device=dev)
        _lib.call('rsp_rot_pass_end', self._h, p(mo), p(acc), p(wts),
                  p(torch.view_as_real(gm) if gm is not None else None), st)
        if mo is not None:
            out.update(model=self._shape_model(mo), acc=self._shape_model(acc),
                       weights=self._shape_model(wts))
        if gm is not None:
            out['g_model'] = self._shape_model(gm)
        return out
