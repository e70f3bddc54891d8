"""FP64 NumPy restatement of the RefSPI-MaxLP reconstruction path (the oracle).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker or
the timed CPU baseline.  The product (``refspi-maxlp_amd/``) never imports it.

Every function follows the reference line by line (citations are ``file:line`` in the
reference tree) and keeps its operation order where the floating-point result depends
on it.  Transcendentals come from NumPy / the C library; CUDA's math functions are not
reproduced.  Where the reference leaves a result undefined the rule chosen here is
stated in the docstring.

Index conventions (kernels.cu:9-13, emc.py:81-84, maxLP.py:80-83): a flat pixel index
is ``t = x*size + y`` with ``x`` the first axis; the centre is ``size//2``.

How each function is pinned (beyond "same formula") is listed in
``tests/test_oracle.py``; functions without an independent pin say "parity unpinned".
"""
from __future__ import annotations

import math

import numpy as np

PI = np.pi


# ----------------------------------------------------------------------------------
# geometry helpers
# ----------------------------------------------------------------------------------

def probmask(size: int) -> np.ndarray:
    """emc.py:89-98 ``_generate_masks``: 2 outside r>=size//2 or r<4, 1 for r<size//8."""
    x, y = np.indices((size, size), dtype='f8')
    x = x.ravel() - size // 2
    y = y.ravel() - size // 2
    rad = np.sqrt(x ** 2 + y ** 2)
    pm = np.zeros(size * size, dtype='i4')
    pm[rad >= size // 2] = 2
    pm[rad < size // 8] = 1
    pm[rad < 4] = 2
    return pm


def state_grid(shiftx, shifty, sphere_dia):
    """emc.py:67-78 ``_generate_states``: ``lo hi n`` triples, meshgrid 'ij', ravel.

    Returns (sx, sy, dia) raveled f8 arrays and num_states i4[3].
    """
    sx = np.linspace(shiftx[0], shiftx[1], int(shiftx[2]))
    sy = np.linspace(shifty[0], shifty[1], int(shifty[2]))
    dd = np.linspace(sphere_dia[0], sphere_dia[1], int(sphere_dia[2]))
    gx, gy, gd = np.meshgrid(sx, sy, dd, indexing='ij')
    num_states = np.array([shiftx[-1], shifty[-1], sphere_dia[-1]]).astype('i4')
    return gx.ravel(), gy.ravel(), gd.ravel(), num_states


def rotation_table(num_rot: int):
    """cos/sin of the EMC angles ``j*pi/num_rot`` (emc.py:192), via the C library."""
    ang = [j * PI / num_rot for j in range(num_rot)]
    return (np.array([math.cos(a) for a in ang]), np.array([math.sin(a) for a in ang]))


# ----------------------------------------------------------------------------------
# kernels.cu restatements
# ----------------------------------------------------------------------------------

def slice_gen_holo(model, shiftx, shifty, diameter, rel_scale, scale, size, bg, log_flag):
    """kernels.cu:44-82 ``slice_gen_holo`` (FP64).

    view = |model + rel*S(s)*exp(i*phase)|^2 * scale + bg, with
    phase = 2*pi*((x-c)*sx + (y-c)*sy)/size and s = sqrt(...)*pi*dia/size (s==0 -> 1e-5).
    ``pow(abs(z), 2.)`` is taken as hypot(re, im)**2.
    """
    x, y = np.indices((size, size), dtype='f8')
    x = x.ravel()
    y = y.ravel()
    cen = math.floor(size / 2.)
    phase = 2. * PI * ((x - cen) * shiftx + (y - cen) * shifty) / size
    ramp_r = np.cos(phase)
    ramp_i = np.sin(phase)
    s = np.sqrt((x - cen) * (x - cen) + (y - cen) * (y - cen)) * PI * diameter / size
    s[s == 0.] = 1.e-5
    sphere = rel_scale * (np.sin(s) - s * np.cos(s)) / (s * s * s)
    model = np.asarray(model, dtype='c16').ravel()
    cr = ramp_r * sphere + model.real
    ci = ramp_i * sphere + model.imag
    view = np.hypot(cr, ci) ** 2
    view = view * scale
    view = view + bg
    if log_flag:
        out = np.full_like(view, -1000.)
        ok = view >= 1.e-20
        out[ok] = np.log(view[ok])
        view = out
    return view


def _rot_coords(size, ac, as_):
    """kernels.cu:19-27: source coordinates and floor indices of the rotated grid."""
    x, y = np.indices((size, size))
    x = x.ravel()
    y = y.ravel()
    cen = size // 2
    xc = (x - cen).astype('f8')
    yc = (y - cen).astype('f8')
    tx = xc * ac - yc * as_ + cen
    ty = xc * as_ + yc * ac + cen
    ix = np.floor(tx).astype(np.int64)
    iy = np.floor(ty).astype(np.int64)
    oob = (ix < 0) | (ix > size - 2) | (iy < 0) | (iy > size - 2)
    return tx, ty, ix, iy, oob


def slice_gen(model, angle, scale, size, bg, log_flag, cos_sin=None):
    """kernels.cu:6-42 ``slice_gen`` (FP64 bilinear rotation, optional log).

    ``cos_sin`` may pass (cos(angle), sin(angle)) computed elsewhere so that the oracle
    and the device use identical rotation constants.
    """
    ac, as_ = cos_sin if cos_sin is not None else (math.cos(angle), math.sin(angle))
    model = np.asarray(model, dtype='f8').ravel()
    tx, ty, ix, iy, oob = _rot_coords(size, ac, as_)
    ok = ~oob
    view = np.full(size * size, -1000. if log_flag else 0.)
    ixo, iyo = ix[ok], iy[ok]
    fx = tx[ok] - ixo
    fy = ty[ok] - iyo
    cx = 1. - fx
    cy = 1. - fy
    v = (model[ixo * size + iyo] * cx * cy +
         model[(ixo + 1) * size + iyo] * fx * cy +
         model[ixo * size + (iyo + 1)] * cx * fy +
         model[(ixo + 1) * size + (iyo + 1)] * fx * fy)
    v = v * scale
    v = v + np.asarray(bg, dtype='f8').ravel()[ok]
    if log_flag:
        lv = np.full_like(v, -1000.)
        good = v >= 1.e-20
        lv[good] = np.log(v[good])
        v = lv
    view[ok] = v
    return view


def slice_merge(view, angle, size, model, mweights, cos_sin=None):
    """kernels.cu:84-114 ``slice_merge``: transpose of slice_gen's bilinear step.

    Accumulates into ``model``/``mweights`` in place (np.add.at replaces atomicAdd).
    """
    ac, as_ = cos_sin if cos_sin is not None else (math.cos(angle), math.sin(angle))
    view = np.asarray(view, dtype='f8').ravel()
    tx, ty, ix, iy, oob = _rot_coords(size, ac, as_)
    ok = ~oob
    t = np.nonzero(ok)[0]
    ixo, iyo = ix[ok], iy[ok]
    fx = tx[ok] - ixo
    fy = ty[ok] - iyo
    cx = 1. - fx
    cy = 1. - fy
    for (dx, dy, w) in ((0, 0, cx * cy), (1, 0, fx * cy), (0, 1, cx * fy), (1, 1, fx * fy)):
        idx = (ixo + dx) * size + (iyo + dy)
        np.add.at(model, idx, view[t] * w)
        np.add.at(mweights, idx, w)


def calc_prob_all(lview, mask, ones, multi, o_acc, m_acc, p_o, p_m, c_m, init, scales):
    """kernels.cu:116-137 ``calc_prob_all``: sequential FP64 sum per frame.

    prob[d] = init*scales[d] + sum_{ones, mask<1} lview[p] + sum_{multi, mask<1} lview[p]*c
    """
    ndata = len(ones)
    prob = np.empty(ndata)
    for d in range(ndata):
        acc = init * scales[d]
        for t in range(o_acc[d], o_acc[d] + ones[d]):
            pix = p_o[t]
            if mask[pix] < 1:
                acc += lview[pix]
        for t in range(m_acc[d], m_acc[d] + multi[d]):
            pix = p_m[t]
            if mask[pix] < 1:
                acc += lview[pix] * c_m[t]
        prob[d] = acc
    return prob


def calc_prob_all_vec(lview, mask, frames):
    """Vectorised form of :func:`calc_prob_all` (same terms, NumPy summation order).

    ``frames`` is a :class:`SparseFramesNP`.  Used where the per-frame Python loop is
    too slow; differs from the sequential sum only by FP64 rounding.
    """
    sel = mask < 1
    lv = np.where(sel, lview, 0.)
    acc = np.zeros(frames.num_data)
    np.add.at(acc, frames.frame_of_one, lv[frames.place_ones])
    np.add.at(acc, frames.frame_of_multi, lv[frames.place_multi] * frames.count_multi)
    return acc


def merge_all(prob_r, ones, multi, o_acc, m_acc, p_o, p_m, c_m, npix):
    """kernels.cu:139-152 ``merge_all``: view[p] += prob_r[d]*count."""
    view = np.zeros(npix)
    for d in range(len(ones)):
        for t in range(o_acc[d], o_acc[d] + ones[d]):
            view[p_o[t]] += prob_r[d]
        for t in range(m_acc[d], m_acc[d] + multi[d]):
            view[p_m[t]] += prob_r[d] * c_m[t]
    return view


def get_f_dt(fobj_t, fref_d):
    """kernels.cu:172-183 ``get_f_dt``: f_dt[d, c] = fobj_t[c] + fref_d[d]."""
    return np.asarray(fref_d)[:, None] + np.asarray(fobj_t)[None, :]


# ----------------------------------------------------------------------------------
# sparse frames (dset.py)
# ----------------------------------------------------------------------------------

class SparseFramesNP:
    """dset.py:16-61 layout: ones/multi counts, exclusive prefix offsets, flat lists."""

    def __init__(self, ones, multi, place_ones, place_multi, count_multi, num_pix, bg=None):
        self.ones = np.asarray(ones, dtype='i4')
        self.multi = np.asarray(multi, dtype='i4')
        self.num_data = len(self.ones)
        # dset.py:32-33, 41-42: roll(cumsum) with element 0 set to 0
        self.ones_accum = np.roll(self.ones.astype(np.int64).cumsum(), 1)
        self.multi_accum = np.roll(self.multi.astype(np.int64).cumsum(), 1)
        if self.num_data:
            self.ones_accum[0] = 0
            self.multi_accum[0] = 0
        self.place_ones = np.asarray(place_ones, dtype='i4')
        self.place_multi = np.asarray(place_multi, dtype='i4')
        self.count_multi = np.asarray(count_multi, dtype='i4')
        self.num_pix = num_pix
        self.bg = np.zeros(num_pix) if bg is None else np.asarray(bg, dtype='f8').ravel()
        # dset.py:46-48
        self.mean_count = float((self.place_ones.shape[0] + self.count_multi.sum()) / self.num_data)
        self.frame_of_one = np.repeat(np.arange(self.num_data), self.ones)
        self.frame_of_multi = np.repeat(np.arange(self.num_data), self.multi)

    def counts(self):
        """dset.py:50-53: per-frame photon totals."""
        c = self.ones.astype('f8').copy()
        np.add.at(c, self.frame_of_multi, self.count_multi)
        return c

    def get_powder(self):
        """dset.py:74-83."""
        powder = np.zeros(self.num_pix)
        np.add.at(powder, self.place_ones, 1)
        np.add.at(powder, self.place_multi, self.count_multi)
        powder /= self.num_data
        return powder

    def dense(self):
        """Dense (N, npix) photon counts (for brute-force checks on tiny inputs)."""
        k = np.zeros((self.num_data, self.num_pix))
        np.add.at(k, (self.frame_of_one, self.place_ones), 1)
        np.add.at(k, (self.frame_of_multi, self.place_multi), self.count_multi)
        return k


# ----------------------------------------------------------------------------------
# EMC E-step (emc.py)
# ----------------------------------------------------------------------------------

def estep_views(model, size, states, bg, num_streams=4):
    """emc.py:166-181 pass 1: unrotated holograms, msums and vscale (rank 0 of 1).

    Returns (views[S, npix], msums[S], vscale) with powder supplied later.
    """
    sx, sy, dia = states
    S = len(sx)
    pm = probmask(size)
    selmask = pm < 1
    views = np.empty((S, size * size))
    msums = np.empty(S)
    sum_views = np.zeros((num_streams, size * size))
    for i in range(S):
        v = slice_gen_holo(model, sx[i], sy[i], dia[i], 1., 1., size, bg, 0)
        views[i] = v
        msums[i] = v[selmask].sum()
        sum_views[i % num_streams] += v
    return views, msums, sum_views


def estep_vscale(powder, sum_views, size, S):
    """emc.py:181: vscale = powder[sel].sum() / sum_views.sum(0)[sel].sum() * S."""
    selmask = probmask(size) < 1
    return powder[selmask].sum() / sum_views.sum(0)[selmask].sum() * S


def estep_logviews(view, size, num_rot, bg, rot=None):
    """emc.py:190-193: the R rotated log-views of one state's hologram."""
    cs, sn = rotation_table(num_rot) if rot is None else rot
    return np.stack([slice_gen(view, j * PI / num_rot, 1., size, bg, 1, (cs[j], sn[j]))
                     for j in range(num_rot)])


def estep_prob(model, size, states, num_rot, frames: SparseFramesNP, scales=None,
               sequential=False):
    """emc.py:159-205 ``_calculate_prob`` (one rank): the full prob[S*R, N] in FP64.

    Returns (prob, msums, vscale).  ``sequential`` uses the literal per-frame loop of
    calc_prob_all (exact reference order); otherwise the vectorised sum.
    """
    bg = frames.bg
    S = len(states[0])
    pm = probmask(size)
    views, msums, sum_views = estep_views(model, size, states, bg)
    vscale = estep_vscale(frames.get_powder(), sum_views, size, S)
    if scales is None:
        scales = np.ones(frames.num_data)
    rot = rotation_table(num_rot)
    prob = np.empty((S * num_rot, frames.num_data))
    for i in range(S):
        lv = estep_logviews(views[i], size, num_rot, bg, rot)
        init = -float(msums[i] * vscale)
        for j in range(num_rot):
            if sequential:
                prob[i * num_rot + j] = calc_prob_all(
                    lv[j], pm, frames.ones, frames.multi, frames.ones_accum,
                    frames.multi_accum, frames.place_ones, frames.place_multi,
                    frames.count_multi, init, scales)
            else:
                prob[i * num_rot + j] = init * scales + calc_prob_all_vec(lv[j], pm, frames)
    return prob, msums, vscale


def get_rmax(prob):
    """emc.py:207-216 ``_get_rmax`` (one rank): max and first-occurrence argmax."""
    return prob.max(0), prob.argmax(axis=0).astype('i4')


def top_margin(prob):
    """Gap between the best and second-best row per frame (0 for exact ties)."""
    srt = np.sort(prob, axis=0)
    if prob.shape[0] < 2:
        return np.full(prob.shape[1], np.inf)
    return srt[-1] - srt[-2]


def unravel_rmax(rmax, num_states, num_rot, shiftx, shifty, sphere_dia):
    """emc.py:218-224 ``_unravel_rmax``: values via unique(grid)[index], angle j*pi/R."""
    isx, isy, idia, iang = np.unravel_index(rmax, tuple(num_states) + (num_rot,))
    sx_vals = np.unique(shiftx)[isx]
    sy_vals = np.unique(shifty)[isy]
    dia_vals = np.unique(sphere_dia)[idia]
    ang_vals = iang * PI / num_rot
    return sx_vals, sy_vals, dia_vals, ang_vals


def log_sum_exp_prob(prob):
    """North-star extension (no reference code): P(h|d) = softmax over hypotheses, FP64."""
    mx = prob.max(0)
    e = np.exp(prob - mx)
    z = e.sum(0)
    return e / z, mx + np.log(z)


# ----------------------------------------------------------------------------------
# MaxLP (maxLP.py)
# ----------------------------------------------------------------------------------

def qvals(size):
    """maxLP.py:79-95 ``_get_qvals``: qx, qy, qrad (0 -> 1e-6), mask, rmask."""
    ind = (np.arange(size, dtype='f8') - size // 2) / size
    qx, qy = np.meshgrid(ind, ind, indexing='ij')
    qx = qx.ravel()
    qy = qy.ravel()
    qrad = np.sqrt(qx ** 2 + qy ** 2)
    qrad[qrad == 0] = 1.e-6
    mask = (qrad > 4 / size) & (qrad < 0.5)
    rmask = (qrad > 0.39) & (qrad < 0.41)
    return qx, qy, qrad, mask, rmask


def pattern(nmax=8):
    """maxLP.py:45-48 ``_gen_pattern``: 33x33 offsets x + iy with arange(-8, 8.5, 0.5)."""
    ind = np.arange(-nmax, nmax + 0.5, 0.5)
    x, y = np.meshgrid(ind, ind, indexing='ij')
    return (x + 1j * y).ravel()


def frame_rot_cos_sin(ang):
    """cos/sin of ``-ang`` (maxLP.py:311 ``_get_rot(-ang)``), via the C library per value."""
    ang = np.asarray(ang, dtype='f8')
    uniq, inv = np.unique(ang, return_inverse=True)
    c = np.array([math.cos(-a) for a in uniq])
    s = np.array([math.sin(-a) for a in uniq])
    return c[inv], s[inv]


def photon_index(t, size, rc, rs, qx, qy):
    """maxLP.py:250-252 + 255-258: detector pixel of model pixel t for each frame.

    rots = [[c, -s], [s, c]] with c = cos(-a), s = sin(-a); the product is evaluated as
    ``c*qx + (-s)*qy`` and ``s*qx + c*qy`` without fused multiply-add, then
    ``*size + size//2``, rint (half to even) and ``astype('i4')``; the flat index is
    ``x*size + y`` with no clamp (maxLP.py:252).  At even sizes x or y can reach ``size``
    for pixels near |q| = 0.5: y = size is then the first pixel of row x+1 (plain
    arithmetic); x = size gives an index >= npix, which the photons matrix's fancy
    indexing reduces modulo its npix columns (cupyx/scipy/sparse/_index.py
    ``_asindices``: ``x % length``).  That last case is parity unpinned (CuPy is not
    importable here), and so is the ulp behaviour of CuPy's matmul at exact half-integers.
    """
    px = (rc * qx[t] + (-rs) * qy[t]) * size + size // 2
    py = (rs * qx[t] + rc * qy[t]) * size + size // 2
    x = np.rint(px).astype('i4').astype(np.int64)
    y = np.rint(py).astype('i4').astype(np.int64)
    return ((x * size + y) % (size * size)).astype(np.int64)


def fref_d(t, qx, qy, qrad, diams, shifts):
    """maxLP.py:238-243 ``get_fref_d``: sphere(pi*qrad*dia) * exp(2*pi*i*(q.s))."""
    sval = PI * qrad[t] * diams
    fref = (np.sin(sval) - sval * np.cos(sval)) / sval ** 3
    theta = (2 * PI) * (qx[t] * shifts[:, 0] + qy[t] * shifts[:, 1])
    return fref * np.cos(theta) + 1j * (fref * np.sin(theta))


class MaxLPOracle:
    """maxLP.py:12-340 ``MaxLPhaser`` restricted to the live path of ``_run_phaser``."""

    def __init__(self, frames: SparseFramesNP, size):
        self.size = size
        self.frames = frames
        self.num_data = frames.num_data
        self.qx, self.qy, self.qrad, self.mask, self.rmask = qvals(size)
        self.pattern = pattern(8)
        self.dense = frames.dense()          # photons (N, npix), duplicates summed

    def set_frames(self, sx, sy, dia, ang):
        """maxLP.py:306-311."""
        self.shifts = np.array([sx, sy], dtype='f8').T
        self.diams = np.asarray(dia, dtype='f8')
        self.rc, self.rs = frame_rot_cos_sin(ang)

    def pixel_constants(self, t):
        """maxLP.py:107-114 ``get_pixel_constants`` (all frames)."""
        tv = photon_index(t, self.size, self.rc, self.rs, self.qx, self.qy)
        k_d = self.dense[np.arange(self.num_data), tv]
        return {'k_d': k_d, 'fref_d': fref_d(t, self.qx, self.qy, self.qrad, self.diams,
                                              self.shifts)}

    @staticmethod
    def rescale_pixel(fobj_t, const):
        """maxLP.py:216-228 ``get_rescale_pixel``."""
        f_dt = get_f_dt(np.atleast_1d(fobj_t), const['fref_d'])
        return const['k_d'].mean() / (np.abs(f_dt) ** 2).mean(0)

    @classmethod
    def logq_pixel(cls, fobj_t, const, rescale=None):
        """maxLP.py:181-197 ``get_logq_pixel``: mean(K log w) - mean(w)."""
        fobj_t = np.atleast_1d(fobj_t)
        if rescale is None:
            rescale = cls.rescale_pixel(fobj_t, const)
        f_dt = get_f_dt(fobj_t, const['fref_d'])
        w_dt = rescale * np.abs(f_dt) ** 2
        with np.errstate(divide='ignore', invalid='ignore'):
            return (const['k_d'][:, None] * np.log(w_dt)).mean(0) - w_dt.mean(0)

    def iterate_pixel_pattern(self, fobj_t, step, const, rescale=None, trace=None):
        """maxLP.py:169-179."""
        vals = self.logq_pixel(fobj_t + self.pattern * step, const, rescale)
        imax = int(np.argmax(vals))
        if trace is not None:
            srt = np.sort(vals[~np.isnan(vals)])
            trace.append((imax, srt[-1] - srt[-2] if srt.size > 1 else np.inf))
        retf = fobj_t + self.pattern[imax] * step
        if imax == self.pattern.size // 2:
            return retf, step / (self.pattern.size ** 0.5 / 5)
        return retf, step

    def run_pixel_pattern(self, fobj_t, t, num_iter=10, rescale=None, trace=None):
        """maxLP.py:154-167 (break returns the pre-step value)."""
        const = self.pixel_constants(t)
        step = np.abs(fobj_t) / 4.
        for _ in range(num_iter):
            fnew, snew = self.iterate_pixel_pattern(fobj_t, step, const, rescale, trace)
            if snew / np.abs(fnew) < 1.e-3:
                break
            fobj_t = fnew
            step = snew
        return fobj_t

    def run_phaser(self, model, sx, sy, dia, ang, pixels=None, traces=None):
        """maxLP.py:304-340 ``_run_phaser`` (phase 1 rmask, rescale, phase 3 mask).

        ``pixels`` restricts phase 3 to a subset (phase 1/2 always run in full);
        ``traces`` (dict) collects per-pixel (argmax, margin) sequences.
        """
        self.set_frames(sx, sy, dia, ang)
        fconv = np.asarray(model, dtype='c16').copy().ravel()
        for t in np.nonzero(self.rmask)[0]:
            tr = [] if traces is not None else None
            fconv[t] = self.run_pixel_pattern(fconv[t], t, num_iter=10, trace=tr)
            if traces is not None:
                traces[('p1', int(t))] = tr
        rescales = np.zeros(self.size ** 2)
        for t in np.nonzero(self.rmask)[0]:
            rescales[t] = self.rescale_pixel(fconv[t], self.pixel_constants(t))[0]
        rescale = np.mean(rescales[self.rmask])
        self.last_rescale = rescale
        todo = np.nonzero(self.mask)[0] if pixels is None else np.asarray(pixels)
        for t in todo:
            if not self.mask[t]:
                continue
            tr = [] if traces is not None else None
            fconv[t] = self.run_pixel_pattern(fconv[t], t, rescale=rescale, num_iter=10,
                                              trace=tr)
            if traces is not None:
                traces[('p3', int(t))] = tr
        return fconv
