"""FP64 NumPy oracle of the rotation-only form: cross term, LL, LSE, merge, adjoint.

TEST INFRASTRUCTURE ONLY (same rule as ``emc_oracle``): only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may use it.

The reference searches (state x rotation) with rotated intensities (emc.py:159-216).
The split-precision, adjoint and streaming workloads instead use per-frame reference
parameters (shift, diameter) and rotations as the only latent variable:

    O_r(p)   = bilinear rotation of the complex model by j*pi/R, re and im separately
               (kernels.cu:6-42 geometry: floor indices, out of bounds -> 0)
    R_d(p)   = rel * S(pi*|p-c|*dia_d/size) * exp(2*pi*i*((x-c)*sx_d + (y-c)*sy_d)/size)
               (the sphere term of kernels.cu:44-82, s == 0 -> 1e-5)
    W_dr(p)  = scale_d * |O_r(p) + R_d(p)|^2
    LL_dr    = sum_{photons p} K_dp * log W_dr(p) - sum_p W_dr(p)
               (log clamp of kernels.cu:37-40: W < 1e-20 -> -1000)

over the scored pixels ``probmask == 0`` (emc.py:89-98).  The dense sum expands as
sum_p W = scale_d * (|O_r|^2 + |R_d|^2 + 2 * Re sum_p O_r(p) conj(R_d(p))), whose last term is
the cross-term GEMM.  P(r|d) is the FP64 softmax of LL over r; the merge is
merge_all + slice_merge (kernels.cu:139-152, 84-114) normalised by
weights * sum_d P * scale.
"""
from __future__ import annotations

import math

import numpy as np

from .emc_oracle import PI, probmask, slice_merge

LOG_FLOOR = -1000.


def scored_pixels(size):
    """Flat indices of the scored annulus (probmask == 0), ascending."""
    return np.nonzero(probmask(size) == 0)[0]


def rot_cos_sin(num_rot, full_turn=False):
    """cos/sin of j*pi/num_rot from the C library (emc.py:192), or of j*2pi/num_rot."""
    step = (2. if full_turn else 1.) * PI / num_rot
    a = [j * step for j in range(num_rot)]
    return np.array([math.cos(v) for v in a]), np.array([math.sin(v) for v in a])


def _bilinear_rows(size, ac, as_, pix):
    """Corner indices and weights of slice_gen's bilinear step for output pixels ``pix``
    (kernels.cu:19-33); out-of-bounds rows get weight 0 and index 0."""
    x, y = pix // size, pix % size
    cen = size // 2
    xc = (x - cen).astype('f8')
    yc = (y - cen).astype('f8')
    tx = xc * ac - yc * as_ + cen
    ty = xc * as_ + yc * ac + cen
    ix = np.floor(tx).astype(np.int64)
    iy = np.floor(ty).astype(np.int64)
    ok = ~((ix < 0) | (ix > size - 2) | (iy < 0) | (iy > size - 2))
    fx, fy = tx - ix, ty - iy
    cx, cy = 1. - fx, 1. - fy
    ixs, iys = np.where(ok, ix, 0), np.where(ok, iy, 0)
    idx = np.stack([ixs * size + iys, (ixs + 1) * size + iys, ixs * size + iys + 1,
                    (ixs + 1) * size + iys + 1], 1)
    w = np.stack([cx * cy, fx * cy, cx * fy, fx * fy], 1) * ok[:, None]
    return idx, w


def rot_slices(model_c, size, num_rot, pix, cos_sin=None):
    """O_r(p) for r < num_rot at pixels ``pix``: (R, len(pix)) complex."""
    model_c = np.asarray(model_c, dtype='c16').ravel()
    rc, rs = cos_sin if cos_sin is not None else rot_cos_sin(num_rot)
    out = np.empty((num_rot, len(pix)), dtype='c16')
    for r in range(num_rot):
        idx, w = _bilinear_rows(size, rc[r], rs[r], pix)
        out[r] = (model_c[idx] * w).sum(1)
    return out


def interp_transpose(g, size, num_rot, pix, cos_sin=None):
    """Adjoint of ``rot_slices``: sum_r Interp_r^T g_r -> complex (size*size,)."""
    rc, rs = cos_sin if cos_sin is not None else rot_cos_sin(num_rot)
    out = np.zeros(size * size, dtype='c16')
    for r in range(num_rot):
        idx, w = _bilinear_rows(size, rc[r], rs[r], pix)
        np.add.at(out, idx.ravel(), (g[r][:, None] * w).ravel())
    return out


def sphere(s):
    return (np.sin(s) - s * np.cos(s)) / (s * s * s)


def sphere_ds(s):
    """dS/ds = sin(s)/s^2 - 3*S(s)/s."""
    return np.sin(s) / (s * s) - 3. * sphere(s) / s


def _ref_geom(size, pix):
    cen = math.floor(size / 2.)
    x, y = (pix // size).astype('f8'), (pix % size).astype('f8')
    rad = np.sqrt((x - cen) * (x - cen) + (y - cen) * (y - cen))
    return x - cen, y - cen, rad


def ref_terms(size, pix, sx, sy, dia, rel_scale=1.):
    """R_d(p): (N, len(pix)) complex, the sphere term of slice_gen_holo per frame."""
    xc, yc, rad = _ref_geom(size, pix)
    sx, sy, dia = (np.asarray(a, dtype='f8')[:, None] for a in (sx, sy, dia))
    phase = 2. * PI * (xc * sx + yc * sy) / size
    s = rad * PI * dia / size
    s = np.where(s == 0., 1.e-5, s)
    return rel_scale * sphere(s) * np.exp(1j * phase)


def ref_derivs(size, pix, sx, sy, dia, rel_scale=1.):
    """dR_d/dsx, dR_d/dsy, dR_d/ddia: three (N, len(pix)) complex arrays."""
    xc, yc, rad = _ref_geom(size, pix)
    R = ref_terms(size, pix, sx, sy, dia, rel_scale)
    dsx = R * (2j * PI * xc / size)
    dsy = R * (2j * PI * yc / size)
    sxa, sya, diaa = (np.asarray(a, dtype='f8')[:, None] for a in (sx, sy, dia))
    phase = 2. * PI * (xc * sxa + yc * sya) / size
    s = rad * PI * diaa / size
    s = np.where(s == 0., 1.e-5, s)
    ddia = rel_scale * sphere_ds(s) * (PI * rad / size) * np.exp(1j * phase)
    return dsx, dsy, ddia


def cross_term(Or, Rd):
    """X[r, d] = Re sum_p O_r(p) conj(R_d(p)) (the GEMM), FP64."""
    return (Or.real @ Rd.real.T) + (Or.imag @ Rd.imag.T)


def dense_photons(frames, pix):
    """K[d, j] = photon count of frame d at pixel pix[j] (outside pix ignored)."""
    K = frames.dense()[:, pix] if hasattr(frames, 'dense') else None
    return np.asarray(K, dtype='f8')


def _logw(W):
    out = np.full_like(W, LOG_FLOOR)
    ok = W >= 1.e-20
    out[ok] = np.log(W[ok])
    return out


def rot_loglik(model_c, size, num_rot, K, sx, sy, dia, scales=None, rel_scale=1.,
               pix=None, cos_sin=None):
    """LL[d, r] directly from W (dense, small sizes): (N, R) FP64, plus the pieces."""
    pix = scored_pixels(size) if pix is None else pix
    N = K.shape[0]
    scales = np.ones(N) if scales is None else np.asarray(scales, dtype='f8')
    Or = rot_slices(model_c, size, num_rot, pix, cos_sin)
    Rd = ref_terms(size, pix, sx, sy, dia, rel_scale)
    Z = Or[None, :, :] + Rd[:, None, :]                  # (N, R, P)
    W = scales[:, None, None] * (Z.real ** 2 + Z.imag ** 2)
    LL = (K[:, None, :] * _logw(W)).sum(2) - W.sum(2)
    return LL, dict(Or=Or, Rd=Rd, W=W, Z=Z)


def rot_loglik_split(model_c, size, num_rot, K, sx, sy, dia, scales=None, rel_scale=1.,
                     pix=None, cos_sin=None):
    """The same LL through the GEMM expansion (sparse log term + norms + 2*cross term)."""
    pix = scored_pixels(size) if pix is None else pix
    N = K.shape[0]
    scales = np.ones(N) if scales is None else np.asarray(scales, dtype='f8')
    Or = rot_slices(model_c, size, num_rot, pix, cos_sin)
    Rd = ref_terms(size, pix, sx, sy, dia, rel_scale)
    no = (Or.real ** 2 + Or.imag ** 2).sum(1)             # (R,)
    nr = (Rd.real ** 2 + Rd.imag ** 2).sum(1)             # (N,)
    X = cross_term(Or, Rd)                                # (R, N)
    LL = np.empty((N, num_rot))
    for d in range(N):
        ph = np.nonzero(K[d])[0]
        Z = Or[:, ph] + Rd[d, ph][None, :]
        W = scales[d] * (Z.real ** 2 + Z.imag ** 2)
        LL[d] = (K[d, ph][None, :] * _logw(W)).sum(1) - scales[d] * (no + nr[d] + 2. * X[:, d])
    return LL, dict(no=no, nr=nr, X=X)


def lse(LL):
    """P[d, r] = softmax_r LL[d, r] and logZ[d] = log sum_r exp LL[d, r], FP64."""
    m = LL.max(1, keepdims=True)
    e = np.exp(LL - m)
    s = e.sum(1, keepdims=True)
    return e / s, (m + np.log(s)).ravel()


def total_loss(LL):
    """-sum_d log((1/R) sum_r exp LL[d, r]) (uniform rotation prior)."""
    _, logz = lse(LL)
    return -(logz - math.log(LL.shape[1])).sum()


def merge_views(P, K):
    """merge_all for every rotation: view[r, j] = sum_d P[d, r] * K[d, j]."""
    return P.T @ K


def merge_model(P, K, scales, size, num_rot, pix=None, cos_sin=None):
    """Merged intensity model (npix): sum_r slice_merge(view_r) / sum_r slice_merge(1) *
    sum_d P[d, r]*scale_d; pixels with zero weight -> 0.  Returns (model, acc, weights)."""
    pix = scored_pixels(size) if pix is None else pix
    rc, rs = cos_sin if cos_sin is not None else rot_cos_sin(num_rot)
    view = merge_views(P, K)
    ws = (P * np.asarray(scales, dtype='f8')[:, None]).sum(0)   # (R,)
    acc = np.zeros(size * size)
    wts = np.zeros(size * size)
    for r in range(num_rot):
        full = np.zeros(size * size)
        full[pix] = view[r]
        ones = np.zeros(size * size)
        ones[pix] = 1.
        a = np.zeros(size * size)
        w = np.zeros(size * size)
        slice_merge(full, 0., size, a, w, cos_sin=(rc[r], rs[r]))
        a1 = np.zeros(size * size)
        w1 = np.zeros(size * size)
        slice_merge(ones, 0., size, a1, w1, cos_sin=(rc[r], rs[r]))
        acc += a
        wts += a1 * ws[r]
    model = np.where(wts > 0, acc / np.where(wts > 0, wts, 1.), 0.)
    return model, acc, wts


def loss_grad(model_c, size, num_rot, K, sx, sy, dia, scales=None, rel_scale=1., pix=None,
              cos_sin=None):
    """Loss = total_loss(LL) and its FP64 gradient.

    Returns (loss, g_model (complex npix: dL/dRe + i dL/dIm), g_sx, g_sy, g_dia (N,)).
    dLL/dW = K/W - 1 (log clamp region: -1); dW/dtheta = 2*scale*Re(conj(Z) dZ/dtheta).
    """
    pix = scored_pixels(size) if pix is None else pix
    N = K.shape[0]
    scales = np.ones(N) if scales is None else np.asarray(scales, dtype='f8')
    LL, pc = rot_loglik(model_c, size, num_rot, K, sx, sy, dia, scales, rel_scale, pix,
                        cos_sin)
    P, _ = lse(LL)
    W, Z = pc['W'], pc['Z']
    ok = W >= 1.e-20
    dW = np.where(ok, K[:, None, :] / np.where(ok, W, 1.), 0.) - 1.     # dLL/dW
    # dL/dW = -P * dLL/dW ; coefficient of dW/dtheta
    c = -P[:, :, None] * dW * (2. * scales[:, None, None])             # (N, R, P)
    gz = c * Z                                                         # complex weights
    g_or = gz.sum(0)                                                   # (R, P)
    g_model = interp_transpose(g_or, size, num_rot, pix, cos_sin)
    dsx, dsy, ddia = ref_derivs(size, pix, sx, sy, dia, rel_scale)
    red = lambda dR: (c * (Z.real * dR.real[:, None, :] +
                           Z.imag * dR.imag[:, None, :])).sum((1, 2))
    return total_loss(LL), g_model, red(dsx), red(dsy), red(ddia)


# ----------------------------------------------------------------------------------
# several object classes: hypotheses h = c * R + r (class-major), one model per class
# ----------------------------------------------------------------------------------

def rot_loglik_multi(models, size, num_rot, K, sx, sy, dia, scales=None, rel_scale=1.,
                     pix=None, cos_sin=None):
    """LL[d, c*R + r] for the class models ``models`` (C, npix): each class block is
    :func:`rot_loglik` of that class's model (the reference term is shared)."""
    return np.concatenate([rot_loglik(m, size, num_rot, K, sx, sy, dia, scales, rel_scale, pix,
                                      cos_sin)[0] for m in np.atleast_2d(models)], axis=1)


def merge_model_multi(P, K, scales, size, num_rot, pix=None, cos_sin=None):
    """Per-class merged intensities: class c uses P[:, c*R:(c+1)*R] in :func:`merge_model`
    (its weights w_r = sum_d P[d, c*R + r] * scale_d).  Returns (C, npix) arrays."""
    C = P.shape[1] // num_rot
    out = [merge_model(P[:, c * num_rot:(c + 1) * num_rot], K, scales, size, num_rot, pix,
                       cos_sin) for c in range(C)]
    return tuple(np.stack(v) for v in zip(*out))


def loss_grad_multi(models, size, num_rot, K, sx, sy, dia, scales=None, rel_scale=1., pix=None,
                    cos_sin=None):
    """Loss -sum_d log((1/(C R)) sum_h exp LL[d, h]) and its FP64 gradient for several
    classes: (loss, g_models (C, npix) complex, g_sx, g_sy, g_dia)."""
    models = np.atleast_2d(models)
    pix = scored_pixels(size) if pix is None else pix
    N = K.shape[0]
    scales = np.ones(N) if scales is None else np.asarray(scales, dtype='f8')
    pieces = [rot_loglik(m, size, num_rot, K, sx, sy, dia, scales, rel_scale, pix, cos_sin)
              for m in models]
    LL = np.concatenate([p[0] for p in pieces], axis=1)
    P, _ = lse(LL)
    dsx, dsy, ddia = ref_derivs(size, pix, sx, sy, dia, rel_scale)
    g_models = []
    g = np.zeros((3, N))
    for c, (_, pc) in enumerate(pieces):
        Pc = P[:, c * num_rot:(c + 1) * num_rot]
        W, Z = pc['W'], pc['Z']
        ok = W >= 1.e-20
        dW = np.where(ok, K[:, None, :] / np.where(ok, W, 1.), 0.) - 1.
        cc = -Pc[:, :, None] * dW * (2. * scales[:, None, None])
        g_models.append(interp_transpose((cc * Z).sum(0), size, num_rot, pix, cos_sin))
        for i, dR in enumerate((dsx, dsy, ddia)):
            g[i] += (cc * (Z.real * dR.real[:, None, :] + Z.imag * dR.imag[:, None, :])).sum((1, 2))
    return total_loss(LL), np.stack(g_models), g[0], g[1], g[2]
