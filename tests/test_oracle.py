"""Pins of the FP64 oracle by closed forms, invariants and brute force (CPU only).

The reference ships no golden vectors or tests, so each oracle function is pinned by
mathematics that a plausible transcription mistake (sign, index order, dropped term)
would break.
"""
import math

import numpy as np
import pytest

from oracle import emc_oracle as O


# ---- geometry (A4, A13) ----------------------------------------------------------

@pytest.mark.parametrize('size,scored,mask,rmask', [(64, 3012, 3156, 220),
                                                     (185, 24912, 26836, 1708),
                                                     (256, 48224, 51380, 3284)])
def test_mask_counts(size, scored, mask, rmask):
    assert (O.probmask(size) == 0).sum() == scored
    _, _, _, m, r = O.qvals(size)
    assert m.sum() == mask and r.sum() == rmask


def test_pattern():
    p = O.pattern(8)
    assert p.size == 1089 and p[544] == 0
    assert p[0] == -8 - 8j and p[-1] == 8 + 8j
    assert p[1] == -8 - 7.5j        # 'ij': the second axis is the imaginary offset


def test_state_grid_order_and_unravel():
    sx, sy, dia, ns = O.state_grid((-1, 1, 3), (-2, 2, 5), (7, 9, 2))
    assert list(ns) == [3, 5, 2]
    # state n = (i_sx*n_sy + i_sy)*n_dia + i_dia
    n = (2 * 5 + 1) * 2 + 1
    assert (sx[n], sy[n], dia[n]) == (1., -1., 9.)
    h = n * 4 + 3
    v = O.unravel_rmax(np.array([h]), ns, 4, sx, sy, dia)
    assert (v[0][0], v[1][0], v[2][0]) == (1., -1., 9.)
    assert v[3][0] == 3 * np.pi / 4


# ---- slice_gen_holo (A8) -----------------------------------------------------------

def test_holo_zero_model_is_sphere_squared():
    size = 21
    bg = np.zeros(size * size)
    v = O.slice_gen_holo(np.zeros(size * size), 1.5, -0.5, 8., 1., 1., size, bg, 0)
    x, y = np.indices((size, size))
    r = np.sqrt((x - 10.) ** 2 + (y - 10.) ** 2).ravel()
    s = r * np.pi * 8. / size
    s[s == 0] = 1e-5
    sph = (np.sin(s) - s * np.cos(s)) / s ** 3
    np.testing.assert_allclose(v, sph ** 2, rtol=1e-12, atol=1e-300)
    # s -> 0 limit of the form factor is 1/3; at the guard value s = 1e-5 the FP64
    # difference sin s - s cos s cancels to ~1e-6 relative accuracy (as in the reference)
    assert abs(v[10 * size + 10] - 1. / 9.) < 1e-5


def test_holo_ramp_and_shift_sign():
    size = 17
    bg = np.zeros(size * size)
    m = np.full(size * size, 0.1 + 0.0j)
    v = O.slice_gen_holo(m, 1., 0., 8., 1., 1., size, bg, 0)
    x, y = np.indices((size, size))
    xc, yc = (x - 8.).ravel(), (y - 8.).ravel()
    s = np.sqrt(xc ** 2 + yc ** 2) * np.pi * 8. / size
    s[s == 0] = 1e-5
    sph = (np.sin(s) - s * np.cos(s)) / s ** 3
    ph = 2 * np.pi * xc / size      # shift along the first (x) axis
    want = (sph * np.cos(ph) + 0.1) ** 2 + (sph * np.sin(ph)) ** 2
    np.testing.assert_allclose(v, want, rtol=1e-12)
    # rel_scale and scale are linear factors, bg adds
    v2 = O.slice_gen_holo(np.zeros(size * size), 0., 0., 8., 2., 3., size, bg + 0.5, 0)
    v1 = O.slice_gen_holo(np.zeros(size * size), 0., 0., 8., 1., 1., size, bg, 0)
    np.testing.assert_allclose(v2, 4 * 3 * v1 + 0.5, rtol=1e-12)


# ---- slice_gen (A9) / slice_merge (A22) -----------------------------------------------

def test_slice_gen_identity_and_rim():
    size = 16
    rng = np.random.default_rng(0)
    img = rng.random(size * size)
    bg = np.zeros(size * size)
    v = O.slice_gen(img, 0., 1., size, bg, 0).reshape(size, size)
    im = img.reshape(size, size)
    np.testing.assert_array_equal(v[:-1, :-1], im[:-1, :-1])
    assert np.all(v[-1, :] == 0) and np.all(v[:, -1] == 0)
    lv = O.slice_gen(img, 0., 1., size, bg, 1).reshape(size, size)
    assert np.all(lv[-1, :] == -1000.)
    np.testing.assert_allclose(lv[:-1, :-1], np.log(im[:-1, :-1]), rtol=1e-15)


def test_slice_gen_quarter_turn_is_permutation():
    size = 15
    c = size // 2
    rng = np.random.default_rng(1)
    img = rng.random(size * size)
    v = O.slice_gen(img, np.pi / 2, 1., size, np.zeros(size * size), 0).reshape(size, size)
    im = img.reshape(size, size)
    # tx = -(y-c)+c, ty = (x-c)+c  -> view[x, y] = model[2c - y, x]
    for x in range(size - 1):
        for y in range(1, size):
            if 0 <= 2 * c - y <= size - 2 and x <= size - 2:
                assert abs(v[x, y] - im[2 * c - y, x]) < 1e-12


def test_slice_gen_partition_of_unity_and_bg_twice():
    size = 21
    const = np.full(size * size, 2.5)
    bg = np.full(size * size, 0.25)
    v = O.slice_gen(const + bg, 0.37, 1., size, bg, 0)
    _, _, _, _, oob = O._rot_coords(size, math.cos(0.37), math.sin(0.37))
    np.testing.assert_allclose(v[~oob], 2.5 + 2 * 0.25, rtol=1e-14)


def test_slice_merge_is_transpose_of_slice_gen():
    size = 19
    rng = np.random.default_rng(2)
    x = rng.random(size * size)
    y = rng.random(size * size)
    bg = np.zeros(size * size)
    for ang in (0.2, 1.3, 2.9):
        gx = O.slice_gen(x, ang, 1., size, bg, 0)
        mt = np.zeros(size * size)
        w = np.zeros(size * size)
        O.slice_merge(y, ang, size, mt, w)
        assert abs(gx @ y - x @ mt) < 1e-10 * abs(gx @ y)
        # weights: merge of ones == column sums of the interpolation matrix
        O.slice_merge(np.zeros(size * size), ang, size, np.zeros(size * size), w2 := np.zeros(size * size))
        _, _, _, _, oob = O._rot_coords(size, math.cos(ang), math.sin(ang))
        assert abs(w2.sum() - (~oob).sum()) < 1e-9


# ---- calc_prob_all (A10) and the E-step -----------------------------------------------

def _frames(seed=0, n=20, size=17):
    rng = np.random.default_rng(seed)
    K = rng.poisson(0.08, (n, size * size))
    K[3] = 0                                    # an empty frame
    ones, multi, po, pm, cm = [], [], [], [], []
    for f in K:
        o = np.nonzero(f == 1)[0]
        m = np.nonzero(f > 1)[0]
        ones.append(len(o)); multi.append(len(m)); po.append(o); pm.append(m); cm.append(f[m])
    cat = lambda l: np.concatenate(l).astype('i4')
    return O.SparseFramesNP(ones, multi, cat(po), cat(pm), cat(cm), size * size), K


def test_calc_prob_all_brute_force():
    fr, K = _frames()
    size = 17
    rng = np.random.default_rng(3)
    lv = rng.normal(size=size * size)
    pm = O.probmask(size)
    sc = rng.random(fr.num_data) + 0.5
    got = O.calc_prob_all(lv, pm, fr.ones, fr.multi, fr.ones_accum, fr.multi_accum,
                          fr.place_ones, fr.place_multi, fr.count_multi, -2., sc)
    want = -2. * sc + K @ np.where(pm < 1, lv, 0.)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.calc_prob_all_vec(lv, pm, fr) - 2. * sc, want, rtol=1e-12,
                               atol=1e-12)
    assert got[3] == -2. * sc[3]                # empty frame: init term only


def test_estep_empty_frame_ties_resolve_to_first_rotation():
    fr, K = _frames(n=6)
    size = 17
    sx, sy, dia, ns = O.state_grid((-1, 1, 2), (0, 0, 1), (7, 8, 2))
    model = np.random.default_rng(4).random(size * size) + 0j
    prob, msums, vscale = O.estep_prob(model, size, (sx, sy, dia), 5, fr, sequential=True)
    mx, rmax = O.get_rmax(prob)
    s_best = np.argmax(-msums * vscale)
    assert rmax[3] == s_best * 5                  # rotation 0 of the best state
    np.testing.assert_allclose(msums.sum() * vscale,
                               len(sx) * fr.get_powder()[O.probmask(size) < 1].sum(), rtol=1e-12)


def test_lse_normalises():
    rng = np.random.default_rng(5)
    prob = rng.normal(size=(30, 7)) * 50
    P, logz = O.log_sum_exp_prob(prob)
    np.testing.assert_allclose(P.sum(0), 1., rtol=1e-14)
    np.testing.assert_allclose(logz, np.log(np.exp(prob - prob.max(0)).sum(0)) + prob.max(0))


# ---- MaxLP (A17-A20) -----------------------------------------------------------------

def test_photon_index_identity_and_quarter_turn():
    size = 25
    qx, qy, _, _, _ = O.qvals(size)
    t = np.arange(size * size)
    c, s = O.frame_rot_cos_sin(np.array([0.]))
    idx = np.array([O.photon_index(tt, size, c, s, qx, qy)[0] for tt in t])
    np.testing.assert_array_equal(idx, t)
    # angle pi/2: rots = R(-pi/2) = [[0, 1], [-1, 0]] -> (x', y') = (y, -x) about the centre
    c, s = O.frame_rot_cos_sin(np.array([np.pi / 2]))
    x, y = np.divmod(t, size)
    want = (y) * size + (2 * (size // 2) - x)
    idx = np.array([O.photon_index(tt, size, c, s, qx, qy)[0] for tt in t])
    np.testing.assert_array_equal(idx, want)


def test_fref_closed_form():
    size = 21
    qx, qy, qrad, _, _ = O.qvals(size)
    t = 5 * size + 13
    dia = np.array([7., 8.5])
    sh = np.array([[0., 0.], [1., -2.]])
    R = O.fref_d(t, qx, qy, qrad, dia, sh)
    s = np.pi * qrad[t] * dia
    f = 3 * (np.sin(s) - s * np.cos(s)) / s ** 3 / 3
    assert abs(R[0] - f[0]) < 1e-15            # zero shift: real reference
    np.testing.assert_allclose(np.abs(R), np.abs(f), rtol=1e-13)
    assert abs(np.angle(R[1]) - np.angle(np.exp(2j * np.pi * (qx[t] * 1. - qy[t] * 2.)) *
                                         np.sign(f[1]))) < 1e-12


def test_logq_single_frame_maximum_and_rescale():
    """N=1, phase-3 form: LL(f) = K log(r|f+R|^2) - r|f+R|^2 peaks where r|f+R|^2 = K."""
    const = {'k_d': np.array([3.]), 'fref_d': np.array([0.2 + 0.1j])}
    r = 2.
    f = np.linspace(0.5, 2., 2001) + 0j
    ll = O.MaxLPOracle.logq_pixel(f, const, rescale=r)
    fbest = f[np.argmax(ll)]
    assert abs(r * abs(fbest + const['fref_d'][0]) ** 2 - 3.) < 1e-2
    # per-candidate rescale makes mean(w) equal mean(K)
    resc = O.MaxLPOracle.rescale_pixel(f[:5], const)
    np.testing.assert_allclose(resc * np.abs(f[:5] + const['fref_d'][0]) ** 2, 3., rtol=1e-14)


def test_pattern_search_break_returns_pre_step_value():
    """maxLP.py:160-165: when the step becomes tiny the pre-step value is returned."""
    fr, K = _frames(n=8, size=17)
    orc = O.MaxLPOracle(fr, 17)
    orc.set_frames(np.zeros(8), np.zeros(8), np.full(8, 8.), np.zeros(8))
    t = int(np.nonzero(orc.mask)[0][10])
    const = orc.pixel_constants(t)
    f0 = 0.3 + 0.2j
    tr = []
    out = orc.run_pixel_pattern(f0, t, num_iter=10, rescale=1.5, trace=tr)
    # replay the trace
    f, step = f0, abs(f0) / 4.
    for imax, _ in tr:
        fn = f + orc.pattern[imax] * step
        sn = step / 6.6 if imax == 544 else step
        if sn / abs(fn) < 1e-3:
            break
        f, step = fn, sn
    assert out == f


def test_photon_index_even_size_wraps_as_the_reference_indexes():
    """maxLP.py:251-252 at an even size: rint can give x or y = size (no clamp).
    size 16, centre 8.  Pixel (11, 15) has q*size = (3, 7), |q| = 0.476 < 0.5 (in mask).
    With rots = [[c, -s], [s, c]], (c, s) = (7, 3)/|(7, 3)|: y' = 3s + 7c + 8 = 15.62 -> 16,
    x' = 3c - 7s + 8 = 8 -> flat index 8*16 + 16 = 144, the first pixel of row 9.
    Pixel (15, 11), (c, s) = (7, -3)/|(7, -3)|: x' = 7c - 3(-s)... = 15.62 -> 16,
    y' = 7s + 3c + 8 = 8 -> 16*16 + 8 = 264 >= 256, which the photons matrix's fancy
    indexing reduces modulo 256 (cupyx _asindices) to 8 = pixel (0, 8)."""
    size = 16
    qx, qy, _, mask, _ = O.qvals(size)
    n = np.hypot(7., 3.)
    t1 = 11 * size + 15
    assert mask[t1]
    got = O.photon_index(t1, size, np.array([7. / n]), np.array([3. / n]), qx, qy)
    assert got[0] == 9 * size + 0
    t2 = 15 * size + 11
    assert mask[t2]
    got = O.photon_index(t2, size, np.array([7. / n]), np.array([-3. / n]), qx, qy)
    assert got[0] == 8
    # odd sizes never reach `size`: every index stays the plain rint of the rotation
    s = 17
    qx, qy, _, mask, _ = O.qvals(s)
    ang = np.linspace(0, np.pi, 181)
    rc, rs = O.frame_rot_cos_sin(ang)
    for t in np.nonzero(mask)[0]:
        px = np.rint((rc * qx[t] - rs * qy[t]) * s + s // 2)
        py = np.rint((rs * qx[t] + rc * qy[t]) * s + s // 2)
        assert px.max() <= s - 1 and py.max() <= s - 1


def test_merge_all_equals_dense_product():
    """kernels.cu:139-152: view[p] = sum_d prob_r[d] * K_dp, i.e. K^T prob_r (brute force
    on the densified frames)."""
    fr, K = _frames(n=25)
    rng = np.random.default_rng(11)
    pr = rng.random(fr.num_data)
    view = O.merge_all(pr, fr.ones, fr.multi, fr.ones_accum, fr.multi_accum, fr.place_ones,
                       fr.place_multi, fr.count_multi, 17 * 17)
    np.testing.assert_allclose(view, K.T @ pr, rtol=1e-13, atol=1e-15)
    assert view[np.nonzero(K.sum(0) == 0)[0]].max(initial=0.) == 0.


def test_rotation_angles_are_a_half_turn():
    """emc.py:192 and 223: rotation r is the angle r*pi/R.  A frame whose photons follow the
    log-view rotated by pi/2 (built with slice_gen at the explicit angle) must take the
    rotation r = R/2 of R = 8 (a table of r*2*pi/R would put pi/2 at r = 2), and its
    unravelled angle is pi/2."""
    size = 21
    rng = np.random.default_rng(12)
    bg = np.zeros(size * size)
    model = 0.2 * (rng.random(size * size) + 1j * rng.random(size * size))
    states = O.state_grid((0.5, 0.5, 1), (-0.5, -0.5, 1), (8., 8., 1))
    sx, sy, dia, ns = states
    holo = O.slice_gen_holo(model, sx[0], sy[0], dia[0], 1., 1., size, bg, 0)
    rot = O.slice_gen(holo, np.pi / 2, 1., size, bg, 0)
    pm = O.probmask(size)
    lam = np.where(pm < 1, rot, 0.)
    K = np.floor(400. * lam / lam.max()).astype(int)
    po = np.repeat(np.arange(size * size), K).astype('i4')
    fr = O.SparseFramesNP([len(po)], [0], po, np.zeros(0, 'i4'), np.zeros(0, 'i4'),
                          size * size)
    prob, _, _ = O.estep_prob(model, size, (sx, sy, dia), 8, fr)
    _, rmax = O.get_rmax(prob)
    assert rmax[0] == 4
    ang = O.unravel_rmax(rmax, ns, 8, sx, sy, dia)[3]
    assert ang[0] == np.pi / 2


def test_run_phaser_composition_single_frame():
    """maxLP.py:304-340 with one frame, where every quantity has a closed form:
    * pixels outside `mask` are returned unchanged;
    * phase 2's rescale is the mean over the rmask pixels of K_t / |f_t + R_t|^2 at the
      phase-1 values f_t (get_rescale_pixel with N = 1);
    * phase 3 maximises K log(r|f+R|^2) - r|f+R|^2 with that global r, whose optimum is the
      circle r|f + R|^2 = K: every searched pixel with K > 0 ends close to it."""
    size = 17
    rng = np.random.default_rng(13)
    K = rng.integers(1, 6, size * size)
    po = np.repeat(np.arange(size * size), K).astype('i4')
    fr = O.SparseFramesNP([len(po)], [0], po, np.zeros(0, 'i4'), np.zeros(0, 'i4'),
                          size * size)
    model = 0.3 * (rng.random(size * size) + 1j * rng.random(size * size)) + 0.2
    orc = O.MaxLPOracle(fr, size)
    args = (np.array([0.5]), np.array([-0.3]), np.array([8.]), np.array([0.]))
    ring = np.nonzero(orc.rmask)[0]
    after1 = orc.run_phaser(model, *args, pixels=[])        # phase 3 over no pixel
    out = orc.run_phaser(model, *args)
    np.testing.assert_array_equal(out[~orc.mask], model[~orc.mask])
    np.testing.assert_array_equal(after1[~orc.rmask], model[~orc.rmask])
    Rt = np.array([O.fref_d(t, orc.qx, orc.qy, orc.qrad, np.array([8.]),
                            np.array([[0.5, -0.3]]))[0] for t in range(size * size)])
    # angle 0: the detector pixel of model pixel t is t itself, so K_t is the frame's count
    resc = np.mean(K[ring] / np.abs(after1[ring] + Rt[ring]) ** 2)
    assert abs(orc.last_rescale - resc) <= 1e-13 * resc
    m = np.nonzero(orc.mask)[0]
    w = orc.last_rescale * np.abs(out[m] + Rt[m]) ** 2
    # the search stops when its step is below 1e-3 of |f|: within a few steps of the circle
    assert np.median(np.abs(w / K[m] - 1.)) < 2e-4
    assert np.max(np.abs(w / K[m] - 1.)) < 3e-3
