"""The oracle's C restatement (oracle/emc_oracle.c) pinned to its NumPy restatement
(oracle/emc_oracle.py), function by function, on small inputs (CPU only).

Both restate the same reference lines in the same operation order; they may differ only
in the last bits of the transcendental functions (C library vs NumPy), so the tests
require equality where no transcendental enters (calc_prob_all, merge_all, the photon
index) and agreement to a few ulps elsewhere.  The NumPy restatement is itself pinned by
the closed forms and brute force of tests/test_oracle.py.
"""
import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import emc_oracle as O


@pytest.fixture(scope='module', autouse=True)
def _built():
    CO.build()
    CO.load()


def _frames(n=40, size=17, seed=0, bg=False):
    rng = np.random.default_rng(seed)
    K = np.zeros((n, size * size), dtype=np.int64)
    ones, multi, po, pm, cm = [], [], [], [], []
    for d in range(n):
        m = rng.integers(0, 30) if d != 3 else 0
        p1 = rng.integers(0, size * size, m)
        k2 = rng.integers(0, 4)
        p2 = rng.integers(0, size * size, k2)
        c2 = rng.integers(2, 5, k2)
        ones.append(m)
        multi.append(k2)
        po.append(p1)
        pm.append(p2)
        cm.append(c2)
        np.add.at(K[d], p1, 1)
        np.add.at(K[d], p2, c2)
    cat = lambda a: np.concatenate(a).astype('i4') if a else np.zeros(0, 'i4')
    b = rng.random(size * size) * 0.01 if bg else None
    return O.SparseFramesNP(ones, multi, cat(po), cat(pm), cat(cm), size * size, bg=b), K


def test_slice_gen_holo_matches_numpy():
    size = 23
    rng = np.random.default_rng(1)
    model = rng.normal(size=size * size) + 1j * rng.normal(size=size * size)
    bg = rng.random(size * size)
    for args in ((0.5, -1.5, 7.3, 1., 1.), (0., 0., 8., 2., 3.)):
        for lf in (0, 1):
            a = O.slice_gen_holo(model, *args, size, bg, lf)
            b = CO.slice_gen_holo(model, *args, size, bg, lf)
            np.testing.assert_allclose(b, a, rtol=4e-15, atol=0)


@pytest.mark.parametrize('size', [16, 21])
def test_slice_gen_matches_numpy(size):
    rng = np.random.default_rng(size)
    img = rng.random(size * size) + 0.01
    bg = rng.random(size * size) * 0.1
    for ang in (0., 0.3, np.pi / 2, 2.9):
        cs = (np.cos(ang), np.sin(ang))
        for lf in (0, 1):
            a = O.slice_gen(img, ang, 1.3, size, bg, lf, cs)
            b = CO.slice_gen(img, ang, 1.3, size, bg, lf, cs)
            if lf:
                np.testing.assert_allclose(b, a, rtol=0, atol=4e-16 * np.abs(a).max())
            else:
                np.testing.assert_array_equal(b, a)      # no transcendental: same bits


def test_slice_merge_matches_numpy():
    size = 19
    rng = np.random.default_rng(2)
    view = rng.random(size * size)
    for ang in (0.2, 1.7):
        cs = (np.cos(ang), np.sin(ang))
        m1, w1 = np.zeros(size * size), np.zeros(size * size)
        m2, w2 = np.zeros(size * size), np.zeros(size * size)
        O.slice_merge(view, ang, size, m1, w1, cs)
        CO.slice_merge(view, ang, size, m2, w2, cs)
        # the same terms, added in another order (NumPy adds corner by corner)
        np.testing.assert_allclose(m2, m1, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(w2, w1, rtol=1e-14, atol=1e-15)


def test_calc_prob_all_and_merge_all_equal_numpy_bit_for_bit():
    fr, K = _frames()
    size = 17
    rng = np.random.default_rng(3)
    lv = rng.normal(size=size * size)
    pm = O.probmask(size)
    sc = rng.random(fr.num_data) + 0.5
    a = O.calc_prob_all(lv, pm, fr.ones, fr.multi, fr.ones_accum, fr.multi_accum,
                        fr.place_ones, fr.place_multi, fr.count_multi, -2.5, sc)
    np.testing.assert_array_equal(CO.calc_prob_all(lv, pm, fr, -2.5, sc), a)
    pr = rng.random(fr.num_data)
    a = O.merge_all(pr, fr.ones, fr.multi, fr.ones_accum, fr.multi_accum, fr.place_ones,
                    fr.place_multi, fr.count_multi, size * size)
    np.testing.assert_array_equal(CO.merge_all(pr, fr, size * size), a)


@pytest.mark.parametrize('bg,scaled', [(False, False), (True, True)])
def test_estep_top2_matches_numpy_prob(bg, scaled):
    size = 17
    fr, _ = _frames(n=30, size=size, seed=7, bg=bg)
    rng = np.random.default_rng(8)
    model = 0.3 * (rng.random(size * size) + 1j * rng.random(size * size))
    states = O.state_grid((-1., 1., 3), (-1., 1., 2), (7., 9., 2))[:3]
    scales = rng.random(fr.num_data) + 0.5 if scaled else None
    R = 6
    prob, msums, vscale = O.estep_prob(model, size, states, R, fr, scales=scales,
                                       sequential=True)
    mx, rmax = O.get_rmax(prob)
    srt = np.sort(prob, axis=0)
    r2, l2, s2, ms2, vs2 = CO.estep_top2(model, size, states, R, fr, scales=scales, nthreads=3)
    np.testing.assert_array_equal(ms2, msums)
    assert vs2 == vscale
    np.testing.assert_allclose(l2, mx, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(s2, srt[-2], rtol=1e-13, atol=1e-12)
    margin = (srt[-1] - srt[-2]) / np.abs(srt[-1])
    bad = np.nonzero(r2 != rmax)[0]
    assert np.all(margin[bad] < 1e-12), (bad, margin[bad])
    # thread count does not change the result (order-independent top-2 merge)
    r3, l3, s3, _, _ = CO.estep_top2(model, size, states, R, fr, scales=scales, nthreads=1)
    np.testing.assert_array_equal(r3, r2)
    np.testing.assert_array_equal(l3, l2)
    np.testing.assert_array_equal(s3, s2)


@pytest.mark.parametrize('size', [21, 64])
def test_photon_index_matches_numpy(size):
    qx, qy, _, mask, _ = O.qvals(size)
    rng = np.random.default_rng(size)
    ang = rng.random(200) * np.pi
    ang[:90] = np.arange(90) * np.pi / 90
    rc, rs = O.frame_rot_cos_sin(ang)
    wraps = 0
    for t in np.nonzero(mask)[0][::7]:
        a = O.photon_index(t, size, rc, rs, qx, qy)
        b = CO.photon_index(t, size, rc, rs)
        np.testing.assert_array_equal(b, a)
    for t in np.nonzero(mask)[0]:
        px = np.rint((rc * qx[t] - rs * qy[t]) * size + size // 2)
        py = np.rint((rs * qx[t] + rc * qy[t]) * size + size // 2)
        wraps += int(((px == size) | (py == size)).sum())
    if size % 2 == 0:
        assert wraps > 0          # the even-size wrap is exercised


def _maxlp_case(size, n, seed):
    rng = np.random.default_rng(seed)
    fr, _ = _frames(n=n, size=size, seed=seed)
    model = 0.05 * (rng.random(size * size) + 1j * rng.random(size * size)) + 0.02
    sx = rng.uniform(-2, 2, n)
    sy = rng.uniform(-2, 2, n)
    dia = rng.uniform(7, 9, n)
    ang = rng.integers(0, 12, n) * np.pi / 12
    return fr, model, sx, sy, dia, ang


@pytest.mark.parametrize('size', [21, 20])
def test_maxlp_run_phaser_matches_numpy(size):
    fr, model, sx, sy, dia, ang = _maxlp_case(size, 24, size)
    orc = O.MaxLPOracle(fr, size)
    sub = np.nonzero(orc.mask)[0][::9]
    traces = {}
    a = orc.run_phaser(model, sx, sy, dia, ang, pixels=sub, traces=traces)
    mc = CO.MaxLPC(fr, size)
    b, resc, info = mc.run_phaser(model, sx, sy, dia, ang, pixels=sub, nthreads=4)
    assert abs(resc - orc.last_rescale) <= 1e-13 * abs(resc)
    chk = np.union1d(np.nonzero(orc.rmask)[0], sub)
    # differences allowed only where a step's top-2 margin is at rounding level
    near = set()
    for (ph, t), tr in traces.items():
        if any(m < 1e-12 * 10 for _, m in tr):
            near.add(t)
    for t in chk:
        if t in near:
            continue
        assert abs(b[t] - a[t]) <= 1e-12 * abs(a[t]), (t, a[t], b[t])
    untouched = np.nonzero(~orc.mask)[0]
    np.testing.assert_array_equal(b[untouched], np.asarray(model)[untouched])


def test_maxlp_logq_matches_numpy():
    size = 21
    fr, model, sx, sy, dia, ang = _maxlp_case(size, 24, 5)
    orc = O.MaxLPOracle(fr, size)
    orc.set_frames(sx, sy, dia, ang)
    mc = CO.MaxLPC(fr, size)
    mc.set_frames(sx, sy, dia, ang)
    t = int(np.nonzero(orc.mask)[0][40])
    const = orc.pixel_constants(t)
    cands = 0.03 + 0.01j + orc.pattern * 0.004
    for resc in (None, 1.7):
        a = O.MaxLPOracle.logq_pixel(cands, const, resc)
        b, k, R = mc.logq(t, cands, resc, nthreads=2)
        np.testing.assert_array_equal(k, const['k_d'])
        np.testing.assert_allclose(R, const['fref_d'], rtol=1e-14, atol=1e-16)
        np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-14)
