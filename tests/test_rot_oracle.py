"""Pins of the rotation-form oracle (oracle/rot_oracle.py), CPU only.

Each function is checked against something other than its own formula: the reference
transcription of slice_gen_holo (emc_oracle), exact index permutations, the adjoint
identity, closed forms, NumPy's logaddexp, mass conservation and central differences.
"""
import math

import numpy as np
import pytest

from oracle import emc_oracle as E
from oracle import rot_oracle as O

SIZE = 17


@pytest.fixture(scope='module')
def setup():
    rng = np.random.default_rng(5)
    n = 6
    model = 0.3 * (rng.random(SIZE * SIZE) + 1j * rng.random(SIZE * SIZE))
    sx, sy = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    dia = rng.uniform(6, 9, n)
    pix = O.scored_pixels(SIZE)
    K = rng.poisson(0.4, (n, len(pix))).astype('f8')
    scales = rng.uniform(0.5, 1.5, n)
    return model, sx, sy, dia, pix, K, scales


def test_rot_slices_identity_and_quarter_turn():
    rng = np.random.default_rng(0)
    m = rng.random(SIZE * SIZE) + 1j * rng.random(SIZE * SIZE)
    allpix = np.arange(SIZE * SIZE)
    o = O.rot_slices(m, SIZE, 1, allpix)[0]
    x, y = allpix // SIZE, allpix % SIZE
    inner = (x <= SIZE - 2) & (y <= SIZE - 2)
    assert np.array_equal(o[inner], m[inner])
    assert np.all(o[~inner] == 0)
    # R = 2: rotation by pi/2 with exact cos/sin maps (x, y) -> (c - (y - c), c + (x - c))
    q = O.rot_slices(m, SIZE, 2, allpix, cos_sin=(np.array([1., 0.]), np.array([0., 1.])))[1]
    c = SIZE // 2
    src_x, src_y = c - (y - c), c + (x - c)
    ok = (src_x >= 0) & (src_x <= SIZE - 2) & (src_y >= 0) & (src_y <= SIZE - 2)
    assert np.array_equal(q[ok], m[src_x[ok] * SIZE + src_y[ok]])


def test_rot_slices_match_slice_gen_on_real_and_imag():
    rng = np.random.default_rng(1)
    m = rng.random(SIZE * SIZE) + 1j * rng.random(SIZE * SIZE)
    allpix = np.arange(SIZE * SIZE)
    R = 5
    rc, rs = O.rot_cos_sin(R)
    o = O.rot_slices(m, SIZE, R, allpix)
    z = np.zeros(SIZE * SIZE)
    for r in range(R):
        re = E.slice_gen(m.real, 0., 1., SIZE, z, 0, cos_sin=(rc[r], rs[r]))
        im = E.slice_gen(m.imag, 0., 1., SIZE, z, 0, cos_sin=(rc[r], rs[r]))
        np.testing.assert_allclose(o[r], re + 1j * im, rtol=0, atol=1e-15)


def test_interp_transpose_is_adjoint():
    rng = np.random.default_rng(2)
    pix = O.scored_pixels(SIZE)
    R = 7
    m = rng.standard_normal(SIZE * SIZE) + 1j * rng.standard_normal(SIZE * SIZE)
    g = rng.standard_normal((R, len(pix))) + 1j * rng.standard_normal((R, len(pix)))
    lhs = np.vdot(g, O.rot_slices(m, SIZE, R, pix))
    rhs = np.vdot(O.interp_transpose(g, SIZE, R, pix), m)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_ref_terms_against_slice_gen_holo():
    """|1+R|^2 - |R|^2 - 1 = 2 Re R and |i+R|^2 - |R|^2 - 1 = 2 Im R, with slice_gen_holo
    (kernels.cu:44-82) giving the |m + R|^2 views."""
    allpix = np.arange(SIZE * SIZE)
    z = np.zeros(SIZE * SIZE)
    for (sx, sy, dia, rel) in ((0.5, -1.25, 7.3, 1.), (-2., 2., 9., 3.), (0., 0., 8., 0.5)):
        R = O.ref_terms(SIZE, allpix, [sx], [sy], [dia], rel)[0]
        v0 = E.slice_gen_holo(np.zeros(SIZE * SIZE), sx, sy, dia, rel, 1., SIZE, z, 0)
        v1 = E.slice_gen_holo(np.ones(SIZE * SIZE), sx, sy, dia, rel, 1., SIZE, z, 0)
        vi = E.slice_gen_holo(1j * np.ones(SIZE * SIZE), sx, sy, dia, rel, 1., SIZE, z, 0)
        np.testing.assert_allclose(np.abs(R) ** 2, v0, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(2 * R.real, v1 - v0 - 1, rtol=0, atol=1e-12)
        np.testing.assert_allclose(2 * R.imag, vi - v0 - 1, rtol=0, atol=1e-12)


def test_ref_derivs_finite_difference():
    pix = O.scored_pixels(SIZE)
    sx, sy, dia = np.array([0.7]), np.array([-1.1]), np.array([7.7])
    dsx, dsy, ddia = O.ref_derivs(SIZE, pix, sx, sy, dia, 2.)
    h = 1e-6
    f = lambda a, b, c: O.ref_terms(SIZE, pix, a, b, c, 2.)
    for got, args_p, args_m in (
            (dsx, (sx + h, sy, dia), (sx - h, sy, dia)),
            (dsy, (sx, sy + h, dia), (sx, sy - h, dia)),
            (ddia, (sx, sy, dia + h), (sx, sy, dia - h))):
        fd = (f(*args_p) - f(*args_m)) / (2 * h)
        np.testing.assert_allclose(got, fd, rtol=1e-6, atol=1e-9)


def test_cross_term_brute_force():
    rng = np.random.default_rng(3)
    Or = rng.standard_normal((3, 11)) + 1j * rng.standard_normal((3, 11))
    Rd = rng.standard_normal((4, 11)) + 1j * rng.standard_normal((4, 11))
    X = O.cross_term(Or, Rd)
    for r in range(3):
        for d in range(4):
            v = sum((Or[r, p] * np.conj(Rd[d, p])).real for p in range(11))
            assert abs(X[r, d] - v) < 1e-12


def test_loglik_closed_forms(setup):
    model, sx, sy, dia, pix, K, scales = setup
    z = np.zeros(SIZE * SIZE)
    # model = 0: LL = sum K log(scale |R|^2) - scale sum |R|^2, |R|^2 from slice_gen_holo
    LL, _ = O.rot_loglik(np.zeros(SIZE * SIZE), SIZE, 3, K, sx, sy, dia, scales)
    for d in range(len(sx)):
        v = E.slice_gen_holo(z, sx[d], sy[d], dia[d], 1., scales[d], SIZE, z, 0)[pix]
        ref = (K[d] * np.log(v)).sum() - v.sum()
        np.testing.assert_allclose(LL[d], ref, rtol=1e-12)
    # no reference (rel 0), r = 0: W = scale |m|^2 on the scored pixels (all interior)
    LL0, _ = O.rot_loglik(model, SIZE, 1, K, sx, sy, dia, scales, rel_scale=0.)
    for d in range(len(sx)):
        w = scales[d] * np.abs(model[pix]) ** 2
        np.testing.assert_allclose(LL0[d, 0], (K[d] * np.log(w)).sum() - w.sum(), rtol=1e-12)


def test_gemm_expansion_equals_direct(setup):
    model, sx, sy, dia, pix, K, scales = setup
    a, _ = O.rot_loglik(model, SIZE, 9, K, sx, sy, dia, scales, rel_scale=1.7)
    b, _ = O.rot_loglik_split(model, SIZE, 9, K, sx, sy, dia, scales, rel_scale=1.7)
    np.testing.assert_allclose(a, b, rtol=1e-12)


def test_log_floor_matches_reference_rule():
    """W < 1e-20 -> log W = -1000 (kernels.cu:37-40)."""
    W = np.array([0., 1e-21, 1e-20, 2.])
    np.testing.assert_array_equal(O._logw(W), [-1000., -1000., math.log(1e-20), math.log(2.)])


def test_lse_properties():
    rng = np.random.default_rng(4)
    LL = rng.standard_normal((5, 8)) * 50 - 3000
    P, logz = O.lse(LL)
    np.testing.assert_allclose(P.sum(1), 1., rtol=1e-14)
    np.testing.assert_allclose(logz, np.logaddexp.reduce(LL, axis=1), rtol=1e-14)
    P2, logz2 = O.lse(LL + 123.)
    np.testing.assert_allclose(P2, P, rtol=1e-12)
    np.testing.assert_allclose(logz2, logz + 123., rtol=1e-14)
    assert O.total_loss(np.zeros((3, 4))) == pytest.approx(0.)


def test_merge_mass_and_constant_counts(setup):
    model, sx, sy, dia, pix, K, scales = setup
    R = 6
    rng = np.random.default_rng(6)
    P = rng.random((len(sx), R))
    P /= P.sum(1, keepdims=True)
    _, acc, wts = O.merge_model(P, K, scales, SIZE, R)
    view = O.merge_views(P, K)
    # bilinear weights sum to 1 per in-bounds pixel, and every scored pixel is in bounds
    assert abs(acc.sum() - view.sum()) < 1e-10 * view.sum()
    # counts proportional to the scale: the merged model is the constant on covered pixels
    Kc = 2.5 * scales[:, None] * np.ones((len(sx), len(pix)))
    m, _, w = O.merge_model(P, Kc, scales, SIZE, R)
    np.testing.assert_allclose(m[w > 0], 2.5, rtol=1e-12)
    assert np.all(m[w == 0] == 0)


def test_loss_grad_central_differences(setup):
    model, sx, sy, dia, pix, K, scales = setup
    R = 4
    f = lambda m, a, b, c: O.total_loss(O.rot_loglik(m, SIZE, R, K, a, b, c, scales, 1.5)[0])
    loss, gm, gsx, gsy, gdia = O.loss_grad(model, SIZE, R, K, sx, sy, dia, scales, 1.5)
    assert loss == pytest.approx(f(model, sx, sy, dia), rel=1e-14)
    h = 1e-6
    rng = np.random.default_rng(7)
    for t in rng.choice(pix, 6, replace=False):
        e = np.zeros(SIZE * SIZE)
        e[t] = h
        fr = (f(model + e, sx, sy, dia) - f(model - e, sx, sy, dia)) / (2 * h)
        fi = (f(model + 1j * e, sx, sy, dia) - f(model - 1j * e, sx, sy, dia)) / (2 * h)
        assert abs(gm[t].real - fr) < 1e-5 * max(1., abs(fr))
        assert abs(gm[t].imag - fi) < 1e-5 * max(1., abs(fi))
    for d in range(len(sx)):
        e = np.zeros(len(sx))
        e[d] = h
        fsx = (f(model, sx + e, sy, dia) - f(model, sx - e, sy, dia)) / (2 * h)
        fsy = (f(model, sx, sy + e, dia) - f(model, sx, sy - e, dia)) / (2 * h)
        fdd = (f(model, sx, sy, dia + e) - f(model, sx, sy, dia - e)) / (2 * h)
        assert abs(gsx[d] - fsx) < 1e-5 * max(1., abs(fsx))
        assert abs(gsy[d] - fsy) < 1e-5 * max(1., abs(fsy))
        assert abs(gdia[d] - fdd) < 1e-5 * max(1., abs(fdd))


def test_multi_class_reduces_to_single(setup):
    model, sx, sy, dia, pix, K, scales = setup
    R = 3
    LL1, _ = O.rot_loglik(model, SIZE, R, K, sx, sy, dia, scales, 1.2)
    LL = O.rot_loglik_multi(model[None], SIZE, R, K, sx, sy, dia, scales, 1.2)
    assert np.array_equal(LL, LL1)
    a = O.loss_grad(model, SIZE, R, K, sx, sy, dia, scales, 1.2)
    b = O.loss_grad_multi(model[None], SIZE, R, K, sx, sy, dia, scales, 1.2)
    assert a[0] == b[0]
    np.testing.assert_allclose(b[1][0], a[1], rtol=1e-13, atol=1e-13 * np.abs(a[1]).max())
    for x, y in zip(a[2:], b[2:]):
        np.testing.assert_allclose(y, x, rtol=1e-13, atol=1e-13 * np.abs(x).max())


def test_multi_class_identical_models_split_probability(setup):
    """Two identical classes: every hypothesis appears twice, so P halves, the loss is
    unchanged (the prior 1/(C R) compensates) and each class gets half the model gradient."""
    model, sx, sy, dia, pix, K, scales = setup
    R = 3
    LL = O.rot_loglik_multi(np.stack([model, model]), SIZE, R, K, sx, sy, dia, scales, 1.2)
    P, _ = O.lse(LL)
    P1, _ = O.lse(LL[:, :R])
    np.testing.assert_allclose(P[:, :R], 0.5 * P1, rtol=1e-13)
    a = O.loss_grad(model, SIZE, R, K, sx, sy, dia, scales, 1.2)
    b = O.loss_grad_multi(np.stack([model, model]), SIZE, R, K, sx, sy, dia, scales, 1.2)
    assert b[0] == pytest.approx(a[0], rel=1e-13)
    for c in range(2):
        np.testing.assert_allclose(b[1][c], 0.5 * a[1], rtol=1e-12,
                                   atol=1e-13 * np.abs(a[1]).max())
    m1, _, _ = O.merge_model(P1, K, scales, SIZE, R)
    mm, _, _ = O.merge_model_multi(P, K, scales, SIZE, R)
    np.testing.assert_allclose(mm[0], m1, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(mm[1], m1, rtol=1e-12, atol=1e-14)


def test_multi_class_loss_grad_central_differences(setup):
    model, sx, sy, dia, pix, K, scales = setup
    R = 2
    rng = np.random.default_rng(9)
    models = np.stack([model, 0.3 * (rng.random(SIZE * SIZE) + 1j * rng.random(SIZE * SIZE))])
    f = lambda ms, a, b, c: O.total_loss(O.rot_loglik_multi(ms, SIZE, R, K, a, b, c, scales, 1.5))
    loss, gm, gsx, gsy, gdia = O.loss_grad_multi(models, SIZE, R, K, sx, sy, dia, scales, 1.5)
    assert loss == pytest.approx(f(models, sx, sy, dia), rel=1e-14)
    h = 1e-6
    for c in range(2):
        for t in rng.choice(pix, 3, replace=False):
            e = np.zeros_like(models)
            e[c, t] = h
            fr = (f(models + e, sx, sy, dia) - f(models - e, sx, sy, dia)) / (2 * h)
            fi = (f(models + 1j * e, sx, sy, dia) - f(models - 1j * e, sx, sy, dia)) / (2 * h)
            assert abs(gm[c, t].real - fr) < 1e-5 * max(1., abs(fr))
            assert abs(gm[c, t].imag - fi) < 1e-5 * max(1., abs(fi))
    e = np.zeros(len(sx))
    e[2] = h
    fdd = (f(models, sx, sy, dia + e) - f(models, sx, sy, dia - e)) / (2 * h)
    assert abs(gdia[2] - fdd) < 1e-5 * max(1., abs(fdd))
    fsx = (f(models, sx + e, sy, dia) - f(models, sx - e, sy, dia)) / (2 * h)
    assert abs(gsx[2] - fsx) < 1e-5 * max(1., abs(fsx))
