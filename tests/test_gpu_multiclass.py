"""Several object classes and streamed frames in the rotation form (GPU).

* Two classes (hypotheses h = c * R + r) against the FP64 oracle's multi-class functions:
  LL 1e-11, P 1e-9, logZ 1e-11 relative (FP64 mode), the first-occurrence argmax exact,
  merge 1e-10, adjoint 1e-9 of the largest entry.
* A streamed pass (frames packed into pinned host memory, chunks of 40 frames copied in on
  a side stream, the last chunk ragged) against the resident plan: per-frame outputs equal
  within 1e-12 relative, merge and gradients within 1e-10 (the accumulation order over
  chunks differs).
"""
import numpy as np
import pytest

from conftest import SMALL

pytestmark = pytest.mark.gpu

R = 12


@pytest.fixture(scope='module')
def setup(small_frames, small_model):
    from oracle import emc_oracle as E
    from oracle import rot_oracle as O
    size = SMALL['size']
    fr = E.SparseFramesNP(small_frames.ones, small_frames.multi, small_frames.place_ones,
                          small_frames.place_multi, small_frames.count_multi,
                          small_frames.num_pix)
    pix = O.scored_pixels(size)
    K = fr.dense()[:, pix]
    rng = np.random.default_rng(21)
    n = fr.num_data
    sx, sy = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    dia, sc = rng.uniform(7, 9, n), rng.uniform(0.6, 1.4, n)
    models = np.stack([0.05 * small_model,
                       0.04 * (rng.random(size * size) + 1j * rng.random(size * size))])
    return dict(size=size, K=K, models=models, sx=sx, sy=sy, dia=dia, sc=sc)


def _plan(frames, mode='fp64', **kw):
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    return RotEM(Dataset(frames=frames), SMALL['size'], R, rel_scale=1.3, mode=mode, classes=2,
                 **kw)


def test_two_classes_match_oracle(small_frames, setup):
    from oracle import rot_oracle as O
    s = setup
    LL = O.rot_loglik_multi(s['models'], s['size'], R, s['K'], s['sx'], s['sy'], s['dia'],
                            s['sc'], 1.3)
    P, logz = O.lse(LL)
    plan = _plan(small_frames)
    assert plan.info()['hypotheses'] == 2 * R
    ll, prob, lz, best = (t.cpu().numpy() for t in plan.estep_device(
        s['models'], s['sx'], s['sy'], s['dia'], s['sc'], want_best=True))
    assert (np.abs(ll - LL) / np.abs(LL)).max() < 1e-11
    np.testing.assert_allclose(prob, P, rtol=1e-9, atol=1e-300)
    np.testing.assert_allclose(lz, logz, rtol=1e-11)
    srt = np.sort(LL, 1)
    assert np.all(srt[:, -1] - srt[:, -2] > 1e-9 * np.abs(srt[:, -1]))   # no near-ties
    assert np.array_equal(best, LL.argmax(1))
    mm, acc, wts = plan.merge(P, s['sc'])
    rm, racc, rw = O.merge_model_multi(P, s['K'], s['sc'], s['size'], R)
    assert mm.shape == (2, s['size'] ** 2)
    np.testing.assert_allclose(acc, racc, rtol=1e-11, atol=1e-13 * np.abs(racc).max())
    np.testing.assert_allclose(mm, rm, rtol=1e-10, atol=1e-13 * np.abs(rm).max())
    ref = O.loss_grad_multi(s['models'], s['size'], R, s['K'], s['sx'], s['sy'], s['dia'],
                            s['sc'], 1.3)
    got = plan.loss_grad(s['models'], s['sx'], s['sy'], s['dia'], s['sc'])
    assert got[0] == pytest.approx(ref[0], rel=1e-11)
    for a, b in zip(got[1:], ref[1:]):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())


@pytest.mark.parametrize('mode', ['fp64', 'split'])
def test_streamed_pass_equals_resident(small_frames, setup, mode):
    import torch
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import PASS_FRAME_GRAD, PASS_MERGE, PASS_MODEL_GRAD, StreamedRotEM
    s = setup
    res = _plan(small_frames, mode)
    ll, prob, lz, best = res.estep_device(s['models'], s['sx'], s['sy'], s['dia'], s['sc'],
                                          want_best=True)
    mm, acc, wts = res.merge_device(prob, s['sc'])
    gm, gx, gy, gd = res.adjoint_device(s['models'], s['sx'], s['sy'], s['dia'], s['sc'], prob)
    st = StreamedRotEM(Dataset(frames=small_frames), SMALL['size'], R, rel_scale=1.3,
                       mode=mode, classes=2, chunk_frames=40)
    assert [c['n'] for c in st.chunks] == [40, 40, 16]
    out = st.run_pass(s['models'], s['sx'], s['sy'], s['dia'], s['sc'],
                      flags=PASS_MERGE | PASS_FRAME_GRAD | PASS_MODEL_GRAD, want_ll=True,
                      want_prob=True)
    torch.cuda.synchronize()
    c = lambda t: t.cpu().numpy()
    np.testing.assert_allclose(c(out['ll']), c(ll), rtol=1e-12)
    np.testing.assert_allclose(c(out['prob']), c(prob), rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(c(out['logz']), c(lz), rtol=1e-12)
    assert np.array_equal(c(out['best']), c(best))
    for a, b in ((out['model'], mm), (out['acc'], acc), (out['weights'], wts),
                 (out['g_model'], gm), (out['g_sx'], gx), (out['g_sy'], gy),
                 (out['g_dia'], gd)):
        np.testing.assert_allclose(c(a), c(b), rtol=0, atol=1e-10 * np.abs(c(b)).max())
    # a second pass without merge/gradients reuses the plan
    out2 = st.run_pass(s['models'], s['sx'], s['sy'], s['dia'], s['sc'])
    np.testing.assert_array_equal(c(out2['logz']), c(out['logz']))


def test_two_classes_split_vs_fp64_at_185():
    """185 x 185, two classes x 180 rotations (full turn), 256 frames of a two-class mixture with the
    generator's objects as models: split-BF16 LL within 1e-5 relative of the FP64 mode, and
    the split argmax equals the FP64 argmax wherever the FP64 top-2 gap exceeds twice the
    frame's largest LL difference between the two modes."""
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    cfg = synth.SynthConfig(size=185, num_data=256, mean_count=300., n_blobs=120,
                            disk_radius=30., beamstop=4., seed=4, class_weights=(0.6, 0.4))
    fr = synth.make_frames(cfg)
    t = fr.truth
    rel = synth.model_scales(cfg)[1]
    models = np.stack([synth.model_scales(cfg, c)[0] * synth.object_ft(cfg, c)
                       for c in range(2)])
    ds = Dataset(frames=fr)
    res = {}
    for mode in ('fp64', 'split'):
        plan = RotEM(ds, 185, 180, rel_scale=rel, mode=mode, classes=2, full_turn=True)
        ll, _, _, best = plan.estep_device(models, t['shifts'][:, 0], t['shifts'][:, 1],
                                           t['diameters'], t['scale'], want_prob=False,
                                           want_best=True)
        res[mode] = (ll.cpu().numpy(), best.cpu().numpy())
        del plan
        torch.cuda.synchronize()
    l64, ls = res['fp64'][0], res['split'][0]
    err = np.abs(ls - l64) / np.abs(l64)
    assert err.max() < 1e-5
    srt = np.sort(l64, 1)
    gap = srt[:, -1] - srt[:, -2]
    bound = 2. * np.abs(ls - l64).max(1)
    sure = gap > bound
    assert sure.mean() > 0.5
    assert np.array_equal(res['split'][1][sure], res['fp64'][1][sure])
    cls = res['fp64'][1] // 180
    print('class of the best hypothesis = generating class for %.3f of the frames' %
          (cls == t['classes']).mean())


def test_full_turn_matches_oracle(small_frames, setup):
    """Rotations over a full turn (j * 2pi / R) in the FP64 mode against the oracle given the
    same angle table."""
    from oracle import rot_oracle as O
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    s = setup
    cs = O.rot_cos_sin(R, full_turn=True)
    LL = O.rot_loglik_multi(s['models'], s['size'], R, s['K'], s['sx'], s['sy'], s['dia'],
                            s['sc'], 1.3, cos_sin=cs)
    plan = RotEM(Dataset(frames=small_frames), SMALL['size'], R, rel_scale=1.3, mode='fp64',
                 classes=2, full_turn=True)
    ll, _, _ = plan.estep(s['models'], s['sx'], s['sy'], s['dia'], s['sc'])
    assert (np.abs(ll - LL) / np.abs(LL)).max() < 1e-11
