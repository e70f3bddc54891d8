"""GPU parity of the rotation form (rsp_rot_*) against the FP64 oracle (oracle/rot_oracle.py).

Tolerances:
* FP64 mode: LL within 1e-11 relative, P within 1e-9 relative.
* FP32 and split-BF16 modes: LL within 1e-5 relative (the stated bound of the split form);
  P within the bound that LL error implies for a softmax, |dlog P| <= 2 max |dLL|.
* merge: identical P in, FP64 sums in another order: 1e-11 relative.
"""
import numpy as np
import pytest

from conftest import SMALL

pytestmark = pytest.mark.gpu

R = 12


def _params(n, seed=4):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), rng.uniform(7, 9, n),
            rng.uniform(0.6, 1.4, n))


@pytest.fixture(scope='module')
def oracle_small(small_frames, small_model):
    from oracle import emc_oracle as E
    from oracle import rot_oracle as O
    size = SMALL['size']
    fr = E.SparseFramesNP(small_frames.ones, small_frames.multi, small_frames.place_ones,
                          small_frames.place_multi, small_frames.count_multi,
                          small_frames.num_pix)
    pix = O.scored_pixels(size)
    K = fr.dense()[:, pix]
    sx, sy, dia, sc = _params(fr.num_data)
    model = 0.05 * small_model
    LL, _ = O.rot_loglik(model, size, R, K, sx, sy, dia, sc, rel_scale=1.3)
    P, logz = O.lse(LL)
    return dict(model=model, sx=sx, sy=sy, dia=dia, sc=sc, LL=LL, P=P, logz=logz, K=K, pix=pix)


def _plan(frames, mode, **kw):
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    return RotEM(Dataset(frames=frames), SMALL['size'], R, rel_scale=1.3, mode=mode, **kw)


@pytest.mark.parametrize('mode', ['fp64', 'fp32', 'split', 'split-stored'])
def test_rot_estep_matches_oracle(small_frames, oracle_small, mode):
    """split: the reference rows generated in the GEMM kernel (the default); split-stored:
    written to HBM and read by the GEMM."""
    o = oracle_small
    plan = _plan(small_frames, mode.split('-')[0])
    if mode == 'split-stored':
        plan.set_ref_rows('stored')
    ll, prob, logz = plan.estep(o['model'], o['sx'], o['sy'], o['dia'], o['sc'])
    err = np.abs(ll - o['LL']) / np.abs(o['LL'])
    if mode == 'fp64':
        assert err.max() < 1e-11
        np.testing.assert_allclose(prob, o['P'], rtol=1e-9, atol=1e-300)
        np.testing.assert_allclose(logz, o['logz'], rtol=1e-11)
    else:
        assert err.max() < 1e-5, err.max()
        dmax = np.abs(ll - o['LL']).max()
        bound = o['P'] * (np.exp(2 * dmax) - 1) + 1e-300
        assert np.all(np.abs(prob - o['P']) <= bound * 1.0001)
    np.testing.assert_allclose(prob.sum(1), 1., rtol=1e-12)


def test_rot_estep_chunking_is_invisible(small_frames, oracle_small):
    o = oracle_small
    a = _plan(small_frames, 'split').estep(o['model'], o['sx'], o['sy'], o['dia'], o['sc'])
    b = _plan(small_frames, 'split', chunk=128).estep(o['model'], o['sx'], o['sy'], o['dia'],
                                                      o['sc'])
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_rot_merge_matches_oracle(small_frames, oracle_small):
    from oracle import rot_oracle as O
    o = oracle_small
    plan = _plan(small_frames, 'fp64')
    _, prob, _ = plan.estep(o['model'], o['sx'], o['sy'], o['dia'], o['sc'])
    model, acc, wts = plan.merge(prob, o['sc'])
    ref, racc, rw = O.merge_model(prob, o['K'], o['sc'], SMALL['size'], R)
    np.testing.assert_allclose(acc, racc, rtol=1e-11, atol=1e-13 * np.abs(racc).max())
    np.testing.assert_allclose(wts, rw, rtol=1e-11, atol=1e-13 * np.abs(rw).max())
    np.testing.assert_allclose(model, ref, rtol=1e-10, atol=1e-13 * np.abs(ref).max())
    assert np.array_equal(model == 0, ref == 0)


def test_rot_split_vs_fp64_at_185(tmp_path):
    """185 x 185 q-grid, 360 rotations, 512 frames: split-BF16 and FP32 LL within 1e-5
    relative of the FP64 mode (the FP64 mode itself matches the oracle above)."""
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    cfg = synth.SynthConfig(size=185, num_data=512, mean_count=300., n_blobs=120,
                            disk_radius=30., beamstop=4., seed=2)
    fr = synth.make_frames(cfg)
    t = fr.truth
    rng = np.random.default_rng(0)
    size = 185
    # a model near the data scale: the object's transform with noise
    model = (rng.standard_normal(size * size) + 1j * rng.standard_normal(size * size)) * 0.3
    ds = Dataset(frames=fr)
    res = {}
    for mode in ('fp64', 'fp32', 'split', 'split-stored'):
        plan = RotEM(ds, size, 360, rel_scale=1., mode=mode.split('-')[0])
        if mode == 'split-stored':
            plan.set_ref_rows('stored')
        res[mode] = plan.estep(model, t['shifts'][:, 0], t['shifts'][:, 1], t['diameters'])
        del plan
        torch.cuda.synchronize()
    ll64 = res['fp64'][0]
    errs = {}
    for mode in ('fp32', 'split', 'split-stored'):
        err = np.abs(res[mode][0] - ll64) / np.abs(ll64)
        errs[mode] = err.max()
        assert err.max() < 1e-5, (mode, err.max())
    print('\nLL relative error against FP64:', errs)


@pytest.mark.parametrize('dias', ['wide', 'zero'])
def test_rot_generated_ref_rows_edge_diameters(small_frames, oracle_small, dias):
    """Generated reference rows over unusual diameters: a wide spread (the sphere table
    spans the largest diameter's range, 25 px here) and frames of diameter 0 (S = 1/3 at
    s = 0, the reference's s == 0 guard): LL within 1e-5 of the FP64 oracle."""
    from oracle import rot_oracle as O
    o = oracle_small
    dia = o['dia'].copy()
    if dias == 'wide':
        dia[::3] = 25.
        dia[1::7] = 0.5
    else:
        dia[::4] = 0.
    size = SMALL['size']
    LL, _ = O.rot_loglik(o['model'], size, R, o['K'], o['sx'], o['sy'], dia, o['sc'],
                         rel_scale=1.3)
    plan = _plan(small_frames, 'split')
    ll, _, _ = plan.estep(o['model'], o['sx'], o['sy'], dia, o['sc'])
    err = np.abs(ll - LL) / np.abs(LL)
    assert err.max() < 1e-5, err.max()


@pytest.mark.parametrize('mode', ['fp64', 'fp32', 'split'])
def test_rot_log_floor_counts_each_photon(small_frames, oracle_small, mode):
    """Frames scaled by 1e-30 put every photon's W = scale |O + R|^2 below the 1e-20 log
    floor (kernels.cu:37-40: -1000 per count), so their LL is -1000 sum K up to a dense term
    of ~1e-26, in the oracle and in every mode (the FP32 forms count floored photons apart
    from their base-2 log sums); the other frames stay within the mode's bound."""
    from oracle import rot_oracle as O
    o = oracle_small
    sc = o['sc'].copy()
    sc[::5] = 1e-30
    LL, _ = O.rot_loglik(o['model'], SMALL['size'], R, o['K'], o['sx'], o['sy'], o['dia'], sc,
                         rel_scale=1.3)
    floor = np.broadcast_to(-1000. * o['K'][::5].sum(1)[:, None], LL[::5].shape)
    assert floor.min() < 0.
    np.testing.assert_allclose(LL[::5], floor, rtol=1e-12, atol=1e-12)
    plan = _plan(small_frames, mode)
    ll, _, _ = plan.estep(o['model'], o['sx'], o['sy'], o['dia'], sc)
    np.testing.assert_allclose(ll[::5], floor, rtol=1e-12, atol=1e-12)
    keep = np.ones(len(sc), bool)
    keep[::5] = False
    err = np.abs(ll[keep] - LL[keep]) / np.abs(LL[keep])
    assert err.max() < (1e-11 if mode == 'fp64' else 1e-5), err.max()
