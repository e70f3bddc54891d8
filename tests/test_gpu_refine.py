"""Gradient refinement of the per-frame reference parameters and the grid search that judges
it (refspi_amd/refine.py), on the GPU.

* grid_search (frames replicated per shift state, one E-step per diameter) equals a brute
  force over the same grid (one E-step per state) in value and first-occurrence state index.
* refine never lowers a frame's log Z (a step is kept only if log Z rose) and raises it for
  most frames; streamed and resident plans give the same refinement.
* Judge: started inside the grid cell of the generating parameters, the refined parameters
  reach at least the grid-best log Z for most frames (the grid is the reference's search,
  refinement is continuous).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZE, R = 65, 72


@pytest.fixture(scope='module')
def work():
    from refspi_amd import synth
    cfg = synth.SynthConfig(size=SIZE, num_data=48, mean_count=300., n_blobs=24,
                            disk_radius=8., beamstop=4., shift_max=2., seed=7,
                            class_weights=(0.6, 0.4))
    fr = synth.make_frames(cfg)
    rel = synth.model_scales(cfg)[1]
    models = np.stack([synth.model_scales(cfg, c)[0] * synth.object_ft(cfg, c)
                       for c in range(2)])
    return cfg, fr, models, rel


def test_grid_search_equals_brute_force(work):
    from refspi_amd.dset import Dataset
    from refspi_amd.refine import grid_search
    from refspi_amd.rotem import RotEM
    cfg, fr, models, rel = work
    grid = dict(shiftx=(-1., 1., 3), shifty=(-1., 1., 2), sphere_dia=(7.5, 8.5, 3))
    sx, sy, dia, lz, st = grid_search(fr, SIZE, R, rel, models, classes=2, full_turn=True,
                                      mode='fp64', **grid)
    plan = RotEM(Dataset(frames=fr), SIZE, R, rel_scale=rel, mode='fp64', classes=2,
                 full_turn=True)
    n = fr.num_data
    gx = np.linspace(-1, 1, 3)
    gy = np.linspace(-1, 1, 2)
    gd = np.linspace(7.5, 8.5, 3)
    allz = []
    for a in gx:
        for b in gy:
            for c in gd:
                _, _, z = plan.estep_device(models, np.full(n, a), np.full(n, b), np.full(n, c),
                                            want_ll=False, want_prob=False)
                allz.append(z.cpu().numpy())
    allz = np.array(allz)                       # (states, frames) in 'ij' order
    np.testing.assert_allclose(lz, allz.max(0), rtol=1e-12)
    assert np.array_equal(st, allz.argmax(0))
    s = allz.argmax(0)
    assert np.array_equal(sx, gx[s // 6]) and np.array_equal(sy, gy[(s // 3) % 2])
    assert np.array_equal(dia, gd[s % 3])


def test_refine_monotone_and_streamed_equals_resident(work):
    from refspi_amd.dset import Dataset
    from refspi_amd.refine import refine
    from refspi_amd.rotem import RotEM, StreamedRotEM
    cfg, fr, models, rel = work
    t = fr.truth
    rng = np.random.default_rng(3)
    n = fr.num_data
    sx0 = t['shifts'][:, 0] + rng.normal(0, 0.3, n)
    sy0 = t['shifts'][:, 1] + rng.normal(0, 0.3, n)
    d0 = t['diameters'] + rng.normal(0, 0.15, n)
    res = RotEM(Dataset(frames=fr), SIZE, R, rel_scale=rel, mode='fp64', classes=2,
                full_turn=True)
    a = refine(res, models, sx0, sy0, d0, t['scale'], steps=5)
    _, _, z0 = res.estep_device(models, sx0, sy0, d0, t['scale'], want_ll=False,
                                want_prob=False)
    z0 = z0.cpu().numpy()
    z1 = a[3].cpu().numpy()
    assert np.all(z1 >= z0)
    assert (z1 > z0).mean() > 0.8
    st = StreamedRotEM(Dataset(frames=fr), SIZE, R, rel_scale=rel, mode='fp64', classes=2,
                       full_turn=True, chunk_frames=20)
    b = refine(st, models, sx0, sy0, d0, t['scale'], steps=5)
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_allclose(x.cpu().numpy(), y.cpu().numpy(), rtol=1e-9, atol=1e-9)


def test_refinement_judged_by_grid_search(work, capsys):
    from refspi_amd.dset import Dataset
    from refspi_amd.refine import grid_search, refine
    from refspi_amd.rotem import RotEM
    cfg, fr, models, rel = work
    t = fr.truth
    grid = dict(shiftx=(-2., 2., 9), shifty=(-2., 2., 9), sphere_dia=(7., 9., 11))
    gsx, gsy, gd, glz, _ = grid_search(fr, SIZE, R, rel, models, scales=t['scale'],
                                       classes=2, full_turn=True, mode='split', **grid)
    plan = RotEM(Dataset(frames=fr), SIZE, R, rel_scale=rel, mode='split', classes=2,
                 full_turn=True)
    # start at the grid point nearest to the generating parameters
    snap = lambda v, lo, hi, k: np.clip(np.round((v - lo) / ((hi - lo) / (k - 1))), 0, k - 1) \
        * ((hi - lo) / (k - 1)) + lo
    sx0 = snap(t['shifts'][:, 0], -2., 2., 9)
    sy0 = snap(t['shifts'][:, 1], -2., 2., 9)
    d0 = snap(t['diameters'], 7., 9., 11)
    sx, sy, dia, lz, hist = refine(plan, models, sx0, sy0, d0, t['scale'], steps=8)
    lz = lz.cpu().numpy()
    reached = lz >= glz - 1e-9 * np.abs(glz)
    # (rotation + pi, -shift) gives the same hologram: error to the nearer twin, per axis
    err = lambda a, b: float(np.sqrt(np.mean(np.minimum(((a - b) ** 2).sum(0),
                                                        ((a + b) ** 2).sum(0))) / 2.))
    with capsys.disabled():
        print('\nrefined log Z >= grid-best log Z for %.3f of the frames; shift rms error vs '
              'truth: start %.3f, grid %.3f, refined %.3f px; history %s' %
              (reached.mean(), err(np.stack([sx0, sy0]), t['shifts'].T),
               err(np.stack([gsx, gsy]), t['shifts'].T),
               err(np.stack([sx.cpu().numpy(), sy.cpu().numpy()]), t['shifts'].T), hist))
    assert reached.mean() > 0.5
