"""The HIP path against the committed golden fixture (tests/golden/golden_small.npz),
without running the oracle: 21x21 q-grid, 24 frames, 27 states x 8 rotations, two
EMC+MaxLP iterations (each started from the fixture's own input model), the kernels.cu
mirrors, the MaxLP photon pull table, and the rotation form (E-step in three precisions,
merge, adjoint).

Bars: argmax bit-exact; maxll, msums, vscale 1e-12 relative; the mirrors 1e-13; MaxLP
1e-9 relative except at pixels whose FP64 pattern-search argmax was a near-tie (margin
< 1e-12 in the fixture); rotation form FP64 1e-11 (LL) and 1e-9 (P, gradients, merge),
FP32 and split-BF16 LL 1e-5 relative.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'golden_small.npz')


@pytest.fixture(scope='module')
def g():
    return dict(np.load(GOLDEN))


class _Frames:
    def __init__(self, g):
        self.num_pix = int(g['size']) ** 2
        self.bg = None
        for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi'):
            setattr(self, k, g[k])


def _grid(g):
    v = g['grid']
    return dict(shiftx=(v[0], v[1], int(v[2])), shifty=(v[3], v[4], int(v[5])),
                sphere_dia=(v[6], v[7], int(v[8])))


def test_golden_mirrors(g):
    import torch
    from refspi_amd import _lib
    from refspi_amd.emc import state_grid
    size, R = int(g['size']), int(g['num_rot'])
    gr = _grid(g)
    sx, sy, dia, _ = state_grid(gr['shiftx'], gr['shifty'], gr['sphere_dia'])
    m = torch.as_tensor(g['model0'], device='cuda')
    bg = torch.zeros(size * size, dtype=torch.float64, device='cuda')
    out = torch.empty(size * size, dtype=torch.float64, device='cuda')
    for k, i in enumerate(g['holo_states']):
        _lib.call('rsp_slice_gen_holo', _lib.ptr(m), sx[i], sy[i], dia[i], 1., 1., size,
                  _lib.ptr(bg), 0, _lib.ptr(out), _lib.stream_handle())
        torch.cuda.synchronize()
        np.testing.assert_allclose(out.cpu().numpy(), g['holo'][k], rtol=1e-13, atol=1e-13)
    view = torch.as_tensor(g['holo'][0], device='cuda')
    for j in range(R):
        _lib.call('rsp_slice_gen', _lib.ptr(view), j * np.pi / R, 1., size, _lib.ptr(bg), 1,
                  _lib.ptr(out), _lib.stream_handle())
        torch.cuda.synchronize()
        np.testing.assert_allclose(out.cpu().numpy(), g['logviews_state0'][j], rtol=1e-13,
                                   atol=1e-12)


@pytest.mark.parametrize('it', [1, 2])
def test_golden_iteration(g, it):
    import torch
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    p = 'it%d_' % it
    size, R = int(g['size']), int(g['num_rot'])
    m_in = g[p + 'model_in']
    emc = EMC(dataset=Dataset(frames=_Frames(g)), size=size, num_rot=R, model=m_in,
              save=False, **_grid(g))
    emc._calculate_prob(torch.as_tensor(m_in, device='cuda'))
    rmax = emc._get_rmax()
    assert np.array_equal(rmax, g[p + 'rmax'])
    np.testing.assert_allclose(emc.maxll, g[p + 'maxll'], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(emc.msums, g[p + 'msums'], rtol=1e-12)
    assert emc.vscale == pytest.approx(float(g[p + 'vscale']), rel=1e-12)
    vals = emc._unravel_rmax(rmax)
    for v, k in zip(vals, ('sx', 'sy', 'dia', 'ang')):
        assert np.array_equal(np.asarray(v), g[p + k]), k
    out = emc.phaser._run_phaser(m_in, *vals)
    assert emc.phaser.rescale == pytest.approx(float(g[p + 'rescale']), rel=1e-12)
    ref = g[p + 'model_out']
    rel = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-300)
    bad = np.nonzero(rel > 1e-9)[0]
    assert np.all(g[p + 'maxlp_margin'][bad] < 1e-12), bad
    if it == 1:
        # the fixture's photon pull table (maxLP.py:250-252 for iteration 1's angles): the
        # GPU pull-back gives every (mask pixel, frame) the count at exactly that pixel
        dense = Dataset(frames=_Frames(g))
        K = np.stack([dense.get_frame(d) for d in range(dense.num_data)])
        rows = np.arange(dense.num_data)
        for t, idx in zip(g['pull_pixels'], g['pull_index']):
            np.testing.assert_array_equal(emc.phaser.get_photons_pixel(int(t)), K[rows, idx])


@pytest.fixture(scope='module')
def rot_plan(g):
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    return RotEM(Dataset(frames=_Frames(g)), int(g['size']), int(g['num_rot']),
                 rel_scale=float(g['rel_scale']), mode='fp64')


def _rot_args(g):
    return g['model0'], g['rot_sx'], g['rot_sy'], g['rot_dia'], g['rot_scales']


@pytest.mark.parametrize('mode', ['fp64', 'fp32', 'split'])
def test_golden_rot_estep(g, rot_plan, mode):
    rot_plan.set_mode(mode)
    try:
        ll, prob, logz = rot_plan.estep(*_rot_args(g))
    finally:
        rot_plan.set_mode('fp64')
    err = np.abs(ll - g['rot_ll']) / np.abs(g['rot_ll'])
    if mode == 'fp64':
        assert err.max() < 1e-11
        np.testing.assert_allclose(prob, g['rot_prob'], rtol=1e-9, atol=1e-300)
        np.testing.assert_allclose(logz, g['rot_logz'], rtol=1e-11)
    else:
        assert err.max() < 1e-5


def test_golden_rot_merge_and_adjoint(g, rot_plan):
    model, acc, wts = rot_plan.merge(g['rot_prob'], g['rot_scales'])
    np.testing.assert_allclose(acc, g['rot_merge_acc'], rtol=1e-11,
                               atol=1e-13 * np.abs(g['rot_merge_acc']).max())
    np.testing.assert_allclose(wts, g['rot_merge_wts'], rtol=1e-11,
                               atol=1e-13 * np.abs(g['rot_merge_wts']).max())
    np.testing.assert_allclose(model, g['rot_merged'], rtol=1e-10,
                               atol=1e-13 * np.abs(g['rot_merged']).max())
    loss, gm, gsx, gsy, gdia = rot_plan.loss_grad(*_rot_args(g))
    assert loss == pytest.approx(float(g['rot_loss']), rel=1e-11)
    for a, k in ((gm, 'rot_g_model'), (gsx, 'rot_g_sx'), (gsy, 'rot_g_sy'), (gdia, 'rot_g_dia')):
        np.testing.assert_allclose(a, g[k], rtol=0, atol=1e-9 * np.abs(g[k]).max(), err_msg=k)
