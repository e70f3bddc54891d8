"""The HIP path against oracle fixtures at the benchmark's own scale and at an even grid.

``golden/cfg185_subset.npz`` (written by ``golden/make_cfg185_subset.py`` with the oracle's
C restatement): 2000 frames of the 185^2 workload over the full 6069 x 360 hypothesis grid
(2,184,840 hypotheses) for two models, and MaxLP over all 1708 ring pixels plus 96 mask
pixels (the 48 brightest and 48 spread over the mask) at the E-step's answer.

``golden/cfg64.npz`` (``golden/make_cfg64.py``): two EMC+MaxLP iterations on a 64^2 grid
(200 frames, 125 states x 90 rotations), where the reference's even-size rules apply: the
-1000 rim of slice_gen and the MaxLP photon index reaching ``size``.

Bars: argmax bit-exact on every frame (a frame may differ only where the oracle's own top
two hypotheses are within 1e-9 relative: an FP64 near-tie); max log-likelihood within
1e-12 relative; MaxLP values within 1e-10 relative except at pixels whose search met a
top-2 candidate margin below 1e-9 relative; ring rescale within 1e-12.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, 'golden')
sys.path.insert(0, GOLDEN)

TIE = 1e-9


def _check_estep(rmax, maxll, fx_rmax, fx_maxll, fx_second, what):
    margin = (fx_maxll - fx_second) / np.abs(fx_maxll)
    bad = np.nonzero(rmax != fx_rmax)[0]
    print('%s: %d frames, %d argmax differences (all at near-ties: %s), oracle near-ties '
          '(< %g): %d' % (what, len(rmax), len(bad), bool(np.all(margin[bad] < TIE)), TIE,
                          int((margin < TIE).sum())))
    assert np.all(margin[bad] < TIE), (bad[:10], margin[bad[:10]])
    ok = rmax == fx_rmax
    np.testing.assert_allclose(maxll[ok], fx_maxll[ok], rtol=1e-12, atol=0)


def _check_maxlp(got, want, margin, what):
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-300)
    bad = rel > 1e-10
    print('%s: %d pixels, largest relative difference %.2e, %d above 1e-10 (near-ties: %d)'
          % (what, len(want), rel.max(), int(bad.sum()), int((margin < TIE).sum())))
    assert np.all(margin[bad] < TIE), (np.nonzero(bad)[0][:10], rel[bad][:10])
    assert bad.sum() <= max(1, len(want) // 100)


@pytest.fixture(scope='module')
def cfg185():
    import make_cfg185_subset as M
    fx = dict(np.load(os.path.join(GOLDEN, 'cfg185_subset.npz')))
    cfg, fr, models = M.inputs()
    assert M.frames_digest(fr) == str(fx['frames_sha256']), 'generator drifted'
    for k, m in models.items():
        assert M.digest(m) == str(fx[k + '_model_sha256'])
    return M, fx, cfg, fr, models


@pytest.mark.parametrize('which', ['object', 'random'])
def test_cfg185_full_grid_estep_equals_oracle(cfg185, which):
    import torch
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    M, fx, cfg, fr, models = cfg185
    emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=M.NUM_ROT,
              model=models[which], save=False, **M.GRID)
    print('\ncase cfg185 %s' % which, flush=True)
    emc._calculate_prob(torch.as_tensor(emc.model, device='cuda'))
    emc._get_rmax()
    st = emc.estep_stats()
    assert st['hypotheses'] == 6069 * 360
    assert abs(emc.vscale - float(fx[which + '_vscale'])) <= 1e-13 * abs(emc.vscale)
    _check_estep(emc.rmax, emc.maxll, fx[which + '_rmax'], fx[which + '_maxll'],
                 fx[which + '_second'], 'cfg185 %s %s' % (which, st))


def test_cfg185_maxlp_equals_oracle(cfg185):
    from refspi_amd.dset import Dataset
    from refspi_amd.maxlp import MaxLPhaser
    M, fx, cfg, fr, models = cfg185
    from oracle import emc_oracle as E
    sx, sy, dia, ns = E.state_grid(M.GRID['shiftx'], M.GRID['shifty'], M.GRID['sphere_dia'])
    vals = E.unravel_rmax(fx['object_rmax'], ns, M.NUM_ROT, sx, sy, dia)
    print('\ncase cfg185 maxlp', flush=True)
    ph = MaxLPhaser(Dataset(frames=fr), size=cfg.size)
    out = ph._run_phaser(models['object'], *vals)
    assert abs(ph.rescale - float(fx['maxlp_rescale'])) <= 1e-12 * abs(ph.rescale)
    chk = fx['maxlp_check']
    _check_maxlp(out[chk], fx['maxlp_values'], fx['maxlp_margin'], 'cfg185 maxlp')
    mask = E.qvals(cfg.size)[3]
    np.testing.assert_array_equal(out[~mask], models['object'][~mask])


@pytest.fixture(scope='module')
def cfg64():
    import make_cfg64 as M
    fx = dict(np.load(os.path.join(GOLDEN, 'cfg64.npz')))
    cfg, fr, model0 = M.inputs()
    assert M.frames_digest(fr) == str(fx['frames_sha256']), 'generator drifted'
    assert M.digest(model0) == str(fx['model0_sha256'])
    return M, fx, cfg, fr, model0


@pytest.mark.parametrize('it', [1, 2])
def test_cfg64_iteration_equals_oracle(cfg64, it):
    """Iteration `it` from the oracle's model of the previous iteration (so that a
    rounding-level difference of one iteration does not carry into the next)."""
    import torch
    from oracle import emc_oracle as E
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    M, fx, cfg, fr, model0 = cfg64
    start = model0 if it == 1 else fx['model_1']
    emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=M.NUM_ROT, model=start,
              save=False, **M.GRID)
    print('\ncase cfg64 iteration %d' % it, flush=True)
    dm = torch.as_tensor(emc.model, device='cuda')
    emc._calculate_prob(dm)
    emc._get_rmax()
    assert abs(emc.vscale - float(fx['vscale_%d' % it])) <= 1e-13 * abs(emc.vscale)
    _check_estep(emc.rmax, emc.maxll, fx['rmax_%d' % it], fx['maxll_%d' % it],
                 fx['second_%d' % it], 'cfg64 iteration %d' % it)
    vals = emc._unravel_rmax(fx['rmax_%d' % it])
    out = emc.phaser._run_phaser(start, *vals)
    assert abs(emc.phaser.rescale - float(fx['rescale_%d' % it])) <= \
        1e-12 * abs(emc.phaser.rescale)
    mask = E.qvals(cfg.size)[3]
    m = np.nonzero(mask)[0]
    _check_maxlp(out[m], fx['model_%d' % it][m], fx['margin_%d' % it][m],
                 'cfg64 maxlp iteration %d' % it)
    np.testing.assert_array_equal(out[~mask], start[~mask])
