"""The joint search at its full hypothesis count on the 185² workload.

2000 frames of the 185² workload (120 blobs, 300 photons per frame, beamstop 4), the full
grid of 17 x 17 shifts x 21 diameters = 6069 states x 360 rotations = 2,184,840
hypotheses: the FP32-table gather with its exactness protocol (the default) equals the
exhaustive FP64 search over every hypothesis of every frame (exact mode) -- argmax and
maximum log-likelihood bit for bit -- for two models: the generator's object (a peaked
landscape, where neighbouring hypotheses nearly tie) and the plan's seeded random start.
The exhaustive FP64 search is itself compared with the oracle on the small grids of
test_gpu_estep.py; at this size the oracle would need hours.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRID = dict(shiftx=(-4., 4., 17), shifty=(-4., 4., 17), sphere_dia=(7., 9., 21))
NUM_ROT = 360


@pytest.fixture(scope='module')
def frames185():
    from refspi_amd import synth
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': 2000, 'seed': 0})
    return cfg, synth.make_frames(cfg)


def _search(fr, size, model, seed, exact):
    import torch
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    emc = EMC(dataset=Dataset(frames=fr), size=size, num_rot=NUM_ROT, model=model, seed=seed,
              save=False, **GRID)
    if exact:
        emc.set_option(1, 1)
    emc._calculate_prob(torch.as_tensor(emc.model, device='cuda'))
    emc._get_rmax()
    return emc.rmax.copy(), emc.maxll.copy(), emc.estep_stats()


@pytest.mark.parametrize('which', ['object', 'random'])
def test_full_grid_fast_equals_exhaustive_fp64(frames185, which):
    from refspi_amd import synth
    cfg, fr = frames185
    model = synth.model_scales(cfg)[0] * synth.object_ft(cfg) if which == 'object' else None
    r_fast, l_fast, st = _search(fr, cfg.size, model, 3, exact=False)
    r_ex, l_ex, st_ex = _search(fr, cfg.size, model, 3, exact=True)
    print('\n%s model: %s' % (which, st))
    assert st['hypotheses'] == 6069 * NUM_ROT
    assert st['decided_fp32'] + st['top2_fp64'] + st['full_fp64'] == fr.num_data
    assert st_ex['full_fp64'] == fr.num_data
    bad = np.nonzero(r_fast != r_ex)[0]
    assert len(bad) == 0, (bad[:10], r_fast[bad[:10]], r_ex[bad[:10]])
    np.testing.assert_array_equal(l_fast, l_ex)
