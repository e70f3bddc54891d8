import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'refspi-maxlp_amd')
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device) and the built library')


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason='no HIP device visible')
    for it in items:
        if 'gpu' in it.keywords:
            it.add_marker(skip)


# small parity workload shared by the tests: 25x25 q-grid, 96 frames, 27 states x 12
SMALL = dict(size=25, num_data=96, mean_count=150., n_blobs=12, disk_radius=5.,
             beamstop=3., dia_mean=8., dia_sd=0.4, shift_max=2., seed=0)
SMALL_GRID = dict(shiftx=(-2., 2., 3), shifty=(-2., 2., 3), sphere_dia=(7., 9., 3), num_rot=12)


@pytest.fixture(scope='session')
def small_frames():
    from refspi_amd import synth
    return synth.make_frames(synth.SynthConfig(**SMALL))


@pytest.fixture(scope='session')
def small_model():
    rng = np.random.default_rng(1)
    n = SMALL['size'] ** 2
    return rng.random(n) + 1j * rng.random(n)
