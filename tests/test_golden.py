"""The oracle reproduces the committed golden fixture (tests/golden/golden_small.npz).

The fixture was written by tests/golden/make_golden.py from the FP64 oracle; re-running the
same computation here must give identical integers (frames, argmax, photon-pull indices)
and floats within 1e-12 relative (NumPy reduction order is the only freedom).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, 'golden', 'golden_small.npz')


@pytest.fixture(scope='module')
def regenerated():
    sys.path.insert(0, os.path.join(HERE, 'golden'))
    try:
        import make_golden
    finally:
        sys.path.pop(0)
    fr, model = make_golden.inputs()
    return make_golden.compute(fr, model)


def test_fixture_is_reproduced(regenerated):
    g = np.load(GOLDEN)
    assert set(g.files) == set(regenerated)
    for k in g.files:
        a, b = g[k], np.asarray(regenerated[k])
        assert a.shape == b.shape, k
        if a.dtype.kind in 'iub':
            assert np.array_equal(a, b), k
        else:
            scale = np.max(np.abs(a)) if a.size else 0.
            np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * scale, err_msg=k)


def test_fixture_invariants():
    g = np.load(GOLDEN)
    size, R = int(g['size']), int(g['num_rot'])
    for it in (1, 2):
        p = 'it%d_' % it
        prob = g[p + 'prob']
        assert np.array_equal(g[p + 'rmax'], prob.argmax(0))
        assert np.array_equal(g[p + 'maxll'], prob.max(0))
        # pixels outside MaxLP's mask keep the input model (maxLP.py:332-336)
        qx = (np.arange(size) - size // 2) / size
        q = np.hypot(qx[:, None], qx[None, :]).ravel()
        out = ~((q > 4 / size) & (q < 0.5))
        assert np.array_equal(g[p + 'model_out'][out], g[p + 'model_in'][out])
    np.testing.assert_allclose(g['rot_prob'].sum(1), 1., rtol=1e-13)
    assert g['rot_loss'] == pytest.approx(-(g['rot_logz'] - np.log(R)).sum(), rel=1e-14)
