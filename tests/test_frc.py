"""FRC (refspi_amd/frc.py, the reference's calc_frc.py metric): closed forms on the CPU."""
import numpy as np
import pytest


def _obj(n=48, seed=0):
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing='ij')
    o = np.zeros((n, n))
    for _ in range(6):
        cx, cy = rng.uniform(-8, 8, 2)
        o += rng.uniform(0.5, 1) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * 1.5 ** 2))
    return o


def test_self_and_negative():
    from refspi_amd.frc import FRC
    o = _obj()
    r, f = FRC(o, o).calc()
    ring = np.zeros_like(f, dtype=bool)
    ring[:len(f) - 1] = True
    nz = f != 0
    np.testing.assert_allclose(f[nz], 1., rtol=1e-12)
    _, g = FRC(o, -o).calc()
    np.testing.assert_allclose(g[nz], -1., rtol=1e-12)
    _, a = FRC(o, -o).calc(do_abs=True)
    np.testing.assert_allclose(a[nz], 1., rtol=1e-12)
    assert np.array_equal(r, np.arange(len(f), dtype='f8'))


def test_scale_invariance_and_complex_input():
    from refspi_amd.frc import FRC
    o = _obj(seed=1)
    fo = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(o)))
    _, a = FRC(o, 3.7 * o).calc()
    _, b = FRC(fo, o).calc()
    nz = a != 0
    np.testing.assert_allclose(a[nz], 1., rtol=1e-12)
    np.testing.assert_allclose(b[nz], 1., rtol=1e-10)


def test_ring_sums_brute_force():
    from refspi_amd.frc import FRC
    o1, o2 = _obj(seed=2), _obj(seed=3)
    fr = FRC(o1, o2)
    r, f = fr.calc(binsize=1.5)
    n = o1.shape[0]
    x, y = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing='ij')
    b = (np.sqrt(x * x + y * y) / 1.5).astype(int)
    for k in (0, 3, 11):
        m = b == k
        num = (fr.fobj1[m] * np.conj(fr.fobj2[m])).sum().real
        den = np.sqrt((abs(fr.fobj1[m]) ** 2).sum() * (abs(fr.fobj2[m]) ** 2).sum())
        assert f[k] == pytest.approx(num / den, rel=1e-12)


def test_rotation_search_finds_the_angle():
    from scipy import ndimage
    from refspi_amd.frc import FRC
    o = _obj(64, seed=4)
    ro = ndimage.rotate(o, -30., order=1, prefilter=False, reshape=False)
    fr = FRC(o, ro)
    _, f0 = fr.calc()
    _, f = fr.calc_rot(num_rot=36)
    assert fr.best_angle == pytest.approx(30.)
    assert f[1:20].mean() > f0[1:20].mean()
