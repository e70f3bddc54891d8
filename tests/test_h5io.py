"""HDF5 files of the reference layout (make_data.py:228-251, emc.py:231-239), CPU only.

Interchange with h5py is checked where an interpreter with h5py exists (this image's
conda Python 3.9); those cases skip elsewhere.
"""
import os
import subprocess

import numpy as np
import pytest

from refspi_amd import h5io, synth
from refspi_amd.dset import Dataset

H5PY_PY = '/opt/conda/bin/python3.9'


def _has_h5py():
    if not os.path.exists(H5PY_PY):
        return False
    return subprocess.run([H5PY_PY, '-c', 'import h5py'], capture_output=True).returncode == 0


needs_h5py = pytest.mark.skipif(not _has_h5py(), reason='no interpreter with h5py')


@pytest.fixture(scope='module')
def frames():
    cfg = synth.SynthConfig(size=21, num_data=40, n_blobs=8, disk_radius=4, seed=3)
    return synth.make_frames(cfg)


def _same_frames(a, b):
    for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi'):
        assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k))), k


def test_photons_round_trip(tmp_path, frames):
    p = str(tmp_path / 'photons.h5')
    h5io.write_photons(p, frames)
    r = h5io.read_photons(p)
    assert r.num_pix == 21 * 21
    _same_frames(r, frames)
    assert r.bg is None or np.array_equal(r.bg, frames.bg)


def test_photons_with_bg_and_empty_frames(tmp_path):
    ones = np.array([0, 2, 0, 1], dtype='i4')
    multi = np.array([1, 0, 0, 0], dtype='i4')
    fr = type('F', (), {})()
    fr.num_pix, fr.ones, fr.multi = 9, ones, multi
    fr.place_ones = np.array([1, 5, 8], dtype='i4')
    fr.place_multi = np.array([4], dtype='i4')
    fr.count_multi = np.array([3], dtype='i4')
    fr.bg = np.linspace(0, 1, 9)
    p = str(tmp_path / 'p.h5')
    h5io.write_photons(p, fr)
    d = Dataset(p, 9)
    _same_frames(d, fr)
    assert np.array_equal(d.bg, fr.bg)
    assert np.array_equal(d.get_frame(0), [0, 0, 0, 0, 3, 0, 0, 0, 0])
    assert np.array_equal(d.get_frame(1), [0, 1, 0, 0, 0, 1, 0, 0, 0])
    assert d.mean_count == pytest.approx(6 / 4)
    with pytest.raises(AttributeError):
        Dataset(p, 10)


def test_dataset_from_file_matches_memory(tmp_path, frames):
    p = str(tmp_path / 'photons.h5')
    h5io.write_photons(p, frames)
    a = Dataset(p, 21 * 21, need_scaling=True)
    b = Dataset(frames=frames, need_scaling=True)
    _same_frames(a, b)
    assert np.array_equal(a.counts, b.counts)
    assert np.array_equal(a.get_powder(), b.get_powder())


def test_fields_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    fields = {'model': rng.random((7, 7)) + 1j * rng.random((7, 7)),
              'angles': rng.random(5), 'diameters': rng.random(5),
              'shifts': rng.random((5, 2)), 'rmax': np.arange(5, dtype='i4')}
    p = str(tmp_path / 'out.h5')
    h5io.write_fields(p, fields)
    back = h5io.read_fields(p, list(fields) + ['absent'])
    assert 'absent' not in back
    for k, v in fields.items():
        assert np.array_equal(back[k], v), k
        assert back[k].shape == v.shape


@needs_h5py
def test_h5py_reads_our_files(tmp_path, frames):
    p = str(tmp_path / 'photons.h5')
    q = str(tmp_path / 'out.h5')
    h5io.write_photons(p, frames)
    m = (np.arange(9.) + 1j * np.arange(9.)[::-1]).reshape(3, 3)
    h5io.write_fields(q, {'model': m, 'shifts': np.ones((4, 2))})
    code = (
        "import h5py, numpy as np, sys\n"
        "f = h5py.File(sys.argv[1], 'r')\n"
        "po = np.hstack(f['place_ones'][:]); cm = np.hstack(f['count_multi'][:])\n"
        "print(int(f['num_pix'][0]), f['ones'][:].sum(), po.sum(), cm.sum(), f['multi'].dtype)\n"
        "g = h5py.File(sys.argv[2], 'r')\n"
        "print(g['model'].dtype, g['model'][1, 2], g['shifts'].shape)\n")
    out = subprocess.run([H5PY_PY, '-c', code, p, q], capture_output=True, text=True,
                         check=True).stdout.split('\n')
    assert out[0].split() == [str(21 * 21), str(int(frames.ones.sum())),
                              str(int(frames.place_ones.sum())),
                              str(int(frames.count_multi.sum())), 'int32']
    assert out[1] == 'complex128 (5+3j) (4, 2)'


@needs_h5py
@pytest.mark.parametrize('with_counts', [True, False])
def test_we_read_h5py_files(tmp_path, frames, with_counts):
    """A photons file written by h5py as make_data.py:245-251 does; without ``ones`` and
    ``multi`` the counts come from the vlen lengths (dset.py:28-31, 37-40)."""
    p = str(tmp_path / 'photons.h5')
    arrs = str(tmp_path / 'arrs.npz')
    np.savez(arrs, ones=frames.ones, multi=frames.multi, po=frames.place_ones,
             pm=frames.place_multi, cm=frames.count_multi)
    code = (
        "import h5py, numpy as np, sys\n"
        "a = np.load(sys.argv[2]); n = len(a['ones'])\n"
        "oo = np.concatenate([[0], np.cumsum(a['ones'])]); mo = np.concatenate([[0], np.cumsum(a['multi'])])\n"
        "f = h5py.File(sys.argv[1], 'w')\n"
        "f['num_pix'] = np.array([441])\n"
        "dt = h5py.special_dtype(vlen=np.dtype('i4'))\n"
        "po = f.create_dataset('place_ones', (n,), dtype=dt); pm = f.create_dataset('place_multi', (n,), dtype=dt)\n"
        "cm = f.create_dataset('count_multi', (n,), dtype=dt)\n"
        "if sys.argv[3] == '1':\n"
        "    f.create_dataset('ones', (n,), dtype='i4')[...] = a['ones']\n"
        "    f.create_dataset('multi', (n,), dtype='i4')[...] = a['multi']\n"
        "for i in range(n):\n"
        "    po[i] = a['po'][oo[i]:oo[i+1]]; pm[i] = a['pm'][mo[i]:mo[i+1]]; cm[i] = a['cm'][mo[i]:mo[i+1]]\n"
        "f['bg'] = np.full((21, 21), 0.25)\n"
        "f.close()\n")
    subprocess.run([H5PY_PY, '-c', code, p, arrs, '1' if with_counts else '0'], check=True)
    d = Dataset(p, 441)
    _same_frames(d, frames)
    assert d.bg.shape == (441,) and np.all(d.bg == 0.25)


def test_missing_library_message(monkeypatch):
    monkeypatch.setattr(h5io, '_h5', None)
    monkeypatch.setattr(h5io, '_CANDIDATES', ['/nonexistent/libhdf5.so'])
    with pytest.raises(RuntimeError, match='libhdf5'):
        h5io.read_photons('x.h5')
