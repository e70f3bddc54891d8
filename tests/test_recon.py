"""Reconstruction driver (recon.py:16-65) and one full iteration from a config file.

CPU tests cover the loop and log format with a stand-in reconstructor; the GPU test runs
EMC from an INI + photons file (the reference's drop-in use) and checks one iteration
against the FP64 oracle (E-step argmax, then MaxLP on a pixel subset).
"""
import io

import numpy as np
import pytest

from conftest import SMALL, SMALL_GRID


class _Stub:
    def __init__(self):
        self.model = np.zeros(4, dtype='c16')
        self.calls = []

    def run_iteration(self, it):
        self.calls.append(it)
        self.model = self.model + (3 + 4j) * (it == 1)


def test_recon_loop_and_log_format():
    from refspi_amd import recon
    r = _Stub()
    log, out = io.StringIO(), io.StringIO()
    hist = recon.run(r, 3, log, out=out)
    assert r.calls == [1, 2, 3]
    lines = log.getvalue().splitlines()
    assert lines[0] == 'Iter  time(s)  change'
    assert len(lines) == 4
    # iteration 1 changes 4 entries by |3+4i| = 5 -> norm 10; later ones change nothing
    assert lines[1].startswith('1     ') and lines[1].endswith(' 1.000000e+01')
    assert lines[2].endswith(' 0.000000e+00')
    assert [h[1] for h in hist] == [10.0, 0.0, 0.0]
    assert 's/iteration on average' in out.getvalue()
    assert recon.log_line(12, 0.5, 2.0) == '12    5.00e-01 2.000000e+00\n'


def test_recon_args_and_devices(tmp_path):
    from refspi_amd import recon
    a = recon.parse_args(['5'])
    assert (a.num_iter, a.config_file, a.devices, a.streams) == (5, 'emc_config.ini', None, 4)
    d = tmp_path / 'devices'
    d.write_text('3\n1\n')
    assert recon.device_from_file(str(d)) == 3
    assert recon.device_from_file(None) == 0


def _write_inputs(tmp, frames):
    from refspi_amd import h5io
    h5io.write_photons(str(tmp / 'photons.h5'), frames)
    g = SMALL_GRID
    (tmp / 'out').mkdir()
    ini = tmp / 'config.ini'
    ini.write_text('[emc]\nsize = %d\nin_photons_file = photons.h5\noutput_folder = out/\n'
                   'log_file = out/EMC.log\nneed_scaling = False\nnum_rot = %d\nnum_modes = 1\n'
                   'sphere_dia = %g %g %d\nshiftx = %g %g %d\nshifty = %g %g %d\n' %
                   ((SMALL['size'], g['num_rot']) + g['sphere_dia'] + g['shiftx'] + g['shifty']))
    return ini


@pytest.mark.gpu
def test_config_file_iteration_matches_oracle(tmp_path, small_frames):
    from oracle import emc_oracle as O
    from refspi_amd import h5io
    from refspi_amd.emc import EMC
    ini = _write_inputs(tmp_path, small_frames)
    rec = EMC(str(ini), seed=11)
    # scale the random model towards the data (as the other parity tests do)
    rec.model = 0.05 * rec.model
    m0 = rec.model.copy()
    rec.run_iteration(1)

    g = SMALL_GRID
    fr = O.SparseFramesNP(small_frames.ones, small_frames.multi, small_frames.place_ones,
                          small_frames.place_multi, small_frames.count_multi,
                          small_frames.num_pix)
    sx, sy, dia, ns = O.state_grid(g['shiftx'], g['shifty'], g['sphere_dia'])
    prob, _, _ = O.estep_prob(m0, SMALL['size'], (sx, sy, dia), g['num_rot'], fr)
    _, rmax = O.get_rmax(prob)
    assert np.array_equal(rec.rmax, rmax)
    vals = O.unravel_rmax(rmax, ns, g['num_rot'], sx, sy, dia)
    orc = O.MaxLPOracle(fr, SMALL['size'])
    sub = np.union1d(np.nonzero(orc.mask)[0][::9], np.nonzero(orc.rmask)[0])
    traces = {}
    ref = orc.run_phaser(m0, *vals, pixels=sub, traces=traces)
    rel = np.abs(rec.model[sub] - ref[sub]) / np.maximum(np.abs(ref[sub]), 1e-300)
    for t in sub[rel > 1e-10]:
        tr = traces.get(('p3', int(t))) or traces.get(('p1', int(t)))
        assert tr is not None and min(m for _, m in tr) < 1e-12, (t, tr)

    out = h5io.read_fields(str(tmp_path / 'out' / 'output_001.h5'),
                           ['model', 'angles', 'diameters', 'shifts'])
    assert np.array_equal(out['model'], rec.model.reshape(SMALL['size'], -1))
    assert np.array_equal(out['angles'], vals[3])
    assert np.array_equal(out['diameters'], vals[2])
    assert np.array_equal(out['shifts'], np.stack([vals[0], vals[1]], 1))


@pytest.mark.gpu
def test_recon_cli_writes_log_and_outputs(tmp_path, small_frames):
    from refspi_amd import recon
    ini = _write_inputs(tmp_path, small_frames)
    recon.main(['2', '-c', str(ini)])
    lines = (tmp_path / 'out' / 'EMC.log').read_text().splitlines()
    assert lines[0] == 'Iter  time(s)  change' and len(lines) == 3
    assert (tmp_path / 'out' / 'output_001.h5').exists()
    assert (tmp_path / 'out' / 'output_002.h5').exists()


def test_dropin_module_names():
    """``from emc import EMC`` etc. (recon.py:13, emc.py:19-20) resolve to this package."""
    import importlib
    import os
    import sys
    from conftest import PKG
    d = os.path.join(PKG, 'dropin')
    sys.path.insert(0, d)
    try:
        from refspi_amd import dset, emc, maxlp
        assert importlib.import_module('emc').EMC is emc.EMC
        assert importlib.import_module('maxLP').MaxLPhaser is maxlp.MaxLPhaser
        assert importlib.import_module('dset').Dataset is dset.Dataset
        from refspi_amd import frc
        assert importlib.import_module('calc_frc').FRC is frc.FRC
    finally:
        sys.path.remove(d)
        for m in ('emc', 'maxLP', 'dset', 'calc_frc'):
            sys.modules.pop(m, None)


def test_recon_writes_frc_log_when_the_solution_is_known():
    """recon's FRC.log: one line per iteration with the FRC against ``solution``."""
    from refspi_amd import recon
    size = 16
    rng = np.random.default_rng(0)
    sol = rng.random((size, size))
    r = _Stub()
    r.model = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(sol))).ravel()
    r.run_iteration = lambda it: r.calls.append(it)
    r.solution = sol
    log, frcf = io.StringIO(), io.StringIO()
    recon.run(r, 2, log, out=io.StringIO(), frcf=frcf)
    lines = frcf.getvalue().splitlines()
    assert lines[0] == 'Iter  FRC' and len(lines) == 3
    # the model is the solution's own transform: FRC 1 at the best (zero) rotation
    assert abs(float(lines[1].split()[1]) - 1.) < 1e-6


@pytest.mark.gpu
def test_frc_improves_over_iterations_64(tmp_path):
    """Five EMC+MaxLP iterations from the seeded random start on a 64^2 photons file with
    its object stored as ``solution`` (make_data.py:164-167): the CLI writes FRC.log
    (calc_frc.py:77-117 against the object), and the reconstruction ends closer to the
    object than after its first iteration."""
    from refspi_amd import recon, synth
    cfg = synth.SynthConfig(**{**synth.CONFIG_CPU64.__dict__, 'num_data': 2000, 'seed': 4})
    synth.write_photons_file(str(tmp_path / 'photons.h5'), cfg)
    (tmp_path / 'out').mkdir()
    ini = tmp_path / 'config.ini'
    ini.write_text('[emc]\nsize = 64\nin_photons_file = photons.h5\noutput_folder = out/\n'
                   'need_scaling = False\nnum_rot = 90\nnum_modes = 1\n'
                   'sphere_dia = 7 9 21\nshiftx = -4 4 17\nshifty = -4 4 17\n')
    print('\ncase frc 64^2 five iterations', flush=True)
    np.random.seed(5)          # the reference's unseeded start, made repeatable
    recon.main(['5', '-c', str(ini)])
    lines = (tmp_path / 'out' / 'FRC.log').read_text().splitlines()
    assert lines[0] == 'Iter  FRC' and len(lines) == 6
    frc = [float(l.split()[1]) for l in lines[1:]]
    print('FRC per iteration:', frc)
    assert frc[-1] > frc[0]
