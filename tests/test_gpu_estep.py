"""GPU parity of the E-step (C-ABI through the facade) against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

from conftest import SMALL, SMALL_GRID

pytestmark = pytest.mark.gpu


def _oracle_setup(frames, model):
    from oracle import emc_oracle as O
    fr = O.SparseFramesNP(frames.ones, frames.multi, frames.place_ones, frames.place_multi,
                          frames.count_multi, frames.num_pix)
    g = SMALL_GRID
    sx, sy, dia, ns = O.state_grid(g['shiftx'], g['shifty'], g['sphere_dia'])
    return O, fr, (sx, sy, dia), ns


def _emc(frames, model, **kw):
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    g = SMALL_GRID
    return EMC(dataset=Dataset(frames=frames), size=SMALL['size'], num_rot=g['num_rot'],
               shiftx=g['shiftx'], shifty=g['shifty'], sphere_dia=g['sphere_dia'],
               model=model, save=False, **kw)


def test_mirror_slice_gen_holo_and_slice_gen(small_model):
    import torch
    from oracle import emc_oracle as O
    from refspi_amd import _lib
    size = SMALL['size']
    bg = np.zeros(size * size)
    m = torch.as_tensor(small_model, device='cuda')
    tbg = torch.as_tensor(bg, device='cuda')
    out = torch.empty(size * size, dtype=torch.float64, device='cuda')
    for (sx, sy, dia, log) in ((0.5, -1.25, 7.3, 0), (-2., 2., 9., 1), (0., 0., 8., 0)):
        _lib.call('rsp_slice_gen_holo', _lib.ptr(m), sx, sy, dia, 1., 1., size, _lib.ptr(tbg),
                  log, _lib.ptr(out), _lib.stream_handle())
        torch.cuda.synchronize()
        ref = O.slice_gen_holo(small_model, sx, sy, dia, 1., 1., size, bg, log)
        np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-13, atol=1e-13)
    view = O.slice_gen_holo(small_model, 1., 0.5, 8., 1., 1., size, bg, 0)
    tv = torch.as_tensor(view, device='cuda')
    for ang in (0., 0.3, np.pi / 2, 2.1):
        _lib.call('rsp_slice_gen', _lib.ptr(tv), ang, 1., size, _lib.ptr(tbg), 1,
                  _lib.ptr(out), _lib.stream_handle())
        torch.cuda.synchronize()
        ref = O.slice_gen(view, ang, 1., size, bg, 1)
        np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-13, atol=1e-12)


def test_mirror_calc_prob_all(small_frames, small_model):
    import torch
    from refspi_amd import _lib
    O, fr, states, ns = _oracle_setup(small_frames, small_model)
    size = SMALL['size']
    pm = O.probmask(size)
    lv = O.slice_gen(O.slice_gen_holo(small_model, 1., -1., 8., 1., 1., size, fr.bg, 0),
                     0.7, 1., size, fr.bg, 1)
    scales = np.linspace(0.5, 1.5, fr.num_data)
    ref = O.calc_prob_all(lv, pm, fr.ones, fr.multi, fr.ones_accum, fr.multi_accum,
                          fr.place_ones, fr.place_multi, fr.count_multi, -3.5, scales)
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device='cuda')
    out = torch.empty(fr.num_data, dtype=torch.float64, device='cuda')
    args = [d(lv), d(pm), fr.num_data, d(fr.ones), d(fr.multi), d(fr.ones_accum),
            d(fr.multi_accum), d(fr.place_ones), d(fr.place_multi), d(fr.count_multi)]
    _lib.call('rsp_calc_prob_all', *[_lib.ptr(a) if not isinstance(a, int) else a for a in args],
              -3.5, _lib.ptr(d(scales)), _lib.ptr(out), _lib.stream_handle())
    torch.cuda.synchronize()
    # same sequential order; only the log() implementations differ
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-13, atol=1e-11)


@pytest.mark.parametrize('exact', [0, 1])
def test_estep_argmax_matches_oracle(small_frames, small_model, exact):
    O, fr, states, ns = _oracle_setup(small_frames, small_model)
    prob, msums, vscale = O.estep_prob(small_model, SMALL['size'], states,
                                       SMALL_GRID['num_rot'], fr, sequential=True)
    mx, rmax = O.get_rmax(prob)
    emc = _emc(small_frames, small_model)
    if exact:
        emc.set_option(1, 1)
    import torch
    emc._calculate_prob(torch.as_tensor(small_model, device='cuda'))
    emc._get_rmax()
    st = emc.estep_stats()
    print('stats', st)
    np.testing.assert_allclose(emc.msums, msums, rtol=1e-13)
    assert abs(emc.vscale - vscale) <= 1e-13 * abs(vscale)
    margin = O.top_margin(prob)
    bad = np.nonzero(emc.rmax != rmax)[0]
    # any disagreement must sit on an FP64 near-tie of the oracle itself
    assert np.all(margin[bad] < 1e-9 * np.abs(mx[bad])), (bad, margin[bad])
    assert len(bad) == 0
    np.testing.assert_allclose(emc.maxll, mx, rtol=1e-12, atol=1e-9)
    if not exact:
        assert st['decided_fp32'] + st['top2_fp64'] + st['full_fp64'] == fr.num_data


def test_estep_deterministic(small_frames, small_model):
    import torch
    emc = _emc(small_frames, small_model)
    m = torch.as_tensor(small_model, device='cuda')
    emc._calculate_prob(m)
    r1, l1 = emc._get_rmax().copy(), emc.maxll.copy()
    emc._calculate_prob(m)
    r2, l2 = emc._get_rmax(), emc.maxll
    assert np.array_equal(r1, r2)
    assert np.array_equal(l1.view('u8'), l2.view('u8'))


@pytest.mark.parametrize('banks', [0, 64])
def test_estep_bank_models_do_not_change_result(small_frames, small_model, banks):
    import torch
    m = torch.as_tensor(small_model, device='cuda')
    base = _emc(small_frames, small_model)
    base._calculate_prob(m)
    r0 = base._get_rmax().copy()
    emc = _emc(small_frames, small_model, bank_count=banks)
    emc._calculate_prob(m)
    assert np.array_equal(emc._get_rmax(), r0)
    np.testing.assert_array_equal(emc.maxll, base.maxll)


def test_estep_frame_regrouping_does_not_change_result(small_frames, small_model):
    """Option 6 (swaps of the lane-group regrouping) changes only the slot layout."""
    import torch
    m = torch.as_tensor(small_model, device='cuda')
    out = []
    for iters in (0, 50, 4000):
        emc = _emc(small_frames, small_model, regroup_iters=iters)
        emc._calculate_prob(m)
        out.append((emc._get_rmax().copy(), emc.maxll.copy()))
    for r, l in out[1:]:
        assert np.array_equal(r, out[0][0])
        np.testing.assert_array_equal(l, out[0][1])


@pytest.mark.parametrize('size,parts', [(185, 2), (128, 1)])
def test_estep_table_layout_does_not_change_result(size, parts):
    """Option 7 (per-tile LDS layout of the table parts) changes only the slot layout and
    the table loads' addresses: argmax and FP64 log-likelihood are identical with and
    without it, on frames numerous enough for full workgroup tiles (185^2: two table
    parts; 128^2: one), and the layout lowers the gather's slot count (one frame per lane
    and the exact, distinct-bank schedule: broadcast-aware slots can already sit at the lane
    floor, and tiles of two frames per lane are twice as many waves)."""
    import types
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': 6000, 'seed': 5,
                               'size': size})
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = types.SimpleNamespace(num_pix=d['num_pix'], bg=None, **{
        k: d[k].cpu().numpy().astype('i4')
        for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')})
    model = synth.model_scales(cfg)[0] * synth.object_ft(cfg)
    grid = dict(shiftx=(-1., 1., 3), shifty=(-1., 1., 3), sphere_dia=(7., 9., 2))
    out = []
    for lay in (0, 1):
        emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=24, model=model, seed=1,
                  save=False, table_layout=lay, shared_slots=0, frames_per_lane=1, **grid)
        emc._calculate_prob(torch.as_tensor(emc.model, device='cuda'))
        emc._get_rmax()
        out.append((emc.rmax.copy(), emc.maxll.copy(), emc.estep_timing(), emc.estep_stats()))
        del emc
    print('\nLDS read bytes per hypothesis: identity %.0f, layout %.0f' %
          (out[0][2]['lds_read_bytes_per_hyp'], out[1][2]['lds_read_bytes_per_hyp']))
    assert out[1][3]['table_parts'] == parts
    assert np.array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[1][2]['lds_read_bytes_per_hyp'] < out[0][2]['lds_read_bytes_per_hyp']


def test_estep_shared_slots_do_not_change_result():
    """Option 8 (broadcast-aware photon slots: lanes that read the same table entry share
    its bank in a slot) changes only the slot layout: argmax and FP64 log-likelihood are
    identical with and without it at 185^2 (two table parts, per-tile layouts on), and it
    lowers the gather's LDS reads."""
    import types
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': 6000, 'seed': 13})
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = types.SimpleNamespace(num_pix=d['num_pix'], bg=None, **{
        k: d[k].cpu().numpy().astype('i4')
        for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')})
    model = synth.model_scales(cfg)[0] * synth.object_ft(cfg)
    grid = dict(shiftx=(-1., 1., 3), shifty=(-1., 1., 3), sphere_dia=(7., 9., 2))
    out = []
    for shared in (0, 1):
        emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=24, model=model, seed=1,
                  save=False, shared_slots=shared, **grid)
        emc._calculate_prob(torch.as_tensor(emc.model, device='cuda'))
        emc._get_rmax()
        out.append((emc.rmax.copy(), emc.maxll.copy(), emc.estep_timing()))
        del emc
    print('\nLDS read bytes per hypothesis: distinct banks %.0f, shared slots %.0f' %
          (out[0][2]['lds_read_bytes_per_hyp'], out[1][2]['lds_read_bytes_per_hyp']))
    assert np.array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[1][2]['lds_read_bytes_per_hyp'] < out[0][2]['lds_read_bytes_per_hyp']


@pytest.mark.parametrize('size,n', [(185, 6000), (128, 3000)])
def test_estep_two_frames_per_lane_do_not_change_result(size, n):
    """Option 9 (whole frames two to a lane, the slot registers holding two 16-bit LDS
    addresses) changes only which lanes and registers hold a frame's photons: argmax and
    FP64 log-likelihood are identical with one and two frames per lane (185^2: two table
    parts; 128^2: one), and the table loads per hypothesis fall (a workgroup of two-frame
    waves holds twice the frames; at these frame counts partial workgroups remain)."""
    import types
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': n, 'seed': 23,
                               'size': size})
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = types.SimpleNamespace(num_pix=d['num_pix'], bg=None, **{
        k: d[k].cpu().numpy().astype('i4')
        for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')})
    model = synth.model_scales(cfg)[0] * synth.object_ft(cfg)
    grid = dict(shiftx=(-1., 1., 3), shifty=(-1., 1., 3), sphere_dia=(7., 9., 2))
    out = []
    for fpl in (1, 2):
        emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=24, model=model, seed=1,
                  save=False, frames_per_lane=fpl, **grid)
        emc._calculate_prob(torch.as_tensor(emc.model, device='cuda'))
        emc._get_rmax()
        out.append((emc.rmax.copy(), emc.maxll.copy(), emc.estep_timing()))
        del emc
    print('\nLDS write bytes per hypothesis (table loads): one frame per lane %.0f, two %.0f'
          % (out[0][2]['lds_write_bytes_per_hyp'], out[1][2]['lds_write_bytes_per_hyp']))
    assert np.array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[1][2]['lds_write_bytes_per_hyp'] < 0.8 * out[0][2]['lds_write_bytes_per_hyp']


def test_estep_class_streams_do_not_change_result():
    """Option 10 (a batch's slot-class gather launches on side streams, forked after the
    table build and joined before the merge) changes only when the classes run: argmax
    and FP64 log-likelihood are identical on the caller's stream alone and with one, two
    and four side streams, over several batches (a batch of 40 hypotheses, so that the
    tables buffer and the partials are reused while streams of the previous batch could
    still be running if the join were missing)."""
    import types
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': 6000, 'seed': 29})
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = types.SimpleNamespace(num_pix=d['num_pix'], bg=None, **{
        k: d[k].cpu().numpy().astype('i4')
        for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')})
    model = synth.model_scales(cfg)[0] * synth.object_ft(cfg)
    grid = dict(shiftx=(-1., 1., 3), shifty=(-1., 1., 3), sphere_dia=(7., 9., 2))
    emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=24, model=model, seed=1,
              save=False, **grid)
    emc.set_option(2, 40)
    m = torch.as_tensor(emc.model, device='cuda')
    out = []
    for cs in (0, 1, 2, 4):
        emc.set_option(10, cs)
        emc._calculate_prob(m)
        emc._get_rmax()
        out.append((emc.rmax.copy(), emc.maxll.copy()))
    for r, ll in out[1:]:
        assert np.array_equal(out[0][0], r)
        np.testing.assert_array_equal(out[0][1], ll)
    with pytest.raises(Exception):
        emc.set_option(10, 5)


def test_estep_bank_models_full_parts():
    """The 64-bank and unscheduled slot layouts on 185^2 frames, whose first table part is
    full (12,512 entries): the zero pads of a full part stay inside its LDS buffer, and
    argmax and FP64 log-likelihood equal the 32-bank plan's bit for bit."""
    import types
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': 3000, 'seed': 9})
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = types.SimpleNamespace(num_pix=d['num_pix'], bg=None, **{
        k: d[k].cpu().numpy().astype('i4')
        for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')})
    model = synth.model_scales(cfg)[0] * synth.object_ft(cfg)
    grid = dict(shiftx=(-1., 1., 3), shifty=(-1., 1., 3), sphere_dia=(7., 9., 2))
    out = {}
    for banks in (32, 64, 0):
        emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=24, model=model, seed=1,
                  save=False, bank_count=banks, **grid)
        emc._calculate_prob(torch.as_tensor(emc.model, device='cuda'))
        emc._get_rmax()
        assert emc.estep_stats()['table_parts'] == 2
        out[banks] = (emc.rmax.copy(), emc.maxll.copy())
        del emc
    for banks in (64, 0):
        assert np.array_equal(out[banks][0], out[32][0]), banks
        np.testing.assert_array_equal(out[banks][1], out[32][1])


def test_estep_batching_and_segments_do_not_change_result(small_frames, small_model):
    import torch
    m = torch.as_tensor(small_model, device='cuda')
    base = _emc(small_frames, small_model)
    base._calculate_prob(m)
    r0 = base._get_rmax().copy()
    for batch, nseg in ((7, 3), (64, 8), (1000, 1)):
        emc = _emc(small_frames, small_model)
        emc.set_option(2, batch)
        emc.set_option(4, nseg)
        emc._calculate_prob(m)
        assert np.array_equal(emc._get_rmax(), r0), (batch, nseg)


@pytest.mark.parametrize('size', [141, 205, 256])
def test_estep_multi_part_tables(size):
    """Scored annuli larger than one LDS part (2, 3 and 4 table parts) match the oracle."""
    import torch
    from oracle import emc_oracle as O
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    cfg = synth.SynthConfig(size=size, num_data=40, mean_count=400., n_blobs=20,
                            disk_radius=8., beamstop=4., shift_max=1., seed=size)
    fr = synth.make_frames(cfg)
    grid = dict(shiftx=(-1., 1., 2), shifty=(0., 0., 1), sphere_dia=(7., 9., 2))
    rng = np.random.default_rng(size)
    model = 0.05 * (rng.random(size * size) + 1j * rng.random(size * size))
    emc = EMC(dataset=Dataset(frames=fr), size=size, num_rot=5, model=model, save=False,
              **grid)
    emc._calculate_prob(torch.as_tensor(model, device='cuda'))
    emc._get_rmax()
    st = emc.estep_stats()
    assert st['table_parts'] == {141: 2, 205: 3, 256: 4}[size]
    ofr = O.SparseFramesNP(fr.ones, fr.multi, fr.place_ones, fr.place_multi, fr.count_multi,
                           fr.num_pix)
    sx, sy, dia, ns = O.state_grid(grid['shiftx'], grid['shifty'], grid['sphere_dia'])
    prob, msums, vscale = O.estep_prob(model, size, (sx, sy, dia), 5, ofr)
    mx, rmax = O.get_rmax(prob)
    margin = O.top_margin(prob)
    bad = np.nonzero(emc.rmax != rmax)[0]
    assert np.all(margin[bad] < 1e-9 * np.abs(mx[bad])), (bad, margin[bad])
    np.testing.assert_allclose(emc.maxll, mx, rtol=1e-12, atol=1e-9)


def test_estep_mixed_split_and_whole_frames():
    """Frames heavy enough to be split over 2-64 lanes come first (their waves hold 32 .. 1
    frames, which shifts the lane groups of the whole-frame waves after them), then
    regrouped whole-frame waves: the argmax still matches the oracle, with the regrouping
    on or off."""
    import torch
    from oracle import emc_oracle as O
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    size = 141
    cfg = synth.SynthConfig(size=size, num_data=300, mean_count=200., n_blobs=20,
                            disk_radius=8., beamstop=4., shift_max=1., seed=11,
                            fluence='lognormal', fluence_sigma=1.0)
    fr = synth.make_frames(cfg)
    n = fr.num_data
    counts = fr.ones + np.bincount(np.repeat(np.arange(n), fr.multi), weights=fr.count_multi,
                                   minlength=n)
    # up to ~13x the median: frames split over 2 .. 64 lanes beside whole frames
    assert counts.max() > 8 * np.median(counts)
    grid = dict(shiftx=(-1., 1., 2), shifty=(0., 0., 1), sphere_dia=(7., 9., 2))
    rng = np.random.default_rng(5)
    model = 0.05 * (rng.random(size * size) + 1j * rng.random(size * size))
    ofr = O.SparseFramesNP(fr.ones, fr.multi, fr.place_ones, fr.place_multi, fr.count_multi,
                           fr.num_pix)
    sx, sy, dia, ns = O.state_grid(grid['shiftx'], grid['shifty'], grid['sphere_dia'])
    prob, msums, vscale = O.estep_prob(model, size, (sx, sy, dia), 5, ofr)
    mx, rmax = O.get_rmax(prob)
    margin = O.top_margin(prob)
    for iters in (None, 0):
        emc = EMC(dataset=Dataset(frames=fr), size=size, num_rot=5, model=model, save=False,
                  regroup_iters=iters, **grid)
        emc._calculate_prob(torch.as_tensor(model, device='cuda'))
        emc._get_rmax()
        bad = np.nonzero(emc.rmax != rmax)[0]
        assert np.all(margin[bad] < 1e-9 * np.abs(mx[bad])), (iters, bad, margin[bad])
        np.testing.assert_allclose(emc.maxll, mx, rtol=1e-12, atol=1e-9)
