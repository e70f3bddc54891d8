"""The joint E-step with per-frame scales (need_scaling) and a background, against the
FP64 oracle.

Reference terms exercised here and nowhere else on the GPU: the init term
``init * scales[d]`` of calc_prob_all (kernels.cu:126) with scales = counts / mean_count
(emc.py:129-132, dset.py:50-53), and the background added twice, once by slice_gen_holo
(kernels.cu:75) and once by slice_gen (kernels.cu:35), read from the photons file's
``bg`` (dset.py:56-60).
"""
import numpy as np
import pytest

from conftest import SMALL, SMALL_GRID

pytestmark = pytest.mark.gpu


def _frames_with_bg(small_frames, level):
    from refspi_amd.synth import SparseFrames
    size = SMALL['size']
    rng = np.random.default_rng(17)
    bg = level * (0.5 + rng.random(size * size)) if level else None
    f = small_frames
    return SparseFrames(f.num_pix, f.ones, f.multi, f.place_ones, f.place_multi,
                        f.count_multi, bg)


@pytest.mark.parametrize('need_scaling,bg_level,exact', [(True, 0., 0), (False, 0.05, 0),
                                                         (True, 0.05, 0), (True, 0.05, 1)])
def test_estep_scales_and_background_match_oracle(small_frames, small_model, need_scaling,
                                                   bg_level, exact):
    import torch
    from oracle import emc_oracle as O
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    fr_in = _frames_with_bg(small_frames, bg_level)
    g = SMALL_GRID
    size = SMALL['size']
    model = 0.2 * small_model
    ds = Dataset(frames=fr_in, need_scaling=need_scaling)
    emc = EMC(dataset=ds, size=size, num_rot=g['num_rot'], shiftx=g['shiftx'],
              shifty=g['shifty'], sphere_dia=g['sphere_dia'], model=model,
              need_scaling=need_scaling, save=False)
    if exact:
        emc.set_option(1, 1)
    if need_scaling:
        assert not np.all(emc.scales == 1.)
    emc._calculate_prob(torch.as_tensor(model, device='cuda'))
    emc._get_rmax()
    ofr = O.SparseFramesNP(fr_in.ones, fr_in.multi, fr_in.place_ones, fr_in.place_multi,
                           fr_in.count_multi, fr_in.num_pix, bg=fr_in.bg)
    # emc.py:129-132: scales = counts / mean_count
    scales = ofr.counts() / ofr.mean_count if need_scaling else None
    if need_scaling:
        np.testing.assert_array_equal(emc.scales, scales)
    sx, sy, dia, ns = O.state_grid(g['shiftx'], g['shifty'], g['sphere_dia'])
    prob, msums, vscale = O.estep_prob(model, size, (sx, sy, dia), g['num_rot'], ofr,
                                       scales=scales, sequential=True)
    mx, rmax = O.get_rmax(prob)
    np.testing.assert_allclose(emc.msums, msums, rtol=1e-13)
    assert abs(emc.vscale - vscale) <= 1e-13 * abs(vscale)
    margin = O.top_margin(prob)
    bad = np.nonzero(emc.rmax != rmax)[0]
    assert np.all(margin[bad] < 1e-9 * np.abs(mx[bad])), (bad, margin[bad])
    np.testing.assert_allclose(emc.maxll, mx, rtol=1e-12, atol=1e-9)
    # the terms matter: without the scales / the background the argmax or LL moves
    if need_scaling:
        p1, _, _ = O.estep_prob(model, size, (sx, sy, dia), g['num_rot'], ofr)
        assert not np.allclose(p1.max(0), mx, rtol=1e-6)
    if bg_level:
        o0 = O.SparseFramesNP(fr_in.ones, fr_in.multi, fr_in.place_ones, fr_in.place_multi,
                              fr_in.count_multi, fr_in.num_pix)
        p0, _, _ = O.estep_prob(model, size, (sx, sy, dia), g['num_rot'], o0, scales=scales)
        assert not np.allclose(p0.max(0), mx, rtol=1e-6)


@pytest.mark.parametrize('bad_count', [0, -3, 1 << 22])
def test_plans_refuse_out_of_range_multi_counts(bad_count):
    """count_multi must lie in [1, 2^22) (a multi photon's count): the E-step plan's
    exactness bound and slot lists, MaxLP's entry keys and the rotation form's packing
    assume it.  Every plan refuses other values with -2 when the frames are loaded."""
    import ctypes as C
    from oracle import emc_oracle as O
    from refspi_amd import _lib
    lib = _lib.load()
    size, R = 21, 4
    i4 = lambda *v: np.array(v, dtype='i4')
    ones, multi = i4(2, 1), i4(1, 1)
    po, pm, cm = i4(50, 60, 70), i4(80, 90), i4(3, bad_count)
    ptr = lambda a: _lib.host_ptr(a, C.c_int32)
    rc, rs = O.rotation_table(R)
    sx = np.zeros(1)
    dia = np.full(1, 8.)
    h = C.c_void_p()
    _lib.check(lib.rsp_emc_create(0, size, R, _lib.host_ptr(rc, C.c_double),
                                  _lib.host_ptr(rs, C.c_double), _lib.host_ptr(sx, C.c_double),
                                  _lib.host_ptr(sx, C.c_double), _lib.host_ptr(dia, C.c_double),
                                  1, C.byref(h)), 'rsp_emc_create')
    try:
        assert lib.rsp_emc_set_frames(h, 2, ptr(ones), ptr(multi), ptr(po), ptr(pm), ptr(cm),
                                      None) == -2
        assert b'count_multi' in lib.rsp_last_error()
    finally:
        lib.rsp_emc_destroy(h)
    m = C.c_void_p()
    _lib.check(lib.rsp_maxlp_create(0, size, C.byref(m)), 'rsp_maxlp_create')
    try:
        assert lib.rsp_maxlp_set_frames(m, 2, ptr(ones), ptr(multi), ptr(po), ptr(pm),
                                        ptr(cm)) == -2
    finally:
        lib.rsp_maxlp_destroy(m)
    r = C.c_void_p()
    _lib.check(lib.rsp_rot_create(0, size, R, _lib.ptr(rc), _lib.ptr(rs), 1., C.byref(r)),
               'rsp_rot_create')
    try:
        assert lib.rsp_rot_set_frames(r, 2, _lib.ptr(ones), _lib.ptr(multi), _lib.ptr(po),
                                      _lib.ptr(pm), _lib.ptr(cm)) == -2
        assert b'count_multi' in lib.rsp_last_error()
    finally:
        lib.rsp_rot_destroy(r)
