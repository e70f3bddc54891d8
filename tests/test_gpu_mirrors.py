"""Every single-kernel mirror of the C-ABI (kernels.cu's entry points) against the oracle,
called through the C-ABI on the GPU: slice_gen_holo, slice_gen, calc_prob_all, merge_all,
slice_merge, get_f_dt."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _d(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device='cuda')


def _frames(n=64, size=24, seed=0):
    from oracle import emc_oracle as O
    rng = np.random.default_rng(seed)
    ones = rng.integers(0, 40, n).astype('i4')
    ones[min(5, n - 1)] = 0
    multi = rng.integers(0, 6, n).astype('i4')
    po = rng.integers(0, size * size, ones.sum()).astype('i4')
    pm = rng.integers(0, size * size, multi.sum()).astype('i4')
    cm = rng.integers(2, 7, multi.sum()).astype('i4')
    return O.SparseFramesNP(ones, multi, po, pm, cm, size * size)


def _frame_args(fr):
    return [_d(fr.ones), _d(fr.multi), _d(fr.ones_accum.astype('i8')),
            _d(fr.multi_accum.astype('i8')), _d(fr.place_ones), _d(fr.place_multi),
            _d(fr.count_multi)]


@pytest.mark.parametrize('size', [24, 25])
def test_slice_gen_holo_and_slice_gen_mirrors(size):
    import torch
    from oracle import emc_oracle as O
    from refspi_amd import _lib
    rng = np.random.default_rng(size)
    model = 0.3 * (rng.random(size * size) + 1j * rng.random(size * size))
    bg = 0.01 * rng.random(size * size)
    out = torch.empty(size * size, dtype=torch.float64, device='cuda')
    tm, tbg = _d(model), _d(bg)
    for (sx, sy, dia, rel, sc, log) in ((0.5, -1.25, 7.3, 1., 1., 0), (-2., 2., 9., 1.3, 2., 1)):
        _lib.call('rsp_slice_gen_holo', _lib.ptr(tm), sx, sy, dia, rel, sc, size,
                  _lib.ptr(tbg), log, _lib.ptr(out), _lib.stream_handle())
        torch.cuda.synchronize()
        ref = O.slice_gen_holo(model, sx, sy, dia, rel, sc, size, bg, log)
        np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-13, atol=1e-14)
    img = O.slice_gen_holo(model, 1., 0.5, 8., 1., 1., size, bg, 0)
    timg = _d(img)
    for ang in (0., 0.3, np.pi / 2, 2.1):
        for log in (0, 1):
            _lib.call('rsp_slice_gen', _lib.ptr(timg), ang, 1.5, size, _lib.ptr(tbg), log,
                      _lib.ptr(out), _lib.stream_handle())
            torch.cuda.synchronize()
            ref = O.slice_gen(img, ang, 1.5, size, bg, log)
            got = out.cpu().numpy()
            if log:
                np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-13)
                # the -1000 rim of kernels.cu:24-25 (last row/column at angle 0)
                if ang == 0.:
                    assert np.all(got.reshape(size, size)[-1] == -1000.)
            else:
                np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


def test_calc_prob_all_mirror():
    import torch
    from oracle import emc_oracle as O
    from refspi_amd import _lib
    size = 24
    fr = _frames(size=size)
    rng = np.random.default_rng(2)
    lv = rng.normal(size=size * size)
    pm = O.probmask(size)
    sc = rng.random(fr.num_data) + 0.5
    ref = O.calc_prob_all(lv, pm, fr.ones, fr.multi, fr.ones_accum, fr.multi_accum,
                          fr.place_ones, fr.place_multi, fr.count_multi, -3.5, sc)
    out = torch.empty(fr.num_data, dtype=torch.float64, device='cuda')
    # device copies held in names for the duration of the call (a temporary's memory could
    # be handed to the next allocation before the kernel runs)
    keep = [_d(lv), _d(pm)] + _frame_args(fr) + [_d(sc)]
    _lib.call('rsp_calc_prob_all', _lib.ptr(keep[0]), _lib.ptr(keep[1]), fr.num_data,
              *[_lib.ptr(a) for a in keep[2:9]], -3.5, _lib.ptr(keep[9]), _lib.ptr(out),
              _lib.stream_handle())
    torch.cuda.synchronize()
    # no transcendental and the reference's order: the same bits
    np.testing.assert_array_equal(out.cpu().numpy(), ref)


def test_merge_all_mirror_equals_sequential_oracle():
    import torch
    from oracle import c_oracle as CO
    from oracle import emc_oracle as O
    from refspi_amd import _lib
    size = 24
    fr = _frames(n=300, size=size, seed=4)
    rng = np.random.default_rng(5)
    pr = rng.random(fr.num_data)
    view0 = rng.random(size * size)
    view = _d(view0.copy())
    tpr, fa = _d(pr), _frame_args(fr)
    _lib.call('rsp_merge_all', _lib.ptr(tpr), fr.num_data, *[_lib.ptr(a) for a in fa],
              _lib.ptr(view), size * size, _lib.stream_handle())
    got = view.cpu().numpy()
    # the sequential loop over frames (ones, then multi) adds each pixel's terms in the
    # order the mirror does: the same bits
    np.testing.assert_array_equal(got, CO.merge_all(pr, fr, size * size, view0))
    np.testing.assert_allclose(got - view0, O.merge_all(
        pr, fr.ones, fr.multi, fr.ones_accum, fr.multi_accum, fr.place_ones, fr.place_multi,
        fr.count_multi, size * size), rtol=1e-13, atol=1e-13)
    # a photon outside [0, npix) is refused before view is touched
    bad = _frames(n=4, size=size, seed=6)
    bad.place_ones[0] = size * size
    view = _d(view0.copy())
    tpr4, fb = _d(pr[:4]), _frame_args(bad)
    rc = _lib.load().rsp_merge_all(_lib.ptr(tpr4), 4, *[_lib.ptr(a) for a in fb],
                                   _lib.ptr(view), size * size, _lib.stream_handle())
    assert rc == -2
    np.testing.assert_array_equal(view.cpu().numpy(), view0)


@pytest.mark.parametrize('size', [24, 25])
def test_slice_merge_mirror(size):
    import torch
    from oracle import emc_oracle as O
    from refspi_amd import _lib
    rng = np.random.default_rng(7)
    view = rng.random(size * size)
    m0 = rng.random(size * size)
    w0 = rng.random(size * size)
    tv = _d(view)
    for ang in (0., 0.4, np.pi / 2, 2.7):
        m, w = _d(m0.copy()), _d(w0.copy())
        _lib.call('rsp_slice_merge', _lib.ptr(tv), ang, size, _lib.ptr(m), _lib.ptr(w),
                  _lib.stream_handle())
        torch.cuda.synchronize()
        rm, rw = m0.copy(), w0.copy()
        O.slice_merge(view, ang, size, rm, rw)
        np.testing.assert_allclose(m.cpu().numpy(), rm, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(w.cpu().numpy(), rw, rtol=1e-14, atol=1e-15)


def test_get_f_dt_mirror():
    import torch
    from oracle import emc_oracle as O
    from refspi_amd import _lib
    rng = np.random.default_rng(8)
    nd, nc = 37, 1089
    fobj = rng.normal(size=nc) + 1j * rng.normal(size=nc)
    fref = rng.normal(size=nd) + 1j * rng.normal(size=nd)
    out = torch.empty(nd * nc, dtype=torch.complex128, device='cuda')
    tf, tr = _d(fobj), _d(fref)
    _lib.call('rsp_get_f_dt', _lib.ptr(tf), _lib.ptr(tr), nd, nc, _lib.ptr(out),
              _lib.stream_handle())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy().reshape(nd, nc), O.get_f_dt(fobj, fref))
