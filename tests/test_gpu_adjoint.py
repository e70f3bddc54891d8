"""GPU adjoint (rsp_rot_adjoint behind refspi_amd.adjoint.RotLoss) against the FP64 NumPy
gradient of oracle/rot_oracle.py and against central differences of the GPU loss.

Tolerances:
* vs the FP64 oracle gradient (same inputs, FP64 forward): 1e-9 relative to the largest
  gradient entry (the two sum in different orders; the loss is a sum over thousands of
  terms of both signs).
* central differences at step h = 1e-3, FP64 forward, 185 x 185, 20,000 frames, 360
  rotations, errors relative to max(|g|, 1e-3 * rms of the sampled gradients): frame
  parameters 1e-3 per entry (3-point difference); object pixels 1e-3 per entry for the
  5-point central difference at step h (nodes +-h, +-2h, truncation O(h^4)), and 1e-3
  norm-wise for the 3-point one D(h) (the random model puts |O + R| near 0 at a few photon
  pixels, where D(h)'s own O(h^2) truncation can exceed 1e-3; there D(2h) is about 4x
  further off, which the test checks).
"""
import math
import sys
import time

import numpy as np
import pytest

from conftest import SMALL

pytestmark = pytest.mark.gpu


def _small_setup(small_frames, small_model, R=12, seed=4):
    from oracle import emc_oracle as E
    from oracle import rot_oracle as O
    size = SMALL['size']
    fr = E.SparseFramesNP(small_frames.ones, small_frames.multi, small_frames.place_ones,
                          small_frames.place_multi, small_frames.count_multi,
                          small_frames.num_pix)
    pix = O.scored_pixels(size)
    K = fr.dense()[:, pix]
    rng = np.random.default_rng(seed)
    n = fr.num_data
    sx, sy = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    dia, sc = rng.uniform(7, 9, n), rng.uniform(0.6, 1.4, n)
    return dict(size=size, K=K, model=0.05 * small_model, sx=sx, sy=sy, dia=dia, sc=sc, R=R)


def _plan(frames, size, R, mode='fp64', rel=1.3, **kw):
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    return RotEM(Dataset(frames=frames), size, R, rel_scale=rel, mode=mode, **kw)


def test_adjoint_matches_fp64_oracle(small_frames, small_model):
    from oracle import rot_oracle as O
    s = _small_setup(small_frames, small_model)
    ref = O.loss_grad(s['model'], s['size'], s['R'], s['K'], s['sx'], s['sy'], s['dia'],
                      s['sc'], 1.3)
    plan = _plan(small_frames, s['size'], s['R'])
    got = plan.loss_grad(s['model'], s['sx'], s['sy'], s['dia'], s['sc'])
    assert got[0] == pytest.approx(ref[0], rel=1e-11)
    for a, b in zip(got[1:], ref[1:]):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())
    # pixels no rotated sample touches have exactly zero gradient on both sides
    assert np.array_equal(got[1] == 0, ref[1] == 0)


def test_adjoint_chunking_is_invisible(small_frames, small_model):
    s = _small_setup(small_frames, small_model)
    a = _plan(small_frames, s['size'], s['R']).loss_grad(s['model'], s['sx'], s['sy'],
                                                         s['dia'], s['sc'])
    b = _plan(small_frames, s['size'], s['R'], chunk=40).loss_grad(s['model'], s['sx'], s['sy'],
                                                                   s['dia'], s['sc'])
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-12 * np.abs(y).max())


def test_autograd_function(small_frames, small_model):
    import torch
    from refspi_amd.adjoint import rot_loss
    s = _small_setup(small_frames, small_model)
    plan = _plan(small_frames, s['size'], s['R'])
    ref = plan.loss_grad(s['model'], s['sx'], s['sy'], s['dia'], s['sc'])
    dev = 'cuda'
    m = torch.view_as_real(torch.as_tensor(s['model'], device=dev)).to(torch.float32)
    m.requires_grad_(True)
    sx, sy, dia = (torch.tensor(s[k], dtype=torch.float32, device=dev, requires_grad=True)
                   for k in ('sx', 'sy', 'dia'))
    loss = rot_loss(plan, m, sx, sy, dia, torch.as_tensor(s['sc'], device=dev))
    (2. * loss).backward()
    assert m.grad.dtype == torch.float32 and m.grad.shape == m.shape
    # FP32 inputs round the model; compare with the FP64 adjoint at the rounded inputs
    mr = torch.view_as_complex(m.detach().double()).cpu().numpy()
    r32 = plan.loss_grad(mr, sx.detach().double().cpu().numpy(),
                         sy.detach().double().cpu().numpy(),
                         dia.detach().double().cpu().numpy(), s['sc'])
    assert float(loss.detach()) == pytest.approx(r32[0], rel=1e-12)
    g = m.grad.double().cpu().numpy()
    np.testing.assert_allclose(g[:, 0] + 1j * g[:, 1], 2 * r32[1], rtol=1e-6,
                               atol=1e-6 * np.abs(r32[1]).max())
    for t, gr in zip((sx, sy, dia), r32[2:]):
        np.testing.assert_allclose(t.grad.double().cpu().numpy(), 2 * gr, rtol=1e-6,
                                   atol=1e-6 * np.abs(gr).max())
    assert abs(r32[0] - ref[0]) < 1e-4 * abs(ref[0])


def _cfg4(num_data):
    from refspi_amd import synth
    cfg = synth.SynthConfig(size=185, num_data=num_data, mean_count=300., n_blobs=120,
                            disk_radius=30., beamstop=4., seed=3)
    fr = synth.make_frames(cfg)
    return cfg, fr


def test_adjoint_vs_oracle_at_185():
    """185 x 185, 360 rotations, 4 frames: GPU adjoint vs the dense FP64 NumPy gradient."""
    from oracle import emc_oracle as E
    from oracle import rot_oracle as O
    cfg, fr = _cfg4(4)
    size, R = 185, 360
    t = fr.truth
    rng = np.random.default_rng(11)
    model = 0.3 * (rng.standard_normal(size * size) + 1j * rng.standard_normal(size * size))
    ofr = E.SparseFramesNP(fr.ones, fr.multi, fr.place_ones, fr.place_multi, fr.count_multi,
                           fr.num_pix)
    pix = O.scored_pixels(size)
    K = ofr.dense()[:, pix]
    sx, sy, dia = t['shifts'][:, 0], t['shifts'][:, 1], t['diameters']
    ref = O.loss_grad(model, size, R, K, sx, sy, dia, None, 1.)
    plan = _plan(fr, size, R, rel=1.)
    got = plan.loss_grad(model, sx, sy, dia)
    assert got[0] == pytest.approx(ref[0], rel=1e-11)
    for a, b in zip(got[1:], ref[1:]):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())


def test_adjoint_central_differences_185_20k():
    """185 x 185, 20,000 frames, 360 rotations: central differences (step 1e-3) of the GPU
    FP64 loss on 64 object pixels (32 real, 32 imaginary parts) and on the shift and
    diameter of 64 frames (perturbed together: each frame's parameters enter only its own
    term of the loss)."""
    import torch
    cfg, fr = _cfg4(20000)
    size, R = 185, 360
    t = fr.truth
    rng = np.random.default_rng(12)
    model = 0.3 * (rng.standard_normal(size * size) + 1j * rng.standard_normal(size * size))
    N = fr.num_data
    sx = t['shifts'][:, 0] + rng.normal(0, 0.3, N)
    sy = t['shifts'][:, 1] + rng.normal(0, 0.3, N)
    dia = t['diameters'] + rng.normal(0, 0.2, N)
    plan = _plan(fr, size, R, rel=1.)
    dev = 'cuda'
    dm = torch.as_tensor(model, device=dev)
    dsx, dsy, ddia = (torch.as_tensor(v, device=dev) for v in (sx, sy, dia))
    t0 = time.time()
    loss, gm, gsx, gsy, gdia = plan.loss_grad(dm, dsx, dsy, ddia)
    torch.cuda.synchronize()
    print('loss + gradient: %.3f s, loss %.6e' % (time.time() - t0, loss), file=sys.stderr)
    h = 1e-3

    def logz(m, a, b, c):
        return plan.estep_device(m, a, b, c, want_ll=False, want_prob=False)[2]

    def total(m):
        return -float((logz(m, dsx, dsy, ddia) - math.log(R)).sum())

    # object pixels: inside the scored annulus' reach, half real and half imaginary parts
    x, y = np.arange(size * size) // size, np.arange(size * size) % size
    c = size // 2
    rad = np.hypot(x - c, y - c)
    cand = np.nonzero((rad > size // 8 + 2) & (rad < c - 3))[0]
    pix = rng.choice(cand, 64, replace=False)
    fd, fdw, an = [], [], []
    t0 = time.time()
    for i, p in enumerate(pix):
        e = torch.zeros(size * size, dtype=torch.complex128, device=dev)
        e[p] = h if i < 32 else 1j * h
        f1p, f1m = total(dm + e), total(dm - e)
        f2p, f2m = total(dm + 2 * e), total(dm - 2 * e)
        fd.append((f1p - f1m) / (2 * h))              # 3-point central, step h
        fdw.append((f2p - f2m) / (4 * h))             # 3-point central, step 2h
        an.append(gm[p].real if i < 32 else gm[p].imag)
    fd, fdw, an = np.array(fd), np.array(fdw), np.array(an)
    # the 5-point central difference at step h (nodes +-h, +-2h; truncation O(h^4)):
    # (8 (f(h) - f(-h)) - (f(2h) - f(-2h))) / 12h
    fd5 = (4. * fd - fdw) / 3.
    print('model FD: %.1f s' % (time.time() - t0), file=sys.stderr)
    floor = 1e-3 * np.sqrt(np.mean(an ** 2))
    rel = lambda v: np.abs(v - an) / np.maximum(np.abs(an), floor)
    err_m, err_w, err_5 = rel(fd), rel(fdw), rel(fd5)
    norm_err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    print('model: 5-point step 1e-3 max rel err %.3e, median %.3e; 3-point step 1e-3 max '
          '%.3e, norm-wise %.3e, entries > 1e-3: %d of 64; 3-point step 2e-3 max %.3e; '
          '|g| rms %.3e' % (err_5.max(), np.median(err_5), err_m.max(), norm_err,
                            (err_m > 1e-3).sum(), err_w.max(), np.sqrt(np.mean(an ** 2))),
          file=sys.stderr)
    frames = rng.choice(N, 64, replace=False)
    sel = torch.zeros(N, dtype=torch.float64, device=dev)
    sel[torch.as_tensor(frames, device=dev)] = h
    errs = {}
    for name, g, which in (('sx', gsx, 0), ('sy', gsy, 1), ('dia', gdia, 2)):
        args_p = [dsx, dsy, ddia]
        args_m = [dsx, dsy, ddia]
        args_p[which] = args_p[which] + sel
        args_m[which] = args_m[which] - sel
        lp = logz(dm, *args_p)
        lm = logz(dm, *args_m)
        fdf = (-(lp - lm) / (2 * h)).cpu().numpy()[frames]
        ga = g[frames]
        fl = 1e-3 * np.sqrt(np.mean(ga ** 2))
        errs[name] = np.abs(fdf - ga) / np.maximum(np.abs(ga), fl)
        print('%s: max rel err %.3e, median %.3e' % (name, errs[name].max(),
                                                      np.median(errs[name])), file=sys.stderr)
    # every entry within 1e-3 for the central difference of fourth order at step 1e-3
    assert err_5.max() < 1e-3
    assert norm_err < 1e-3
    # where the 3-point difference at step 1e-3 misses 1e-3, its error is the step's
    # truncation: it grows about 4x at step 2e-3 (O(h^2))
    bad = err_m > 1e-3
    assert np.all(err_w[bad] > 2.5 * err_m[bad])
    for name, e in errs.items():
        assert e.max() < 1e-3, name


def test_split_frame_gradients_close_to_fp64():
    """Option 6: the frame-gradient GEMM Obar = P O in split BF16 on MFMA gives the frame
    gradients of the FP64 adjoint to a small norm-wise relative error (185^2, 2048 frames,
    360 rotations, parameters off the truth); the model gradient does not use Obar and is
    unchanged bit for bit."""
    import torch
    cfg, fr = _cfg4(2048)
    size, R = 185, 360
    t = fr.truth
    rng = np.random.default_rng(21)
    model = 0.3 * (rng.standard_normal(size * size) + 1j * rng.standard_normal(size * size))
    N = fr.num_data
    sx = t['shifts'][:, 0] + rng.normal(0, 0.3, N)
    sy = t['shifts'][:, 1] + rng.normal(0, 0.3, N)
    dia = t['diameters'] + rng.normal(0, 0.2, N)
    plan = _plan(fr, size, R, rel=1.)
    dev = 'cuda'
    dm = torch.as_tensor(model, device=dev)
    dsx, dsy, ddia = (torch.as_tensor(v, device=dev) for v in (sx, sy, dia))
    _, prob, _ = plan.estep_device(dm, dsx, dsy, ddia, want_ll=False)
    ref = [v.cpu().numpy() for v in plan.adjoint_device(dm, dsx, dsy, ddia, None, prob)]
    plan.set_frame_grad('split')
    got = [v.cpu().numpy() for v in plan.adjoint_device(dm, dsx, dsy, ddia, None, prob)]
    np.testing.assert_array_equal(got[0], ref[0])
    for a, b in zip(got[1:], ref[1:]):
        rel_err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print('frame gradient relative error %.3g' % rel_err, file=sys.stderr)
        assert rel_err < 1e-4   # measured 1.4e-6


def test_split_model_gradient_gemm_close_to_fp64():
    """Option 7: the model-gradient GEMM GT = Q^T B in split BF16 on MFMA (each sub-chunk's
    frames accumulated in FP32, sub-chunks added in FP64) gives the FP64 adjoint's model
    gradient to a small norm-wise relative error (185^2, 5000 frames over two sub-chunks,
    360 rotations); the frame gradients do not use GT and are unchanged bit for bit."""
    import torch
    cfg, fr = _cfg4(5000)
    size, R = 185, 360
    t = fr.truth
    rng = np.random.default_rng(22)
    model = 0.3 * (rng.standard_normal(size * size) + 1j * rng.standard_normal(size * size))
    N = fr.num_data
    sx = t['shifts'][:, 0] + rng.normal(0, 0.3, N)
    sy = t['shifts'][:, 1] + rng.normal(0, 0.3, N)
    dia = t['diameters'] + rng.normal(0, 0.2, N)
    plan = _plan(fr, size, R, rel=1.)
    dev = 'cuda'
    dm = torch.as_tensor(model, device=dev)
    dsx, dsy, ddia = (torch.as_tensor(v, device=dev) for v in (sx, sy, dia))
    _, prob, _ = plan.estep_device(dm, dsx, dsy, ddia, want_ll=False)
    ref = [v.cpu().numpy() for v in plan.adjoint_device(dm, dsx, dsy, ddia, None, prob)]
    plan.set_model_grad('split')
    got = [v.cpu().numpy() for v in plan.adjoint_device(dm, dsx, dsy, ddia, None, prob)]
    for a, b in zip(got[1:], ref[1:]):
        np.testing.assert_array_equal(a, b)
    rel_err = np.linalg.norm(got[0] - ref[0]) / np.linalg.norm(ref[0])
    worst = np.abs(got[0] - ref[0]).max() / np.abs(ref[0]).max()
    print('model gradient relative error %.3g (largest entry error / largest entry %.3g)' %
          (rel_err, worst), file=sys.stderr)
    assert rel_err < 1e-4 and worst < 1e-4
    plan.set_model_grad('fp64')
    back = plan.adjoint_device(dm, dsx, dsy, ddia, None, prob)[0].cpu().numpy()
    np.testing.assert_array_equal(back, ref[0])
