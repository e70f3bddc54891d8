"""Benchmark: one EMC+MaxLP iteration per step on one MI355X.

Workload (default): 185x185 q-grid, 100,000 synthetic frames at 300 photons/frame
(120 Gaussian blobs in a 30 px disk, 370 px box, beamstop 4 px, sphere reference with
diameter N(8, 0.4) px and shift U(-4, 4) px carrying 5x the object photons, uniform
in-plane rotation, Poisson sampling), joint search over 17x17x21 = 6069
(shift_x, shift_y, diameter) states x 360 rotations, then the MaxLP object update.

Metric: frame x rotation x state log-likelihood evaluations per second of whole
iterations, N * R * S / t_iteration.

Prints one JSON line with the metric, the live roofline of the dominant kernel (the
sparse log-likelihood gather, LDS-bound) and of MaxLP's pattern search (FP64 VALU), the FRC
of the reconstruction against the generator's object, and a CPU baseline (the oracle's C
restatement timed on a bounded sample of the same workload on this host, extrapolated).
"""
import argparse
import json
import os
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, 'refspi-maxlp_amd')]

HBM_PEAK_GBS = 8000.
LDS_PEAK_GBS = 256 * 128 * 2.4   # 256 CUs x 128 B/clk x 2.4 GHz


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_dataset(args):
    import torch
    from refspi_amd import synth
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': args.frames,
                               'seed': args.seed})
    t = time.time()
    d = synth.make_frames_torch(cfg, device='cuda')
    torch.cuda.synchronize()
    fr = types.SimpleNamespace(
        num_pix=d['num_pix'], bg=None,
        **{k: d[k].cpu().numpy().astype('i4') for k in
           ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')})
    log('generated %d frames in %.1f s (%.1f photons/frame)' %
        (args.frames, time.time() - t,
         (fr.ones.sum() + fr.count_multi.sum()) / args.frames))
    return cfg, fr


def host_facts():
    """The host the CPU baseline ran on."""
    model = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(cpu_model=model, affinity_cores=len(os.sched_getaffinity(0)),
                cpu_count=os.cpu_count(),
                env={k: os.environ.get(k) for k in ('OMP_NUM_THREADS', 'OPENBLAS_NUM_THREADS',
                                                    'MKL_NUM_THREADS')})


def _sub_frames(fr, n):
    from oracle import emc_oracle as O
    no = int(fr.ones[:n].sum())
    nm = int(fr.multi[:n].sum())
    return O.SparseFramesNP(fr.ones[:n], fr.multi[:n], fr.place_ones[:no], fr.place_multi[:nm],
                            fr.count_multi[:nm], fr.num_pix)


def cpu_baseline(fr, size, num_rot, grid, n_frames, n_states, n_pixels):
    """The oracle's C restatement (FP64, the reference's operation order, OpenMP threads) on
    a bounded sample of the same workload: the E-step of n_frames frames x n_states states
    (spread over the grid) x num_rot rotations, and MaxLP phase 1 / phase 3 on n_pixels
    ring / mask pixels of those frames.  A whole iteration of the full workload is
    extrapolated linearly (E-step in frames x states, MaxLP in frames x pixels)."""
    from oracle import c_oracle as CO
    from oracle import emc_oracle as O
    CO.load()
    nth = CO.threads()
    N = fr.ones.shape[0]
    ofr = _sub_frames(fr, n_frames)
    sx, sy, dia, ns = O.state_grid(*grid)
    S = len(sx)
    sel = np.linspace(0, S - 1, n_states).astype(int)
    rng = np.random.default_rng(0)
    model = rng.random(size * size) + 1j * rng.random(size * size)
    t = time.time()
    rmax, _, _, _, _ = CO.estep_top2(model, size, (sx[sel], sy[sel], dia[sel]), num_rot, ofr,
                                     nthreads=nth)
    t_e = time.time() - t
    # MaxLP at the sample's answer
    st = rmax // num_rot
    vals = (sx[sel][st], sy[sel][st], dia[sel][st], (rmax % num_rot) * np.pi / num_rot)
    mc = CO.MaxLPC(ofr, size)
    mc.set_frames(*vals)
    ring = np.nonzero(mc.rmask)[0]
    mask = np.nonzero(mc.mask)[0]
    pr = ring[np.linspace(0, len(ring) - 1, n_pixels).astype(int)]
    pm = mask[np.linspace(0, len(mask) - 1, n_pixels).astype(int)]
    t = time.time()
    f1, resc, _, _ = mc.pixels(model[pr], pr, None, nth)
    t_m1 = time.time() - t
    t = time.time()
    mc.pixels(model[pm], pm, float(np.mean(resc)), nth)
    t_m3 = time.time() - t
    est = t_e * (N / n_frames) * (S / n_states)
    msp = (t_m1 * len(ring) + t_m3 * len(mask)) / n_pixels * (N / n_frames)
    t_iter = est + msp
    return dict(value=N * num_rot * S / t_iter, unit='frame*rotation*state evals/s', cores=nth,
                kind='port', extrapolated=True,
                seconds_per_iteration_extrapolated=t_iter,
                estep_evals_per_s=n_frames * n_states * num_rot / t_e,
                sample='oracle C restatement (FP64, OpenMP, %d threads): E-step of %d frames x '
                       '%d states x %d rotations of the same workload (%.1f s) and MaxLP on %d '
                       'ring + %d mask pixels of those frames (%.1f s); a full iteration '
                       '(%d frames, %d states, %d ring + %d mask pixels) extrapolated '
                       'linearly: E-step %.0f s + MaxLP %.0f s' %
                       (nth, n_frames, n_states, num_rot, t_e, n_pixels, n_pixels,
                        t_m1 + t_m3, N, S, len(ring), len(mask), est, msp),
                host=host_facts())


def cpu_config1(args):
    """The reference's CPU configuration in full with the oracle's C restatement: 64^2,
    2000 frames (150 photons/frame), 6069 states x 90 rotations, 3 EMC+MaxLP iterations
    from a seeded random model."""
    from oracle import c_oracle as CO
    from oracle import emc_oracle as O
    from refspi_amd import synth
    CO.load()
    nth = CO.threads()
    cfg = synth.SynthConfig(**{**synth.CONFIG_CPU64.__dict__, 'seed': args.seed})
    fr = synth.make_frames(cfg)
    ofr = O.SparseFramesNP(fr.ones, fr.multi, fr.place_ones, fr.place_multi, fr.count_multi,
                           fr.num_pix)
    sx, sy, dia, ns = O.state_grid((-4., 4., 17), (-4., 4., 17), (7., 9., 21))
    R = 90
    rng = np.random.default_rng(args.seed)
    model = rng.random(cfg.size ** 2) + 1j * rng.random(cfg.size ** 2)
    rng0 = model.copy()
    mc = CO.MaxLPC(ofr, cfg.size)
    its = []
    for it in range(3):
        t = time.time()
        rmax, _, _, _, _ = CO.estep_top2(model, cfg.size, (sx, sy, dia), R, ofr, nthreads=nth)
        t1 = time.time()
        vals = O.unravel_rmax(rmax, ns, R, sx, sy, dia)
        model, rescale, _ = mc.run_phaser(model, *vals, nthreads=nth)
        t2 = time.time()
        its.append(dict(estep_s=t1 - t, maxlp_s=t2 - t1, rescale=rescale))
        log('cpu config 1 iteration %d: E-step %.1f s, MaxLP %.1f s' % (it + 1, t1 - t, t2 - t1))
    tot = sum(i['estep_s'] + i['maxlp_s'] for i in its)
    out = dict(workload='64x64 q-grid, 2000 frames, 6069 states x 90 rotations, 3 EMC+MaxLP '
                        'iterations (the reference\'s CPU configuration), oracle C '
                        'restatement FP64, %d threads' % nth,
               iterations=its, seconds_per_iteration=tot / 3,
               value=cfg.num_data * R * len(sx) / (tot / 3),
               unit='frame*rotation*state evals/s', host=host_facts())
    # the same three iterations on the GPU (same frames, grid and start model)
    import torch
    if torch.cuda.is_available():
        from refspi_amd.dset import Dataset
        from refspi_amd.emc import EMC
        emc = EMC(dataset=Dataset(frames=fr), size=cfg.size, num_rot=R,
                  shiftx=(-4., 4., 17), shifty=(-4., 4., 17), sphere_dia=(7., 9., 21),
                  model=rng0, save=False)
        gi = []
        for it in range(3):
            torch.cuda.synchronize()
            t = time.time()
            emc.run_iteration()
            torch.cuda.synchronize()
            gi.append(time.time() - t)
        out['gpu_seconds_per_iteration'] = gi
        out['gpu_note'] = ('the same 3 iterations through the HIP path (iteration 1 includes '
                           'the plans\' set-up)')
    return out


def frc_report(model, cfg):
    """FRC (calc_frc.py:77-117) of a reconstruction against the generator's object
    transform, with the best in-plane rotation; mean over rings 4 .. size//2."""
    import contextlib
    from refspi_amd import frc, synth
    truth = (synth.model_scales(cfg)[0] * synth.object_ft(cfg)).reshape(cfg.size, cfg.size)
    f = frc.FRC(np.asarray(model).reshape(cfg.size, cfg.size), truth)
    with contextlib.redirect_stdout(sys.stderr):
        r, c = f.calc_rot(num_rot=360, radmin=4)
    cen = cfg.size // 2
    return dict(mean_4_to_edge=float(c[4:cen].mean()), best_angle_deg=float(f.best_angle),
                ring_frc=[round(float(v), 4) for v in c[:cen]],
                note='FRC of the model after the run against the generator\'s object '
                     'transform (calc_frc.py:77-117, best of 360 rotations)')


def _frames_from_torch(d):
    return types.SimpleNamespace(
        num_pix=d['num_pix'], bg=None,
        **{k: d[k].cpu().numpy().astype('i4') for k in
           ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')})


def split_leg(args):
    """Rotation-form E-step (cross-term GEMM + sparse log term + FP64 log-sum-exp) at
    256x256, 200,000 frames (log-normal fluence, sigma 0.5, mean 400 photons), 720
    rotations: split-BF16 (MFMA) vs plain FP32 (rocBLAS) vs FP64, the LL error of each
    against the FP64 mode, and the time of each."""
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    cfg = synth.SynthConfig(**{**synth.CONFIG_256.__dict__, 'num_data': args.split_frames,
                               'seed': args.seed})
    R = args.split_rot
    t = time.time()
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = _frames_from_torch(d)
    alpha, rel = synth.model_scales(cfg)
    model = torch.as_tensor(alpha * synth.object_ft(cfg), device='cuda')
    sx, sy = d['shifts'][:, 0].contiguous(), d['shifts'][:, 1].contiguous()
    dia, sc = d['diameters'], d['scale']
    plan = RotEM(Dataset(frames=fr), cfg.size, R, rel_scale=rel, mode='split', full_turn=True)
    plan.set_timing(True)
    info = plan.info()
    log('split leg: %d frames generated, plan built in %.1f s %s' %
        (cfg.num_data, time.time() - t, info))
    N, K = cfg.num_data, info['scored_pixels']
    out = dict(workload='rotation-form E-step, %dx%d q-grid, %d frames, log-normal fluence '
                        '(sigma %.1f, mean %d photons), %d rotations (full turn), %d scored '
                        'pixels' %
                        (cfg.size, cfg.size, N, cfg.fluence_sigma, cfg.mean_count, R, K))
    # Re(O conj(R)) needs 4 real flops per (r, d, pixel): a real GEMM of M = R, N frames,
    # K = 2 x scored pixels (re/im interleaved), 2 M N K flops
    useful = 4. * R * N * K
    res = {}
    # split: the reference rows generated inside the GEMM kernel (the default);
    # split_stored: written to HBM as bf16 hi/lo rows and read by the GEMM
    for mode in ('fp64', 'fp32', 'split_stored', 'split'):
        plan.set_mode(mode.split('_')[0])
        plan.set_ref_rows('stored' if mode == 'split_stored' else 'generated')
        plan.estep_device(model, sx, sy, dia, sc, want_ll=False, want_prob=False)  # warm-up
        torch.cuda.synchronize()
        best = None
        for _ in range(args.split_reps if mode != 'fp64' else 1):
            t = time.time()
            ll, _, logz = plan.estep_device(model, sx, sy, dia, sc, want_ll=True,
                                            want_prob=False)
            torch.cuda.synchronize()
            el = time.time() - t
            tm = plan.timing()
            if best is None or el < best[0]:
                best = (el, tm)
        el, tm = best
        res[mode] = ll
        gemm_s = tm['gemm_ms'] / 1e3
        out[mode] = dict(estep_s=el, gemm_ms=tm['gemm_ms'], slices_ms=tm['slices_ms'],
                         ref_rows_ms=tm['ref_rows_ms'], logterm_lse_ms=tm['logterm_lse_ms'],
                         gemm_tflops=useful / gemm_s / 1e12,
                         frame_rotation_evals_per_s=N * R / el)
        log('split leg %s: %.3f s %s' % (mode, el, tm))
    ll64 = res['fp64']
    for mode in ('fp32', 'split_stored', 'split'):
        err = ((res[mode] - ll64).abs() / ll64.abs()).max().item()
        out[mode]['ll_max_rel_err_vs_fp64'] = err
    out['split_speedup_vs_fp32'] = dict(estep=out['fp32']['estep_s'] / out['split']['estep_s'],
                                        gemm=out['fp32']['gemm_ms'] / out['split']['gemm_ms'])
    out['split_stored_speedup_vs_fp32'] = dict(
        estep=out['fp32']['estep_s'] / out['split_stored']['estep_s'],
        gemm=out['fp32']['gemm_ms'] / out['split_stored']['gemm_ms'])
    out['split']['ref_rows'] = ('generated in the GEMM kernel (k_gemm_split_gen); ref_rows_ms '
                                'is the |R_d|^2 kernel')
    out['gemm'] = dict(M=R, N=N, K=2 * K, flops=useful,
                       note='cross term Re sum_p O_r(p) conj(R_d(p)) as a real GEMM over '
                            're/im-interleaved rows (K = 2 x scored pixels): 2 M N K flops; '
                            'split-BF16 executes 3 BF16 products per term (3x these flops on '
                            'MFMA)')
    for mode in ('split', 'split_stored'):
        out[mode]['mfma_bf16_tflops_executed'] = 3. * useful / (out[mode]['gemm_ms'] / 1e3) / 1e12
    out['peaks_tflops'] = dict(bf16_dense=2500., fp32_vector=157.3, fp64=78.6)
    del plan, res
    torch.cuda.empty_cache()
    return out


def adjoint_leg(args):
    """Loss + gradient of the rotation form (rsp_rot_estep then rsp_rot_adjoint) at
    185x185, 20,000 frames, 360 rotations, gradients with respect to the complex model and
    every frame's shift and diameter."""
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.rotem import RotEM
    cfg = synth.SynthConfig(**{**synth.CONFIG_185.__dict__, 'num_data': args.adj_frames,
                               'seed': args.seed})
    R = 360
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = _frames_from_torch(d)
    alpha, rel = synth.model_scales(cfg)
    model = torch.as_tensor(alpha * synth.object_ft(cfg), device='cuda')
    sx, sy = d['shifts'][:, 0].contiguous(), d['shifts'][:, 1].contiguous()
    dia, sc = d['diameters'], d['scale']
    plan = RotEM(Dataset(frames=fr), cfg.size, R, rel_scale=rel, mode='fp64', full_turn=True)
    out = dict(workload='loss + gradient (model, shifts, diameters), %dx%d q-grid, %d frames, '
                        '%d rotations' % (cfg.size, cfg.size, cfg.num_data, R))
    for mode in ('fp64', 'split', 'split_gemms'):
        # split_gemms: split-BF16 forward and the adjoint's two dense products (Obar = P A,
        # GT = Q^T B) in split BF16 on MFMA as well
        plan.set_mode('split' if mode == 'split_gemms' else mode)
        plan.set_frame_grad('split' if mode == 'split_gemms' else 'fp64')
        plan.set_model_grad('split' if mode == 'split_gemms' else 'fp64')
        plan.loss_grad(model, sx, sy, dia, sc)       # warm-up (allocations)
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            t = time.time()
            _, prob, logz = plan.estep_device(model, sx, sy, dia, sc, want_ll=False)
            torch.cuda.synchronize()
            t1 = time.time()
            plan.adjoint_device(model, sx, sy, dia, sc, prob)
            torch.cuda.synchronize()
            ts.append((time.time() - t, t1 - t, time.time() - t1))
        best = min(ts)
        out['forward_' + mode] = dict(total_s=best[0], forward_s=best[1], adjoint_s=best[2])
        log('adjoint leg (%s forward): %s' % (mode, best))
    del plan
    torch.cuda.empty_cache()
    return out


def stream_leg(args):
    """Two object classes (60/40, independent blob seeds) at 256x256, 500,000 frames
    (log-normal fluence, sigma 0.7, mean 400 photons) packed in pinned host memory and
    streamed in 50,000-frame chunks; 720 rotations (full turn) x 2 classes; split-BF16
    cross term.  One E-step + merge pass, then adjoint-driven (shift, diameter) refinement
    passes; the refinement is judged on a subset by the reference's grid search; peak HBM
    is sampled after every chunk."""
    import torch
    from refspi_amd import synth
    from refspi_amd.dset import Dataset
    from refspi_amd.refine import grid_search, refine
    from refspi_amd.rotem import PASS_MERGE, StreamedRotEM
    free0, total = torch.cuda.mem_get_info()
    peak = [total - free0]

    def sample(*_):
        f, _t = torch.cuda.mem_get_info()
        peak[0] = max(peak[0], total - f)

    cfg = synth.SynthConfig(**{**synth.CONFIG_256.__dict__, 'num_data': args.stream_frames,
                               'fluence_sigma': 0.7, 'class_weights': (0.6, 0.4),
                               'seed': args.seed})
    R, C = 720, 2
    t = time.time()
    d = synth.make_frames_torch(cfg, device='cuda')
    fr = _frames_from_torch(d)
    truth = {k: d[k].cpu().numpy() for k in ('shifts', 'diameters', 'scale', 'classes')}
    del d
    torch.cuda.empty_cache()
    rel = synth.model_scales(cfg)[1]
    models = torch.as_tensor(np.stack([synth.model_scales(cfg, c)[0] * synth.object_ft(cfg, c)
                                       for c in range(C)]), device='cuda')
    t_gen = time.time() - t
    t = time.time()
    plan = StreamedRotEM(Dataset(frames=fr), cfg.size, R, rel_scale=rel, mode='split',
                         classes=C, full_turn=True, chunk_frames=args.stream_chunk)
    t_pack = time.time() - t
    N = cfg.num_data
    log('stream leg: %d frames generated in %.1f s, packed into %d chunks (%.2f GB pinned) in '
        '%.1f s' % (N, t_gen, len(plan.chunks), plan.host_bytes / 1e9, t_pack))
    rng = np.random.default_rng(1)
    sx0 = truth['shifts'][:, 0] + rng.normal(0, 0.25, N)
    sy0 = truth['shifts'][:, 1] + rng.normal(0, 0.25, N)
    d0 = truth['diameters'] + rng.normal(0, 0.1, N)
    sc = truth['scale']
    plan.run_pass(models, sx0, sy0, d0, sc, flags=PASS_MERGE)     # warm-up (allocations)
    torch.cuda.synchronize()
    sample()
    t = time.time()
    out = plan.run_pass(models, sx0, sy0, d0, sc, flags=PASS_MERGE, on_chunk=sample)
    torch.cuda.synchronize()
    t_pass = time.time() - t
    sample()
    cls = (out['best'].cpu().numpy() // R)
    t = time.time()
    sx, sy, dia, lz, hist = refine(plan, models, sx0, sy0, d0, sc, steps=args.refine_steps,
                                   on_step=lambda i, h: (sample(), log('refine step', i, h)))
    torch.cuda.synchronize()
    t_ref = time.time() - t
    sample()
    # log Z at the generator's true parameters: refinement maximises log Z, so it should end
    # at or above the truth wherever the likelihood peaks near it
    lz_truth = plan.run_pass(models, truth['shifts'][:, 0], truth['shifts'][:, 1],
                             truth['diameters'], sc, want_best=False)['logz']
    lzr = lz.cpu().numpy()
    lzt = lz_truth.cpu().numpy()
    res = dict(workload='two classes (60/40) at 256x256, %d frames (log-normal fluence sigma '
                        '0.7, mean 400) streamed from pinned host memory in %d-frame chunks, '
                        '%d rotations (full turn) x %d classes, split-BF16 cross term' %
                        (N, args.stream_chunk, R, C),
               estep_merge_pass_s=t_pass,
               frame_hypothesis_evals_per_s=N * R * C / t_pass,
               refine_steps=args.refine_steps, refine_s=t_ref,
               refine_s_per_step=t_ref / max(1, args.refine_steps),
               refine_history=hist,
               class_of_best_hypothesis_correct=float((cls == truth['classes']).mean()),
               refined_logz_ge_truth_logz=float((lzr >= lzt - 1e-9 * np.abs(lzt)).mean()),
               median_logz_gain_over_truth=float(np.median(lzr - lzt)),
               peak_hbm_gb=peak[0] / 1e9, hbm_total_gb=total / 1e9,
               pinned_host_gb=plan.host_bytes / 1e9, chunks=len(plan.chunks))
    del plan
    torch.cuda.empty_cache()
    # the judge: the reference's grid over (shift, diameter) on a subset of frames
    n = args.judge_frames
    sub = types.SimpleNamespace(num_pix=fr.num_pix, bg=None)
    oa = np.concatenate([[0], np.cumsum(fr.ones, dtype='i8')])
    ma = np.concatenate([[0], np.cumsum(fr.multi, dtype='i8')])
    sub.ones, sub.multi = fr.ones[:n], fr.multi[:n]
    sub.place_ones = fr.place_ones[:oa[n]]
    sub.place_multi, sub.count_multi = fr.place_multi[:ma[n]], fr.count_multi[:ma[n]]
    t = time.time()
    gsx, gsy, gd, glz, _ = grid_search(sub, cfg.size, R, rel, models, scales=sc[:n],
                                       classes=C, full_turn=True, mode='split')
    t_judge = time.time() - t
    lzs = lz[:n].cpu().numpy()
    err = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2)))
    ts = truth['shifts'][:n].T

    def serr(a, b):
        # (rotation + pi, -shift) gives the same hologram: the error is taken to the nearer
        # of the two twins
        e = np.minimum(((a - b) ** 2).sum(0), ((a + b) ** 2).sum(0))
        return float(np.sqrt(np.mean(e) / 2.))

    res['judge'] = dict(
        frames=n, grid='shift -4..4 (17) x -4..4 (17), diameter 7..9 (21): 6069 states',
        seconds=t_judge,
        refined_logz_ge_grid_best=float((lzs >= glz - 1e-9 * np.abs(glz)).mean()),
        median_logz_gain_over_grid=float(np.median(lzs - glz)),
        shift_rms_err_px=dict(start=serr(np.stack([sx0[:n], sy0[:n]]), ts),
                              grid=serr(np.stack([gsx, gsy]), ts),
                              refined=serr(np.stack([sx[:n].cpu().numpy(),
                                                     sy[:n].cpu().numpy()]), ts),
                              note='per axis, to the nearer of the twins s and -s'),
        dia_rms_err_px=dict(start=err(d0[:n], truth['diameters'][:n]),
                            grid=err(gd, truth['diameters'][:n]),
                            refined=err(dia[:n].cpu().numpy(), truth['diameters'][:n])))
    log('stream leg:', res)
    return res


def main():
    # the JSON line is the only output on stdout: the library's progress prints go to stderr
    import contextlib
    with contextlib.redirect_stdout(sys.stderr):
        out = run()
    print(json.dumps(out), flush=True)


def run():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--frames', type=int, default=100_000)
    ap.add_argument('--num-rot', type=int, default=360)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--banks', type=int, default=None,
                    help='LDS bank model of the photon scheduler (0 = none, 32, 64)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-frames', type=int, default=2000)
    ap.add_argument('--cpu-states', type=int, default=256)
    ap.add_argument('--cpu-pixels', type=int, default=64)
    ap.add_argument('--stream-frames', type=int, default=500_000)
    ap.add_argument('--stream-chunk', type=int, default=50_000)
    ap.add_argument('--refine-steps', type=int, default=3)
    ap.add_argument('--judge-frames', type=int, default=512)
    ap.add_argument('--workload', choices=('all', 'joint', 'split', 'adjoint', 'stream',
                                           'cpu-config1'),
                    default='all',
                    help='joint: the EMC+MaxLP iteration (the headline value); split: the '
                         'rotation-form split-BF16 vs FP32 E-step; adjoint: loss + gradient; '
                         'cpu-config1: the 64^2 CPU configuration in full on the host cores')
    ap.add_argument('--split-frames', type=int, default=200_000)
    ap.add_argument('--split-rot', type=int, default=720)
    ap.add_argument('--split-reps', type=int, default=3)
    ap.add_argument('--adj-frames', type=int, default=20_000)
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only copy 1 GiB device-to-device (HBM counter calibration) and exit')
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device='cuda')
        b = torch.ones(1 << 28, dtype=torch.float32, device='cuda')
        torch.cuda.synchronize()
        a.copy_(b)
        torch.cuda.synchronize()
        return {'copy_bytes_read': 1 << 30, 'copy_bytes_written': 1 << 30}
    if args.gpus != 1:
        log('bench.py runs on one GPU only: this build has no multi-GPU path '
            '(asked for --gpus %d)' % args.gpus)
        sys.exit(2)
    if args.workload == 'cpu-config1':
        return {'cpu_config1': cpu_config1(args)}
    import torch
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
    if args.workload in ('split', 'adjoint', 'stream'):
        leg = dict(split=split_leg, adjoint=adjoint_leg, stream=stream_leg)[args.workload](args)
        return {args.workload: leg}
    grid = ((-4., 4., 17), (-4., 4., 17), (7., 9., 21))
    cfg, fr = make_dataset(args)
    size = cfg.size
    t = time.time()
    emc = EMC(dataset=Dataset(frames=fr), size=size, num_rot=args.num_rot, shiftx=grid[0],
              shifty=grid[1], sphere_dia=grid[2], seed=args.seed, save=False,
              bank_count=args.banks)
    emc._make_plan()
    emc.set_option(5, 1)
    log('plans built in %.1f s' % (time.time() - t))
    S = len(emc.shiftx)
    H = S * args.num_rot
    for i in range(args.warmup):
        t = time.time()
        emc.run_iteration()
        log('warmup %d: %.2f s %s %s' % (i, time.time() - t, emc.timings, emc.estep_stats()))
    torch.cuda.synchronize()
    gather_ms = tables_ms = 0.
    batches = 0
    lds_rw = 0.
    mx_products = mx_ms = 0.
    t0 = time.time()
    for i in range(args.steps):
        ts = time.time()
        emc.run_iteration()
        tm = emc.estep_timing()
        gather_ms += tm['gather_ms']
        tables_ms += tm['tables_ms']
        batches += tm['batches']
        lds_rw += H * (tm['lds_read_bytes_per_hyp'] + tm['lds_write_bytes_per_hyp'])
        mi = emc.phaser.info()
        mx_products += mi['entry_candidate_products']
        mx_ms += mi['pattern_ms']
        log('step %d: %.2f s %s gather %.1f ms tables %.1f ms' %
            (i, time.time() - ts, emc.timings, tm['gather_ms'], tm['tables_ms']))
    torch.cuda.synchronize()
    el = time.time() - t0
    t_iter = el / args.steps
    n = fr.ones.shape[0]
    value = n * args.num_rot * S / t_iter
    # dominant kernel: the sparse log-likelihood gather.  Its photons stay in VGPRs across
    # hypotheses and the tables stream through LDS, so the binding resource is LDS: the
    # algorithmic work per frame x hypothesis is one 4-byte LDS read per scored photon.
    gather_s = gather_ms / 1e3
    avg_launch_s = gather_s / max(batches, 1)       # one launch = one batch's slot classes
    pm = emc.probmask
    scored = (pm[fr.place_ones] == 0).sum() + fr.count_multi[pm[fr.place_multi] == 0].sum()
    lds_bytes_per_step = H * 4. * scored
    lds_gbs = lds_bytes_per_step * args.steps / gather_s / 1e9
    lds_exec_gbs = lds_rw / gather_s / 1e9
    # the sparse records calc_prob_all re-reads per hypothesis (4 ones + 8 multi bytes)
    rec_bytes = 4. * fr.ones.sum() + 8. * fr.multi.sum()
    rec_gbs = H * rec_bytes * args.steps / gather_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, 'profiles', 'gather_traffic_185.json')
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get('bytes_per_launch')
    hbm_gbs = traffic / avg_launch_s / 1e9 if traffic else None
    mx_tflops = 6. * mx_products / (mx_ms / 1e3) / 1e12 if mx_ms > 0 else None
    out = dict(
        metric='frame x rotation evals/s per EMC+MaxLP iteration at 185^2 (x states, joint '
               'search)',
        value=value, unit='frame*rotation*state evals/s', n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=t_iter * 1e3, higher_is_better=True, scaling='weak',
        vs_baseline=None, dtype='f32 tables + f64 accumulation/f64 MaxLP', data='synthetic',
        config=dict(workload='185x185 q-grid, %d frames, %d photons/frame, %d states x %d '
                             'rotations joint search + MaxLP' %
                             (n, cfg.mean_count, S, args.num_rot),
                    frames=n, states=S, rotations=args.num_rot, size=size,
                    estep_s=emc.timings.get('estep'), maxlp_s=emc.timings.get('maxlp'),
                    estep_stats=emc.estep_stats()),
        roofline=dict(bound='lds', kernel='k_gather', unit='GB/s',
                      achieved=lds_gbs, peak=LDS_PEAK_GBS, frac=lds_gbs / LDS_PEAK_GBS,
                      traffic=traffic, launches=batches,
                      avg_launch_ms=gather_ms / max(batches, 1),
                      what='one 4-byte LDS read per scored photon per frame x hypothesis '
                           '(%.0f scored photons per step-hypothesis), / gather time' % scored,
                      peak_what='256 CUs x 128 B/clk (32 banks x 4 B, the ds_read_b32 ceiling) '
                                'x 2.4 GHz; measured conflict-free ds_read_b32: 24.8-27 of 32 '
                                'lanes per clock.  Wider reads reach more bytes (ds_read_b64 '
                                '178, b128 233 B/clk/CU measured) but one address per lane '
                                'and clock: a gather of single 4-byte values is bound by '
                                'the 32 lanes per clock of this figure',
                      lds_hw_read_peak_gbs=256 * 256 * 2.4,
                      executed_gbs=lds_exec_gbs, executed_frac=lds_exec_gbs / LDS_PEAK_GBS,
                      executed_what='every slot read (bank-schedule zero pads included) + the '
                                    'table loads\' LDS writes',
                      hbm_measured_gbs=hbm_gbs,
                      hbm_measured_frac=hbm_gbs / HBM_PEAK_GBS if hbm_gbs else None,
                      hbm_what='traffic (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, '
                               'calibrated by a 1 GiB copy) / avg launch',
                      record_equivalent_gbs=rec_gbs,
                      record_what='sparse records (4 ones + 8 multi bytes) per frame x '
                                  'hypothesis, as calc_prob_all re-reads them, / gather time: '
                                  'a rate of useful work, not a bandwidth drawn'),
        maxlp_roofline=dict(bound='valu_fp64', kernel='k_pattern', unit='TFLOP/s',
                            achieved=mx_tflops, peak=78.6,
                            frac=mx_tflops / 78.6 if mx_tflops else None,
                            ms_per_step=mx_ms / args.steps,
                            what='6 FP64 operations per entry x candidate product (c + R_d, '
                                 '|.|^2, the running product) / pattern-search time, both '
                                 'phases; %d products per step' % (mx_products / args.steps)),
    )
    out['frc'] = frc_report(emc.model, cfg)
    out['frc']['iterations'] = args.warmup + args.steps
    if args.workload == 'all':
        del emc
        torch.cuda.empty_cache()
        out['split_precision'] = split_leg(args)
        out['adjoint'] = adjoint_leg(args)
        out['streamed_two_class'] = stream_leg(args)
    if not args.no_cpu_baseline:
        out['cpu_baseline'] = cpu_baseline(fr, size, args.num_rot, grid, args.cpu_frames,
                                           args.cpu_states, args.cpu_pixels)
    return out


if __name__ == '__main__':
    main()
