"""Benchmark: one EMC+MaxLP iteration per step on one MI355X.

Workload (default): 185x185 q-grid, 100,000 synthetic frames at 300 photons/frame
(120 Gaussian blobs in a 30 px disk, 370 px box, beamstop 4 px, sphere reference with
diameter N(8, 0.4) px and shift U(-4, 4) px carrying 5x the object photons, uniform
in-plane rotation, Poisson sampling), joint search over 17x17x21 = 6069
(shift_x, shift_y, diameter) states x 360 rotations, then the MaxLP object update.

Metric: frame x rotation x state log-likelihood evaluations per second of whole
iterations, N * R * S / t_iteration.

Prints one JSON line with the metric, the live roofline of the dominant kernel (the
sparse log-likelihood gather) and a CPU baseline (the NumPy FP64 oracle timed on a
bounded sample of the same workload on this host).
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


def cpu_baseline(fr, size, num_rot, grid, n_frames, n_states):
    """NumPy FP64 oracle on a bounded sample: n_frames x n_states x num_rot hypotheses."""
    from oracle import emc_oracle as O
    no = int(fr.ones[:n_frames].sum())
    nm = int(fr.multi[:n_frames].sum())
    ofr = O.SparseFramesNP(fr.ones[:n_frames], fr.multi[:n_frames], fr.place_ones[:no],
                           fr.place_multi[:nm], fr.count_multi[:nm], fr.num_pix)
    sx, sy, dia, _ = O.state_grid(*grid)
    sel = np.linspace(0, len(sx) - 1, n_states).astype(int)
    rng = np.random.default_rng(0)
    model = rng.random(size * size) + 1j * rng.random(size * size)
    t = time.time()
    O.estep_prob(model, size, (sx[sel], sy[sel], dia[sel]), num_rot, ofr)
    el = time.time() - t
    return dict(value=n_frames * num_rot * n_states / el, unit='frame*rotation*state evals/s',
                cores=1, kind='port',
                sample='E-step only (hologram, rotated log-views, sparse LL, argmax) of %d '
                       'frames x %d states x %d rotations of the same workload, NumPy FP64 '
                       'single thread, %.1f s; MaxLP excluded' %
                       (n_frames, n_states, num_rot, el))


def main():
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
    ap.add_argument('--cpu-frames', type=int, default=1000)
    ap.add_argument('--cpu-states', type=int, default=10)
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
        print(json.dumps({'copy_bytes_read': 1 << 30, 'copy_bytes_written': 1 << 30}))
        return
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(asked for --gpus %d)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from refspi_amd.dset import Dataset
    from refspi_amd.emc import EMC
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
    t0 = time.time()
    for i in range(args.steps):
        ts = time.time()
        emc.run_iteration()
        tm = emc.estep_timing()
        gather_ms += tm['gather_ms']
        tables_ms += tm['tables_ms']
        batches += tm['batches']
        log('step %d: %.2f s %s gather %.1f ms tables %.1f ms' %
            (i, time.time() - ts, emc.timings, tm['gather_ms'], tm['tables_ms']))
    torch.cuda.synchronize()
    el = time.time() - t0
    t_iter = el / args.steps
    n = fr.ones.shape[0]
    value = n * args.num_rot * S / t_iter
    # dominant kernel: the sparse log-likelihood gather.  Algorithmic bytes per frame x
    # hypothesis evaluation = the frame's sparse record, 4*ones + 8*multi bytes.
    rec_bytes = 4. * fr.ones.sum() + 8. * fr.multi.sum()
    alg_bytes_per_step = H * rec_bytes
    gather_s = gather_ms / 1e3
    achieved = alg_bytes_per_step * args.steps / gather_s / 1e9
    pm = emc.probmask
    scored = (pm[fr.place_ones] == 0).sum() + fr.count_multi[pm[fr.place_multi] == 0].sum()
    lds_bytes = H * 4. * scored
    lds_gbs = lds_bytes * args.steps / gather_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, 'profiles', 'gather_traffic_185.json')
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get('bytes_per_launch')
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
        roofline=dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s',
                      frac=achieved / HBM_PEAK_GBS, traffic=traffic,
                      kernel='k_gather', launches=batches,
                      avg_launch_ms=gather_ms / max(batches, 1),
                      note='algorithmic bytes = H x sum_d(4 ones_d + 8 multi_d) per step; '
                           'photons stay in VGPRs across hypotheses so this exceeds HBM '
                           'traffic; the binding resource is LDS (see lds)'),
        lds=dict(achieved=lds_gbs, peak=LDS_PEAK_GBS, unit='GB/s', frac=lds_gbs / LDS_PEAK_GBS,
                 what='scored-photon gathers (4 B each) per step / gather time'),
    )
    if not args.no_cpu_baseline:
        out['cpu_baseline'] = cpu_baseline(fr, size, args.num_rot, grid, args.cpu_frames,
                                           args.cpu_states)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
