"""Write the golden fixture ``golden_small.npz`` with the FP64 oracle (oracle/).

The reference cannot run in this build's environment (it needs CuPy, h5py, mpi4py and
nvcc, none present, and it ships no golden vectors or tests), so these vectors come from
the oracle, a NumPy FP64 restatement of the reference whose functions are pinned
independently (closed forms, brute force, exact permutations, adjoint identities,
central differences).  The fixture holds inputs and expected outputs so that GPU tests
check the HIP path without running the oracle, and CPU tests guard the oracle against
regressions.

Workload: 21x21 q-grid, 24 frames (Gaussian-blob object, sphere reference, beamstop 3,
seed 0), 3x3x3 states (shift -1..1, diameter 7..9), 8 rotations, two EMC+MaxLP
iterations; the rotation form at rel_scale 1.3 with per-frame scales.

Usage: ``python tests/golden/make_golden.py`` (from the repository root).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, 'refspi-maxlp_amd')]

OUT = os.path.join(HERE, 'golden_small.npz')

SIZE, N, R = 21, 24, 8
GRID = dict(shiftx=(-1., 1., 3), shifty=(-1., 1., 3), sphere_dia=(7., 9., 3))
REL = 1.3


def inputs():
    from refspi_amd import synth
    cfg = synth.SynthConfig(size=SIZE, num_data=N, mean_count=120., n_blobs=8, disk_radius=4.,
                            beamstop=3., shift_max=1., seed=0)
    fr = synth.make_frames(cfg)
    rng = np.random.default_rng(1)
    model = 0.05 * (rng.random(SIZE * SIZE) + 1j * rng.random(SIZE * SIZE))
    return fr, model


def compute(fr, model):
    from oracle import emc_oracle as E
    from oracle import rot_oracle as O
    ofr = E.SparseFramesNP(fr.ones, fr.multi, fr.place_ones, fr.place_multi, fr.count_multi,
                           SIZE * SIZE)
    sx, sy, dia, ns = E.state_grid(GRID['shiftx'], GRID['shifty'], GRID['sphere_dia'])
    out = dict(ones=fr.ones, multi=fr.multi, place_ones=fr.place_ones,
               place_multi=fr.place_multi, count_multi=fr.count_multi, model0=model,
               grid=np.array(GRID['shiftx'] + GRID['shifty'] + GRID['sphere_dia']),
               size=SIZE, num_rot=R, rel_scale=REL)
    # kernels.cu:44-82 / 6-42 one by one (slice_gen_holo, slice_gen log views)
    bg = np.zeros(SIZE * SIZE)
    holo_states = np.array([0, 13, 26])
    out['holo_states'] = holo_states
    out['holo'] = np.stack([E.slice_gen_holo(model, sx[i], sy[i], dia[i], 1., 1., SIZE, bg, 0)
                            for i in holo_states])
    out['logviews_state0'] = E.estep_logviews(out['holo'][0], SIZE, R, bg)
    # two EMC + MaxLP iterations (emc.py:134-157)
    m = model
    for it in (1, 2):
        prob, msums, vscale = E.estep_prob(m, SIZE, (sx, sy, dia), R, ofr, sequential=True)
        mx, rmax = E.get_rmax(prob)
        vals = E.unravel_rmax(rmax, ns, R, sx, sy, dia)
        orc = E.MaxLPOracle(ofr, SIZE)
        traces = {}
        m_new = orc.run_phaser(m, *vals, traces=traces)
        margin = np.full(SIZE * SIZE, np.inf)
        for (_, t), tr in traces.items():
            if tr:
                margin[t] = min(margin[t], min(mg for _, mg in tr))
        p = 'it%d_' % it
        out.update({p + 'model_in': m, p + 'prob': prob, p + 'msums': msums,
                    p + 'vscale': vscale, p + 'maxll': mx, p + 'rmax': rmax,
                    p + 'estep_margin': E.top_margin(prob), p + 'sx': vals[0],
                    p + 'sy': vals[1], p + 'dia': vals[2], p + 'ang': vals[3],
                    p + 'rescale': orc.last_rescale, p + 'model_out': m_new,
                    p + 'maxlp_margin': margin})
        if it == 1:
            # maxLP.py:250-252 photon pull table for the frames' angles of iteration 1
            qx, qy, _, mask, _ = E.qvals(SIZE)
            rc, rs = E.frame_rot_cos_sin(vals[3])
            tpix = np.nonzero(mask)[0]
            out['pull_pixels'] = tpix
            out['pull_index'] = np.stack([E.photon_index(t, SIZE, rc, rs, qx, qy)
                                          for t in tpix]).astype('i4')
        m = m_new
    # rotation form: per-frame reference parameters, rotations latent
    rng = np.random.default_rng(4)
    rsx, rsy = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    rdia, rsc = rng.uniform(7, 9, N), rng.uniform(0.6, 1.4, N)
    pix = O.scored_pixels(SIZE)
    K = ofr.dense()[:, pix]
    LL, _ = O.rot_loglik(model, SIZE, R, K, rsx, rsy, rdia, rsc, REL)
    P, logz = O.lse(LL)
    merged, acc, wts = O.merge_model(P, K, rsc, SIZE, R)
    loss, gm, gsx, gsy, gdia = O.loss_grad(model, SIZE, R, K, rsx, rsy, rdia, rsc, REL)
    out.update(rot_sx=rsx, rot_sy=rsy, rot_dia=rdia, rot_scales=rsc, rot_ll=LL, rot_prob=P,
               rot_logz=logz, rot_merged=merged, rot_merge_acc=acc, rot_merge_wts=wts,
               rot_loss=loss, rot_g_model=gm, rot_g_sx=gsx, rot_g_sy=gsy, rot_g_dia=gdia)
    return out


def main():
    fr, model = inputs()
    out = compute(fr, model)
    np.savez_compressed(OUT, **out)
    print('wrote %s (%d arrays, %.0f KB)' % (OUT, len(out), os.path.getsize(OUT) / 1024))


if __name__ == '__main__':
    main()
