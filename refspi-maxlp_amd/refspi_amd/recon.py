"""Command-line reconstruction driver (mirror of the reference's recon.py:16-65).

    python -m refspi_amd.recon NUM_ITER [-c emc_config.ini] [-d devices] [-s streams]

Runs ``NUM_ITER`` EMC + MaxLP iterations on one GPU, writes ``output_%03d.h5`` per
iteration and ``EMC.log`` (header ``Iter  time(s)  change``, one line per iteration with
the wall time and the L2 norm of the model change, recon.py:45-60) into the config's
``output_folder``; when the photons file holds the object (``solution``), also ``FRC.log``
with each iteration's FRC against it (calc_frc.py:77-117).  A devices file (one device number per line) selects the GPU of
rank 0 (recon.py:30-40); there is one rank.
"""
from __future__ import annotations

import argparse
import os.path as op
import sys
import time

import numpy as np


def parse_args(argv=None):
    p = argparse.ArgumentParser(description='In-plane rotation EMC')
    p.add_argument('num_iter', type=int, help='Number of iterations')
    p.add_argument('-c', '--config_file', default='emc_config.ini',
                   help='Path to configuration file (default: emc_config.ini)')
    p.add_argument('-d', '--devices', help='Path to devices file')
    p.add_argument('-s', '--streams', type=int, default=4,
                   help='Number of streams to use (default=4)')
    return p.parse_args(argv)


def device_from_file(path, rank=0):
    """recon.py:35-38: the rank's line of the devices file."""
    if path is None:
        return 0
    with open(path) as f:
        return int(f.readlines()[rank].strip())


def log_line(it, seconds, change):
    """recon.py:56 line format."""
    return '%-6d%-.2e %e\n' % (it, seconds, change)


def run(recon, num_iter, logf=None, out=sys.stdout, frcf=None):
    """recon.py:43-63 loop; returns the per-iteration (time, change) list.  With ``frcf``
    and a ``solution`` on the reconstructor (the object stored in the photons file), each
    iteration's FRC against it (calc_frc.py:77-117, best of 360 rotations, mean over rings
    4 .. size//2) is written there as ``iteration FRC``."""
    solution = getattr(recon, 'solution', None)
    if logf is not None:
        logf.write('Iter  time(s)  change\n')
        logf.flush()
    if frcf is not None and solution is not None:
        frcf.write('Iter  FRC\n')
        frcf.flush()
    hist = []
    avgtime, numavg = 0., 0
    for i in range(num_iter):
        m0 = np.array(recon.model, copy=True)
        stime = time.time()
        recon.run_iteration(i + 1)
        etime = time.time()
        sys.stderr.write('\r%d/%d (%f s)\n' % (i + 1, num_iter, etime - stime))
        norm = float(np.linalg.norm(np.asarray(recon.model).ravel() - m0.ravel()))
        if logf is not None:
            logf.write(log_line(i + 1, etime - stime, norm))
            logf.flush()
        print('Change from last iteration: ', norm, file=out)
        hist.append((etime - stime, norm))
        if frcf is not None and solution is not None:
            from .frc import model_frc
            frcf.write('%-6d%.6f\n' % (i + 1, model_frc(recon.model, solution)[0]))
            frcf.flush()
        if i > 0:
            avgtime += etime - stime
            numavg += 1
    if numavg > 0:
        print('\n%.4e s/iteration on average' % (avgtime / numavg), file=out)
    return hist


def main(argv=None):
    args = parse_args(argv)
    from .emc import EMC
    dev = device_from_file(args.devices)
    if args.devices is None:
        print('Running on default device 0')
    recon = EMC(args.config_file, num_streams=args.streams, device=dev)
    with open(op.join(recon.output_folder, 'EMC.log'), 'w') as logf:
        if recon.solution is not None:
            with open(op.join(recon.output_folder, 'FRC.log'), 'w') as frcf:
                run(recon, args.num_iter, logf, frcf=frcf)
        else:
            run(recon, args.num_iter, logf)


if __name__ == '__main__':
    main()
