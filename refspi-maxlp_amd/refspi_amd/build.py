"""Build the gfx950 shared library ``librefspi_amd.so`` in-tree with hipcc.

Usage: ``python -m refspi_amd.build`` (from ``refspi-maxlp_amd/``) or ``build()``.
Every HIP source is compiled for gfx950 only, without fused multiply-add contraction
(the oracle's rounding is reproduced bit for bit where indices depend on it) and with
packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                      # refspi-maxlp_amd/
CSRC = os.path.join(ROOT, 'csrc')
INCLUDE = os.path.join(os.path.dirname(ROOT), 'include')
BUILD_DIR = os.path.join(ROOT, '_build')
LIB_PATH = os.path.join(PKG_DIR, 'librefspi_amd.so')

SOURCES = ['rsp_status.cpp', 'mirror.hip', 'estep.hip', 'maxlp.hip', 'xterm.hip', 'rotem.hip']

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', INCLUDE, '-I', CSRC, '-Wno-unused-result']


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + FLAGS + ['-x', 'hip', '-c', os.path.join(CSRC, src), '-o', obj]
    if src.endswith('.cpp'):
        cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = '\n'.join(l for l in r.stderr.splitlines() if 'packed-fp32-ops' not in l)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, err))
    return obj


def build(verbose=False):
    """Compile every source and link the shared library (always rebuilds)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB_PATH + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC'] + objs + [
        '-L/opt/rocm/lib', '-Wl,-rpath,/opt/rocm/lib', '-lrocblas', '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stderr)
    os.replace(tmp, LIB_PATH)
    if verbose:
        print('built', LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
