"""EMC reconstructor on one MI355X (mirror of the reference's emc.py:23-254 ``EMC``).

Same construction (``EMC(config_file, num_streams=4)``), attributes (``model``, ``size``,
``num_rot``, ``shiftx``/``shifty``/``sphere_dia``, ``num_states``, ``probmask``,
``scales``, ``rmax``, ``output_folder``, ``phaser``) and per-iteration methods
(``run_iteration``, ``_calculate_prob``, ``_get_rmax``, ``_unravel_rmax``,
``save_output``).  The hypothesis search of emc.py:159-216 runs as one fused GPU
E-step (csrc/estep.hip) which never materialises the (S*R, N) ``prob`` array; instead it
yields ``rmax`` and ``maxll`` (the FP64 log-likelihood of the chosen hypothesis), the
quantities emc.py:207-216 extracts from ``prob``.
"""
from __future__ import annotations

import configparser
import ctypes as C
import math
import os
import os.path as op
import sys
import time

import numpy as np

from . import _lib
from .dset import Dataset
from .maxlp import MaxLPhaser


def state_grid(shiftx, shifty, sphere_dia):
    """emc.py:67-78: ``lo hi n`` triples -> raveled 'ij' meshgrid and num_states."""
    sx = np.linspace(shiftx[0], shiftx[1], int(shiftx[2]))
    sy = np.linspace(shifty[0], shifty[1], int(shifty[2]))
    dd = np.linspace(sphere_dia[0], sphere_dia[1], int(sphere_dia[2]))
    gx, gy, gd = np.meshgrid(sx, sy, dd, indexing='ij')
    num_states = np.array([shiftx[-1], shifty[-1], sphere_dia[-1]]).astype('i4')
    return gx.ravel(), gy.ravel(), gd.ravel(), num_states


def probmask(size):
    """emc.py:89-98."""
    x, y = np.indices((size, size), dtype='f8')
    rad = np.sqrt((x.ravel() - size // 2) ** 2 + (y.ravel() - size // 2) ** 2)
    pm = np.zeros(size * size, dtype='i4')
    pm[rad >= size // 2] = 2
    pm[rad < size // 8] = 1
    pm[rad < 4] = 2
    return pm


def rotation_table(num_rot):
    """cos/sin of j*pi/num_rot (emc.py:192) from the C library."""
    ang = [j * np.pi / num_rot for j in range(num_rot)]
    return np.array([math.cos(a) for a in ang]), np.array([math.sin(a) for a in ang])


class EMC:
    """emc.py:23-55.

    ``config_file`` is the reference's INI (``[emc]`` section: size, in_photons_file,
    output_folder, log_file, need_scaling, num_rot, num_modes, sphere_dia, shiftx,
    shifty).  Keyword arguments override it for programmatic use: ``dataset`` (a
    Dataset), ``size``, ``num_rot``, ``shiftx``/``shifty``/``sphere_dia`` (lo hi n),
    ``need_scaling``, ``output_folder``, ``model`` (initial c16 model), ``seed``
    (seeded random model; the reference's is unseeded), ``device``.
    """

    def __init__(self, config_file=None, num_streams=4, *, dataset=None, size=None,
                 num_rot=None, shiftx=None, shifty=None, sphere_dia=None, need_scaling=None,
                 output_folder=None, model=None, seed=None, device=0, save=True,
                 bank_count=None, regroup_iters=None, table_layout=None, shared_slots=None,
                 frames_per_lane=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('EMC needs a GPU (HIP device)')
        self.num_streams = num_streams
        self.device = device
        self.rank, self.num_proc = 0, 1
        self.save = save
        self.bank_count = bank_count
        self.regroup_iters = regroup_iters
        self.table_layout = table_layout
        self.shared_slots = shared_slots
        self.frames_per_lane = frames_per_lane
        config = configparser.ConfigParser()
        start_dir = '.'
        if config_file is not None:
            if not config.read(config_file):
                raise FileNotFoundError(config_file)
            start_dir = op.dirname(config_file)
        self._parse_params(config, start_dir, size, num_rot, need_scaling, output_folder)
        self._generate_states(config, shiftx, shifty, sphere_dia)
        self._generate_masks()
        self._parse_data(dataset)
        self.phaser = MaxLPhaser(self.dset, size=self.size, device=device)
        self._init_model(model, seed)
        self._plan = None
        self.rmax = None
        self.maxll = None
        self.timings = {}

    # -- emc.py:57-65
    def _parse_params(self, config, start_dir, size, num_rot, need_scaling, output_folder):
        has = config.has_section('emc')
        self.size = size if size is not None else config.getint('emc', 'size')
        self.num_modes = config.getint('emc', 'num_modes', fallback=1) if has else 1
        self.num_rot = num_rot if num_rot is not None else config.getint('emc', 'num_rot')
        self.photons_file = (op.join(start_dir, config.get('emc', 'in_photons_file'))
                             if has and config.has_option('emc', 'in_photons_file') else None)
        if output_folder is not None:
            self.output_folder = output_folder
        elif has and config.has_option('emc', 'output_folder'):
            self.output_folder = op.join(start_dir, config.get('emc', 'output_folder'))
        else:
            self.output_folder = '.'
        self.log_file = (op.join(start_dir, config.get('emc', 'log_file'))
                         if has and config.has_option('emc', 'log_file') else None)
        self.need_scaling = (need_scaling if need_scaling is not None else
                             (config.getboolean('emc', 'need_scaling', fallback=False)
                              if has else False))

    # -- emc.py:67-87 (the unused sphere_ramps/sphere_intens arrays are not built)
    def _generate_states(self, config, shiftx=None, shifty=None, sphere_dia=None):
        tup = lambda v, key: tuple(float(s) for s in
                                   (v if v is not None else config.get('emc', key).split()))
        dia = tup(sphere_dia, 'sphere_dia')
        sx = tup(shiftx, 'shiftx')
        sy = tup(shifty, 'shifty')
        self.shiftx, self.shifty, self.sphere_dia, self.num_states = state_grid(sx, sy, dia)
        print(self.num_states.prod(), 'sampled states')

    # -- emc.py:89-98
    def _generate_masks(self):
        self.probmask = probmask(self.size)

    # -- emc.py:109-118
    def _parse_data(self, dataset=None):
        stime = time.time()
        self.dset = dataset if dataset is not None else \
            Dataset(self.photons_file, self.size ** 2, self.need_scaling)
        if self.dset.num_pix != self.size ** 2:
            raise AttributeError('Number of pixels in photons file does not match')
        self.powder = self.dset.get_powder()
        # the generator's object, when the photons file holds it (make_data.py:164-167);
        # used only for the FRC quality report of recon
        self.solution = None
        if dataset is None and self.photons_file is not None:
            from . import h5io
            self.solution = h5io.read_fields(self.photons_file, ['solution']).get('solution')
        etime = time.time()
        print('%d frames with %.3f photons/frame (%.3f s) (%.2f MB)' %
              (self.dset.num_data, self.dset.mean_count, etime - stime, self.dset.mem / 1024 ** 2))
        sys.stdout.flush()

    # -- emc.py:120-132
    def _init_model(self, model=None, seed=None):
        if model is not None:
            self.model = np.ascontiguousarray(np.asarray(model, dtype='c16').ravel())
        else:
            rng = np.random.default_rng(seed) if seed is not None else np.random
            self.model = rng.random(self.size ** 2) + 1j * rng.random(self.size ** 2)
            if self.save and op.isdir(self.output_folder):
                np.save(op.join(self.output_folder, 'model_000.npy'), self.model)
        if self.need_scaling:
            self.scales = self.dset.photon_counts() / self.dset.mean_count
        else:
            self.scales = np.ones(self.dset.num_data, dtype='f8')

    def _make_plan(self):
        lib = _lib.load()
        rc, rs = rotation_table(self.num_rot)
        h = C.c_void_p()
        arrs = [np.ascontiguousarray(a, dtype='f8') for a in
                (rc, rs, self.shiftx, self.shifty, self.sphere_dia)]
        _lib.check(lib.rsp_emc_create(self.device, self.size, self.num_rot,
                                      *[_lib.host_ptr(a, C.c_double) for a in arrs],
                                      len(self.shiftx), C.byref(h)), 'rsp_emc_create')
        if self.bank_count is not None:   # LDS bank model of the photon scheduler
            _lib.check(lib.rsp_emc_set_option(h, 3, int(self.bank_count)), 'rsp_emc_set_option')
        if self.regroup_iters is not None:   # swaps of the frame regrouping per window
            _lib.check(lib.rsp_emc_set_option(h, 6, int(self.regroup_iters)),
                       'rsp_emc_set_option')
        if self.table_layout is not None:   # per-tile LDS layout of the table parts
            _lib.check(lib.rsp_emc_set_option(h, 7, int(bool(self.table_layout))),
                       'rsp_emc_set_option')
        if self.shared_slots is not None:   # lanes reading one entry share its bank slot
            _lib.check(lib.rsp_emc_set_option(h, 8, int(bool(self.shared_slots))),
                       'rsp_emc_set_option')
        if self.frames_per_lane is not None:   # whole frames two to a lane (2) or one
            _lib.check(lib.rsp_emc_set_option(h, 9, int(self.frames_per_lane)),
                       'rsp_emc_set_option')
        d = self.dset
        bg = np.ascontiguousarray(d.bg, dtype='f8')
        _lib.check(lib.rsp_emc_set_frames(
            h, d.num_data, _lib.host_ptr(d.ones, C.c_int32), _lib.host_ptr(d.multi, C.c_int32),
            _lib.host_ptr(d.place_ones, C.c_int32), _lib.host_ptr(d.place_multi, C.c_int32),
            _lib.host_ptr(d.count_multi, C.c_int32), _lib.host_ptr(bg, C.c_double)),
            'rsp_emc_set_frames')
        self._plan = h

    def set_option(self, key, value):
        if self._plan is None:
            self._make_plan()
        _lib.call('rsp_emc_set_option', self._plan, int(key), int(value))

    def __del__(self):
        h = getattr(self, '_plan', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.rsp_emc_destroy(h)
            self._plan = None

    # -- emc.py:134-157
    def run_iteration(self, iternum=None):
        """One EMC+MaxLP iteration: E-step argmax, unravel, MaxLP, output."""
        import torch
        dev = 'cuda:%d' % self.device
        dmodel = torch.as_tensor(self.model.ravel(), device=dev)
        t0 = time.time()
        self._calculate_prob(dmodel)
        self._get_rmax()
        t1 = time.time()
        sx_vals, sy_vals, dia_vals, ang_vals = self._unravel_rmax(self.rmax)
        self.phaser.run_device(dmodel, sx_vals, sy_vals, dia_vals, ang_vals)
        torch.cuda.synchronize()
        t2 = time.time()
        self.model = dmodel.cpu().numpy()
        self.timings = dict(estep=t1 - t0, maxlp=t2 - t1)
        if self.save:
            self.save_output(self.model, iternum)

    # -- emc.py:159-205 (fused; drange chunking is not needed on one 288 GB GPU)
    def _calculate_prob(self, dmodel, views=None, drange=None):
        import torch
        if drange is not None:
            raise NotImplementedError('drange: the fused E-step covers all frames')
        if self._plan is None:
            self._make_plan()
        dev = dmodel.device
        n = self.dset.num_data
        S = len(self.shiftx)
        scales = torch.as_tensor(np.ascontiguousarray(self.scales, dtype='f8'), device=dev)
        self._rmax_dev = torch.empty(n, dtype=torch.int32, device=dev)
        self._maxll_dev = torch.empty(n, dtype=torch.float64, device=dev)
        self._msums_dev = torch.empty(S, dtype=torch.float64, device=dev)
        vs = C.c_double(0.)
        _lib.call('rsp_emc_estep', self._plan, _lib.ptr(dmodel.contiguous()), _lib.ptr(scales),
                  _lib.ptr(self._rmax_dev), _lib.ptr(self._maxll_dev), _lib.ptr(self._msums_dev),
                  C.byref(vs), _lib.stream_handle())
        self.vscale = vs.value

    def estep_stats(self):
        out = np.zeros(8, dtype='i8')
        _lib.call('rsp_emc_stats', self._plan, _lib.host_ptr(out, C.c_int64))
        keys = ('frames', 'hypotheses', 'decided_fp32', 'top2_fp64', 'full_fp64',
                'slots_per_lane', 'waves_per_group', 'table_parts')
        return dict(zip(keys, (int(v) for v in out)))

    def estep_timing(self):
        out = np.zeros(6, dtype='f8')
        _lib.call('rsp_emc_timing', self._plan, _lib.host_ptr(out, C.c_double))
        return dict(tables_ms=float(out[0]), gather_ms=float(out[1]), batches=int(out[2]),
                    batch=int(out[3]), lds_read_bytes_per_hyp=float(out[4]),
                    lds_write_bytes_per_hyp=float(out[5]))

    # -- emc.py:207-216 (one rank)
    def _get_rmax(self):
        self.rmax = self._rmax_dev.cpu().numpy().astype('i4')
        self.maxll = self._maxll_dev.cpu().numpy()
        self.msums = self._msums_dev.cpu().numpy()
        return self.rmax

    # -- emc.py:218-224
    def _unravel_rmax(self, rmax):
        sx, sy, dia, ang = np.unravel_index(rmax, tuple(self.num_states) + (self.num_rot,))
        sx_vals = np.unique(self.shiftx)[sx]
        sy_vals = np.unique(self.shifty)[sy]
        dia_vals = np.unique(self.sphere_dia)[dia]
        ang_vals = ang * np.pi / self.num_rot
        return sx_vals, sy_vals, dia_vals, ang_vals

    # -- emc.py:226-239
    def save_output(self, model, iternum, intens=None):
        if iternum is None:
            np.save(op.join(self.output_folder, 'model.npy'), self.model)
            return
        from . import h5io
        sx_vals, sy_vals, dia_vals, ang_vals = self._unravel_rmax(self.rmax)
        fields = {'model': self.model.reshape((self.size,) * 2),
                  'angles': ang_vals, 'diameters': dia_vals,
                  'shifts': np.array([sx_vals, sy_vals]).T}
        if intens is not None:
            fields['intens'] = np.asarray(intens)
        h5io.write_fields(op.join(self.output_folder, 'output_%.3d.h5' % iternum), fields)
