"""ctypes binding of ``librefspi_amd.so`` (the gfx950 C-ABI declared in include/refspi_amd.h).

The library is loaded from this package directory.  There is no fallback: if the library
is missing or fails to load, every entry point raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'librefspi_amd.so')

_lib = None

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p

# name: (restype, argtypes)
_SIGS = {
    'rsp_last_error': (C.c_char_p, []),
    'rsp_version': (C.c_int, []),
    'rsp_slice_gen': (C.c_int, [vp, C.c_double, C.c_double, C.c_int64, vp, C.c_int64, vp, vp]),
    'rsp_slice_gen_holo': (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int64, vp, C.c_int64, vp, vp]),
    'rsp_calc_prob_all': (C.c_int, [vp, vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp, C.c_double,
                                    vp, vp, vp]),
    'rsp_get_f_dt': (C.c_int, [vp, vp, C.c_int64, C.c_int64, vp, vp]),
    'rsp_slice_merge': (C.c_int, [vp, C.c_double, C.c_int64, vp, vp, vp]),
    'rsp_merge_all': (C.c_int, [vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int64, vp]),
    'rsp_emc_create': (C.c_int, [C.c_int, C.c_int64, C.c_int64, f64p, f64p, f64p, f64p, f64p,
                                 C.c_int64, C.POINTER(vp)]),
    'rsp_emc_set_frames': (C.c_int, [vp, C.c_int64, i32p, i32p, i32p, i32p, i32p, f64p]),
    'rsp_emc_set_option': (C.c_int, [vp, C.c_int, C.c_int64]),
    'rsp_emc_estep': (C.c_int, [vp, vp, vp, vp, vp, vp, f64p, vp]),
    'rsp_emc_stats': (C.c_int, [vp, i64p]),
    'rsp_emc_timing': (C.c_int, [vp, f64p]),
    'rsp_emc_schedule_check': (C.c_int, [C.c_int64, C.c_uint64, i64p]),
    'rsp_emc_destroy': (C.c_int, [vp]),
    'rsp_maxlp_create': (C.c_int, [C.c_int, C.c_int64, C.POINTER(vp)]),
    'rsp_maxlp_set_frames': (C.c_int, [vp, C.c_int64, i32p, i32p, i32p, i32p, i32p]),
    'rsp_maxlp_run': (C.c_int, [vp, vp, vp, vp, vp, vp, vp, f64p, vp]),
    'rsp_maxlp_photons_pixel': (C.c_int, [vp, C.c_int64, vp, vp]),
    'rsp_maxlp_logq_pixel': (C.c_int, [vp, C.c_int64, vp, C.c_int64, C.c_double, vp, vp]),
    'rsp_maxlp_info': (C.c_int, [vp, i64p]),
    'rsp_maxlp_destroy': (C.c_int, [vp]),
    'rsp_xterm_gemm': (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.c_int64, vp, vp, vp, vp]),
    'rsp_xterm_gemm_static': (C.c_int, [C.c_int64, C.c_int64, C.c_int64, vp, vp, vp, vp]),
    'rsp_rot_create': (C.c_int, [C.c_int, C.c_int64, C.c_int64, vp, vp, C.c_double, vp]),
    'rsp_rot_set_frames': (C.c_int, [vp, C.c_int64, vp, vp, vp, vp, vp]),
    'rsp_rot_set_option': (C.c_int, [vp, C.c_int, C.c_int64]),
    'rsp_rot_estep': (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    'rsp_rot_merge': (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    'rsp_rot_estep_best': (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    'rsp_rot_adjoint': (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    'rsp_rot_pack': (C.c_int, [vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    'rsp_rot_pass_begin': (C.c_int, [vp, vp, C.c_int, C.c_int64, vp]),
    'rsp_rot_pass_chunk': (C.c_int, [vp, C.c_int64] + [vp] * 18),
    'rsp_rot_pass_end': (C.c_int, [vp, vp, vp, vp, vp, vp]),
    'rsp_rot_timing': (C.c_int, [vp, vp]),
    'rsp_rot_info': (C.c_int, [vp, i64p]),
    'rsp_rot_destroy': (C.c_int, [vp]),
}


def load():
    """Load the library (once) and attach the signatures; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('librefspi_amd.so not built (expected at %s); run '
                           '`python -m refspi_amd.build` in refspi-maxlp_amd/' % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names the C-ABI declares (for the export check)."""
    return list(_SIGS)


class RspError(RuntimeError):
    pass


def check(rc, what=''):
    if rc != 0:
        msg = load().rsp_last_error().decode(errors='replace')
        raise RspError('%s failed (status %d): %s' % (what, rc, msg))


def call(name, *args):
    check(getattr(load(), name)(*args), name)


def ptr(t):
    """Device (or host) address of a torch tensor or numpy array as c_void_p."""
    if t is None:
        return None
    if hasattr(t, 'data_ptr'):
        return C.c_void_p(t.data_ptr())
    return t.ctypes.data_as(C.c_void_p)


def host_ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
