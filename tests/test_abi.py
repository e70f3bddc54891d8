"""The C-ABI library loads here (no GPU calls) and exports what include/refspi_amd.h declares."""
import os
import re

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, 'include', 'refspi_amd.h')).read()
    return sorted(set(re.findall(r'^\s*(?:const\s+)?\w+\*?\s+\**(rsp_\w+)\s*\(', src, re.M)))


def test_header_declares_expected_entry_points():
    names = _declared()
    for n in ('rsp_slice_gen', 'rsp_slice_gen_holo', 'rsp_calc_prob_all', 'rsp_get_f_dt',
              'rsp_emc_create', 'rsp_emc_estep', 'rsp_maxlp_run', 'rsp_last_error'):
        assert n in names


def test_library_exports_every_declared_symbol():
    import ctypes
    from refspi_amd import _lib
    lib = _lib.load()
    for n in _declared():
        assert hasattr(lib, n), n
        getattr(lib, n)
    assert lib.rsp_version() >= 100
    assert set(_declared()) == set(_lib.exported_symbols())


def test_argument_errors_without_gpu():
    """Null / size checks return negative codes before touching the device."""
    import ctypes as C
    from refspi_amd import _lib
    lib = _lib.load()
    assert lib.rsp_slice_gen(None, 0., 1., 16, None, 0, None, None) == -1
    assert b'null' in lib.rsp_last_error()
    assert lib.rsp_calc_prob_all(None, None, -1, None, None, None, None, None, None, None, 0.,
                                 None, None, None) == -2
    assert lib.rsp_emc_estep(None, None, None, None, None, None, None, None) == -1
    h = C.c_void_p()
    assert lib.rsp_emc_create(0, 4, 1, None, None, None, None, None, 1, C.byref(h)) == -1


def test_rotation_form_argument_errors_without_gpu():
    """The rotation-form entry points reject null plans / arguments before any device call."""
    import ctypes as C
    from refspi_amd import _lib
    lib = _lib.load()
    assert lib.rsp_rot_set_option(None, 1, 0) == -1
    assert lib.rsp_rot_estep(None, None, None, None, None, None, None, None, None, None) == -1
    assert lib.rsp_rot_estep_best(None, *([None] * 10)) == -1
    assert lib.rsp_rot_merge(None, None, None, None, None, None, None) == -1
    assert lib.rsp_rot_adjoint(None, *([None] * 11)) == -1
    n = C.c_int64()
    assert lib.rsp_rot_pack(None, 1, None, None, None, None, None, None, None, None, None,
                            None, None, C.byref(n)) == -1
    assert lib.rsp_rot_pass_begin(None, None, 0, 1, None) == -1
    assert lib.rsp_rot_pass_chunk(None, 1, *([None] * 18)) == -1
    assert lib.rsp_rot_pass_end(None, None, None, None, None, None) == -1
    assert lib.rsp_rot_create(0, 64, 8, None, None, 1., None) == -1
    assert b'null' in lib.rsp_last_error()
    out = C.c_void_p()
    cs = (C.c_double * 8)()
    assert lib.rsp_rot_create(0, 4, 8, cs, cs, 1., C.byref(out)) == -2       # size < 9
    assert lib.rsp_rot_create(0, 64, 0, cs, cs, 1., C.byref(out)) == -2      # no rotations
    assert lib.rsp_rot_create(0, 64, 2000, cs, cs, 1., C.byref(out)) == -2   # > 1536
    assert lib.rsp_xterm_gemm(1, -1, 1, 1, None, None, None, None) == -2

