"""HDF5 files of the reference, written and read through the HDF5 C library (ctypes).

* photons file (make_data.py:228-251, read by dset.py:23-60 and maxLP.py:50-71):
  ``num_pix`` (i8[1]), ``place_ones``/``place_multi``/``count_multi`` (vlen i4 per frame),
  ``ones``/``multi`` (i4 per frame), optional ``bg`` (f8).
* output file (emc.py:231-239): ``model`` (complex (size, size)), ``angles``,
  ``diameters`` (f8 N), ``shifts`` (f8 N x 2), optional ``intens``.

Complex arrays use h5py's layout (a compound of two f8 fields "r" and "i"), so files are
interchangeable with h5py.  The library is ``libhdf5`` (1.10 API; hid_t = int64).
"""
from __future__ import annotations

import ctypes as C
import os
import types

import numpy as np

_CANDIDATES = [os.environ.get('RSP_LIBHDF5', ''), 'libhdf5.so', 'libhdf5.so.103',
               '/opt/conda/lib/libhdf5.so.103', '/opt/conda/lib/libhdf5.so',
               '/usr/lib/x86_64-linux-gnu/libhdf5_serial.so']
_h5 = None

hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_ulonglong
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2
H5T_COMPOUND = 6


class hvl_t(C.Structure):
    _fields_ = [('len', C.c_size_t), ('p', C.c_void_p)]


def _lib():
    global _h5
    if _h5 is not None:
        return _h5
    err = None
    for name in _CANDIDATES:
        if not name:
            continue
        try:
            lib = C.CDLL(name)
            break
        except OSError as e:  # try the next location
            err = e
    else:
        raise RuntimeError('libhdf5 not found (set RSP_LIBHDF5): %s' % err)
    sig = {
        'H5open': (herr_t, []),
        'H5Fcreate': (hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
        'H5Fopen': (hid_t, [C.c_char_p, C.c_uint, hid_t]),
        'H5Fclose': (herr_t, [hid_t]),
        'H5Screate_simple': (hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        'H5Sclose': (herr_t, [hid_t]),
        'H5Sget_simple_extent_ndims': (C.c_int, [hid_t]),
        'H5Sget_simple_extent_dims': (C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        'H5Dcreate2': (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        'H5Dopen2': (hid_t, [hid_t, C.c_char_p, hid_t]),
        'H5Dwrite': (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        'H5Dread': (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        'H5Dget_space': (hid_t, [hid_t]),
        'H5Dget_type': (hid_t, [hid_t]),
        'H5Dclose': (herr_t, [hid_t]),
        'H5Dvlen_reclaim': (herr_t, [hid_t, hid_t, hid_t, C.c_void_p]),
        'H5Tcreate': (hid_t, [C.c_int, C.c_size_t]),
        'H5Tinsert': (herr_t, [hid_t, C.c_char_p, C.c_size_t, hid_t]),
        'H5Tvlen_create': (hid_t, [hid_t]),
        'H5Tget_class': (C.c_int, [hid_t]),
        'H5Tclose': (herr_t, [hid_t]),
        'H5Lexists': (C.c_int, [hid_t, C.c_char_p, hid_t]),
    }
    for n, (r, a) in sig.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    if lib.H5open() < 0:
        raise RuntimeError('H5open failed')
    _h5 = lib
    return lib


def _t(name):
    return hid_t.in_dll(_lib(), name).value


def _check(v, what):
    if v < 0:
        raise RuntimeError('HDF5 call failed: %s' % what)
    return v


def _complex_type():
    h5 = _lib()
    t = _check(h5.H5Tcreate(H5T_COMPOUND, 16), 'H5Tcreate')
    dbl = _t('H5T_NATIVE_DOUBLE_g')
    _check(h5.H5Tinsert(t, b'r', 0, dbl), 'H5Tinsert r')
    _check(h5.H5Tinsert(t, b'i', 8, dbl), 'H5Tinsert i')
    return t


def _space(shape):
    dims = (hsize_t * max(1, len(shape)))(*shape)
    return _check(_lib().H5Screate_simple(len(shape), dims, None), 'H5Screate_simple')


def _write(fid, name, arr):
    h5 = _lib()
    arr = np.ascontiguousarray(arr)
    close_t = False
    if arr.dtype == np.complex128:
        t = _complex_type()
        close_t = True
    elif arr.dtype == np.float64:
        t = _t('H5T_NATIVE_DOUBLE_g')
    elif arr.dtype == np.int32:
        t = _t('H5T_NATIVE_INT_g')
    elif arr.dtype == np.int64:
        t = _t('H5T_NATIVE_LLONG_g')
    else:
        raise TypeError('unsupported dtype %s' % arr.dtype)
    sp = _space(arr.shape)
    d = _check(h5.H5Dcreate2(fid, name.encode(), t, sp, 0, 0, 0), 'H5Dcreate2 ' + name)
    _check(h5.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), 'H5Dwrite ' + name)
    h5.H5Dclose(d)
    h5.H5Sclose(sp)
    if close_t:
        h5.H5Tclose(t)


def _write_vlen_i4(fid, name, lists):
    h5 = _lib()
    t = _check(h5.H5Tvlen_create(_t('H5T_NATIVE_INT_g')), 'H5Tvlen_create')
    n = len(lists)
    keep = [np.ascontiguousarray(a, dtype=np.int32) for a in lists]
    buf = (hvl_t * max(1, n))()
    for i, a in enumerate(keep):
        buf[i].len = a.size
        buf[i].p = a.ctypes.data if a.size else None
    sp = _space((n,))
    d = _check(h5.H5Dcreate2(fid, name.encode(), t, sp, 0, 0, 0), 'H5Dcreate2 ' + name)
    _check(h5.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.addressof(buf)), 'H5Dwrite ' + name)
    h5.H5Dclose(d)
    h5.H5Sclose(sp)
    h5.H5Tclose(t)


def write_fields(path, fields):
    """Write a dict of NumPy arrays (f8, i4, i8, c16) to a new HDF5 file."""
    h5 = _lib()
    fid = _check(h5.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, 0, 0), 'H5Fcreate ' + path)
    try:
        for k, v in fields.items():
            v = np.asarray(v)
            if v.dtype == np.float32:
                v = v.astype('f8')
            _write(fid, k, v)
    finally:
        h5.H5Fclose(fid)


def write_photons(path, frames, extra=None):
    """Photons file in the layout of make_data.py:228-251 (+ optional bg), plus the
    generator-truth arrays in ``extra`` (``true_angles``, ``scale``, ... make_data.py:258-275)."""
    h5 = _lib()
    ones = np.asarray(frames.ones, dtype=np.int32)
    multi = np.asarray(frames.multi, dtype=np.int32)
    oo = np.concatenate([[0], np.cumsum(ones, dtype=np.int64)])
    mo = np.concatenate([[0], np.cumsum(multi, dtype=np.int64)])
    po = np.asarray(frames.place_ones, dtype=np.int32)
    pm = np.asarray(frames.place_multi, dtype=np.int32)
    cm = np.asarray(frames.count_multi, dtype=np.int32)
    n = len(ones)
    fid = _check(h5.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, 0, 0), 'H5Fcreate ' + path)
    try:
        bg = getattr(frames, 'bg', None)
        if bg is not None:
            _write(fid, 'bg', np.asarray(bg, dtype='f8'))
        _write(fid, 'num_pix', np.array([frames.num_pix], dtype=np.int64))
        _write_vlen_i4(fid, 'place_ones', [po[oo[i]:oo[i + 1]] for i in range(n)])
        _write_vlen_i4(fid, 'place_multi', [pm[mo[i]:mo[i + 1]] for i in range(n)])
        _write_vlen_i4(fid, 'count_multi', [cm[mo[i]:mo[i + 1]] for i in range(n)])
        _write(fid, 'ones', ones)
        _write(fid, 'multi', multi)
        for k, v in (extra or {}).items():
            _write(fid, k, np.asarray(v))
    finally:
        h5.H5Fclose(fid)


def _read(fid, name):
    h5 = _lib()
    d = _check(h5.H5Dopen2(fid, name.encode(), H5P_DEFAULT), 'H5Dopen2 ' + name)
    sp = h5.H5Dget_space(d)
    nd = h5.H5Sget_simple_extent_ndims(sp)
    dims = (hsize_t * max(1, nd))()
    h5.H5Sget_simple_extent_dims(sp, dims, None)
    shape = tuple(int(dims[i]) for i in range(nd))
    ft = h5.H5Dget_type(d)
    cls = h5.H5Tget_class(ft)
    h5.H5Tclose(ft)
    try:
        if cls == H5T_COMPOUND:
            out = np.empty(shape, dtype=np.complex128)
            t = _complex_type()
            _check(h5.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), name)
            h5.H5Tclose(t)
        elif cls == 1:      # H5T_FLOAT
            out = np.empty(shape, dtype=np.float64)
            _check(h5.H5Dread(d, _t('H5T_NATIVE_DOUBLE_g'), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                              out.ctypes.data), name)
        elif cls == 0:      # H5T_INTEGER
            out = np.empty(shape, dtype=np.int64)
            _check(h5.H5Dread(d, _t('H5T_NATIVE_LLONG_g'), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                              out.ctypes.data), name)
        elif cls == 9:      # H5T_VLEN (of integers)
            n = shape[0] if shape else 1
            buf = (hvl_t * max(1, n))()
            t = h5.H5Tvlen_create(_t('H5T_NATIVE_INT_g'))
            _check(h5.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.addressof(buf)), name)
            lens = np.array([buf[i].len for i in range(n)], dtype=np.int64)
            flat = np.empty(int(lens.sum()), dtype=np.int32)
            pos = 0
            for i in range(n):
                if buf[i].len:
                    C.memmove(flat.ctypes.data + 4 * pos, buf[i].p, 4 * buf[i].len)
                pos += buf[i].len
            h5.H5Dvlen_reclaim(t, sp, H5P_DEFAULT, C.addressof(buf))
            h5.H5Tclose(t)
            out = (lens, flat)
        else:
            raise TypeError('unsupported HDF5 class %d for %s' % (cls, name))
    finally:
        h5.H5Sclose(sp)
        h5.H5Dclose(d)
    return out


def read_fields(path, names):
    """The datasets ``names`` that exist in the file (missing ones are left out)."""
    h5 = _lib()
    fid = _check(h5.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT), 'H5Fopen ' + path)
    try:
        out = {}
        for n in names:
            if h5.H5Lexists(fid, n.encode(), H5P_DEFAULT) > 0:
                out[n] = _read(fid, n)
        return out
    finally:
        h5.H5Fclose(fid)


def read_photons(path):
    """dset.py:23-60: the sparse frames of a photons file (missing ones/multi recomputed
    from the vlen lengths, missing bg -> None)."""
    f = read_fields(path, ['num_pix', 'place_ones', 'place_multi', 'count_multi', 'ones',
                           'multi', 'bg'])
    lo, po = f['place_ones']
    lm, pm = f['place_multi']
    lc, cm = f['count_multi']
    ones = f['ones'].astype(np.int32) if 'ones' in f else lo.astype(np.int32)
    multi = f['multi'].astype(np.int32) if 'multi' in f else lm.astype(np.int32)
    return types.SimpleNamespace(num_pix=int(np.asarray(f['num_pix']).ravel()[0]), ones=ones,
                                 multi=multi, place_ones=po, place_multi=pm, count_multi=cm,
                                 bg=f.get('bg'))
