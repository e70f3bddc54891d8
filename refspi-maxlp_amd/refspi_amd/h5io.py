"""HDF5 files of the reference: photons input (make_data.py:228-251) and output_NNN.h5."""
from __future__ import annotations


def read_photons(path):
    raise NotImplementedError('photons HDF5 reader not available yet')


def write_fields(path, fields):
    raise NotImplementedError('output HDF5 writer not available yet')
