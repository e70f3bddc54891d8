"""Sparse-photon dataset (mirror of the reference's dset.py:5-83 ``Dataset``).

Holds the flat ``ones``/``multi``/``place_ones``/``place_multi``/``count_multi`` layout on
the host (NumPy) -- the GPU plans copy what they need into their own device layouts.
"""
from __future__ import annotations

import numpy as np


class Dataset:
    """dset.py:16-61.  Build from a photons HDF5 file or from in-memory arrays.

    Args:
        photons_file: path of a photons HDF5 file (make_data.py:228-251 layout), or None
        num_pix: expected number of pixels (AttributeError on mismatch, dset.py:24-25)
        need_scaling: compute per-frame photon totals ``counts`` (dset.py:50-53)
        frames: alternatively a ``synth.SparseFrames``-like object with the flat arrays
    """

    def __init__(self, photons_file=None, num_pix=None, need_scaling=False, frames=None):
        self.powder = None
        self.photons_file = photons_file
        if frames is None:
            from . import h5io
            frames = h5io.read_photons(photons_file)
        if num_pix is not None and int(frames.num_pix) != int(num_pix):
            raise AttributeError('Number of pixels in photons file does not match')
        self.num_pix = int(frames.num_pix)
        self.ones = np.ascontiguousarray(frames.ones, dtype='i4')
        self.multi = np.ascontiguousarray(frames.multi, dtype='i4')
        self.num_data = len(self.ones)
        self.place_ones = np.ascontiguousarray(frames.place_ones, dtype='i4')
        self.place_multi = np.ascontiguousarray(frames.place_multi, dtype='i4')
        self.count_multi = np.ascontiguousarray(frames.count_multi, dtype='i4')
        if len(self.place_ones) != int(self.ones.sum()) or \
                len(self.place_multi) != int(self.multi.sum()) or \
                len(self.count_multi) != len(self.place_multi):
            raise ValueError('inconsistent sparse photon arrays')
        # dset.py:32-33, 41-42 (exclusive prefix offsets)
        self.ones_accum = np.zeros(self.num_data, dtype='i8')
        self.multi_accum = np.zeros(self.num_data, dtype='i8')
        if self.num_data > 1:
            self.ones_accum[1:] = np.cumsum(self.ones, dtype='i8')[:-1]
            self.multi_accum[1:] = np.cumsum(self.multi, dtype='i8')[:-1]
        # dset.py:46-48
        self.mean_count = float((self.place_ones.shape[0] + self.count_multi.sum())
                                / self.num_data)
        if need_scaling:
            self.counts = self.photon_counts()
        bg = getattr(frames, 'bg', None)
        self.bg = np.zeros(self.num_pix) if bg is None else np.asarray(bg, 'f8').ravel()
        self.mem = sum(a.nbytes for a in (self.ones, self.multi, self.place_ones,
                                          self.place_multi, self.count_multi))

    def photon_counts(self):
        """dset.py:50-53: total photons per frame (f8)."""
        c = self.ones.astype('f8')
        fm = np.repeat(np.arange(self.num_data), self.multi)
        np.add.at(c, fm, self.count_multi)
        return c

    def get_frame(self, num):
        """dset.py:63-72: dense photon counts of one frame."""
        o_a, m_a = self.ones_accum[num], self.multi_accum[num]
        frame = np.zeros(self.num_pix, dtype='i4')
        frame[self.place_ones[o_a:o_a + self.ones[num]]] += 1
        frame[self.place_multi[m_a:m_a + self.multi[num]]] += \
            self.count_multi[m_a:m_a + self.multi[num]]
        return frame

    def get_powder(self):
        """dset.py:74-83: mean photons per pixel."""
        if self.powder is not None:
            return self.powder
        powder = np.zeros(self.num_pix)
        np.add.at(powder, self.place_ones, 1)
        np.add.at(powder, self.place_multi, self.count_multi)
        powder /= self.num_data
        self.powder = powder
        return powder
