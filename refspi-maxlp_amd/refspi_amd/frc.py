"""Fourier ring correlation of two 2-D objects (the reference's quality metric, calc_frc.py).

Same interface as the reference's ``FRC`` (calc_frc.py:5-140): objects are given in real
space (real arrays) or as centred Fourier transforms (complex arrays); ``calc`` returns
the per-ring correlation Re sum F1 conj(F2) / sqrt(sum |F1|^2 sum |F2|^2) over integer
radial bins of width ``binsize`` (calc_frc.py:33-75), and ``calc_rot`` searches in-plane
rotations of the second object (linear interpolation about the array centre, the
scipy.ndimage.rotate call of calc_frc.py:99-102) for the largest mean FRC over
[radmin, centre) (calc_frc.py:77-117).  Host-side validation, not part of the per-iteration
path.
"""
from __future__ import annotations

import sys

import numpy as np


def _centred_fft(obj):
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(obj)))


def _centred_ifft_real(fobj):
    return np.real(np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(fobj))))


class FRC:
    """calc_frc.py:5-140."""

    def __init__(self, obj1=None, obj2=None, verbose=False):
        self.verbose = verbose
        self.obj1 = self.obj2 = self.fobj1 = self.fobj2 = None
        self.binrad = None
        if obj1 is not None:
            self.set_obj1(obj1)
        if obj2 is not None:
            self.set_obj2(obj2)

    def set_obj1(self, obj):
        self.obj1, self.fobj1 = self._pair(obj)

    def set_obj2(self, obj):
        self.obj2, self.fobj2 = self._pair(obj)

    @staticmethod
    def _pair(obj):
        """(real-space object, centred transform) from either representation."""
        obj = np.asarray(obj)
        if np.iscomplexobj(obj):
            return _centred_ifft_real(obj), obj
        return obj, _centred_fft(obj)

    def _rings(self, binsize):
        n0, n1 = self.obj1.shape
        x = np.arange(n0, dtype='f8')[:, None] - n0 // 2
        y = np.arange(n1, dtype='f8')[None, :] - n1 // 2
        self.binrad = (np.sqrt(x * x + y * y) / binsize).astype('i4')
        return int(self.binrad.max()) + 2

    def calc(self, binsize=1., do_abs=False, rsize=None, fobj2=None):
        """(radii, FRC) between obj1 and obj2 (or the transform ``fobj2``)."""
        if self.obj1 is None or self.obj2 is None:
            raise AttributeError('Need to first set obj1 and obj2 to be compared')
        if self.obj1.shape != self.obj2.shape:
            raise ValueError('Both objects must have same shape: %s vs %s' %
                             (self.obj1.shape, self.obj2.shape))
        if rsize is None:
            rsize = self._rings(binsize)
        f2 = self.fobj2 if fobj2 is None else fobj2
        b = self.binrad.ravel()
        cross = self.fobj1.ravel() * np.conj(f2.ravel())
        num = np.bincount(b, cross.real, rsize) + 1j * np.bincount(b, cross.imag, rsize)
        den = np.sqrt(np.bincount(b, np.abs(self.fobj1.ravel()) ** 2, rsize) *
                      np.bincount(b, np.abs(f2.ravel()) ** 2, rsize))
        out = np.zeros(rsize)
        ok = den > 0
        out[ok] = (np.abs(num[ok]) if do_abs else num[ok].real) / den[ok]
        return np.arange(rsize, dtype='f8') * binsize, out

    def calc_rot(self, binsize=1., num_rot=1800, do_abs=False, angle=None, radmin=0):
        """Best FRC over rotations of obj2 by angle in degrees (all of 360 * i / num_rot, or
        the given ``angle``).  Returns (radii, FRC); the best angle is kept in
        ``best_angle``."""
        from scipy import ndimage
        cen = self.fobj2.shape[-1] // 2
        rsize = self._rings(binsize)
        angles = [angle] if angle is not None else np.arange(num_rot) / num_rot * 360.
        best = (-np.inf, None, None)
        rot = np.empty_like(self.fobj2)
        for i, a in enumerate(angles):
            rot.real = ndimage.rotate(self.fobj2.real, a, order=1, prefilter=False, reshape=False)
            rot.imag = ndimage.rotate(self.fobj2.imag, a, order=1, prefilter=False, reshape=False)
            rvals, cc = self.calc(binsize, do_abs, rsize=rsize, fobj2=rot)
            score = cc[radmin:cen].mean()
            if score > best[0]:
                best = (score, a, cc)
            if self.verbose:
                sys.stderr.write('\r%d/%d' % (i + 1, len(angles)))
        if self.verbose:
            sys.stderr.write('\n')
        self.best_angle = best[1]
        print('Best correlation for %.3f degrees' % best[1])
        return np.arange(rsize, dtype='f8') * binsize, best[2]


def model_frc(model, solution, num_rot=360, radmin=4):
    """Quality of a reconstruction: the FRC of ``model`` (the complex q-space model, flat
    or square) against ``solution`` (real-space object or centred transform), with the best
    of ``num_rot`` in-plane rotations (calc_frc.py:77-117).  Returns (mean FRC over rings
    radmin .. size//2, per-ring FRC, best angle in degrees)."""
    import contextlib
    import io
    sol = np.asarray(solution)
    size = sol.shape[-1]
    f = FRC(np.asarray(model).reshape(size, size), sol)
    with contextlib.redirect_stdout(io.StringIO()):
        _, cc = f.calc_rot(num_rot=num_rot, radmin=radmin)
    cen = size // 2
    return float(cc[radmin:cen].mean()), cc[:cen], float(f.best_angle)
