"""Seeded synthetic holographic SPI frames in the reference's sparse-photon layout.

The workload recipe (Gaussian-blob object, spherical reference, in-plane rotation,
Poisson sampling) follows the reference generator's forward model (make_data.py:199-333:
hologram -> beamstop/edge mask -> normalise to the mean count -> Poisson -> sparse
``place_ones``/``place_multi``/``count_multi``), with the object and fluence recipes of
the benchmark workloads:

* object: ``n_blobs`` Gaussian blobs (sigma 1.5 px, amplitudes U(0.5, 1)) at uniform
  positions in a disk of ``disk_radius`` px inside a ``2*size`` px real-space box; its
  Fourier transform is evaluated analytically on the central ``size x size`` q-grid
  (identical to FFT-then-crop up to the negligible aliasing of a 1.5 px Gaussian);
* rotation: the object is rotated by U[0, 2pi) per frame (F_obj(Rq));
* reference: a uniform sphere, diameter ~ N(dia_mean, dia_sd) px and shift
  U(-shift_max, shift_max) px per axis, in slice_gen_holo's convention
  (kernels.cu:58-69), scaled to carry ``ref_factor`` times the object's photons;
* fluence: constant, or log-normal with unit mean and the given sigma.

Two back ends share the recipe: NumPy (small parity inputs, bit-reproducible across
machines for a given seed) and torch (large inputs on the GPU, its own random stream).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class SynthConfig:
    size: int = 64
    num_data: int = 2000
    mean_count: float = 150.
    n_blobs: int = 40
    disk_radius: float = 12.
    sigma: float = 1.5
    beamstop: float = 3.
    dia_mean: float = 8.
    dia_sd: float = 0.4
    shift_max: float = 4.
    ref_factor: float = 5.
    fluence: str = 'constant'          # or 'lognormal'
    fluence_sigma: float = 0.5
    seed: int = 0
    class_weights: tuple = ()          # () = one class; else the mixing fractions

    @property
    def box(self):
        return 2 * self.size


# Benchmark workloads (frames x q-grid x rotations); see the bench and tests.
CONFIG_CPU64 = SynthConfig()
CONFIG_185 = SynthConfig(size=185, num_data=100_000, mean_count=300., n_blobs=120,
                         disk_radius=30., beamstop=4.)
CONFIG_256 = SynthConfig(size=256, num_data=200_000, mean_count=400., n_blobs=120,
                         disk_radius=30., beamstop=4., fluence='lognormal',
                         fluence_sigma=0.5)


@dataclass
class SparseFrames:
    """Flat sparse-photon arrays (dset.py:23-60 layout) plus generator truth."""
    num_pix: int
    ones: np.ndarray
    multi: np.ndarray
    place_ones: np.ndarray
    place_multi: np.ndarray
    count_multi: np.ndarray
    bg: np.ndarray | None = None
    truth: dict = field(default_factory=dict)

    @property
    def num_data(self):
        return len(self.ones)


def _blobs(cfg: SynthConfig, rng):
    """Blob centres (relative to the box centre) and amplitudes."""
    n = cfg.n_blobs
    r = cfg.disk_radius * np.sqrt(rng.random(n))
    phi = 2 * np.pi * rng.random(n)
    cen = np.stack([r * np.cos(phi), r * np.sin(phi)], 1)
    amp = 0.5 + 0.5 * rng.random(n)
    return cen, amp


def _class_blobs(cfg: SynthConfig, rng):
    """Blobs of every class: class 0 from the main stream ``rng`` (so one class equals the
    single-class generator), class c > 0 from its own stream seeded (seed, c)."""
    out = [_blobs(cfg, rng)]
    for c in range(1, max(1, len(cfg.class_weights))):
        out.append(_blobs(cfg, np.random.default_rng([cfg.seed, c])))
    return out


def _class_labels(cfg: SynthConfig, n):
    """Per-frame class (int) drawn with the mixing fractions from stream (seed, 'class')."""
    if len(cfg.class_weights) <= 1:
        return np.zeros(n, dtype='i4')
    w = np.asarray(cfg.class_weights, dtype='f8')
    rng = np.random.default_rng([cfg.seed, 0x636c])
    return rng.choice(len(w), size=n, p=w / w.sum()).astype('i4')


def _grid(size):
    x = np.arange(size, dtype='f8') - size // 2
    return x


def _reference_ft(cfg: SynthConfig, dia, shx, shy):
    """Sphere FT times shift ramp in slice_gen_holo's convention (kernels.cu:58-69)."""
    size = cfg.size
    g = _grid(size)
    xx, yy = np.meshgrid(g, g, indexing='ij')
    rad = np.sqrt(xx * xx + yy * yy)
    s = rad[None] * np.pi * dia[:, None, None] / size
    s[s == 0] = 1.e-5
    sph = (np.sin(s) - s * np.cos(s)) / (s * s * s)
    phase = 2 * np.pi * (xx[None] * shx[:, None, None] + yy[None] * shy[:, None, None]) / size
    return sph * np.exp(1j * phase)


def detector_mask(size, beamstop):
    g = _grid(size)
    xx, yy = np.meshgrid(g, g, indexing='ij')
    rad = np.sqrt(xx * xx + yy * yy)
    return ((rad >= beamstop) & (rad < size // 2)).astype('f8')


def _object_ft_of(cfg: SynthConfig, cen, amp):
    """Unrotated transform of blobs (cen, amp) on the q-grid, (size, size) complex."""
    q = _grid(cfg.size) / cfg.box
    qq = q[:, None] ** 2 + q[None, :] ** 2
    env = 2 * np.pi * cfg.sigma ** 2 * np.exp(-2 * np.pi ** 2 * cfg.sigma ** 2 * qq)
    A = np.exp(-2j * np.pi * q[:, None] * cen[None, :, 0]) * amp[None, :]
    B = np.exp(-2j * np.pi * q[:, None] * cen[None, :, 1])
    return env * (A @ B.T)


def _amplitudes(cfg: SynthConfig, blobs):
    """Global amplitudes of the forward model: alpha_c scales class c's object transform
    so that it carries mean_count / (1 + ref_factor) photons over the detector mask, and
    rel scales the reference sphere so that one of the mean diameter carries ref_factor
    times that (one constant for all frames, as the reference generator's rel_scale,
    make_data.py:312)."""
    dmask = detector_mask(cfg.size, cfg.beamstop)
    pobj = cfg.mean_count / (1. + cfg.ref_factor)
    alpha = np.array([np.sqrt(pobj / (np.abs(_object_ft_of(cfg, c, a)) ** 2 * dmask).sum())
                      for c, a in blobs])
    sph = _reference_ft(cfg, np.array([cfg.dia_mean]), np.zeros(1), np.zeros(1))[0]
    rel = np.sqrt(cfg.ref_factor * pobj / (np.abs(sph) ** 2 * dmask).sum())
    return alpha, float(rel)


def make_frames(cfg: SynthConfig, batch: int = 256) -> SparseFrames:
    """NumPy generator (deterministic for a given seed)."""
    rng = np.random.default_rng(cfg.seed)
    size, n = cfg.size, cfg.num_data
    blobs = _class_blobs(cfg, rng)
    angles = 2 * np.pi * rng.random(n)
    dia = cfg.dia_mean + cfg.dia_sd * rng.standard_normal(n)
    shifts = rng.uniform(-cfg.shift_max, cfg.shift_max, (n, 2))
    if cfg.fluence == 'lognormal':
        fl = np.exp(cfg.fluence_sigma * rng.standard_normal(n) - 0.5 * cfg.fluence_sigma ** 2)
    else:
        fl = np.ones(n)
    q = _grid(size) / cfg.box                       # cycles per real-space pixel
    qq = q[:, None] ** 2 + q[None, :] ** 2
    env = 2 * np.pi * cfg.sigma ** 2 * np.exp(-2 * np.pi ** 2 * cfg.sigma ** 2 * qq)
    dmask = detector_mask(size, cfg.beamstop)
    labels = _class_labels(cfg, n)
    cen_all = np.stack([b[0] for b in blobs])          # (C, n_blobs, 2)
    amp_all = np.stack([b[1] for b in blobs])          # (C, n_blobs)
    alpha, rel = _amplitudes(cfg, blobs)
    scale = np.empty(n)
    ones, multi, po, pm, cm = [], [], [], [], []
    for b0 in range(0, n, batch):
        b1 = min(n, b0 + batch)
        ca, sa = np.cos(angles[b0:b1]), np.sin(angles[b0:b1])
        cen = cen_all[labels[b0:b1]]                   # (b, n_blobs, 2)
        amp = amp_all[labels[b0:b1]]
        # R^T c for each frame: centres rotated by -angle
        u = ca[:, None] * cen[:, :, 0] + sa[:, None] * cen[:, :, 1]
        v = -sa[:, None] * cen[:, :, 0] + ca[:, None] * cen[:, :, 1]
        A = np.exp(-2j * np.pi * q[None, :, None] * u[:, None, :]) * amp[:, None, :]
        B = np.exp(-2j * np.pi * q[None, :, None] * v[:, None, :])
        fobj = alpha[labels[b0:b1]][:, None, None] * env[None] * \
            np.matmul(A, B.transpose(0, 2, 1))
        fref = rel * _reference_ft(cfg, dia[b0:b1], shifts[b0:b1, 0], shifts[b0:b1, 1])
        inten = np.abs(fobj + fref) ** 2 * dmask
        # each frame carries mean_count * fluence photons (make_data.py:319)
        scale[b0:b1] = cfg.mean_count * fl[b0:b1] / inten.sum((1, 2))
        inten *= scale[b0:b1, None, None]
        frames = rng.poisson(inten).reshape(b1 - b0, -1)
        for f in frames:
            o = np.nonzero(f == 1)[0]
            m = np.nonzero(f > 1)[0]
            ones.append(len(o))
            multi.append(len(m))
            po.append(o)
            pm.append(m)
            cm.append(f[m])
    cat = lambda l: np.concatenate(l).astype('i4') if l else np.zeros(0, 'i4')
    return SparseFrames(size * size, np.array(ones, 'i4'), np.array(multi, 'i4'), cat(po),
                        cat(pm), cat(cm), None,
                        dict(angles=angles, diameters=dia, shifts=shifts, fluence=fl,
                             scale=scale, blob_centres=cen_all[0], blob_amps=amp_all[0],
                             classes=labels, alpha=alpha, rel=rel))


def make_frames_torch(cfg: SynthConfig, device='cuda', batch: int = 1024):
    """torch generator for large workloads (float32 intensities, torch RNG stream).

    Returns a dict of int32/int64 torch tensors on ``device`` in the flat sparse layout
    (ones, multi, place_ones, place_multi, count_multi) plus truth tensors.
    """
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(cfg.seed)
    size, n = cfg.size, cfg.num_data
    f32 = torch.float32
    rng = np.random.default_rng(cfg.seed)
    blobs = _class_blobs(cfg, rng)
    cen_all = torch.tensor(np.stack([b[0] for b in blobs]), dtype=f32, device=device)
    amp_all = torch.tensor(np.stack([b[1] for b in blobs]), dtype=f32, device=device)
    labels = torch.as_tensor(_class_labels(cfg, n), device=device, dtype=torch.int64)
    alpha_np, rel = _amplitudes(cfg, blobs)
    alpha = torch.tensor(alpha_np, dtype=f32, device=device)
    scale_l = []
    angles = 2 * np.pi * torch.rand(n, generator=gen, device=device, dtype=torch.float64)
    dia = cfg.dia_mean + cfg.dia_sd * torch.randn(n, generator=gen, device=device,
                                                  dtype=torch.float64)
    shifts = (torch.rand(n, 2, generator=gen, device=device, dtype=torch.float64) * 2 - 1) \
        * cfg.shift_max
    if cfg.fluence == 'lognormal':
        fl = torch.exp(cfg.fluence_sigma * torch.randn(n, generator=gen, device=device,
                                                       dtype=torch.float64)
                       - 0.5 * cfg.fluence_sigma ** 2)
    else:
        fl = torch.ones(n, device=device, dtype=torch.float64)
    g = torch.arange(size, device=device, dtype=f32) - size // 2
    q = g / cfg.box
    qq = q[:, None] ** 2 + q[None, :] ** 2
    env = 2 * np.pi * cfg.sigma ** 2 * torch.exp(-2 * np.pi ** 2 * cfg.sigma ** 2 * qq)
    xx, yy = torch.meshgrid(g, g, indexing='ij')
    rad = torch.sqrt(xx * xx + yy * yy)
    dmask = ((rad >= cfg.beamstop) & (rad < size // 2)).to(f32)
    ones_l, multi_l, po_l, pm_l, cm_l = [], [], [], [], []
    for b0 in range(0, n, batch):
        b1 = min(n, b0 + batch)
        a = angles[b0:b1].to(f32)
        ca, sa = torch.cos(a), torch.sin(a)
        cen = cen_all[labels[b0:b1]]
        amp = amp_all[labels[b0:b1]][:, None, :]
        u = ca[:, None] * cen[:, :, 0] + sa[:, None] * cen[:, :, 1]
        v = -sa[:, None] * cen[:, :, 0] + ca[:, None] * cen[:, :, 1]
        pa = -2 * np.pi * q[None, :, None] * u[:, None, :]
        pb = -2 * np.pi * q[None, :, None] * v[:, None, :]
        Ar, Ai = torch.cos(pa) * amp, torch.sin(pa) * amp
        Br = torch.cos(pb).transpose(1, 2).contiguous()
        Bi = torch.sin(pb).transpose(1, 2).contiguous()
        al = alpha[labels[b0:b1]][:, None, None] * env
        fr = al * (torch.bmm(Ar, Br) - torch.bmm(Ai, Bi))
        fi = al * (torch.bmm(Ar, Bi) + torch.bmm(Ai, Br))
        d = dia[b0:b1].to(f32)[:, None, None]
        s = rad[None] * np.pi * d / size
        s = torch.where(s == 0, torch.full_like(s, 1.e-5), s)
        sph = (torch.sin(s) - s * torch.cos(s)) / (s * s * s)
        sh = shifts[b0:b1].to(f32)
        ph = 2 * np.pi * (xx[None] * sh[:, 0, None, None] + yy[None] * sh[:, 1, None, None]) / size
        rr, ri = rel * sph * torch.cos(ph), rel * sph * torch.sin(ph)
        inten = ((fr + rr) ** 2 + (fi + ri) ** 2) * dmask
        sc = cfg.mean_count * fl[b0:b1].to(f32) / inten.sum((1, 2))
        scale_l.append(sc.to(torch.float64))
        inten = inten * sc[:, None, None]
        k = torch.poisson(inten, generator=gen).reshape(b1 - b0, -1).to(torch.int32)
        one = k == 1
        mul = k > 1
        ones_l.append(one.sum(1, dtype=torch.int32))
        multi_l.append(mul.sum(1, dtype=torch.int32))
        po_l.append(torch.nonzero(one)[:, 1].to(torch.int32))
        mi = torch.nonzero(mul)
        pm_l.append(mi[:, 1].to(torch.int32))
        cm_l.append(k[mi[:, 0], mi[:, 1]])
    return dict(num_pix=size * size, ones=torch.cat(ones_l), multi=torch.cat(multi_l),
                place_ones=torch.cat(po_l), place_multi=torch.cat(pm_l),
                count_multi=torch.cat(cm_l), angles=angles, diameters=dia, shifts=shifts,
                fluence=fl, scale=torch.cat(scale_l), classes=labels.to(torch.int32),
                alpha=alpha_np, rel=rel)


PRESETS = {'cpu64': CONFIG_CPU64, '185': CONFIG_185, '256': CONFIG_256}


def object_ft(cfg: SynthConfig, cls: int = 0):
    """The generator's unrotated transform of object class ``cls`` on the q-grid
    (complex, size*size, flat index x*size + y), before its amplitude factor, from the same
    blob draws as :func:`make_frames` / :func:`make_frames_torch` for ``cfg.seed``."""
    rng = np.random.default_rng(cfg.seed)
    cen, amp = _class_blobs(cfg, rng)[cls]
    return _object_ft_of(cfg, cen, amp).ravel()


def model_scales(cfg: SynthConfig, cls: int = 0):
    """(alpha_cls, rel): the generator's amplitude of class ``cls``'s object transform and
    of the reference sphere.  With model = alpha * object_ft(cfg, cls), rel_scale = rel and
    per-frame scales = truth['scale'], the rotation form's W is the generator's intensity
    (up to the interpolation of the rotated model)."""
    rng = np.random.default_rng(cfg.seed)
    alpha, rel = _amplitudes(cfg, _class_blobs(cfg, rng))
    return float(alpha[cls]), rel


def solution(cfg: SynthConfig, cls: int = 0):
    """The generator's object as a real-space size x size image (the ``solution`` dataset
    of make_data.py:164-167): the real part of the centred inverse transform of its scaled
    q-space transform."""
    f = (model_scales(cfg, cls)[0] * object_ft(cfg, cls)).reshape(cfg.size, cfg.size)
    return np.real(np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(f))))


def write_photons_file(path, cfg: SynthConfig):
    """Generate ``cfg`` with the NumPy back end and write a photons file in the reference
    layout (make_data.py:228-275: sparse photons, ``num_pix``, ``scale``, ``true_*``, and the
    object as ``solution``, make_data.py:164-167)."""
    from . import h5io
    fr = make_frames(cfg)
    t = fr.truth
    h5io.write_photons(path, fr, extra={'scale': t['fluence'], 'true_shifts': t['shifts'],
                                        'true_diameters': t['diameters'],
                                        'true_angles': t['angles'],
                                        'solution': solution(cfg)})
    return fr


def main(argv=None):
    import argparse
    import dataclasses
    ap = argparse.ArgumentParser(description='Synthetic holographic SPI photons file')
    ap.add_argument('out', help='output HDF5 photons file')
    ap.add_argument('--preset', choices=sorted(PRESETS), default='cpu64')
    ap.add_argument('--num-data', type=int, default=None)
    ap.add_argument('--seed', type=int, default=None)
    a = ap.parse_args(argv)
    cfg = PRESETS[a.preset]
    if a.num_data is not None:
        cfg = dataclasses.replace(cfg, num_data=a.num_data)
    if a.seed is not None:
        cfg = dataclasses.replace(cfg, seed=a.seed)
    fr = write_photons_file(a.out, cfg)
    print('%d frames, %.1f photons/frame -> %s' %
          (fr.num_data, (fr.place_ones.size + fr.count_multi.sum()) / fr.num_data, a.out))


if __name__ == '__main__':
    main()
