"""Synthetic input fields on the host (numpy).

``correlated_field`` restates generate_correlated_field_cpp
(material_generator.hpp:121-258): mt19937(seed) words -> Box-Muller normals
(one raw word per uniform, u1 redrawn while <= 1e-15, :121-128) -> ortho FFT
-> Gaussian filter exp(-2 pi^2 sigma^2 |f|^2) with f = fftfreq(n) (:160-176,
:196-219) -> inverse FFT, real part -> standardise with the population std ->
scale -> clip.  numpy's FFT is pocketfft, the reference's FFT library, and
numpy's MT19937 with legacy seeding produces std::mt19937's words.
"""
import numpy as np


def mt19937_words(seed):
    bg = np.random.MT19937()
    bg._legacy_seeding(int(seed) & 0xFFFFFFFF)
    return bg


def box_muller_normals(seed, n):
    """n draws of box_muller_normal (material_generator.hpp:121-128).  u1 is
    redrawn only for a zero word (word / 2^32 <= 1e-15 means word == 0), so
    unless a zero word comes up the draws are the word pairs in order."""
    bg = mt19937_words(seed)
    w = bg.random_raw(2 * n).astype(np.float64)
    u1 = w[0::2] / 4294967296.0
    if not (u1 <= 1e-15).any():
        u2 = w[1::2] / 4294967296.0
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return _box_muller_sequential(seed, n)


def _box_muller_sequential(seed, n):
    bg = mt19937_words(seed)
    out = np.empty(n)
    buf = bg.random_raw(2 * n + 64).astype(np.float64)
    pos = 0
    scale = 4294967296.0   # rng.max() - rng.min() + 1
    for i in range(n):
        if pos + 3 > buf.size:
            buf = np.concatenate([buf[pos:], bg.random_raw(2 * n + 64).astype(np.float64)])
            pos = 0
        u1 = buf[pos] / scale
        u2 = buf[pos + 1] / scale
        pos += 2
        while u1 <= 1e-15:
            if pos >= buf.size:
                buf = bg.random_raw(64).astype(np.float64)
                pos = 0
            u1 = buf[pos] / scale
            pos += 1
        out[i] = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return out


def correlated_field(shape, mean, std, corr, seed, clip_min=None, clip_max=None, noise=None):
    """Gaussian random field on a periodic grid, flattened in the reference's
    voxel order (x slowest).  noise (N values) replaces the seeded draws."""
    shape = tuple(int(s) for s in shape)
    N = int(np.prod(shape))
    if noise is not None:
        field = np.asarray(noise, dtype=np.float64).reshape(shape)
    else:
        field = box_muller_normals(seed, N).reshape(shape)
    hat = np.fft.fftn(field, norm="ortho")
    f2 = np.zeros(shape)
    for ax, n in enumerate(shape):
        f = np.fft.fftfreq(n)
        sh = [1] * len(shape)
        sh[ax] = n
        f2 = f2 + (f * f).reshape(sh)
    hat = hat * np.exp(-2.0 * np.pi * np.pi * corr * corr * f2)
    field = np.fft.ifftn(hat, norm="ortho").real.reshape(-1)
    m = field.mean()
    sd = np.sqrt(((field - m) ** 2).sum() / N)
    if sd < 1e-15:
        sd = 1.0
    field = mean + std * ((field - m) / sd)
    if clip_min is not None:
        field = np.maximum(field, clip_min)
    if clip_max is not None:
        field = np.minimum(field, clip_max)
    return field
