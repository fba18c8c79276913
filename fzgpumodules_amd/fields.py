"""Seeded synthetic scientific fields for the benchmark workloads.

All generators work in float64 and cast to float32 at the end; the random
draws come from numpy's PCG64 ``default_rng(seed)``, so a given seed gives the
same field on every machine with this image.

Workloads (sizes and recipes):
  sinusoid_1d      float32[2^24]: sum of 8 sinusoids (freqs U[1e-5,1e-2] cycles/
                   sample, amplitudes U[0.1,1], phases U[0,2pi]) + N(0, 1e-3^2)
  cloud_2d         float32[1800][3600]: Gaussian random field with power ~ k^-3,
                   affinely mapped so the 30th percentile -> 0 and max -> 1, then
                   clamped to [0,1] (about 30% exact zeros)
  turbulence_3d    float32[512][512][512]: Gaussian random field with power
                   ~ k^-5/3 plus 1% seeded spikes multiplied by 100
  power_law_field  float32 field of any shape with power ~ k^expo
"""
from __future__ import annotations

import hashlib

import numpy as np

BASE_SEED = 20261202


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sinusoid_1d(n: int = 1 << 24, seed: int = BASE_SEED + 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    f = rng.uniform(1e-5, 1e-2, 8)
    a = rng.uniform(0.1, 1.0, 8)
    ph = rng.uniform(0.0, 2 * np.pi, 8)
    i = np.arange(n, dtype=np.float64)
    x = np.zeros(n, dtype=np.float64)
    for k in range(8):
        x += a[k] * np.sin(2 * np.pi * f[k] * i + ph[k])
    x += rng.normal(0.0, 1e-3, n)
    return x.astype(np.float32)


def power_law_field(shape, expo: float, seed: int) -> np.ndarray:
    """Gaussian random field with isotropic power spectrum ~ k^expo (DC = 0),
    normalised to unit standard deviation (float64)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    F = np.fft.rfftn(noise)
    k2 = None
    for ax, n in enumerate(shape):
        if ax == len(shape) - 1:
            kk = np.fft.rfftfreq(n)
        else:
            kk = np.fft.fftfreq(n)
        sh = [1] * len(shape)
        sh[ax] = kk.size
        kk = kk.reshape(sh) ** 2
        k2 = kk if k2 is None else k2 + kk
    k2[(0,) * len(shape)] = 1.0
    amp = k2 ** (expo / 4.0)  # |k|^(expo/2)
    amp[(0,) * len(shape)] = 0.0
    F *= amp
    del amp, k2
    g = np.fft.irfftn(F, s=shape, axes=tuple(range(len(shape))))
    sd = g.std()
    return g / (sd if sd > 0 else 1.0)


def cloud_2d(ny: int = 1800, nx: int = 3600, seed: int = BASE_SEED + 2) -> np.ndarray:
    g = power_law_field((ny, nx), -3.0, seed)
    q30 = np.quantile(g, 0.30)
    x = (g - q30) / (g.max() - q30)
    return np.clip(x, 0.0, 1.0).astype(np.float32)


def turbulence_3d(n: int = 512, seed: int = BASE_SEED + 3, spike_frac: float = 0.01) -> np.ndarray:
    g = power_law_field((n, n, n), -5.0 / 3.0, seed).astype(np.float32)
    rng = np.random.default_rng(seed + 1000)
    k = int(round(spike_frac * g.size))
    idx = rng.choice(g.size, size=k, replace=False)
    flat = g.reshape(-1)
    flat[idx] *= np.float32(100.0)
    return g


def smooth_1d(n: int) -> np.ndarray:
    """50 sin(0.01 i) + 20 cos(0.003 i), the reference tests' smooth signal
    (tests/helpers/fz_test_utils.h:184-196)."""
    i = np.arange(n, dtype=np.float64)
    return (50.0 * np.sin(i * 0.01) + 20.0 * np.cos(i * 0.003)).astype(np.float32)


# ---------------------------------------------------------------------------
# Chain-selector workload: fields of five regimes cut into 256 KiB chunks
# (2D 256 x 256, 3D 16 x 64 x 64 as (nz, ny, nx)); every chunk's bound is
# 1e-3 of its parent field's max |x|.
# ---------------------------------------------------------------------------
REGIMES = ("smooth_k4", "white_noise", "sparse80", "piecewise", "poisson20")
CHUNK_2D = (256, 256)
CHUNK_3D = (16, 64, 64)


def regime_field(regime: int, shape, seed: int) -> np.ndarray:
    """One parent field of regime `regime` (index into REGIMES)."""
    rng = np.random.default_rng(seed)
    if regime == 0:  # smooth Gaussian random field, power ~ k^-4
        g = power_law_field(shape, -4.0, seed)
    elif regime == 1:  # white noise, sigma 1
        g = rng.standard_normal(shape)
    elif regime == 2:  # 80% exact zeros, the rest N(0, 1)
        g = rng.standard_normal(shape) * (rng.random(shape) >= 0.8)
    elif regime == 3:  # piecewise constant: random axis-aligned boxes
        g = np.zeros(shape)
        for _ in range(48):
            sl = []
            for n in shape:
                a, b = sorted(rng.integers(0, n + 1, 2))
                sl.append(slice(a, max(b, a + 1)))
            g[tuple(sl)] = rng.standard_normal()
    elif regime == 4:  # integer-valued Poisson, lambda 20
        g = rng.poisson(20.0, shape).astype(np.float64)
    else:
        raise ValueError("regime must be 0..4")
    return g.astype(np.float32)


def cut_chunks(field: np.ndarray) -> np.ndarray:
    """Chunks of a 2D (ny, nx) or 3D (nz, ny, nx) field, shape (n, *chunk)."""
    cs = CHUNK_2D if field.ndim == 2 else CHUNK_3D
    grid = [s // c for s, c in zip(field.shape, cs)]
    f = field[tuple(slice(0, g * c) for g, c in zip(grid, cs))]
    if field.ndim == 2:
        f = f.reshape(grid[0], cs[0], grid[1], cs[1]).transpose(0, 2, 1, 3)
    else:
        f = f.reshape(grid[0], cs[0], grid[1], cs[1], grid[2], cs[2]).transpose(0, 2, 4, 1, 3, 5)
    return np.ascontiguousarray(f.reshape(-1, *cs))


def selector_set(n_fields: int, seed: int, dims=(2, 3), fields_shape_2d=(512, 512), fields_shape_3d=(32, 128, 128)):
    """(chunks_2d, eb_2d, regime_2d, chunks_3d, eb_3d, regime_3d): parent
    fields drawn round-robin over the regimes, cut into chunks, each chunk
    with bound 1e-3 * max |parent|."""
    out = {}
    for nd in dims:
        cs, ebs, rg = [], [], []
        for i in range(n_fields):
            r = i % len(REGIMES)
            shape = fields_shape_2d if nd == 2 else fields_shape_3d
            f = regime_field(r, shape, seed * 7919 + 31 * i + nd)
            vb = float(np.max(np.abs(f)))
            eb = np.float32(np.float32(1e-3) * np.float32(vb)) if vb > 0 else np.float32(1e-3)
            ch = cut_chunks(f)
            cs.append(ch)
            ebs.append(np.full(ch.shape[0], eb, np.float32))
            rg.append(np.full(ch.shape[0], r, np.int32))
        out[nd] = (np.concatenate(cs), np.concatenate(ebs), np.concatenate(rg))
    return out
