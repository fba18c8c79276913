"""CPU restatement of the transform stages (numpy).

TEST INFRASTRUCTURE: this module is a checker for the HIP path.  Only
tests/ may import it.

  zigzag / negabinary   modules/transforms/zigzag/zigzag.h:70-80,
                        modules/transforms/negabinary/negabinary.h
  diff_fwd / diff_inv   modules/predictors/diff/diff.cu:22-45 (forward),
                        :75-105 + :148-162 (inverse: cub::BlockScan<T, 512>
                        per 512-element tile plus the running tile prefix)
  lorenzo_int_*         modules/predictors/lorenzo/lorenzo_stage.cu:14-357

The inverse difference's floating-point association is that of CUB's
raking BlockScan (a dependency absent from the reference tree: the CUB
shipped with the CUDA toolkit the reference builds against; restated from
its published algorithm, block_scan_raking.cuh): 32 raking threads each sum
16 consecutive elements left to right, a warp Kogge-Stone scan (offsets
1, 2, 4, 8, 16; the lower lane's value on the left) combines the 32 partial
sums, and each raking thread adds its 16 elements left to right onto the
exclusive partial (segment 0: onto nothing).  The reference then adds the
running sum of the chunk's earlier tile aggregates (diff.cu:97-102).
Integer arithmetic wraps (numpy fixed-width dtypes).
"""
from __future__ import annotations

import numpy as np

_UNSIGNED = {np.dtype(np.int8): np.uint8, np.dtype(np.int16): np.uint16, np.dtype(np.int32): np.uint32,
             np.dtype(np.int64): np.uint64}


def _u(dt):
    return np.dtype(_UNSIGNED[np.dtype(dt)])


def zigzag_enc(x: np.ndarray) -> np.ndarray:
    """zigzag.h:70: (U(x) << 1) ^ U(x >> (W-1))."""
    w = 8 * x.dtype.itemsize
    u = _u(x.dtype)
    return ((x.astype(u) << u.type(1)) ^ (x >> (w - 1)).astype(u)).astype(u)


def zigzag_dec(u: np.ndarray, dtype) -> np.ndarray:
    """zigzag.h:80: (u >> 1) ^ U(-(S)(u & 1))."""
    s = np.dtype(dtype)
    one = u.dtype.type(1)
    return ((u >> one) ^ (-(u & one).astype(s)).astype(u.dtype)).astype(s)


def _mask(u):
    return u.type(int("AA" * u.itemsize, 16))


def negabinary_enc(x: np.ndarray) -> np.ndarray:
    """negabinary.h: (U(n) + M) ^ M, M = 0xAA..AA."""
    u = _u(x.dtype)
    m = _mask(u)
    return ((x.astype(u) + m) ^ m).astype(u)


def negabinary_dec(u: np.ndarray, dtype) -> np.ndarray:
    """negabinary.h: S((u ^ M) - M)."""
    m = _mask(u.dtype)
    return ((u ^ m) - m).astype(np.dtype(dtype))


def diff_fwd(x: np.ndarray, chunk_elems: int = 0, negabinary: bool = False) -> np.ndarray:
    """diff.cu:22-45: out[i] = in[i] - in[i-1]; chunk starts (and 0) as is."""
    d = x.copy()
    if x.size > 1:
        with np.errstate(over="ignore"):
            d[1:] = x[1:] - x[:-1]
    if chunk_elems:
        d[::chunk_elems] = x[::chunk_elems]
    elif x.size:
        d[0] = x[0]
    return negabinary_enc(d) if negabinary else d


def _raking_tile_scan(t: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Inclusive scans of tiles t[k, 512] in the raking association;
    returns (scan, aggregate per tile)."""
    seg = t.reshape(t.shape[0], 32, 16)
    part = seg[:, :, 0].copy()
    for j in range(1, 16):
        part = part + seg[:, :, j]
    inc = part.copy()
    for o in (1, 2, 4, 8, 16):
        nxt = inc.copy()
        nxt[:, o:] = inc[:, :-o] + inc[:, o:]
        inc = nxt
    out = np.empty_like(seg)
    run = seg[:, :, 0].copy()
    run[:, 1:] = inc[:, :-1] + seg[:, 1:, 0]
    out[:, :, 0] = run
    for j in range(1, 16):
        run = run + seg[:, :, j]
        out[:, :, j] = run
    return out.reshape(t.shape[0], 512), inc[:, 31].copy()


def diff_inv(d: np.ndarray, dtype, chunk_elems: int = 0, negabinary: bool = False) -> np.ndarray:
    """Cumulative sum per chunk (diff.cu:75-162) in the reference's association."""
    x = negabinary_dec(d, dtype) if negabinary else d.astype(np.dtype(dtype))
    n = x.size
    out = np.empty_like(x)
    if n == 0:
        return out
    if np.issubdtype(x.dtype, np.integer):
        c = chunk_elems or n
        for b in range(0, n, c):
            out[b:b + c] = np.cumsum(x[b:b + c], dtype=x.dtype)
        return out
    c = chunk_elems or n
    zero = x.dtype.type(0)
    for b in range(0, n, c):
        part = x[b:b + c]
        m = part.size
        nt = -(-m // 512)
        tiles = np.zeros(nt * 512, x.dtype)
        tiles[:m] = part
        scan, agg = _raking_tile_scan(tiles.reshape(nt, 512))
        prefix = zero
        res = np.empty(nt * 512, x.dtype)
        for k in range(nt):
            res[512 * k:512 * k + 512] = scan[k] + prefix
            prefix = prefix + agg[k]
        out[b:b + m] = res[:m]
    return out


def lorenzo_int_fwd(x: np.ndarray, dims) -> np.ndarray:
    """lorenzo_stage.cu: 1D deltas in 256-element blocks; 2D/3D inclusion-
    exclusion over the whole field with zeros outside."""
    nx, ny, nz = dims
    dt = x.dtype
    w = x.astype(np.int64)
    if ny == 1 and nz == 1:
        prev = np.zeros_like(w)
        prev[1:] = w[:-1]
        prev[::256] = 0
        return (w - prev).astype(dt)
    a = w[: nx * ny * nz].reshape(nz, ny, nx)
    p = np.zeros((nz + 1, ny + 1, nx + 1), np.int64)
    p[1:, 1:, 1:] = a
    if nz == 1:
        q = p[1]
        r = q[1:, 1:] - q[1:, :-1] - q[:-1, 1:] + q[:-1, :-1]
    else:
        r = (p[1:, 1:, 1:] - p[1:, 1:, :-1] - p[1:, :-1, 1:] - p[:-1, 1:, 1:]
             + p[1:, :-1, :-1] + p[:-1, 1:, :-1] + p[:-1, :-1, 1:] - p[:-1, :-1, :-1])
    out = np.zeros_like(x)
    out[: nx * ny * nz] = r.reshape(-1).astype(dt)
    return out


def lorenzo_int_inv(d: np.ndarray, dims) -> np.ndarray:
    """Inclusive prefix sums per 256-element block (1D) or along x, y, z."""
    nx, ny, nz = dims
    dt = d.dtype
    w = d.astype(np.int64)
    if ny == 1 and nz == 1:
        out = np.empty_like(w)
        for b in range(0, w.size, 256):
            out[b:b + 256] = np.cumsum(w[b:b + 256])
        return out.astype(dt)
    a = w[: nx * ny * nz].reshape(nz, ny, nx)
    a = np.cumsum(np.cumsum(np.cumsum(a, axis=2), axis=1), axis=0)
    out = np.zeros_like(d)
    out[: nx * ny * nz] = a.reshape(-1).astype(dt)
    return out
