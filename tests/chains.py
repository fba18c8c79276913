"""Pipeline-level expectations composed from the oracle (test helper).

Each function restates what the reference pipeline produces for one stage
chain: the input zero-padded to the stages' alignment
(src/pipeline/compressor_exec.cpp:35-54), the error bound resolved as in
lorenzo_quant.cu:612-644, the stages applied in order, and the pipeline
outputs concatenated as [u32 n][u64 size x n] + 16-byte-aligned slots
(src/pipeline/pipeline_utils.h:36-46).  Slot order: the deepest stage's
output first (the coder stream), then the others in stage and port order
(Lorenzo's outlier_errors, outlier_indices, outlier_count); pads are zero.
"""
from __future__ import annotations

import math

import numpy as np

import oracle as orc


def align(n: int, a: int) -> int:
    return (n + a - 1) // a * a


def concat(parts) -> np.ndarray:
    parts = [np.ascontiguousarray(p).view(np.uint8).ravel() for p in parts]
    n = len(parts)
    hdr = bytearray(align(4 + 8 * n, 16))
    hdr[0:4] = np.uint32(n).tobytes()
    for i, p in enumerate(parts):
        hdr[4 + 8 * i: 12 + 8 * i] = np.uint64(p.size).tobytes()
    out = [np.frombuffer(bytes(hdr), np.uint8)]
    for p in parts:
        out.append(p)
        pad = align(p.size, 16) - p.size
        if pad:
            out.append(np.zeros(pad, np.uint8))
    return np.concatenate(out)


def split(blob: np.ndarray):
    n = int(blob[:4].view(np.uint32)[0])
    sizes = blob[4:4 + 8 * n].view(np.uint64).astype(np.int64)
    off = align(4 + 8 * n, 16)
    parts = []
    for s in sizes:
        parts.append(blob[off:off + s])
        off += align(int(s), 16)
    return parts


def resolve_eb(x_padded: np.ndarray, mode: int, user_eb: float):
    """(abs_eb, ebx2_r, ebx2) in float32 (mode 0 ABS, 1 REL, 2 NOA)."""
    eb = np.float32(user_eb)
    if mode != 0:
        vb = np.float32(orc.value_base(x_padded, mode))
        if vb > 0:
            eb = np.float32(eb * vb)
    return eb, np.float32(1) / (np.float32(2) * eb), np.float32(2) * eb


def max_outliers(n: int, cap: float) -> int:
    return math.ceil(n * float(np.float32(cap)))


def lorenzo_part(x, dims, mode, user_eb, radius, cap, zigzag, code_bytes, alignment):
    """Padded input, Lorenzo outputs and the resolved bound."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    padded = align(x.nbytes, alignment) // 4
    xp = np.zeros(padded, np.float32)
    xp[: x.size] = x
    abs_eb, r, ebx2 = resolve_eb(xp, mode, user_eb)
    nx, ny, nz = dims
    if ny == 1 and nz == 1:
        nx = padded
    m = max_outliers(padded, cap)
    codes, err, idx, cnt = orc.lorenzo_fwd(xp, (nx, ny, nz), r, radius, zigzag, code_bytes, m)
    return xp, codes, err, idx, cnt, abs_eb, ebx2


def lorenzo_bitshuffle_rze(x, dims, mode=1, user_eb=1e-3, radius=32768, cap=0.1, zigzag=True):
    """Config-2 chain: LorenzoQuant -> Bitshuffle(16 KiB, ew 2) -> RZE."""
    xp, codes, err, idx, cnt, abs_eb, ebx2 = lorenzo_part(x, dims, mode, user_eb, radius, cap, zigzag, 2, 16384)
    bs = orc.bitshuffle_enc(codes.view(np.uint8), 16384, 2)
    rze = orc.rze_enc(bs)
    blob = concat([rze, err, idx, np.array([cnt], np.uint32)])
    return blob, dict(xp=xp, codes=codes, err=err, idx=idx, cnt=cnt, abs_eb=abs_eb, ebx2=ebx2)


def lorenzo_rle(x, user_eb=1e-4, radius=32768, cap=0.2, zigzag=False):
    """Config-1 chain: 1D LorenzoQuant (ABS) -> RLE<uint16_t>."""
    xp, codes, err, idx, cnt, abs_eb, ebx2 = lorenzo_part(x, (0, 1, 1), 0, user_eb, radius, cap, zigzag, 2, 1)
    rle = orc.rle_enc(codes)
    blob = concat([rle, err, idx, np.array([cnt], np.uint32)])
    return blob, dict(xp=xp, codes=codes, err=err, idx=idx, cnt=cnt, abs_eb=abs_eb, ebx2=ebx2)


def lorenzo_huffman(x, dims, mode=1, user_eb=1e-4, radius=512, cap=0.2, zigzag=True, bklen=1024):
    """Config-3 chain: LorenzoQuant -> Huffman (coarse, bklen 2*radius)."""
    xp, codes, err, idx, cnt, abs_eb, ebx2 = lorenzo_part(x, dims, mode, user_eb, radius, cap, zigzag, 2, 1)
    huf = orc.huffman_encode(codes, bklen)
    blob = concat([huf, err, idx, np.array([cnt], np.uint32)])
    return blob, dict(xp=xp, codes=codes, err=err, idx=idx, cnt=cnt, abs_eb=abs_eb, ebx2=ebx2)


def resolve_eb64(x_padded: np.ndarray, mode: int, user_eb: float):
    """(abs_eb, ebx2_r, ebx2) for FP64 input: the value base reduced in
    double and cast to float, abs_eb = (double)user_eb * (double)base
    (lorenzo_quant.cu:612-644 with TInput = double)."""
    eb = float(np.float32(user_eb))
    abs_eb = eb
    if mode != 0:
        vb = float(np.float32(orc.value_base(x_padded, mode)))
        if vb > 0:
            abs_eb = eb * vb
    return abs_eb, 1.0 / (2.0 * abs_eb), 2.0 * abs_eb


def lorenzo_part64(x, dims, mode, user_eb, radius, cap, zigzag, code_bytes, alignment):
    """lorenzo_part for FP64 input (LorenzoQuantStage<double, TCode>)."""
    x = np.ascontiguousarray(x, np.float64).ravel()
    padded = align(x.nbytes, alignment) // 8
    xp = np.zeros(padded, np.float64)
    xp[: x.size] = x
    abs_eb, r, ebx2 = resolve_eb64(xp, mode, user_eb)
    nx, ny, nz = dims
    if ny == 1 and nz == 1:
        nx = padded
    m = max_outliers(padded, cap)
    codes, err, idx, cnt = orc.lorenzo_fwd(xp, (nx, ny, nz), r, radius, zigzag, code_bytes, m)
    return xp, codes, err, idx, cnt, abs_eb, ebx2


def lorenzo64_chain(x, dims, coder, mode=1, user_eb=1e-3, radius=512, cap=0.2, zigzag=True, code_bytes=2,
                    bklen=1024):
    """FP64 LorenzoQuant -> coder ("rze": Bitshuffle(16 KiB, 2) + RZE;
    "huffman": Huffman coarse; "rle": RLE), the pipeline's blob."""
    alignment = 16384 if coder == "rze" else 1
    xp, codes, err, idx, cnt, abs_eb, ebx2 = lorenzo_part64(x, dims, mode, user_eb, radius, cap, zigzag,
                                                            code_bytes, alignment)
    if coder == "rze":
        head = orc.rze_enc(orc.bitshuffle_enc(codes.view(np.uint8), 16384, 2))
    elif coder == "huffman":
        head = orc.huffman_encode(codes, bklen)
    else:
        head = orc.rle_enc(codes)
    blob = concat([head, err, idx, np.array([cnt], np.uint32)])
    return blob, dict(xp=xp, codes=codes, err=err, idx=idx, cnt=cnt, abs_eb=abs_eb, ebx2=ebx2)


def reconstruct(parts_info, dims, zigzag=True):
    """Oracle inverse of the Lorenzo part (padded length)."""
    p = parts_info
    nx, ny, nz = dims
    n = p["codes"].size
    if ny == 1 and nz == 1:
        nx = n
    return orc.lorenzo_inv(p["codes"], (nx, ny, nz), p["err"], p["idx"], p["cnt"], p["ebx2"], zigzag)


def tile_key(idx, dims):
    """Sort key of the canonical outlier/patch order: tile-major, then the
    element's position inside its tile (1D 1024, 2D 32x32, 3D 8x8x8)."""
    idx = np.asarray(idx, np.int64)
    nx, ny, nz = dims
    if ny <= 1 and nz <= 1:
        return idx
    if nz <= 1:
        y, x = idx // nx, idx % nx
        gx = (nx + 31) // 32
        return ((y // 32) * gx + x // 32) * 1024 + (y % 32) * 32 + x % 32
    x = idx % nx
    y = (idx // nx) % ny
    z = idx // (nx * ny)
    gx, gy = (nx + 7) // 8, (ny + 7) // 8
    return (((z // 8) * gy + y // 8) * gx + x // 8) * 512 + (z % 8) * 64 + (y % 8) * 8 + x % 8


def strict_patches(x, recon, abs_eb, dims):
    """Elements the oracle's inverse reconstructs outside eb, in canonical
    order, with their original values (the strict-mode patch list)."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    err = np.abs(recon[: x.size].astype(np.float64) - x.astype(np.float64))
    bad = np.nonzero(~(err <= float(abs_eb)))[0]
    bad = bad[np.argsort(tile_key(bad, dims), kind="stable")]
    return bad.astype(np.uint32), x[bad]
