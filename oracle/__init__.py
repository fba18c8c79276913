"""ctypes binding of the CPU oracle (oracle/fz_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
(``fzgpumodules_amd`` / ``libfzh.so``) never imports this module.

Every wrapper mirrors one ``orc_*`` entry point; the C source cites the
reference file:line each one restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_phf.so")

_lib = None
_ref = None


def build() -> None:
    """Compile the oracle (and, when the reference tree exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def ref_phf():
    """The reference's own Huffman codebook builder, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        _ref = C.CDLL(_REF_PATH)
        _ref.ref_phf_book.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        _ref.ref_phf_book.restype = None
    return _ref


_P = C.c_void_p
_Z = C.c_size_t


def _declare(L):
    sig = {
        "orc_lorenzo_fwd_f32": (C.c_uint32, [_P, _Z, _Z, _Z, _Z, C.c_float, C.c_uint32, C.c_int, C.c_int, _P, _P, _P, _Z]),
        "orc_lorenzo_inv_f32": (None, [_P, _Z, _Z, _Z, _Z, _P, _P, C.c_uint32, _Z, C.c_float, C.c_int, C.c_int, _P]),
        "orc_value_base_f32": (C.c_float, [_P, _Z, C.c_int]),
        "orc_lorenzo_fwd_f64": (C.c_uint32, [_P, _Z, _Z, _Z, _Z, C.c_double, C.c_uint32, C.c_int, C.c_int, _P, _P, _P, _Z]),
        "orc_lorenzo_inv_f64": (None, [_P, _Z, _Z, _Z, _Z, _P, _P, C.c_uint32, _Z, C.c_double, C.c_int, C.c_int, _P]),
        "orc_value_base_f64": (C.c_float, [_P, _Z, C.c_int]),
        "orc_bitshuffle_enc": (None, [_P, _P, _Z, C.c_uint32, C.c_int]),
        "orc_bitshuffle_dec": (None, [_P, _P, _Z, C.c_uint32, C.c_int]),
        "orc_rze_enc": (_Z, [_P, _Z, _P]),
        "orc_rze_dec": (_Z, [_P, _Z, _P, _Z]),
        "orc_rle_enc": (_Z, [_P, _Z, C.c_int, _P]),
        "orc_rle_dec": (_Z, [_P, _Z, C.c_int, _P, _Z]),
        "orc_histogram": (None, [_P, _Z, C.c_int, C.c_uint32, _P]),
        "orc_huffman_book": (None, [_P, C.c_uint32, C.c_int, _P, _P]),
        "orc_huffman_revbook_bytes": (_Z, [C.c_uint32, C.c_int]),
        "orc_huffman_encode": (_Z, [_P, _Z, C.c_int, C.c_uint32, _P, _Z]),
        "orc_huffman_encode_mode": (_Z, [_P, _Z, C.c_int, C.c_uint32, C.c_int, _P, _Z]),
        "orc_huffman_encode_bound": (_Z, [_Z, C.c_uint32, C.c_int]),
        "orc_huffman_decode": (C.c_int, [_P, _Z, C.c_int, _P, _Z]),
        "orc_quant_abs_fwd_f32": (C.c_uint32, [_P, _Z, C.c_float, C.c_float, C.c_uint32, C.c_int, C.c_int, _P, _P, _P, _Z]),
        "orc_quant_abs_inv_f32": (None, [_P, _Z, C.c_float, C.c_int, C.c_int, _P, _P, C.c_uint32, _Z, _P]),
        "orc_quant_rel_params": (None, [C.c_float, _P]),
        "orc_quant_rel_fwd_f32": (C.c_uint32, [_P, _Z, C.c_float, C.c_float, C.c_uint32, C.c_int, _P, _P, _P, _Z]),
        "orc_quant_rel_inv_f32": (None, [_P, _Z, C.c_float, C.c_int, _P, _P, C.c_uint32, _Z, _P]),
        "orc_bitpack_nbits_for_max": (C.c_int, [C.c_uint32, C.c_int]),
        "orc_bitpack_enc": (_Z, [_P, _Z, C.c_int, C.c_int, _P]),
        "orc_bitpack_dec": (None, [_P, _Z, C.c_int, C.c_int, _P]),
        "orc_crc32": (C.c_uint32, [_P, _Z]),
        "orc_chain_l2d_bs_rze": (_Z, [_P, _Z, _Z, _Z, C.c_float, C.c_uint32, C.c_int, _P, _P, _P, _P, _P, _Z, _P]),
        "orc_num_threads": (C.c_int, []),
        "orc_set_num_threads": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


_CODE_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}


# ---------------------------------------------------------------------------
# Lorenzo
# ---------------------------------------------------------------------------
def lorenzo_fwd(x: np.ndarray, dims, ebx2_r: float, radius: int, zigzag: bool,
                code_bytes: int = 2, max_outliers: int | None = None, n: int | None = None):
    """Returns (codes, err, idx, count).  ``dims`` = (nx, ny, nz).  A float64
    ``x`` runs the FP64 path (double arithmetic, double outlier errors);
    anything else is taken as float32."""
    f64 = np.asarray(x).dtype == np.float64
    ft = np.float64 if f64 else np.float32
    x = np.ascontiguousarray(x, dtype=ft).ravel()
    n = x.size if n is None else n
    nx, ny, nz = dims
    if max_outliers is None:
        max_outliers = n
    codes = np.zeros(n, dtype=_CODE_DT[code_bytes])
    err = np.zeros(max(max_outliers, 1), dtype=ft)
    idx = np.zeros(max(max_outliers, 1), dtype=np.uint32)
    fn = lib().orc_lorenzo_fwd_f64 if f64 else lib().orc_lorenzo_fwd_f32
    cnt = fn(_ptr(x), n, nx, ny, nz, ft(ebx2_r), radius, int(zigzag), code_bytes, _ptr(codes), _ptr(err),
             _ptr(idx), max_outliers)
    k = min(cnt, max_outliers)
    return codes, err[:k].copy(), idx[:k].copy(), int(cnt)


def lorenzo_inv(codes: np.ndarray, dims, err, idx, count: int, ebx2: float, zigzag: bool,
                max_outliers: int | None = None) -> np.ndarray:
    """Inverse; float64 outlier errors select the FP64 path (float64 output)."""
    codes = np.ascontiguousarray(codes)
    cb = codes.dtype.itemsize
    n = codes.size
    nx, ny, nz = dims
    f64 = np.asarray(err).dtype == np.float64
    ft = np.float64 if f64 else np.float32
    err = np.ascontiguousarray(err, dtype=ft)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    if max_outliers is None:
        max_outliers = err.size
    out = np.zeros(n, dtype=ft)
    e = err if err.size else np.zeros(1, ft)
    i = idx if idx.size else np.zeros(1, np.uint32)
    fn = lib().orc_lorenzo_inv_f64 if f64 else lib().orc_lorenzo_inv_f32
    fn(_ptr(codes), n, nx, ny, nz, _ptr(e), _ptr(i), count, min(max_outliers, err.size), ft(ebx2), int(zigzag), cb,
       _ptr(out))
    return out


def value_base(x: np.ndarray, mode: int) -> float:
    """REL/NOA value base (float), in the input's precision for float64."""
    if np.asarray(x).dtype == np.float64:
        x = np.ascontiguousarray(x).ravel()
        return float(np.float32(lib().orc_value_base_f64(_ptr(x), x.size, mode)))
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    return float(np.float32(lib().orc_value_base_f32(_ptr(x), x.size, mode)))


# ---------------------------------------------------------------------------
# Bitshuffle / RZE / RLE
# ---------------------------------------------------------------------------
def bitshuffle_enc(b: np.ndarray, block_size: int = 16384, elem_width: int = 2) -> np.ndarray:
    b = np.ascontiguousarray(b).view(np.uint8).ravel()
    out = np.empty_like(b)
    lib().orc_bitshuffle_enc(_ptr(b), _ptr(out), b.size, block_size, elem_width)
    return out


def bitshuffle_dec(b: np.ndarray, block_size: int = 16384, elem_width: int = 2) -> np.ndarray:
    b = np.ascontiguousarray(b).view(np.uint8).ravel()
    out = np.empty_like(b)
    lib().orc_bitshuffle_dec(_ptr(b), _ptr(out), b.size, block_size, elem_width)
    return out


def rze_enc(b: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(b).view(np.uint8).ravel()
    nc = (b.size + 16383) // 16384
    # A partial last chunk may grow by up to ~1/7 (thresholds use the full
    # 16 KiB chunk, rze_stage.cu:369-431).
    out = np.zeros(b.size + 8 + 4 * nc + (b.size % 16384) // 7 + 64, dtype=np.uint8)
    sz = lib().orc_rze_enc(_ptr(b), b.size, _ptr(out))
    return out[:sz].copy()


def rze_dec(s: np.ndarray, out_cap: int) -> np.ndarray:
    s = np.ascontiguousarray(s).view(np.uint8).ravel()
    out = np.zeros(max(out_cap, 1), dtype=np.uint8)
    sz = lib().orc_rze_dec(_ptr(s), s.size, _ptr(out), out_cap)
    if sz == C.c_size_t(-1).value:
        raise ValueError("malformed RZE stream")
    return out[:sz].copy()


def rle_enc(v: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(v)
    eb = v.dtype.itemsize
    out = np.zeros(4 + ((v.size * eb + 3) & ~3) + 4 * v.size + 4, dtype=np.uint8)
    sz = lib().orc_rle_enc(_ptr(v), v.size, eb, _ptr(out))
    return out[:sz].copy()


def rle_dec(s: np.ndarray, dtype, cap: int) -> np.ndarray:
    s = np.ascontiguousarray(s).view(np.uint8).ravel()
    out = np.zeros(max(cap, 1), dtype=dtype)
    n = lib().orc_rle_dec(_ptr(s), s.size, np.dtype(dtype).itemsize, _ptr(out), cap)
    if n == C.c_size_t(-1).value:
        raise ValueError("malformed RLE stream")
    return out[:n].copy()


# ---------------------------------------------------------------------------
# Histogram / Huffman
# ---------------------------------------------------------------------------
def histogram(v: np.ndarray, bklen: int) -> np.ndarray:
    v = np.ascontiguousarray(v)
    f = np.zeros(bklen, dtype=np.uint32)
    lib().orc_histogram(_ptr(v), v.size, v.dtype.itemsize, bklen, _ptr(f))
    return f


def huffman_book(freq: np.ndarray, elem_bytes: int = 2):
    freq = np.ascontiguousarray(freq, dtype=np.uint32)
    bklen = freq.size
    bk4 = np.zeros(bklen, dtype=np.uint32)
    rev = np.zeros(lib().orc_huffman_revbook_bytes(bklen, elem_bytes), dtype=np.uint8)
    lib().orc_huffman_book(_ptr(freq), bklen, elem_bytes, _ptr(bk4), _ptr(rev))
    return bk4, rev


def ref_huffman_book(freq: np.ndarray, elem_bytes: int = 2):
    """The reference's compiled codebook builder (None if oracle/_ref absent)."""
    R = ref_phf()
    if R is None:
        return None
    freq = np.ascontiguousarray(freq, dtype=np.uint32).copy()
    bklen = freq.size
    bk4 = np.zeros(bklen, dtype=np.uint32)
    rev = np.zeros(256 + elem_bytes * bklen, dtype=np.uint8)
    R.ref_phf_book(_ptr(freq), bklen, elem_bytes, _ptr(bk4), _ptr(rev))
    return bk4, rev


def huffman_encode(v: np.ndarray, bklen: int, fine: bool = False) -> np.ndarray:
    v = np.ascontiguousarray(v)
    eb = v.dtype.itemsize
    cap = lib().orc_huffman_encode_bound(v.size, bklen, eb)
    out = np.zeros(cap, dtype=np.uint8)
    sz = lib().orc_huffman_encode_mode(_ptr(v), v.size, eb, bklen, int(bool(fine)), _ptr(out), cap)
    if sz == 0 and v.size:
        raise ValueError("symbol out of range")
    return out[:sz].copy()


def huffman_decode(s: np.ndarray, dtype, n: int) -> np.ndarray:
    s = np.ascontiguousarray(s).view(np.uint8).ravel()
    out = np.zeros(max(n, 1), dtype=dtype)
    rc = lib().orc_huffman_decode(_ptr(s), s.size, np.dtype(dtype).itemsize, _ptr(out), n)
    if rc != 0:
        raise ValueError("malformed Huffman stream")
    return out[:n].copy()


# ---------------------------------------------------------------------------
# Quantizer / Bitpack / CRC
# ---------------------------------------------------------------------------
def quant_abs_fwd(x, ebx2_r, radius, zigzag, code_bytes=2, threshold=np.inf, max_outliers=None):
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    if max_outliers is None:
        max_outliers = x.size
    codes = np.zeros(x.size, dtype=_CODE_DT[code_bytes])
    val = np.zeros(max(max_outliers, 1), np.float32)
    idx = np.zeros(max(max_outliers, 1), np.uint32)
    cnt = lib().orc_quant_abs_fwd_f32(_ptr(x), x.size, np.float32(ebx2_r), np.float32(threshold), radius,
                                      int(zigzag), code_bytes, _ptr(codes), _ptr(val), _ptr(idx), max_outliers)
    k = min(cnt, max_outliers)
    return codes, val[:k].copy(), idx[:k].copy(), int(cnt)


def quant_abs_inv(codes, ebx2, zigzag, vals, idx, count):
    codes = np.ascontiguousarray(codes)
    out = np.zeros(codes.size, np.float32)
    vals = np.ascontiguousarray(vals, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    v = vals if vals.size else np.zeros(1, np.float32)
    i = idx if idx.size else np.zeros(1, np.uint32)
    lib().orc_quant_abs_inv_f32(_ptr(codes), codes.size, np.float32(ebx2), int(zigzag), codes.dtype.itemsize,
                                _ptr(v), _ptr(i), count, vals.size, _ptr(out))
    return out


def quant_rel_params(eps: float) -> np.ndarray:
    """[log2eb, 1/log2eb, 1+eps, 1/(1+eps)] as float32 (quantizer.cu:563-568)."""
    p = np.zeros(4, np.float32)
    lib().orc_quant_rel_params(np.float32(eps), _ptr(p))
    return p


def quant_rel_fwd(x, eps, radius, code_bytes=2, threshold=np.inf, max_outliers=None):
    """Quantizer REL forward: (codes, outlier values, indices, count)."""
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    if max_outliers is None:
        max_outliers = x.size
    codes = np.zeros(x.size, dtype=_CODE_DT[code_bytes])
    val = np.zeros(max(max_outliers, 1), np.float32)
    idx = np.zeros(max(max_outliers, 1), np.uint32)
    cnt = lib().orc_quant_rel_fwd_f32(_ptr(x), x.size, np.float32(eps), np.float32(threshold), radius, code_bytes,
                                      _ptr(codes), _ptr(val), _ptr(idx), max_outliers)
    k = min(cnt, max_outliers)
    return codes, val[:k].copy(), idx[:k].copy(), int(cnt)


def quant_rel_inv(codes, eps, vals, idx, count):
    codes = np.ascontiguousarray(codes)
    out = np.zeros(codes.size, np.float32)
    vals = np.ascontiguousarray(vals, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    v = vals if vals.size else np.zeros(1, np.float32)
    i = idx if idx.size else np.zeros(1, np.uint32)
    lib().orc_quant_rel_inv_f32(_ptr(codes), codes.size, np.float32(eps), codes.dtype.itemsize, _ptr(v), _ptr(i),
                                count, vals.size, _ptr(out))
    return out


def bitpack_nbits_for_max(m: int, elem_bytes: int) -> int:
    return lib().orc_bitpack_nbits_for_max(m, elem_bytes)


def bitpack_enc(v, nbits):
    v = np.ascontiguousarray(v)
    out = np.zeros((v.size * nbits + 7) // 8 + 1, np.uint8)
    sz = lib().orc_bitpack_enc(_ptr(v), v.size, v.dtype.itemsize, nbits, _ptr(out))
    return out[:sz].copy()


def bitpack_dec(b, n, dtype, nbits):
    b = np.ascontiguousarray(b, np.uint8)
    out = np.zeros(max(n, 1), dtype)
    lib().orc_bitpack_dec(_ptr(b), n, np.dtype(dtype).itemsize, nbits, _ptr(out))
    return out[:n].copy()


def crc32(b) -> int:
    b = np.ascontiguousarray(np.frombuffer(bytes(b), np.uint8) if isinstance(b, (bytes, bytearray)) else b).view(np.uint8)
    return int(lib().orc_crc32(_ptr(b) if b.size else 0, b.size))


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(n)


def num_threads() -> int:
    return lib().orc_num_threads()
