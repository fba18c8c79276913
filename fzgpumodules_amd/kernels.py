"""Host-array convenience wrappers over the C ABI (one call = one GPU round trip).

Each function uploads numpy inputs, calls one ``fzh_*`` entry point through
ctypes and downloads the outputs.  They exist for tests and examples; the
hot path in production is the C ABI / fz::Pipeline operating on device
pointers directly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._fzh import DeviceBuffer, call, lib, workspace

_CODE_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}


def lorenzo_fwd(x, dims, ebx2_r, radius, zigzag, code_bytes=2, max_outliers=None, n=None,
                d_params=None):
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    n = x.size if n is None else n
    nx, ny, nz = dims
    if max_outliers is None:
        max_outliers = n
    din = DeviceBuffer.from_array(x)
    codes = DeviceBuffer(n * code_bytes)
    err = DeviceBuffer(max(max_outliers, 1) * 4)
    idx = DeviceBuffer(max(max_outliers, 1) * 4)
    cnt = DeviceBuffer(4)
    ws = workspace(lib().fzh_lorenzo_quant_fwd_ws_bytes(n, nx, ny, nz))
    call("fzh_lorenzo_quant_fwd_f32", None, din.ptr, n, nx, ny, nz, np.float32(ebx2_r),
         d_params, radius, int(zigzag), code_bytes, codes.ptr, err.ptr, idx.ptr, cnt.ptr,
         max_outliers, ws.ptr, ws.nbytes)
    c = int(cnt.download(np.uint32, 1)[0])
    k = min(c, max_outliers)
    return (codes.download(_CODE_DT[code_bytes], n), err.download(np.float32, k),
            idx.download(np.uint32, k), c)


def lorenzo_inv(codes, dims, err, idx, count, ebx2, zigzag, max_outliers=None):
    codes = np.ascontiguousarray(codes)
    n = codes.size
    nx, ny, nz = dims
    err = np.ascontiguousarray(err, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    if max_outliers is None:
        max_outliers = err.size
    dc = DeviceBuffer.from_array(codes)
    de = DeviceBuffer.from_array(err if err.size else np.zeros(1, np.float32))
    di = DeviceBuffer.from_array(idx if idx.size else np.zeros(1, np.uint32))
    dn = DeviceBuffer.from_array(np.array([count], np.uint32))
    out = DeviceBuffer(n * 4)
    ws = workspace(lib().fzh_lorenzo_quant_inv_ws_bytes(n, nx, ny, nz, max_outliers))
    call("fzh_lorenzo_quant_inv_f32", None, dc.ptr, n, nx, ny, nz, de.ptr, di.ptr, dn.ptr,
         max_outliers, np.float32(ebx2), int(zigzag), codes.dtype.itemsize, out.ptr, ws.ptr, ws.nbytes)
    return out.download(np.float32, n)


def value_base_params(x, mode, user_eb):
    x = np.ascontiguousarray(x, np.float32).ravel()
    din = DeviceBuffer.from_array(x)
    par = DeviceBuffer(16)
    ws = workspace(lib().fzh_value_base_ws_bytes(x.size))
    call("fzh_value_base_f32", None, din.ptr, x.size, mode, np.float32(user_eb), par.ptr, ws.ptr, ws.nbytes)
    return par.download(np.float32, 4)


def lorenzo_fwd64(x, dims, ebx2_r, radius, zigzag, code_bytes=2, max_outliers=None, n=None, d_params=None):
    """FP64 Lorenzo forward (fzh_lorenzo_quant_fwd_f64): codes, double outlier
    errors, indices, total count."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    n = x.size if n is None else n
    nx, ny, nz = dims
    if max_outliers is None:
        max_outliers = n
    din = DeviceBuffer.from_array(x if x.size else np.zeros(1))
    codes = DeviceBuffer(max(n, 1) * code_bytes)
    err = DeviceBuffer(max(max_outliers, 1) * 8)
    idx = DeviceBuffer(max(max_outliers, 1) * 4)
    cnt = DeviceBuffer(4)
    ws = workspace(lib().fzh_lorenzo_quant_fwd_f64_ws_bytes(n, nx, ny, nz))
    call("fzh_lorenzo_quant_fwd_f64", None, din.ptr, n, nx, ny, nz, float(ebx2_r), d_params, radius, int(zigzag),
         code_bytes, codes.ptr, err.ptr, idx.ptr, cnt.ptr, max_outliers, ws.ptr, ws.nbytes)
    c = int(cnt.download(np.uint32, 1)[0])
    k = min(c, max_outliers)
    return (codes.download(_CODE_DT[code_bytes], n), err.download(np.float64, k),
            idx.download(np.uint32, k), c)


def lorenzo_inv64(codes, dims, err, idx, count, ebx2, zigzag, max_outliers=None):
    """FP64 Lorenzo inverse (fzh_lorenzo_quant_inv_f64)."""
    codes = np.ascontiguousarray(codes)
    n = codes.size
    nx, ny, nz = dims
    err = np.ascontiguousarray(err, np.float64)
    idx = np.ascontiguousarray(idx, np.uint32)
    if max_outliers is None:
        max_outliers = err.size
    dc = DeviceBuffer.from_array(codes)
    de = DeviceBuffer.from_array(err if err.size else np.zeros(1, np.float64))
    di = DeviceBuffer.from_array(idx if idx.size else np.zeros(1, np.uint32))
    dn = DeviceBuffer.from_array(np.array([count], np.uint32))
    out = DeviceBuffer(n * 8)
    ws = workspace(lib().fzh_lorenzo_quant_inv_f64_ws_bytes(n, nx, ny, nz, max_outliers))
    call("fzh_lorenzo_quant_inv_f64", None, dc.ptr, n, nx, ny, nz, de.ptr, di.ptr, dn.ptr, max_outliers,
         float(ebx2), int(zigzag), codes.dtype.itemsize, out.ptr, ws.ptr, ws.nbytes)
    return out.download(np.float64, n)


def value_base_params64(x, mode, user_eb):
    """fzh_value_base_f64: [value_base, abs_eb, ebx2_r, ebx2] as doubles."""
    x = np.ascontiguousarray(x, np.float64).ravel()
    din = DeviceBuffer.from_array(x)
    par = DeviceBuffer(32)
    ws = workspace(lib().fzh_value_base_f64_ws_bytes(x.size))
    call("fzh_value_base_f64", None, din.ptr, x.size, mode, np.float32(user_eb), par.ptr, ws.ptr, ws.nbytes)
    return par.download(np.float64, 4)


def bitshuffle(b, block_size=16384, elem_width=2, inverse=False):
    b = np.ascontiguousarray(b).view(np.uint8).ravel()
    din = DeviceBuffer.from_array(b)
    dout = DeviceBuffer(b.size)
    call("fzh_bitshuffle_dec" if inverse else "fzh_bitshuffle_enc", None, din.ptr, dout.ptr, b.size,
         block_size, elem_width)
    return dout.download(np.uint8, b.size)


def rze_enc(b, shuffle_width=0):
    """RZE stream of b; with shuffle_width 1/2 the fused Bitshuffle(16 KiB)+RZE."""
    b = np.ascontiguousarray(b).view(np.uint8).ravel()
    din = DeviceBuffer.from_array(b)
    cap = lib().fzh_rze_enc_bound(b.size)
    dout = DeviceBuffer(cap)
    dsz = DeviceBuffer(8)
    ws = workspace(lib().fzh_rze_enc_ws_bytes(b.size))
    if shuffle_width:
        call("fzh_rze_enc_shuffled", None, din.ptr, b.size, shuffle_width, dout.ptr, cap, dsz.ptr, ws.ptr, ws.nbytes)
    else:
        call("fzh_rze_enc", None, din.ptr, b.size, dout.ptr, cap, dsz.ptr, ws.ptr, ws.nbytes)
    sz = int(dsz.download(np.uint64, 1)[0])
    return dout.download(np.uint8, sz)


def rze_dec(s, orig_bytes, shuffle_width=0):
    s = np.ascontiguousarray(s).view(np.uint8).ravel()
    din = DeviceBuffer.from_array(s)
    dout = DeviceBuffer(orig_bytes)
    st = DeviceBuffer.zeros(4)
    ws = workspace(lib().fzh_rze_dec_ws_bytes(orig_bytes))
    if shuffle_width:
        call("fzh_rze_dec_shuffled", None, din.ptr, s.size, dout.ptr, orig_bytes, shuffle_width, st.ptr, ws.ptr,
             ws.nbytes)
    else:
        call("fzh_rze_dec", None, din.ptr, s.size, dout.ptr, orig_bytes, st.ptr, ws.ptr, ws.nbytes)
    if int(st.download(np.int32, 1)[0]) != 0:
        raise ValueError("malformed RZE stream")
    return dout.download(np.uint8, orig_bytes)


def rle_enc(v):
    v = np.ascontiguousarray(v)
    eb = v.dtype.itemsize
    din = DeviceBuffer.from_array(v)
    cap = lib().fzh_rle_enc_bound(v.size, eb)
    dout = DeviceBuffer(cap)
    dsz = DeviceBuffer(8)
    ws = workspace(lib().fzh_rle_enc_ws_bytes(v.size))
    call("fzh_rle_enc", None, din.ptr, v.size, eb, dout.ptr, cap, dsz.ptr, ws.ptr, ws.nbytes)
    return dout.download(np.uint8, int(dsz.download(np.uint64, 1)[0]))


def rle_dec(s, dtype, n):
    s = np.ascontiguousarray(s).view(np.uint8).ravel()
    eb = np.dtype(dtype).itemsize
    din = DeviceBuffer.from_array(s)
    dout = DeviceBuffer(max(n * eb, 16))
    st = DeviceBuffer.zeros(4)
    ws = workspace(lib().fzh_rle_dec_ws_bytes(n))
    call("fzh_rle_dec", None, din.ptr, s.size, eb, dout.ptr, n, st.ptr, ws.ptr, ws.nbytes)
    if int(st.download(np.int32, 1)[0]) != 0:
        raise ValueError("malformed RLE stream")
    return dout.download(dtype, n)


def histogram(v, bklen):
    v = np.ascontiguousarray(v)
    din = DeviceBuffer.from_array(v)
    f = DeviceBuffer(bklen * 4)
    call("fzh_histogram", None, din.ptr, v.size, v.dtype.itemsize, bklen, f.ptr)
    return f.download(np.uint32, bklen)


def huffman_book(freq, elem_bytes=2):
    freq = np.ascontiguousarray(freq, np.uint32)
    bk4 = np.zeros(freq.size, np.uint32)
    rev = np.zeros(lib().fzh_huffman_revbook_bytes(freq.size, elem_bytes), np.uint8)
    call("fzh_huffman_build_book", freq.ctypes.data, freq.size, elem_bytes, bk4.ctypes.data, rev.ctypes.data)
    return bk4, rev


def huffman_book_device(freq, elem_bytes=2, n_expected=None):
    """The codebook built on the device from a histogram: (bk4, revbook, status)."""
    freq = np.ascontiguousarray(freq, np.uint32)
    if n_expected is None:
        n_expected = int(freq.sum(dtype=np.uint64))
    df = DeviceBuffer.from_array(freq)
    dbk = DeviceBuffer(freq.size * 4)
    nrev = lib().fzh_huffman_revbook_bytes(freq.size, elem_bytes)
    drev = DeviceBuffer(nrev)
    st = DeviceBuffer.zeros(4)
    call("fzh_huffman_build_book_device", None, df.ptr, freq.size, elem_bytes, int(n_expected),
         dbk.ptr, drev.ptr, st.ptr)
    return dbk.download(np.uint32, freq.size), drev.download(np.uint8, nrev), int(st.download(np.int32, 1)[0])


def huffman_encode(v, bklen, fine=False):
    """Histogram on the GPU, codebook on the host, encode on the GPU (fine:
    1024-symbol partitions when no codeword exceeds 8 bits, as HuffmanStage's
    Fine mode)."""
    v = np.ascontiguousarray(v)
    eb = v.dtype.itemsize
    freq = histogram(v, bklen)
    if int(freq.sum(dtype=np.uint64)) != v.size:
        raise ValueError("symbol out of range")
    bk4, rev = huffman_book(freq, eb)
    din = DeviceBuffer.from_array(v)
    dbk = DeviceBuffer.from_array(bk4)
    drev = DeviceBuffer.from_array(rev)
    cap = lib().fzh_huffman_enc_bound(v.size, bklen, eb)
    dout = DeviceBuffer(cap)
    dsz = DeviceBuffer(8)
    ws = workspace(lib().fzh_huffman_enc_ws_bytes(v.size))
    fn = "fzh_huffman_encode_fine" if fine and int((bk4 >> 27).max()) <= 8 else "fzh_huffman_encode"
    call(fn, None, din.ptr, v.size, eb, bklen, dbk.ptr, drev.ptr, dout.ptr, cap, dsz.ptr, ws.ptr, ws.nbytes)
    return dout.download(np.uint8, int(dsz.download(np.uint64, 1)[0]))


def huffman_decode(s, dtype, n):
    s = np.ascontiguousarray(s).view(np.uint8).ravel()
    din = DeviceBuffer.from_array(s)
    dout = DeviceBuffer(max(n * np.dtype(dtype).itemsize, 16))
    st = DeviceBuffer.zeros(4)
    call("fzh_huffman_decode", None, din.ptr, s.size, np.dtype(dtype).itemsize, dout.ptr, n, st.ptr)
    if int(st.download(np.int32, 1)[0]) != 0:
        raise ValueError("malformed Huffman stream")
    return dout.download(dtype, n)


def quant_abs_fwd(x, ebx2_r, radius, zigzag, code_bytes=2, threshold=np.inf, max_outliers=None):
    x = np.ascontiguousarray(x, np.float32).ravel()
    if max_outliers is None:
        max_outliers = x.size
    din = DeviceBuffer.from_array(x)
    codes = DeviceBuffer(x.size * code_bytes)
    val = DeviceBuffer(max(max_outliers, 1) * 4)
    idx = DeviceBuffer(max(max_outliers, 1) * 4)
    cnt = DeviceBuffer(4)
    ws = workspace(lib().fzh_quant_abs_fwd_ws_bytes(x.size))
    call("fzh_quant_abs_fwd_f32", None, din.ptr, x.size, np.float32(ebx2_r), np.float32(threshold), radius,
         int(zigzag), code_bytes, codes.ptr, val.ptr, idx.ptr, cnt.ptr, max_outliers, ws.ptr, ws.nbytes)
    c = int(cnt.download(np.uint32, 1)[0])
    k = min(c, max_outliers)
    return codes.download(_CODE_DT[code_bytes], x.size), val.download(np.float32, k), idx.download(np.uint32, k), c


def quant_abs_inv(codes, ebx2, zigzag, vals, idx, count):
    codes = np.ascontiguousarray(codes)
    dc = DeviceBuffer.from_array(codes)
    dv = DeviceBuffer.from_array(np.ascontiguousarray(vals, np.float32) if len(vals) else np.zeros(1, np.float32))
    di = DeviceBuffer.from_array(np.ascontiguousarray(idx, np.uint32) if len(idx) else np.zeros(1, np.uint32))
    dn = DeviceBuffer.from_array(np.array([count], np.uint32))
    out = DeviceBuffer(codes.size * 4)
    call("fzh_quant_abs_inv_f32", None, dc.ptr, codes.size, np.float32(ebx2), int(zigzag), codes.dtype.itemsize,
         dv.ptr, di.ptr, dn.ptr, len(vals), out.ptr)
    return out.download(np.float32, codes.size)


def quant_rel_fwd(x, eps, radius, code_bytes=2, threshold=np.inf, max_outliers=None):
    """Quantizer REL forward (fzh_quant_rel_fwd_f32)."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    if max_outliers is None:
        max_outliers = x.size
    din = DeviceBuffer.from_array(x)
    codes = DeviceBuffer(x.size * code_bytes)
    val = DeviceBuffer(max(max_outliers, 1) * 4)
    idx = DeviceBuffer(max(max_outliers, 1) * 4)
    cnt = DeviceBuffer(4)
    ws = workspace(lib().fzh_quant_abs_fwd_ws_bytes(x.size))
    call("fzh_quant_rel_fwd_f32", None, din.ptr, x.size, np.float32(eps), np.float32(threshold), radius, code_bytes,
         codes.ptr, val.ptr, idx.ptr, cnt.ptr, max_outliers, ws.ptr, ws.nbytes)
    c = int(cnt.download(np.uint32, 1)[0])
    k = min(c, max_outliers)
    return codes.download(_CODE_DT[code_bytes], x.size), val.download(np.float32, k), idx.download(np.uint32, k), c


def quant_rel_inv(codes, eps, vals, idx, count):
    """Quantizer REL inverse (fzh_quant_rel_inv_f32)."""
    codes = np.ascontiguousarray(codes)
    dc = DeviceBuffer.from_array(codes)
    dv = DeviceBuffer.from_array(np.ascontiguousarray(vals, np.float32) if len(vals) else np.zeros(1, np.float32))
    di = DeviceBuffer.from_array(np.ascontiguousarray(idx, np.uint32) if len(idx) else np.zeros(1, np.uint32))
    dn = DeviceBuffer.from_array(np.array([count], np.uint32))
    out = DeviceBuffer(codes.size * 4)
    call("fzh_quant_rel_inv_f32", None, dc.ptr, codes.size, np.float32(eps), codes.dtype.itemsize, dv.ptr, di.ptr,
         dn.ptr, len(vals), out.ptr)
    return out.download(np.float32, codes.size)


def scan_u32(v, wrap=False, clear=False, in_place=False):
    """Exclusive scan through fzh_scan_u32: (prefixes, total, input after the call)."""
    v = np.ascontiguousarray(v, np.uint32)
    din = DeviceBuffer.from_array(v if v.size else np.zeros(1, np.uint32))
    dout = din if in_place else DeviceBuffer(max(v.nbytes, 16))
    tot = DeviceBuffer(16)
    ws = workspace(lib().fzh_scan_ws_bytes(v.size))
    flags = (1 if wrap else 0) | (2 if clear else 0)
    call("fzh_scan_u32", None, din.ptr, dout.ptr, v.size, tot.ptr, flags, ws.ptr, ws.nbytes)
    return dout.download(np.uint32, v.size), int(tot.download(np.uint32, 1)[0]), din.download(np.uint32, v.size)


def max_u32(v):
    v = np.ascontiguousarray(v)
    din = DeviceBuffer.from_array(v)
    m = DeviceBuffer(4)
    ws = workspace(lib().fzh_max_u32_ws_bytes())
    call("fzh_max_u32", None, din.ptr, v.size, v.dtype.itemsize, m.ptr, ws.ptr, ws.nbytes)
    return int(m.download(np.uint32, 1)[0])


def bitpack_enc(v, nbits):
    v = np.ascontiguousarray(v)
    din = DeviceBuffer.from_array(v)
    nout = (v.size * nbits + 7) // 8
    dout = DeviceBuffer(max(nout, 16))
    call("fzh_bitpack_enc", None, din.ptr, v.size, v.dtype.itemsize, nbits, dout.ptr)
    return dout.download(np.uint8, nout)


def bitpack_dec(b, n, dtype, nbits):
    b = np.ascontiguousarray(b, np.uint8)
    din = DeviceBuffer.from_array(b if b.size else np.zeros(1, np.uint8))
    dout = DeviceBuffer(max(n * np.dtype(dtype).itemsize, 16))
    call("fzh_bitpack_dec", None, din.ptr, n, np.dtype(dtype).itemsize, nbits, dout.ptr)
    return dout.download(dtype, n)
