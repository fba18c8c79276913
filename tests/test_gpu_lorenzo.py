"""Parity of the HIP Lorenzo predict+quantize kernels against the oracle.

Codes, outlier lists and outlier counts must be bit-exact (same canonical
outlier order), the inverse output must be bit-exact (same FP32 addition
tree), and the reconstruction must stay within the error bound.
Reference behaviour: modules/fused/lorenzo_quant/lorenzo_quant.cu,
lorenzo_quant_nd.cu.
"""
import zlib

import numpy as np
import pytest

from conftest import eb_params

pytestmark = pytest.mark.gpu

CASES = [
    # name, dims (nx, ny, nz), n (None = nx*ny*nz)
    ("1d_ragged", (5000, 1, 1), None),
    ("1d_tiles", (4096, 1, 1), None),
    ("1d_tiny", (3, 1, 1), None),
    ("2d_partial", (100, 70, 1), None),
    ("2d_exact", (64, 64, 1), None),
    ("2d_padded", (37, 33, 1), 37 * 33 + 333),
    ("3d_partial", (37, 29, 21), None),
    ("3d_exact", (16, 16, 16), None),
    ("3d_padded", (9, 9, 9), 729 + 100),
]


def make_field(rng, n_cover, n, spikes=True):
    x = (np.cumsum(rng.normal(0, 1, n_cover)) * 0.01).astype(np.float32)
    if spikes and n_cover > 10:
        x[rng.integers(0, n_cover, max(1, n_cover // 200))] *= 1000
    out = np.zeros(n, np.float32)
    out[:n_cover] = x
    return out


@pytest.mark.parametrize("name,dims,n", CASES)
@pytest.mark.parametrize("zigzag", [False, True])
@pytest.mark.parametrize("code_bytes,radius", [(2, 32768), (2, 512), (1, 64)])
def test_forward_and_inverse_bit_exact(gpu, orc, name, dims, n, zigzag, code_bytes, radius):
    rng = np.random.default_rng(zlib.crc32(f"{name}-{zigzag}-{code_bytes}-{radius}".encode()))
    nx, ny, nz = dims
    cover = nx * ny * nz
    n = cover if n is None else n
    x = make_field(rng, cover, n)
    ebx2_r, ebx2 = eb_params(1e-3)
    cap = max(1, int(np.ceil(n * 0.2)))
    oc, oe, oi, ocnt = orc.lorenzo_fwd(x, dims, ebx2_r, radius, zigzag, code_bytes, cap, n)
    gc, ge, gi, gcnt = gpu.lorenzo_fwd(x, dims, ebx2_r, radius, zigzag, code_bytes, cap, n)
    assert gcnt == ocnt
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(ge.view(np.uint32), oe.view(np.uint32))
    # inverse: bit-exact output
    oy = orc.lorenzo_inv(oc, dims, oe, oi, ocnt, ebx2, zigzag, cap)
    gy = gpu.lorenzo_inv(gc, dims, ge, gi, gcnt, ebx2, zigzag, cap)
    np.testing.assert_array_equal(gy[:cover].view(np.uint32), oy[:cover].view(np.uint32))


@pytest.mark.parametrize("dims", [(5000, 1, 1), (100, 70, 1), (37, 29, 21)])
def test_inverse_accepts_any_outlier_order(gpu, orc, dims):
    """The reference emits outliers in atomic order; decode must not depend on it."""
    rng = np.random.default_rng(7)
    n = int(np.prod(dims))
    x = make_field(rng, n, n)
    ebx2_r, ebx2 = eb_params(1e-3)
    c, e, i, cnt = orc.lorenzo_fwd(x, dims, ebx2_r, 512, True, 2, n)
    assert cnt > 0
    perm = rng.permutation(cnt)
    y_ref = orc.lorenzo_inv(c, dims, e, i, cnt, ebx2, True, n)
    y_gpu = gpu.lorenzo_inv(c, dims, e[perm], i[perm], cnt, ebx2, True, n)
    np.testing.assert_array_equal(y_gpu.view(np.uint32), y_ref.view(np.uint32))


@pytest.mark.parametrize("name,dims,n", CASES + [("2d_large", (1000, 700, 1), None), ("1d_large", (300000, 1, 1), None)])
def test_inverse_short_outlier_lists(gpu, orc, name, dims, n):
    """Lists of at most 256 outliers are read by every tile directly (no
    binning launches); lists beyond that are binned.  Both give the oracle's
    bits, for a capacity equal to the count (as a stored stream has it), for
    a larger capacity, and in any outlier order."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    nx, ny, nz = dims
    cover = nx * ny * nz
    n = cover if n is None else n
    x = make_field(rng, cover, n)
    ebx2_r, ebx2 = eb_params(1e-3)
    c, e, i, cnt = orc.lorenzo_fwd(x, dims, ebx2_r, 512, True, 2, n, n)
    for k in sorted({min(cnt, 1), min(cnt, 255), min(cnt, 256), min(cnt, 257), cnt}):
        ek, ik = e[:k], i[:k]
        for cap in sorted({k, max(k, 256), max(k, 257)}):
            want = orc.lorenzo_inv(c, dims, ek, ik, k, ebx2, True, cap)
            got = gpu.lorenzo_inv(c, dims, ek, ik, k, ebx2, True, cap)
            np.testing.assert_array_equal(got[:cover].view(np.uint32), want[:cover].view(np.uint32))
        perm = rng.permutation(k)
        got = gpu.lorenzo_inv(c, dims, ek[perm], ik[perm], k, ebx2, True, k)
        np.testing.assert_array_equal(got[:cover].view(np.uint32), want[:cover].view(np.uint32))


@pytest.mark.parametrize("n", [1024 * 16384, 1024 * 16384 + 1])
def test_outlier_compaction_both_ways(gpu, orc, n):
    """Up to 16384 tiles the place kernel sums the outlier counts before its
    tiles itself; one tile more and the counts are scanned first.  Both give
    the oracle's outlier list, in the same order."""
    rng = np.random.default_rng(n)
    x = make_field(rng, n, n)
    ebx2_r, ebx2 = eb_params(1e-3)
    cap = n // 5
    oc, oe, oi, ocnt = orc.lorenzo_fwd(x, (n, 1, 1), ebx2_r, 512, True, 2, cap, n)
    gc, ge, gi, gcnt = gpu.lorenzo_fwd(x, (n, 1, 1), ebx2_r, 512, True, 2, cap, n)
    assert gcnt == ocnt > 256
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(ge.view(np.uint32), oe.view(np.uint32))


@pytest.mark.parametrize("name,dims,n", CASES + [("2d_large", (1000, 700, 1), None)])
@pytest.mark.parametrize("npatch", [0, 1, 200, 256, 257, 1000])
def test_strict_inverse_matches_inverse_then_patches(gpu, orc, name, dims, n, npatch):
    """fzh_lorenzo_quant_inv_strict_f32 == fzh_lorenzo_quant_inv_f32 then
    fzh_lorenzo_apply_patches, bit for bit: patches folded into the inverse
    launch (short lists) or applied after it (long ones), including patch
    indices past the tiles of a padded field."""
    from fzgpumodules_amd._fzh import DeviceBuffer, call, lib, workspace
    rng = np.random.default_rng(zlib.crc32(f"{name}-{npatch}".encode()))
    nx, ny, nz = dims
    cover = nx * ny * nz
    n = cover if n is None else n
    x = make_field(rng, cover, n)
    ebx2_r, ebx2 = eb_params(1e-3)
    c, e, i, cnt = orc.lorenzo_fwd(x, dims, ebx2_r, 512, True, 2, n, n)
    k = min(cnt, 200)
    pidx = np.sort(rng.choice(n, size=min(npatch, n), replace=False)).astype(np.uint32)
    pval = rng.standard_normal(pidx.size).astype(np.float32)
    dc = DeviceBuffer.from_array(c)
    de, di = _dev_or_dummy(e[:k], np.float32), _dev_or_dummy(i[:k], np.uint32)
    dn = DeviceBuffer.from_array(np.array([k], np.uint32))
    pv, pi = _dev_or_dummy(pval, np.float32), _dev_or_dummy(pidx, np.uint32)
    pn = DeviceBuffer.from_array(np.array([pidx.size], np.uint32))
    ws = workspace(lib().fzh_lorenzo_quant_inv_ws_bytes(n, nx, ny, nz, k))
    a, b = DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    call("fzh_memset", a.ptr, 0, 4 * n)
    call("fzh_memset", b.ptr, 0, 4 * n)
    call("fzh_lorenzo_quant_inv_f32", None, dc.ptr, n, nx, ny, nz, de.ptr, di.ptr, dn.ptr, k, np.float32(ebx2), 1, 2,
         a.ptr, ws.ptr, ws.nbytes)
    call("fzh_lorenzo_apply_patches", None, a.ptr, pv.ptr, pi.ptr, pn.ptr, pidx.size, n)
    call("fzh_lorenzo_quant_inv_strict_f32", None, dc.ptr, n, nx, ny, nz, de.ptr, di.ptr, dn.ptr, k,
         np.float32(ebx2), 1, 2, pv.ptr, pi.ptr, pn.ptr, pidx.size, b.ptr, ws.ptr, ws.nbytes)
    np.testing.assert_array_equal(b.download(np.uint32, n), a.download(np.uint32, n))
    want = orc.lorenzo_inv(c, dims, e[:k], i[:k], k, ebx2, True, k)
    want[pidx] = pval
    np.testing.assert_array_equal(b.download(np.uint32, n)[:cover], want[:cover].view(np.uint32))


def _dev_or_dummy(a, dt):
    from fzgpumodules_amd._fzh import DeviceBuffer
    a = np.ascontiguousarray(a, dt)
    return DeviceBuffer.from_array(a if a.size else np.zeros(1, dt))


@pytest.mark.parametrize("dims", [(1 << 20, 1, 1), (1024, 768, 1), (96, 80, 64)])
def test_error_bound_round_trip(gpu, dims):
    """Values stay where float32 resolves the bound (ulp(max|x|) << eb)."""
    rng = np.random.default_rng(11)
    n = int(np.prod(dims))
    x = make_field(rng, n, n, spikes=False)
    x[rng.integers(0, n, max(1, n // 500))] += np.float32(3.0)  # outlier deltas, |x| stays < 16
    eb = np.float32(1e-3)
    ebx2_r, ebx2 = eb_params(eb)
    c, e, i, cnt = gpu.lorenzo_fwd(x, dims, ebx2_r, 32768, True, 2, int(np.ceil(n * 0.2)))
    y = gpu.lorenzo_inv(c, dims, e, i, cnt, ebx2, True, int(np.ceil(n * 0.2)))
    err = np.abs(y.astype(np.float64) - x.astype(np.float64))
    # reference tolerance is 1.01*eb (tests/stages/test_lorenzo.cpp:189)
    assert err.max() <= eb * 1.01


def test_outlier_capacity_overflow_counts_all(gpu, orc):
    """Count reports every outlier; only `capacity` of them are stored."""
    rng = np.random.default_rng(3)
    x = rng.normal(0, 100, 4096).astype(np.float32)  # almost every delta is an outlier
    ebx2_r, _ = eb_params(1e-3)
    oc, oe, oi, ocnt = orc.lorenzo_fwd(x, (4096, 1, 1), ebx2_r, 8, False, 2, 100)
    gc, ge, gi, gcnt = gpu.lorenzo_fwd(x, (4096, 1, 1), ebx2_r, 8, False, 2, 100)
    assert gcnt == ocnt > 100
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gc, oc)


@pytest.mark.parametrize("mode", [1, 2])
def test_value_base_on_device(gpu, orc, mode):
    rng = np.random.default_rng(5)
    x = (rng.normal(3, 2, 1_000_003)).astype(np.float32)
    user_eb = np.float32(1e-3)
    p = gpu.value_base_params(x, mode, user_eb)
    vb = np.float32(orc.value_base(x, mode))
    abs_eb = user_eb * vb
    assert p[0] == vb
    assert p[1] == abs_eb
    assert p[2] == np.float32(1) / (np.float32(2) * abs_eb)
    assert p[3] == np.float32(2) * abs_eb


def test_forward_reads_device_eb(gpu, orc):
    """REL mode path: the kernel takes ebx2_r from the device parameter block."""
    rng = np.random.default_rng(9)
    x = make_field(rng, 200 * 150, 200 * 150)
    from fzgpumodules_amd import _fzh
    p = gpu.value_base_params(x, 1, np.float32(1e-3))
    dpar = _fzh.DeviceBuffer.from_array(p)
    gc, ge, gi, gcnt = gpu.lorenzo_fwd(x, (200, 150, 1), np.float32(0), 32768, True, 2, x.size,
                                       d_params=dpar.ptr + 8)
    oc, oe, oi, ocnt = orc.lorenzo_fwd(x, (200, 150, 1), p[2], 32768, True, 2, x.size)
    np.testing.assert_array_equal(gc, oc)
    assert gcnt == ocnt


@pytest.mark.parametrize("dims", [(5000, 1, 1), (100, 70, 1), (37, 29, 21)])
def test_inverse_in_two_parts_on_two_streams(gpu, orc, dims):
    """fzh_lorenzo_quant_inv_bin on a second stream, then _apply after an
    event join: bit-exact with the oracle, as the one-call inverse."""
    import ctypes as C
    from fzgpumodules_amd._fzh import DeviceBuffer, call, lib, workspace
    rng = np.random.default_rng(11)
    nx, ny, nz = dims
    n = nx * ny * nz
    x = make_field(rng, n, n)
    ebx2_r, ebx2 = eb_params(1e-3)
    c, e, i, cnt = orc.lorenzo_fwd(x, dims, ebx2_r, 512, True, 2, n)
    assert cnt > 0
    y_ref = orc.lorenzo_inv(c, dims, e, i, cnt, ebx2, True, n)
    dc = DeviceBuffer.from_array(c)
    de = DeviceBuffer.from_array(e)
    di = DeviceBuffer.from_array(i)
    dn = DeviceBuffer.from_array(np.array([cnt], np.uint32))
    out = DeviceBuffer(4 * n)
    ws = workspace(lib().fzh_lorenzo_quant_inv_ws_bytes(n, nx, ny, nz, n))
    s1, s2 = C.c_void_p(), C.c_void_p()
    call("fzh_stream_create", C.byref(s1))
    call("fzh_stream_create", C.byref(s2))
    try:
        for rep in range(2):  # the scan clears the per-tile counters for the next call
            call("fzh_lorenzo_quant_inv_bin", s2.value, n, nx, ny, nz, de.ptr, di.ptr, dn.ptr, n, ws.ptr, ws.nbytes)
            call("fzh_stream_sync", s2.value)
            call("fzh_lorenzo_quant_inv_apply_f32", s1.value, dc.ptr, n, nx, ny, nz, de.ptr, di.ptr, dn.ptr, n,
                 np.float32(ebx2), 1, 2, out.ptr, ws.ptr, ws.nbytes)
            call("fzh_stream_sync", s1.value)
            y = out.download(np.float32, n)
            np.testing.assert_array_equal(y.view(np.uint32), y_ref.view(np.uint32))
    finally:
        call("fzh_stream_destroy", s1.value)
        call("fzh_stream_destroy", s2.value)
