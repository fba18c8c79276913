"""The device exclusive scan (fzh_scan_u32) against numpy: the offsets
primitive under RZE chunk offsets, outlier compaction and binning, and
Huffman cell offsets.  Covers both launch shapes (one workgroup up to 65536
elements, three launches above), pass and segment boundaries, saturation,
wrap-around, in-place use and clearing of the input."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 7, 1023, 1024, 1025, 8191, 8192, 8193, 12656, 16383, 16384, 16385, 32769, 65535, 65536, 65537,
         100003, (1 << 20) + 5]


def expected(v, wrap):
    c = np.cumsum(v.astype(np.uint64), dtype=np.uint64)
    ex = np.concatenate([np.zeros(1, np.uint64), c[:-1]]) if v.size else np.zeros(0, np.uint64)
    tot = int(c[-1]) if v.size else 0
    if wrap:
        return (ex & np.uint64(0xFFFFFFFF)).astype(np.uint32), tot & 0xFFFFFFFF
    return np.minimum(ex, np.uint64(0xFFFFFFFF)).astype(np.uint32), min(tot, 0xFFFFFFFF)


@pytest.mark.parametrize("n", SIZES)
def test_scan_sizes(gpu, n):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 2000, n, dtype=np.uint32)
    got, tot, after = gpu.scan_u32(v)
    want, wtot = expected(v, False)
    assert np.array_equal(got, want)
    assert tot == wtot
    assert np.array_equal(after, v)


@pytest.mark.parametrize("n", [300, 70001])
@pytest.mark.parametrize("wrap", [False, True])
def test_scan_past_2_32(gpu, n, wrap):
    """Prefixes cross 2^32 part way: saturate at 0xFFFFFFFF, or wrap."""
    rng = np.random.default_rng(7 + n)
    hi = (1 << 27) if n == 300 else (1 << 20)
    v = rng.integers(0, hi, n, dtype=np.uint32)
    v[n // 2] = 0xFFFFFFF0  # one huge value in the middle
    got, tot, _ = gpu.scan_u32(v, wrap=wrap)
    want, wtot = expected(v, wrap)
    assert np.array_equal(got, want)
    assert tot == wtot
    if not wrap:
        assert got[-1] == 0xFFFFFFFF


@pytest.mark.parametrize("n", [5000, 65536, 200000])
def test_scan_in_place(gpu, n):
    v = np.random.default_rng(n).integers(0, 70000, n, dtype=np.uint32)
    got, tot, _ = gpu.scan_u32(v, in_place=True)
    want, wtot = expected(v, False)
    assert np.array_equal(got, want)
    assert tot == wtot


@pytest.mark.parametrize("n", [9000, 300000])
@pytest.mark.parametrize("wrap", [False, True])
def test_scan_clear(gpu, n, wrap):
    v = np.random.default_rng(3 * n).integers(0, 5000, n, dtype=np.uint32)
    got, tot, after = gpu.scan_u32(v, wrap=wrap, clear=True)
    want, wtot = expected(v, wrap)
    assert np.array_equal(got, want)
    assert tot == wtot
    assert not after.any()


def test_scan_rejects(gpu):
    from fzgpumodules_amd._fzh import DeviceBuffer, FzhError, call
    d = DeviceBuffer(64)
    t = DeviceBuffer(16)
    with pytest.raises(FzhError):
        call("fzh_scan_u32", None, d.ptr, d.ptr, 8, t.ptr, 2, None, 0)  # clear in place
    with pytest.raises(FzhError):
        call("fzh_scan_u32", None, d.ptr, d.ptr, 8, t.ptr, 4, None, 0)  # unknown flag
    with pytest.raises(FzhError):
        call("fzh_scan_u32", None, d.ptr, d.ptr, 100000, t.ptr, 0, None, 0)  # no workspace


def test_kernel_timing(gpu):
    """fzh_kernel_timing_*: one event pair per launch of a named kernel."""
    import ctypes as C
    from fzgpumodules_amd._fzh import call
    tot, cnt = C.c_double(), C.c_uint64()
    call("fzh_kernel_timing_enable", 1)
    try:
        v = np.random.default_rng(1).integers(0, 9, 40000, dtype=np.uint32)
        gpu.scan_u32(v)
        gpu.scan_u32(v)
        call("fzh_kernel_timing_get", b"scan_small_kernel", C.byref(tot), C.byref(cnt))
        assert cnt.value == 2 and tot.value > 0
        call("fzh_kernel_timing_get", b"no_such_kernel", C.byref(tot), C.byref(cnt))
        assert cnt.value == 0 and tot.value == 0
    finally:
        call("fzh_kernel_timing_enable", 0)
    gpu.scan_u32(v)
    call("fzh_kernel_timing_get", b"scan_small_kernel", C.byref(tot), C.byref(cnt))
    assert cnt.value == 0  # disabled: nothing recorded, totals cleared
