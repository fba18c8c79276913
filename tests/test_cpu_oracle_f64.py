"""Pins the oracle's FP64 Lorenzo path (orc_lorenzo_fwd_f64 / _inv_f64 /
orc_value_base_f64), which restates LorenzoQuantStage<double, u16|u32>
(modules/fused/lorenzo_quant/lorenzo_quant.h:335-339; kernels
lorenzo_quant.cu:262-364, lorenzo_quant_nd.cu:29-105,199-274 run with
TInput = double).  The reference holds no FP64 golden vectors, so the pins
are:
  * an independent vectorised numpy statement of the forward (tile-local
    differences with zero boundaries, |delta| < radius, (TCode)(int)delta or
    zigzag, outliers tile-major and in element order inside a tile),
    compared bit for bit on 1D/2D/3D, ragged and zero-padded shapes;
  * the inverse against the closed form: every delta is an integer, so each
    partial sum of the scan trees is exact and the inverse returns
    round(x * ebx2_r) * ebx2 exactly, whatever the association;
  * hand-derived known answers (halves rounded away from zero in double, a
    value whose float32 rounding would differ, the radius test, zigzag of
    both widths, u32 saturation);
  * agreement with the golden-pinned FP32 path where both are exact;
  * the value base against numpy in double, cast to float.
"""
import numpy as np
import pytest

CODE_DT = {2: np.uint16, 4: np.uint32}
SIGNED = {2: np.int16, 4: np.int32}


def rha(v):
    """round() of C (halves away from zero), exact in float64."""
    v = np.asarray(v, np.float64)
    t = np.trunc(v)
    return t + np.where(np.abs(v - t) >= 0.5, np.sign(v), 0.0)


def enc(d, zz, cb):
    q = np.clip(d, -2147483648.0, 2147483647.0).astype(np.int64).astype(np.int32)
    s = q.astype(SIGNED[cb])  # (SCode)(int)delta, two's complement wrap
    u = s.view(CODE_DT[cb])
    if not zz:
        return u
    w = 8 * np.dtype(CODE_DT[cb]).itemsize
    return ((u << 1) ^ (s >> (w - 1)).view(CODE_DT[cb])).astype(CODE_DT[cb])


def shift(a, axis, tile):
    """a shifted by one along `axis`, zero where the index is a tile start."""
    out = np.zeros_like(a)
    sl_dst = [slice(None)] * a.ndim
    sl_src = [slice(None)] * a.ndim
    sl_dst[axis] = slice(1, None)
    sl_src[axis] = slice(0, -1)
    out[tuple(sl_dst)] = a[tuple(sl_src)]
    idx = [slice(None)] * a.ndim
    idx[axis] = slice(0, None, tile)
    out[tuple(idx)] = 0.0
    return out


def fwd_numpy(x, dims, r, radius, zz, cb, n):
    """Codes, outlier errors, indices (canonical order) for FP64 input."""
    nx, ny, nz = dims
    rad = float(np.array(radius).astype(CODE_DT[cb]))
    codes = np.zeros(n, CODE_DT[cb])
    if nz > 1:
        s = rha(x[:nx * ny * nz].reshape(nz, ny, nx) * r)
        a = s - shift(s, 0, 8)
        b = a - shift(a, 2, 8)
        d = b - shift(b, 1, 8)
        zz_, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        gx, gy = -(-nx // 8), -(-ny // 8)
        tile = ((zz_ // 8) * gy + yy // 8) * gx + xx // 8
        local = (zz_ % 8) * 64 + (yy % 8) * 8 + xx % 8
    elif ny > 1:
        s = rha(x[:nx * ny].reshape(ny, nx) * r)
        t = s - shift(s, 0, 32)
        d = t - shift(t, 1, 32)
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        tile = (yy // 32) * (-(-nx // 32)) + xx // 32
        local = (yy % 32) * 32 + xx % 32
    else:
        s = rha(x[:n] * r)
        d = s - shift(s, 0, 1024)
        g = np.arange(n)
        tile, local = g // 1024, g % 1024
    d, tile, local = d.ravel(), tile.ravel(), local.ravel()
    q = np.abs(d) < rad
    cover = d.size
    codes[:cover] = np.where(q, enc(np.where(q, d, 0.0), zz, cb), 0)
    gid = np.arange(cover)
    o = np.nonzero(~q)[0]
    order = np.lexsort((local[o], tile[o]))
    return codes, d[o][order], gid[o][order].astype(np.uint32)


def field(rng, n, spikes=True):
    x = np.cumsum(rng.normal(0, 1, n)) * 0.01
    if spikes and n > 10:
        x[rng.integers(0, n, max(1, n // 100))] *= 1000
    return x


SHAPES = [((5000, 1, 1), None), ((3, 1, 1), None), ((4096, 1, 1), None), ((100, 70, 1), None),
          ((37, 33, 1), 37 * 33 + 333), ((37, 29, 21), None), ((9, 9, 9), 729 + 100), ((16, 16, 16), None)]


@pytest.mark.parametrize("dims,n", SHAPES)
@pytest.mark.parametrize("zz", [False, True])
@pytest.mark.parametrize("cb,radius", [(2, 512), (2, 32768), (4, 512), (4, 1 << 20)])
def test_forward_matches_numpy_statement(orc, dims, n, zz, cb, radius):
    nx, ny, nz = dims
    cover = nx * ny * nz
    n = cover if n is None else n
    rng = np.random.default_rng(cover + 7 * cb + zz)
    x = np.zeros(n)
    x[:cover] = field(rng, cover)
    r = 1.0 / (2.0 * 1e-3)
    c, e, i, cnt = orc.lorenzo_fwd(x, dims, r, radius, zz, cb, n, n)
    assert e.dtype == np.float64
    nc, ne, ni = fwd_numpy(x, dims, r, radius, zz, cb, n)
    assert cnt == ne.size
    np.testing.assert_array_equal(c, nc)
    np.testing.assert_array_equal(i, ni)
    np.testing.assert_array_equal(e.view(np.uint64), ne.view(np.uint64))


@pytest.mark.parametrize("dims,n", SHAPES)
@pytest.mark.parametrize("zz", [False, True])
@pytest.mark.parametrize("cb", [2, 4])
def test_inverse_closed_form(orc, dims, n, zz, cb):
    """Integer deltas: every partial sum exact, so x_hat = round(x r) * ebx2."""
    nx, ny, nz = dims
    cover = nx * ny * nz
    n = cover if n is None else n
    rng = np.random.default_rng(cover + cb)
    x = np.zeros(n)
    x[:cover] = field(rng, cover)
    eb = 1e-3 * (1 + rng.random())
    r, ebx2 = 1.0 / (2.0 * eb), 2.0 * eb
    c, e, i, cnt = orc.lorenzo_fwd(x, dims, r, 512, zz, cb, n, n)
    y = orc.lorenzo_inv(c, dims, e, i, cnt, ebx2, zz, n)
    assert y.dtype == np.float64
    want = rha(x[:cover] * r) * ebx2
    np.testing.assert_array_equal(y[:cover].view(np.uint64), want.view(np.uint64))
    assert np.max(np.abs(y[:cover] - x[:cover])) <= eb * (1 + 1e-9)


def test_known_answer_1d():
    import oracle
    # eb = 0.1 -> r = 5: x r = 0, 1.25, 2.5, 5, -18.5 -> s = 0, 1, 3, 5, -19
    # deltas 0, 1, 2, 2, -24; radius 16 makes the last an outlier
    x = np.array([0.0, 0.25, 0.5, 1.0, -3.7])
    r = 1.0 / (2.0 * 0.1)
    assert x[4] * r == -18.5
    c, e, i, cnt = oracle.lorenzo_fwd(x, (5, 1, 1), r, 16, False, 2)
    np.testing.assert_array_equal(c, [0, 1, 2, 2, 0])
    assert cnt == 1 and list(i) == [4] and list(e) == [-24.0]
    c, _, _, _ = oracle.lorenzo_fwd(x, (5, 1, 1), r, 16, True, 2)
    np.testing.assert_array_equal(c, [0, 2, 4, 4, 0])
    c, _, _, _ = oracle.lorenzo_fwd(x, (5, 1, 1), r, 32, True, 4)
    np.testing.assert_array_equal(c, [0, 2, 4, 4, 47])  # zigzag(-24) = 47


def test_double_rounding_not_float():
    """2.5 - 1e-10 rounds to 2 in double; as a float32 it is 2.5 and would
    round to 3 -- the FP64 path must not pass through float."""
    import oracle
    x = np.array([2.5 - 1e-10, 0.0])
    c, _, _, _ = oracle.lorenzo_fwd(x, (2, 1, 1), 1.0, 512, False, 2)
    np.testing.assert_array_equal(c, [2, np.uint16(-2 & 0xFFFF)])
    c32, _, _, _ = oracle.lorenzo_fwd(x.astype(np.float32), (2, 1, 1), 1.0, 512, False, 2)
    assert c32[0] == 3


def test_u32_codes_saturate_beyond_int_range():
    """With u32 codes and radius 2^32-1, |delta| >= 2^31 is quantizable and
    (int)delta saturates (as the GPUs' conversions do)."""
    import oracle
    x = np.array([3e9, 0.0, -3e9, 5.0])
    c, _, _, cnt = oracle.lorenzo_fwd(x, (4, 1, 1), 1.0, 0xFFFFFFFF, False, 4)
    assert cnt == 0
    # deltas 3e9, -3e9, -3e9, 3e9 + 5
    np.testing.assert_array_equal(c, [0x7FFFFFFF, 0x80000000, 0x80000000, 0x7FFFFFFF])
    c, _, _, _ = oracle.lorenzo_fwd(x, (4, 1, 1), 1.0, 0xFFFFFFFF, True, 4)
    np.testing.assert_array_equal(c, [0xFFFFFFFE, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFE])


@pytest.mark.parametrize("dims", [(3000, 1, 1), (70, 50, 1), (20, 17, 11)])
@pytest.mark.parametrize("zz", [False, True])
def test_agrees_with_fp32_path_where_both_are_exact(orc, dims, zz):
    """Eighths times r = 4 are exact in both precisions: codes, outlier
    indices and values equal those of the golden-pinned FP32 path."""
    n = int(np.prod(dims))
    rng = np.random.default_rng(n)
    x = np.cumsum(rng.integers(-40, 41, n)) / 8.0
    x[rng.integers(0, n, n // 50)] *= 100
    c64, e64, i64, n64 = orc.lorenzo_fwd(x, dims, 4.0, 512, zz, 2, n)
    c32, e32, i32, n32 = orc.lorenzo_fwd(x.astype(np.float32), dims, np.float32(4.0), 512, zz, 2, n)
    assert n64 == n32 > 0
    np.testing.assert_array_equal(c64, c32)
    np.testing.assert_array_equal(i64, i32)
    np.testing.assert_array_equal(e64, e32.astype(np.float64))


@pytest.mark.parametrize("mode", [1, 2])
def test_value_base_f64(orc, mode):
    rng = np.random.default_rng(mode)
    x = rng.normal(0, 1, 100001) * 1e7 + 3.0
    x[5] = -4.6e7
    want = np.float32(np.max(np.abs([x.min(), x.max()]))) if mode == 1 else np.float32(x.max() - x.min())
    assert orc.value_base(x, mode) == float(want)
    # the base is reduced in double and cast once: with float32 inputs the
    # maximum would round to 1 + 2^-23 and the base would be 2^-23
    y = np.array([1.0, 1.0 + 3 * 2.0 ** -25])
    assert orc.value_base(y, 2) == 3 * 2.0 ** -25
    assert orc.value_base(y.astype(np.float32), 2) == 2.0 ** -23
