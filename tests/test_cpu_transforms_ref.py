"""Pins of the transform oracle (oracle/transforms_ref.py): the reference's
own known answers (tests/stages/test_zigzag.cpp:32-80,
test_negabinary.cpp:34-69, test_negabinary_stage.cpp:168-183,
test_difference.cpp:88-351), exhaustive round trips, the number-system
definitions, and closed forms of the prefix sums."""
import numpy as np
import pytest

from oracle import transforms_ref as T


def test_zigzag_known_answers():
    x = np.array([0, -1, 1, -2, 2, -3, 3], np.int32)
    assert T.zigzag_enc(x).tolist() == [0, 1, 2, 3, 4, 5, 6]
    ext = np.array([np.iinfo(np.int32).max, np.iinfo(np.int32).min], np.int32)
    assert T.zigzag_enc(ext).tolist() == [np.iinfo(np.int32).max * 2, np.iinfo(np.uint32).max]
    for dt in (np.int8, np.int16):
        assert T.zigzag_enc(np.array([0, -1, 1, -2, 2], dt)).tolist() == [0, 1, 2, 3, 4]


@pytest.mark.parametrize("dt", [np.int8, np.int16])
def test_zigzag_and_negabinary_exhaustive(dt):
    info = np.iinfo(dt)
    x = np.arange(info.min, info.max + 1, dtype=np.int64).astype(dt)
    z = T.zigzag_enc(x)
    assert np.array_equal(T.zigzag_dec(z, dt), x)
    assert np.unique(z).size == x.size  # a bijection
    nb = T.negabinary_enc(x)
    assert np.array_equal(T.negabinary_dec(nb, dt), x)
    assert np.unique(nb).size == x.size


def test_negabinary_known_answers():
    x = np.array([0, 1, -1, 2, -2], np.int32)
    assert T.negabinary_enc(x).tolist() == [0, 1, 3, 6, 2]


@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.int64])
def test_negabinary_is_base_minus_two(dt):
    """The code's W bits are base -2 digits of the value: sum b_k (-2)^k
    equals x modulo 2^W, and equals x itself wherever x is representable
    with W digits."""
    rng = np.random.default_rng(1)
    info = np.iinfo(dt)
    w = 8 * np.dtype(dt).itemsize
    x = rng.integers(info.min, info.max, 2000, dtype=np.int64, endpoint=True).astype(dt)
    u = T.negabinary_enc(x)
    hi = sum(4 ** k for k in range((w + 1) // 2))          # 1 + 4 + 16 + ...
    lo = -sum(2 * 4 ** k for k in range(w // 2))           # -(2 + 8 + 32 + ...)
    for xv, uv in zip(x.tolist(), u.tolist()):
        val = sum(((uv >> k) & 1) * (-2) ** k for k in range(w))
        assert (val - xv) % (1 << w) == 0
        if lo <= xv <= hi:
            assert val == xv


def test_difference_known_answers():
    x = np.array([10, 11, 12, 13, 20, 21, 22, 23, 30, 31, 32, 33], np.int32)
    d = T.diff_fwd(x, chunk_elems=4)
    assert d.tolist() == [10, 1, 1, 1, 20, 1, 1, 1, 30, 1, 1, 1]
    assert np.array_equal(T.diff_inv(d, np.int32, chunk_elems=4), x)
    f = np.array([1, 3, 6, 10, 15, 21, 28], np.float32)
    df = T.diff_fwd(f)
    assert df.tolist() == [1, 2, 3, 4, 5, 6, 7]
    assert np.array_equal(T.diff_inv(df, np.float32), f)


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.int32, np.int64])
@pytest.mark.parametrize("chunk", [0, 512, 4096])
def test_difference_integer_round_trip(dt, chunk):
    rng = np.random.default_rng(3)
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, 10000, dtype=np.int64, endpoint=True).astype(dt)
    assert np.array_equal(T.diff_inv(T.diff_fwd(x, chunk), dt, chunk), x)


@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.int64])
def test_difference_negabinary_round_trip(dt):
    rng = np.random.default_rng(4)
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, 5000, dtype=np.int64, endpoint=True).astype(dt)
    d = T.diff_fwd(x, 1024, negabinary=True)
    assert d.dtype == np.dtype(T._UNSIGNED[np.dtype(dt)])
    assert np.array_equal(T.diff_inv(d, dt, 1024, negabinary=True), x)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_difference_float_inverse_exact_sums(dt):
    """Integer-valued floats sum exactly in any association: the inverse is
    the exact cumulative sum (f64 reference), per chunk."""
    rng = np.random.default_rng(5)
    d = rng.integers(-1000, 1000, 9000).astype(dt)
    for chunk in (0, 512, 2048):
        got = T.diff_inv(d, dt, chunk)
        c = chunk or d.size
        want = np.concatenate([np.cumsum(d[b:b + c].astype(np.float64)) for b in range(0, d.size, c)])
        assert np.array_equal(got.astype(np.float64), want)


def test_difference_float_association_is_not_sequential():
    """The raking association differs from a left-to-right running sum on
    general floats (so the order is actually exercised): for some input the
    two disagree, and the raking one equals an element-by-element restatement
    of one tile."""
    rng = np.random.default_rng(6)
    d = (rng.standard_normal(512) * 10 ** rng.uniform(-3, 3, 512)).astype(np.float32)
    rak = T.diff_inv(d, np.float32)
    seq = np.empty_like(d)
    acc = np.float32(0)
    for i, v in enumerate(d):
        acc = np.float32(acc + v) if i else v
        seq[i] = acc
    assert not np.array_equal(rak, seq)
    # scalar restatement of one tile
    part = [d[16 * s] for s in range(32)]
    for s in range(32):
        for j in range(1, 16):
            part[s] = np.float32(part[s] + d[16 * s + j])
    inc = list(part)
    for o in (1, 2, 4, 8, 16):
        inc = [np.float32(inc[s - o] + inc[s]) if s >= o else inc[s] for s in range(32)]
    want = np.empty(512, np.float32)
    for s in range(32):
        run = d[16 * s] if s == 0 else np.float32(inc[s - 1] + d[16 * s])
        want[16 * s] = np.float32(run + np.float32(0))
        for j in range(1, 16):
            run = np.float32(run + d[16 * s + j])
            want[16 * s + j] = np.float32(run + np.float32(0))
    assert np.array_equal(rak, want)


@pytest.mark.parametrize("dims", [(1000, 1, 1), (64, 64, 1), (37, 23, 1), (16, 16, 16), (9, 7, 5)])
@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.int64])
def test_integer_lorenzo_round_trip(dims, dt):
    rng = np.random.default_rng(7)
    n = int(np.prod(dims))
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, n, dtype=np.int64, endpoint=True).astype(dt)
    d = T.lorenzo_int_fwd(x, dims)
    assert np.array_equal(T.lorenzo_int_inv(d, dims), x)


def test_integer_lorenzo_closed_forms():
    # a constant field: the only nonzero 2D/3D delta is at the origin; 1D at every block start
    c = np.full(64 * 32, 7, np.int32)
    d2 = T.lorenzo_int_fwd(c, (64, 32, 1))
    assert d2[0] == 7 and not d2[1:].any()
    c3 = np.full(8 * 8 * 8, 5, np.int32)
    d3 = T.lorenzo_int_fwd(c3, (8, 8, 8))
    assert d3[0] == 5 and not d3[1:].any()
    c1 = np.full(1000, 3, np.int32)
    d1 = T.lorenzo_int_fwd(c1, (1000, 1, 1))
    assert (d1[::256] == 3).all() and d1.sum() == 3 * 4
    # a ramp in x: delta 1 along the first row, 0 elsewhere except the first column
    r = np.tile(np.arange(40, dtype=np.int32), 30)
    dr = T.lorenzo_int_fwd(r, (40, 30, 1)).reshape(30, 40)
    assert (dr[0, 1:] == 1).all() and dr[0, 0] == 0 and not dr[1:].any()
