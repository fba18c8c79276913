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
    general floats, so the order is actually exercised."""
    rng = np.random.default_rng(6)
    d = (rng.standard_normal(512) * 10 ** rng.uniform(-3, 3, 512)).astype(np.float32)
    rak = T.diff_inv(d, np.float32)
    seq = np.empty_like(d)
    acc = np.float32(0)
    for i, v in enumerate(d):
        acc = np.float32(acc + v) if i else v
        seq[i] = acc
    assert not np.array_equal(rak, seq)


# Known answers of the inverse difference's float association
# (diff.cu:75-103: cub::BlockScan<float, 512>::InclusiveSum per 512-element
# tile -- raking: 32 segments of 16 summed left to right, a Kogge-Stone scan
# of the segment sums with offsets 1, 2, 4, 8, 16, each segment's elements
# added left to right onto the previous segments' sum -- then
# `scan + prefix` with prefix = 0 + the earlier tiles' aggregates, left to
# right).  B = 2^24: B + 1 rounds to B, B + 3 and B + 5 round to B + 4 (ties
# to even), B + 2 and B + 4 are exact.  Every expected value below is worked
# out by hand from those rules; each case differs from what a plausible
# wrong association gives (noted per case).
B = np.float32(2.0 ** 24)


def _inv(d):
    return T.diff_inv(np.asarray(d, np.float32), np.float32)


def test_difference_association_serial_inside_a_segment():
    # segment 0 = [B, 1, 1]: B + 1 = B, then B + 1 = B (a tree inside the
    # segment would give (1 + 1) + B = B + 2 at index 2)
    d = np.zeros(512, np.float32)
    d[0:3] = [B, 1, 1]
    out = _inv(d)
    assert np.array_equal(out, np.full(512, B, np.float32))


def test_difference_association_kogge_stone_over_segments():
    # segment sums p = [B, 1, 1, 0, ...] (d[0] = B, d[16] = d[32] = 1).
    # Offset 1: [B, B, 2, 1, 0 ...]; offset 2: [B, B, B+2, B, 2, 1, 0 ...];
    # offset 4: [B, B, B+2, B, B+2, B, B+2, B, 2, 1, 0 ...]; offsets 8 and 16
    # carry the same pattern on, so inc[s] = B for s = 0, 1 and every odd s,
    # B + 2 for every even s >= 2.  Segment s >= 1 outputs inc[s-1] + its
    # elements: s = 1, 2: B + 1 = B; s >= 3: B + 2 when s is odd, else B.
    # (A sequential scan gives B everywhere; the exact sums are B + 2.)
    d = np.zeros(512, np.float32)
    d[0], d[16], d[32] = B, 1, 1
    want = np.full(512, B, np.float32)
    for s in range(3, 32, 2):
        want[16 * s:16 * s + 16] = B + 2
    assert np.array_equal(_inv(d), want)


def test_difference_association_segment_elements_after_the_prefix():
    # segment 1 = [1, 1]: its outputs are (inc[0] + 1) = B, then B + 1 = B;
    # inc[1] = B + 2 (exact), so segments 2.. output B + 2.  (Scanning the
    # segment locally and adding the prefix last gives B + 2 at index 17.)
    d = np.zeros(512, np.float32)
    d[0], d[16], d[17] = B, 1, 1
    want = np.full(512, B + 2, np.float32)
    want[:32] = B
    assert np.array_equal(_inv(d), want)


def test_difference_association_tile_prefix():
    # tiles: [B, 0..] [1, 1, 0..] [1, 0..] [1, 0..] [0..]
    # tile 1: local scan [1, 2, 2 ..] + prefix B -> [B, B+2, B+2 ..];
    # prefix B + 2; tile 2: 1 + (B+2) = B+3 -> B+4; prefix (B+2) + 1 -> B+4;
    # tile 3: 1 + (B+4) = B+5 -> B+4; prefix -> B+4; tile 4: B+4.
    # (A running sum over the whole chunk stays at B; updating the prefix
    # before adding it gives B+4 already at index 512.)
    d = np.zeros(5 * 512, np.float32)
    d[0] = B
    d[512], d[513] = 1, 1
    d[1024] = 1
    d[1536] = 1
    want = np.empty_like(d)
    want[:512] = B
    want[512] = B
    want[513:1024] = B + 2
    want[1024:] = B + 4
    assert np.array_equal(_inv(d), want)
    # chunks restart the prefix: the same tiles as chunks of 512 elements
    # are each scanned on their own
    got = T.diff_inv(d, np.float32, 512)
    want_c = np.zeros_like(d)
    want_c[:512] = B
    want_c[512] = 1
    want_c[513:1024] = 2
    want_c[1024:1536] = 1
    want_c[1536:2048] = 1
    assert np.array_equal(got, want_c)


def test_difference_association_negative_zero():
    # every tile adds its prefix (+0.0 for the first): -0.0 + 0.0 = +0.0
    got = _inv(np.full(1024, -0.0, np.float32))
    assert np.all(got == 0) and not np.any(np.signbit(got))
    got = T.diff_inv(np.full(1024, -0.0, np.float32), np.float32, 512)
    assert not np.any(np.signbit(got))


def test_difference_association_known_answers_reject_mutants(monkeypatch):
    """The known answers above fail an oracle whose segment sums are combined
    by a sequential scan instead of Kogge-Stone, and one whose tiles add the
    prefix inside the scan (the mistakes the cases are built to catch)."""
    real = T._raking_tile_scan

    def sequential_partials(t):
        seg = t.reshape(t.shape[0], 32, 16)
        part = seg[:, :, 0].copy()
        for j in range(1, 16):
            part = part + seg[:, :, j]
        inc = part.copy()
        for s_ in range(1, 32):
            inc[:, s_] = inc[:, s_ - 1] + part[:, s_]
        out = np.empty_like(seg)
        run = seg[:, :, 0].copy()
        run[:, 1:] = inc[:, :-1] + seg[:, 1:, 0]
        out[:, :, 0] = run
        for j in range(1, 16):
            run = run + seg[:, :, j]
            out[:, :, j] = run
        return out.reshape(t.shape[0], 512), inc[:, 31].copy()

    monkeypatch.setattr(T, "_raking_tile_scan", sequential_partials)
    with pytest.raises(AssertionError):
        test_difference_association_kogge_stone_over_segments()
    monkeypatch.setattr(T, "_raking_tile_scan", real)
    test_difference_association_kogge_stone_over_segments()


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
