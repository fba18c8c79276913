"""Pins the CPU oracle (CPU only; no GPU).

Each check is independent of the oracle's own formulas: the reference's
known-answer tests, a compiled copy of the reference's Huffman codebook
builder (fixtures in tests/golden/), closed forms on integer-valued inputs
where every floating-point step is exact, brute-force definitions written
here with numpy, library routines (zlib's CRC32) and round-trip invariants.
"""
import os
import zlib

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# CRC32 (compressor_io.cpp:25-52)
# ---------------------------------------------------------------------------
def test_crc32_check_value(orc):
    assert orc.crc32(b"123456789") == 0xCBF43926


def test_crc32_matches_zlib(orc):
    rng = np.random.default_rng(1)
    for n in (0, 1, 7, 8, 9, 1000, 65537):
        b = rng.integers(0, 256, n).astype(np.uint8)
        assert orc.crc32(b) == zlib.crc32(b.tobytes())


# ---------------------------------------------------------------------------
# Lorenzo: closed forms on integer data (eb = 0.5 -> ebx2_r = 1, s = x)
# ---------------------------------------------------------------------------
def zigzag16(d):
    d = d.astype(np.int64)
    return (((d << 1) ^ (d >> 15)) & 0xFFFF).astype(np.uint16)


def tile_deltas_1d(s, tile=1024):
    d = s.astype(np.int64).copy()
    d[1:] -= s[:-1].astype(np.int64)
    d[::tile] = s[::tile]  # chain restarts at every tile start (lorenzo_quant.cu:308-312)
    return d


@pytest.mark.parametrize("zz", [False, True])
def test_lorenzo_1d_closed_form_all_int16_deltas(orc, zz):
    # every int16 delta value appears once, inside tiles, in shuffled order
    rng = np.random.default_rng(7)
    want = rng.permutation(np.arange(-32768, 32768, dtype=np.int64))
    n = want.size
    s = np.zeros(n, np.int64)
    for t0 in range(0, n, 1024):
        s[t0:t0 + 1024] = np.cumsum(want[t0:t0 + 1024])
    assert np.abs(s).max() < 2 ** 24  # exactly representable in float32
    x = s.astype(np.float32)
    codes, err, idx, cnt = orc.lorenzo_fwd(x, (n, 1, 1), np.float32(1.0), 32768, zz, 2, n)
    d = tile_deltas_1d(s)
    np.testing.assert_array_equal(d, want)
    ok = np.abs(d) < 32768
    exp = np.where(ok, zigzag16(d) if zz else (d & 0xFFFF).astype(np.uint16), 0)
    np.testing.assert_array_equal(codes, exp)
    assert cnt == int((~ok).sum()) == 1
    np.testing.assert_array_equal(idx, np.nonzero(~ok)[0])
    np.testing.assert_array_equal(err, d[~ok].astype(np.float32))


def test_lorenzo_1d_radius_and_u8(orc):
    rng = np.random.default_rng(8)
    s = rng.integers(-300, 300, 5000)
    x = s.astype(np.float32)
    codes, err, idx, cnt = orc.lorenzo_fwd(x, (x.size, 1, 1), np.float32(1.0), 64, True, 1, x.size)
    d = tile_deltas_1d(s)
    ok = np.abs(d) < 64
    np.testing.assert_array_equal(codes, np.where(ok, zigzag16(d) & 0xFF, 0).astype(np.uint8))
    np.testing.assert_array_equal(idx, np.nonzero(~ok)[0])
    assert cnt == int((~ok).sum())


def lorenzo_nd_deltas(s, tile):
    """Inclusion-exclusion inside tiles of shape `tile` (zero outside the tile)."""
    s = s.astype(np.int64)
    d = s.copy()
    nd = s.ndim
    for ax in range(nd):
        prev = np.roll(d, 1, axis=ax)
        coord = np.arange(s.shape[ax]) % tile[ax]
        sh = [1] * nd
        sh[ax] = -1
        prev = np.where(coord.reshape(sh) == 0, 0, prev)
        d = d - prev
    return d


@pytest.mark.parametrize("shape", [(50, 70), (32, 32), (65, 33)])
def test_lorenzo_2d_closed_form(orc, shape):
    ny, nx = shape
    rng = np.random.default_rng(ny * nx)
    s = rng.integers(-4000, 4000, shape)
    x = s.astype(np.float32)
    n = x.size + 100  # padded buffer: codes past ny*nx are zero
    xp = np.zeros(n, np.float32)
    xp[: x.size] = x.ravel()
    codes, err, idx, cnt = orc.lorenzo_fwd(xp, (nx, ny, 1), np.float32(1.0), 2048, True, 2, n)
    d = lorenzo_nd_deltas(s, (32, 32))
    ok = np.abs(d) < 2048
    exp = np.where(ok, zigzag16(d), 0).astype(np.uint16).ravel()
    np.testing.assert_array_equal(codes[: x.size], exp)
    assert not codes[x.size:].any()
    assert cnt == int((~ok).sum())
    # order: tile-major (ty*gx+tx), row-major inside the tile
    gx = (nx + 31) // 32
    lin = np.nonzero((~ok).ravel())[0]
    yy, xx = lin // nx, lin % nx
    key = ((yy // 32) * gx + xx // 32) * 1024 + (yy % 32) * 32 + xx % 32
    np.testing.assert_array_equal(idx, lin[np.argsort(key, kind="stable")])
    np.testing.assert_array_equal(err, d.ravel()[idx].astype(np.float32))


@pytest.mark.parametrize("shape", [(9, 10, 17), (8, 8, 8), (3, 20, 5)])
def test_lorenzo_3d_closed_form(orc, shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(nz * ny * nx)
    s = rng.integers(-1000, 1000, shape)
    x = s.astype(np.float32)
    codes, err, idx, cnt = orc.lorenzo_fwd(x, (nx, ny, nz), np.float32(1.0), 512, False, 2, x.size)
    d = lorenzo_nd_deltas(s, (8, 8, 8))
    ok = np.abs(d) < 512
    np.testing.assert_array_equal(codes, np.where(ok, d & 0xFFFF, 0).astype(np.uint16).ravel())
    assert cnt == int((~ok).sum())
    assert set(idx.tolist()) == set(np.nonzero((~ok).ravel())[0].tolist())


@pytest.mark.parametrize("dims", [(3000, 1, 1), (70, 50, 1), (17, 10, 9)])
def test_lorenzo_inverse_closed_form(orc, dims):
    """x_hat = fl(q * ebx2) with q the exact integer prefix of the codes."""
    nx, ny, nz = dims
    shape = tuple(v for v in (nz, ny, nx) if v > 1) if ny > 1 else (nx,)
    rng = np.random.default_rng(nx)
    s = rng.integers(-20000, 20000, shape)
    x = s.astype(np.float32)
    codes, err, idx, cnt = orc.lorenzo_fwd(x, dims, np.float32(1.0), 32768, True, 2, x.size)
    ebx2 = np.float32(0.37)
    y = orc.lorenzo_inv(codes, dims, err, idx, cnt, ebx2, True)
    np.testing.assert_array_equal(y, (s.astype(np.float32) * ebx2).ravel())


def test_lorenzo_round_trip_error_bound(orc):
    from conftest import eb_params

    rng = np.random.default_rng(5)
    for dims, shape in (((100000, 1, 1), (100000,)), ((300, 200, 1), (200, 300)), ((40, 30, 20), (20, 30, 40))):
        x = np.cumsum(rng.normal(0, 1e-2, int(np.prod(shape)))).astype(np.float32)
        eb = np.float32(1e-3)
        r, e2 = eb_params(eb)
        codes, err, idx, cnt = orc.lorenzo_fwd(x, dims, r, 32768, True, 2, x.size)
        y = orc.lorenzo_inv(codes, dims, err, idx, cnt, e2, True)
        # reference tolerance, tests/stages/test_lorenzo.cpp:189
        assert np.abs(y.astype(np.float64) - x.astype(np.float64)).max() <= eb * 1.01


def test_value_base_textbook(orc):
    rng = np.random.default_rng(2)
    x = (rng.normal(0, 3, 12345) + 1.5).astype(np.float32)
    assert orc.value_base(x, 1) == float(np.float32(max(abs(x.min()), abs(x.max()))))
    assert orc.value_base(x, 2) == float(np.float32(x.max() - x.min()))


# ---------------------------------------------------------------------------
# Bitshuffle (bitshuffle_stage.cu:299-349; known answers test_bitshuffle_stage.cpp)
# ---------------------------------------------------------------------------
def bitshuffle_def(block: np.ndarray, ew: int) -> np.ndarray:
    """Definition of the bit-plane layout.  ew 1/2 (ballot kernel,
    bitshuffle_stage.cu:299-349): with N = block/ew elements and npp = N/32,
    word (W-1-b)*npp + g holds bit b of elements 32g..32g+31, element 32g+e
    at bit e.  ew 4 (32x32 butterfly, :66-170): the same words, element 32g+e
    at bit 31-e.  ew 8 (64x64 butterfly, :172-293): 64-bit words, u64
    (63-b)*(N/64) + g holds bit b of elements 64g..64g+63, element 64g+e at
    bit 63-e.  Plane order MSB first in all cases."""
    W = 8 * ew
    v = block.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[ew]).astype(np.uint64)
    G = 64 if ew == 8 else 32
    npp = v.size // G
    e = np.arange(G, dtype=np.uint64)
    pos = (np.uint64(G - 1) - e) if ew >= 4 else e
    out = np.zeros(W * npp, np.uint64)
    for b in range(W):
        bits = ((v >> np.uint64(b)) & np.uint64(1)).reshape(npp, G)
        acc = np.zeros(npp, np.uint64)
        for j in range(G):
            acc |= bits[:, j] << pos[j]
        out[(W - 1 - b) * npp:(W - b) * npp] = acc
    return (out if ew == 8 else out.astype(np.uint32)).view(np.uint8)


@pytest.mark.parametrize("ew", [1, 2, 4, 8])
def test_bitshuffle_definition(orc, ew):
    rng = np.random.default_rng(ew)
    bs = 1024 * ew * 2
    b = rng.integers(0, 256, bs * 3 + 77).astype(np.uint8)
    enc = orc.bitshuffle_enc(b, bs, ew)
    for k in range(3):
        np.testing.assert_array_equal(enc[k * bs:(k + 1) * bs], bitshuffle_def(b[k * bs:(k + 1) * bs], ew))
    np.testing.assert_array_equal(enc[3 * bs:], b[3 * bs:])  # tail copied verbatim
    np.testing.assert_array_equal(orc.bitshuffle_dec(enc, bs, ew), b)


def test_bitshuffle_known_answers(orc):
    ones = np.full(16384, 0xFF, np.uint8)
    np.testing.assert_array_equal(orc.bitshuffle_enc(ones, 16384, 4), ones)
    msb = np.full(1024, 0x80000000, np.uint32)
    w = orc.bitshuffle_enc(msb.view(np.uint8), 4096, 4).view(np.uint32)
    assert (w[:32] == 0xFFFFFFFF).all() and (w[32:] == 0).all()


# ---------------------------------------------------------------------------
# RZE (rze_stage.cu; known answer test_rze_stage.cpp:197-214)
# ---------------------------------------------------------------------------
def test_rze_all_zero_chunk(orc):
    s = orc.rze_enc(np.zeros(16384, np.uint8))
    assert s.size == 16  # 8-byte header + 4-byte size + 2-byte tag, padded to 4
    assert s[:4].view(np.uint32)[0] == 16384 and s[4:8].view(np.uint32)[0] == 1


def test_rze_round_trips(orc):
    rng = np.random.default_rng(3)
    n = 16384 * 5 + 1234
    cases = [
        np.zeros(n, np.uint8),
        rng.integers(0, 256, n).astype(np.uint8),
        ((rng.random(n) < 0.03) * rng.integers(1, 256, n)).astype(np.uint8),
        np.repeat(rng.integers(0, 3, n // 100 + 1).astype(np.uint8), 100)[:n],
    ]
    for b in cases:
        s = orc.rze_enc(b)
        assert s.size % 4 == 0
        np.testing.assert_array_equal(orc.rze_dec(s, b.size), b)
    raw = orc.rze_enc(cases[1][:16384])
    assert raw[8:12].view(np.uint32)[0] & 0x80000000  # incompressible chunk stored raw


# RZE chunk and stream format restated with numpy primitives, independently
# of the oracle's loops (rze_stage.cu:135-224 ZE/RE helpers, :324-445 the
# levels, raw-fallback thresholds and size tag, :774-843 the stream):
#   ZE:  nonzero bytes in order; bitmap bit i of byte i/8 = (byte i != 0)
#   RE:  byte i kept iff it differs from byte i-1, byte -1 := 0; bitmap of
#        the kept flags, bit i of byte i/8
#   chunk: [ZE data][RE(bm1) data][RE(bm2) data][RE(bm3) data][bm4, 4 B]
#          [in_size u16 LE]; all-zero chunk -> the 2-byte tag alone;
#          incompressible -> the input verbatim, size | 0x8000_0000
#   stream: [orig u32][nchunks u32][size word u32 x nchunks][chunks], zero
#           padded to 4 bytes
# The `mutant` switches restate plausible mistakes (MSB-first bitmaps, a
# repeat-elimination seed of 0xFF, the bitmap levels written in reverse);
# the tests check that each one changes the known answers.
# ---------------------------------------------------------------------------
RZE_CS = 16384


def _bits_to_bytes(flags, mutant):
    return np.packbits(flags, bitorder="big" if mutant == "msb_first" else "little")


def _re_def(b, mutant=None):
    seed = 0xFF if mutant == "seed_ff" else 0
    prev = np.concatenate([np.array([seed], np.uint8), b[:-1]])
    keep = b != prev
    return b[keep], _bits_to_bytes(keep, mutant)


def rze_chunk_def(x, mutant=None):
    """(stored bytes, size word, branch taken)."""
    x = np.asarray(x, np.uint8)
    n = x.size
    tag = np.array([n & 0xFF, (n >> 8) & 0xFF], np.uint8)
    nz = x != 0
    if not nz.any():
        return tag, 2, "zero"

    def raw(why):
        return x.copy(), 0x80000000 | n, why

    avail = RZE_CS - 2
    data = x[nz]
    bm1 = _bits_to_bytes(nz, mutant)
    w = data.size
    if w >= avail or w > avail - bm1.size:
        return raw("ze")
    r1, bm2 = _re_def(bm1, mutant)
    w += r1.size
    if w >= avail:
        return raw("re1")
    if w > avail - bm2.size:
        return raw("bm2")
    r2, bm3 = _re_def(bm2, mutant)
    w += r2.size
    if w >= avail:
        return raw("re2")
    if w >= avail - 4 or w > avail - 4 - bm3.size:
        return raw("bm3")
    r3, bm4 = _re_def(bm3, mutant)
    w += r3.size
    if w >= avail - 4:
        return raw("re3")
    bm4p = np.zeros(4, np.uint8)
    bm4p[: bm4.size] = bm4
    levels = [r1, r2, r3] if mutant != "levels_reversed" else [r3, r2, r1]
    out = np.concatenate([data] + levels + [bm4p, tag])
    return out, out.size, "packed"


def rze_stream_def(b, mutant=None):
    b = np.asarray(b, np.uint8)
    nc = -(-b.size // RZE_CS)
    parts = [rze_chunk_def(b[c * RZE_CS:(c + 1) * RZE_CS], mutant) for c in range(nc)]
    head = np.array([b.size, nc] + [p[1] for p in parts], np.uint32).view(np.uint8)
    s = np.concatenate([head] + [p[0] for p in parts])
    return np.concatenate([s, np.zeros(-s.size % 4, np.uint8)])


def _values_for(keep, pop_total):
    """Bytes v with v[i] != v[i-1] (v[-1] = 0) exactly where keep[i], and
    sum(popcount(v)) = pop_total where the run structure allows it (the
    returned total says what was reached)."""
    keep = np.asarray(keep, bool)
    starts = np.nonzero(keep)[0]
    forced_zero = not keep[0]  # a first run equal to the seed 0
    if forced_zero:
        starts = np.r_[0, starts]
    lens = np.diff(np.r_[starts, keep.size])
    c = np.full(starts.size, 7)
    if forced_zero:
        c[0] = 0
    total = int((c * lens).sum())

    def lowest(r):  # a run must differ from the run before it (run -1 = 0)
        return 1 if (r == 0 or c[r - 1] == 0) else 0

    # lower run popcounts (a run may drop to 0 unless a neighbour is 0) or
    # raise them to 8 (0xFF, never on two adjacent runs); longest runs first
    # so the final steps are fine grained
    order = np.argsort(-lens, kind="stable")
    for r in order:
        if forced_zero and r == 0:
            continue
        while total > pop_total and c[r] > lowest(r) and total - lens[r] >= pop_total:
            if c[r] == 1 and r + 1 < c.size and c[r + 1] == 0:
                break
            c[r] -= 1
            total -= lens[r]
        if total < pop_total and c[r] == 7 and total + lens[r] <= pop_total:
            if (r == 0 or c[r - 1] != 8) and (r + 1 == c.size or c[r + 1] != 8):
                c[r] = 8
                total += lens[r]
    pal = {k: [v for v in range(256) if bin(v).count("1") == k] for k in range(9)}
    vals = np.zeros(starts.size, np.uint8)
    prev = 0
    for r in range(starts.size):
        choices = [v for v in pal[int(c[r])] if v != prev] if not (forced_zero and r == 0) else [0]
        vals[r] = choices[0]  # the smallest: set bits low, so lower levels start at byte 0
        prev = int(vals[r])
    return np.repeat(vals, lens), total


def _chunk_with_counts(nr4_bits, nr3, nr2, nz, rng):
    """A full 16 KiB chunk whose level counts are (nz, nr2, nr3, nr4): the
    bitmaps are chosen top-down (bm3 from the bm4 flags, bm2 from bm3, bm1
    from bm2) and the chunk's nonzero bytes follow bm1."""
    bm4 = np.zeros(32, bool)
    bm4[:nr4_bits] = True
    bm3, p3 = _values_for(bm4, nr3)
    bm2, p2 = _values_for(np.unpackbits(bm3, bitorder="little").astype(bool), nr2)
    bm1, p1 = _values_for(np.unpackbits(bm2, bitorder="little").astype(bool), nz)
    mask = np.unpackbits(bm1, bitorder="little").astype(bool)
    x = np.where(mask, rng.integers(1, 256, RZE_CS), 0).astype(np.uint8)
    return x, (p1, p2, p3, nr4_bits)


def test_rze_known_answers(orc):
    """Hand-derived chunks: ZE + every RE level, and a repeat-heavy bitmap."""
    x = np.zeros(24, np.uint8)
    x[[2, 3, 4]] = 5
    x[16], x[23] = 7, 9
    # data 05 05 05 07 09 | bm1 = 1C 00 81, all kept | bm2 = 07 | bm3 = 01 |
    # bm4 = 01 00 00 00 | tag 24
    chunk = [5, 5, 5, 7, 9, 0x1C, 0x00, 0x81, 0x07, 0x01, 1, 0, 0, 0, 24, 0]
    want = np.array([24, 0, 0, 0, 1, 0, 0, 0, 16, 0, 0, 0] + chunk, np.uint8)
    np.testing.assert_array_equal(rze_stream_def(x), want)
    np.testing.assert_array_equal(orc.rze_enc(x), want)
    y = np.full(64, 0x11, np.uint8)
    # bm1 = FF x 8 -> RE keeps byte 0 only: FF | bm2 = 01 | bm3 = 01
    chunk = [0x11] * 64 + [0xFF, 0x01, 0x01, 1, 0, 0, 0, 64, 0]
    want = np.array([64, 0, 0, 0, 1, 0, 0, 0, 73, 0, 0, 0] + chunk + [0, 0, 0], np.uint8)
    np.testing.assert_array_equal(rze_stream_def(y), want)
    np.testing.assert_array_equal(orc.rze_enc(y), want)
    # each plausible mistake changes one of the two known answers
    for m in ("msb_first", "seed_ff", "levels_reversed"):
        assert not (np.array_equal(rze_stream_def(x, m), orc.rze_enc(x))
                    and np.array_equal(rze_stream_def(y, m), orc.rze_enc(y))), m


def test_rze_threshold_boundaries(orc):
    """Every raw-fallback branch of rze_stage.cu:369-431, each at the exact
    count where it starts to apply, against the definition."""
    rng = np.random.default_rng(77)
    cases = []
    # ZE level: nz = avail - bm1_bytes = 14334 passes, 14335 falls back
    for w in (14333, 14334, 14335):
        x = np.zeros(RZE_CS, np.uint8)
        x[:w] = rng.integers(1, 256, w)
        cases.append((x, "ze" if w > 14334 else "packed"))
    # (bm4 flags, nr3, nr2, nz): the cumulative write position meets each
    # later threshold exactly (16382 = avail, 16126 = avail - 256 bm2 bytes,
    # 16378 = avail - 4, 16346 = avail - 4 - 32 bm3 bytes)
    for spec, branch in (((2, 1, 2048, 14334), "re1"),           # 14334 + 2048 = 16382
                         ((31, 220, 1792, 14334), "packed"),     # 16126, 16346, 16377: every test passes
                         ((32, 224, 1793, 14334), "bm2"),        # 16127 > 16126
                         ((32, 224, 1792, 14330), "re3"),        # 16122 + 224 + 32 = 16378
                         ((31, 224, 1792, 14330), "packed"),     # 16377 + 6 = 16383 bytes
                         ((32, 224, 1792, 14331), "bm3")):       # 16347 > 16346
        x, got = _chunk_with_counts(*spec, rng)
        assert got == (spec[3], spec[2], spec[1], spec[0]), (spec, got)
        cases.append((x, branch))
    for x, branch in cases:
        s, size, why = rze_chunk_def(x)
        assert why == branch, (why, branch)
        enc = orc.rze_enc(x)
        np.testing.assert_array_equal(enc, rze_stream_def(x))
        assert enc[8:12].view(np.uint32)[0] == size
        if branch == "packed":
            np.testing.assert_array_equal(orc.rze_dec(enc, x.size), x)


def test_rze_matches_definition(orc):
    rng = np.random.default_rng(78)
    n = 16384 * 6 + 5000
    cases = [
        np.zeros(n, np.uint8),
        rng.integers(0, 256, n).astype(np.uint8),
        ((rng.random(n) < 0.03) * rng.integers(1, 256, n)).astype(np.uint8),
        ((rng.random(n) < 0.6) * rng.integers(1, 256, n)).astype(np.uint8),
        np.repeat(rng.integers(0, 3, n // 100 + 1).astype(np.uint8), 100)[:n],
        np.tile(np.array([0, 0, 1, 0], np.uint8), n // 4 + 1)[:n],
        rng.integers(0, 256, 16384 + 15000).astype(np.uint8),   # raw partial last chunk
        np.r_[np.zeros(16384, np.uint8), np.full(300, 3, np.uint8)],
    ]
    for b in cases:
        np.testing.assert_array_equal(orc.rze_enc(b), rze_stream_def(b))


# ---------------------------------------------------------------------------
# RLE (rle.cu; known answer test_rle.cpp:120-145)
# ---------------------------------------------------------------------------
def rle_def(v):
    starts = np.r_[0, np.nonzero(v[1:] != v[:-1])[0] + 1] if v.size else np.zeros(0, np.int64)
    lengths = np.diff(np.r_[starts, v.size]).astype(np.uint32)
    vals = v[starts]
    vb = vals.tobytes()
    vb += b"\0" * ((-len(vb)) % 4)
    return np.frombuffer(np.uint32(len(starts)).tobytes() + vb + lengths.tobytes(), np.uint8)


def test_rle_known_answer(orc):
    v = np.array([7, 7, 7, 42, 42, 1], np.uint16)
    s = orc.rle_enc(v)
    assert s[:4].view(np.uint32)[0] == 3
    np.testing.assert_array_equal(s, rle_def(v))
    np.testing.assert_array_equal(orc.rle_dec(s, np.uint16, 6), v)


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32])
def test_rle_definition(orc, dt):
    rng = np.random.default_rng(np.dtype(dt).itemsize)
    v = np.repeat(rng.integers(0, 5, 3000), rng.integers(1, 9, 3000)).astype(dt)
    s = orc.rle_enc(v)
    np.testing.assert_array_equal(s, rle_def(v))
    np.testing.assert_array_equal(orc.rle_dec(s, dt, v.size), v)


# ---------------------------------------------------------------------------
# Histogram / Huffman
# ---------------------------------------------------------------------------
def test_histogram_bruteforce(orc):
    rng = np.random.default_rng(4)
    v = rng.integers(0, 1500, 100000).astype(np.uint16)
    f = orc.histogram(v, 1024)
    np.testing.assert_array_equal(f, np.bincount(v[v < 1024], minlength=1024))  # >= bklen skipped


def _fixtures():
    z = np.load(os.path.join(HERE, "golden", "huffman_books.npz"), allow_pickle=False)
    return z, [str(n) for n in z["names"]]


def test_huffman_codebook_matches_reference_fixtures(orc):
    z, names = _fixtures()
    assert len(names) >= 20
    for name in names:
        f = z[f"{name}.freq"]
        eb = int(z[f"{name}.eb"][0])
        bk4, rev = orc.huffman_book(f, eb)
        np.testing.assert_array_equal(bk4, z[f"{name}.bk4"], err_msg=name)
        np.testing.assert_array_equal(rev, z[f"{name}.rev"], err_msg=name)


def test_huffman_codebook_matches_compiled_reference(orc):
    if orc.ref_phf() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(11)
    for k in range(40):
        bk = int(rng.choice([256, 1024, 4096]))
        f = (rng.integers(0, 1000, bk) * (rng.random(bk) < rng.uniform(0.05, 1))).astype(np.uint32)
        if f.sum() == 0:
            continue
        a, b = orc.huffman_book(f, 2), orc.ref_huffman_book(f, 2)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


def test_huffman_codebook_invariants(orc):
    z, names = _fixtures()
    for name in names:
        bk4 = z[f"{name}.bk4"].astype(np.uint64)
        f = z[f"{name}.freq"]
        lens = (bk4 >> np.uint64(27)).astype(np.int64)  # bit count in the top 5 bits
        used = (f > 0) & (lens > 0)
        if used.sum() < 2 or lens.max() > 27:
            continue  # >27-bit codes are cut off by the reference (hf_bk.cc:113): no prefix code
        assert (2.0 ** -lens[used]).sum() <= 1.0 + 1e-12, name  # Kraft
        codes = [(int(c) & ((1 << 27) - 1), int(l)) for c, l in zip(bk4[used], lens[used])]
        strs = sorted(format(c, f"0{l}b") for c, l in codes)
        for a, b in zip(strs, strs[1:]):
            assert not b.startswith(a), name  # prefix-free


@pytest.mark.parametrize("bklen", [256, 1024])
def test_huffman_stream_round_trip(orc, bklen):
    rng = np.random.default_rng(bklen)
    v = np.minimum(np.abs(rng.normal(0, 20, 100000)), bklen - 1).astype(np.uint16)
    s = orc.huffman_encode(v, bklen)
    np.testing.assert_array_equal(orc.huffman_decode(s, np.uint16, v.size), v)
    assert not s[64:128].any()  # header filler bytes zero in this build


# PHF stream restated with numpy primitives, independently of the oracle's
# loops (hf_kernels.cu:73-122 packing, hf_hl.cc:50-67 metadata, hf_buf.cc:
# 151-176 the merge, hf.h:38-44 the header):
#   codeword w = bk4[symbol]: length = w >> 27, code = low 27 bits
#   partition p (sublen symbols): the codes' bits, most significant first,
#   concatenated and cut into u32 cells whose first bit is bit 31;
#   par_nbit[p] = its bit count; par_entry = exclusive prefix of the cell
#   counts; stream = [header 64 B][64 zero B][revbook][par_nbit][par_entry]
#   [cells].  `mutant="lsb_first"` packs each cell from bit 0 instead.
# ---------------------------------------------------------------------------
def huffman_stream_def(v, bk4, rev, bklen, sublen, mutant=None):
    v = np.asarray(v)
    words = bk4[v.astype(np.int64)].astype(np.uint64)
    lens = (words >> np.uint64(27)).astype(np.int64)
    codes = words & np.uint64(0x07FFFFFF)
    pardeg = -(-v.size // sublen)
    nbits, cells = [], []
    for p in range(pardeg):
        L, cw = lens[p * sublen:(p + 1) * sublen], codes[p * sublen:(p + 1) * sublen]
        M = (cw[:, None] >> np.arange(26, -1, -1, dtype=np.uint64)[None, :]) & np.uint64(1)
        bits = M[np.arange(27)[None, :] >= 27 - L[:, None]].astype(np.uint8)
        packed = np.packbits(bits, bitorder="big" if mutant != "lsb_first" else "little")
        packed = np.concatenate([packed, np.zeros(-packed.size % 4, np.uint8)])
        cells.append(np.frombuffer(packed.tobytes(), ">u4").astype("<u4"))
        nbits.append(bits.size)
    ncell = np.array([c.size for c in cells], np.int64)
    entry = np.r_[0, np.cumsum(ncell)[:-1]].astype(np.uint32)
    tot_cells = int(ncell.sum())
    h = np.zeros(128, np.uint8)
    h[0:2] = np.frombuffer(np.uint16(bklen).tobytes(), np.uint8)
    h[4:12] = np.array([sublen, pardeg], np.int32).view(np.uint8)
    h[16:40] = np.array([v.size, sum(nbits), tot_cells], np.uint64).view(np.uint8)
    e = np.cumsum([0, 128, rev.size, 4 * pardeg, 4 * pardeg, 4 * tot_cells]).astype(np.uint32)
    h[40:64] = e.view(np.uint8)
    return np.concatenate([h, rev, np.array(nbits, np.uint32).view(np.uint8), entry.view(np.uint8)]
                          + [c.view(np.uint8) for c in cells])


def test_huffman_stream_known_answer(orc):
    """Four symbols with frequencies 4, 2, 1, 1: canonical codes 1, 01, 000,
    001 (the compiled reference builder's book), so 0 1 2 3 0 0 1 0 packs
    to the 14 bits 1 01 000 001 1 1 01 1 = cell 0xA0EC0000."""
    v = np.array([0, 1, 2, 3, 0, 0, 1, 0], np.uint16)
    f = orc.histogram(v, 4)
    bk4, rev = orc.huffman_book(f, 2)
    assert [(int(w) >> 27, int(w) & 0x7FFFFFF) for w in bk4] == [(1, 1), (2, 1), (3, 0), (3, 1)]
    ref = orc.ref_huffman_book(f, 2)
    if ref is not None:
        np.testing.assert_array_equal(ref[0], bk4)
    s = orc.huffman_encode(v, 4)
    ent = s[40:64].view(np.uint32)
    assert ent.tolist() == [0, 128, 128 + rev.size, 132 + rev.size, 136 + rev.size, 140 + rev.size]
    assert s[ent[2]:ent[3]].view(np.uint32).tolist() == [14]
    assert s[ent[3]:ent[4]].view(np.uint32).tolist() == [0]
    assert s[ent[4]:ent[5]].view(np.uint32).tolist() == [0xA0EC0000]
    np.testing.assert_array_equal(s, huffman_stream_def(v, bk4, rev, 4, 768))
    assert not np.array_equal(s, huffman_stream_def(v, bk4, rev, 4, 768, mutant="lsb_first"))


@pytest.mark.parametrize("case", ["coarse_1024", "coarse_256_u8", "fine_256_u8", "long_codes"])
def test_huffman_stream_matches_definition(orc, case):
    rng = np.random.default_rng(len(case))
    if case == "coarse_1024":
        v = np.minimum(np.abs(rng.normal(0, 30, 50000)), 1023).astype(np.uint16)
        bklen, fine, sub = 1024, False, 768
    elif case == "coarse_256_u8":
        v = np.minimum(np.abs(rng.laplace(0, 6, 40001)), 255).astype(np.uint8)
        bklen, fine, sub = 256, False, 768
    elif case == "fine_256_u8":
        # every symbol used, every codeword 8 bits: the fine layout applies
        v = np.random.default_rng(21).integers(0, 256, 1024 * 50 + 77).astype(np.uint8)
        bklen, fine, sub = 256, True, 1024
    else:  # geometric frequencies: codewords well past 16 bits
        v = np.minimum(rng.geometric(0.45, 200000) - 1, 1023).astype(np.uint16)
        bklen, fine, sub = 1024, False, 768
    f = orc.histogram(v, bklen)
    bk4, rev = orc.huffman_book(f, v.dtype.itemsize)
    if case == "long_codes":
        assert int((bk4[f > 0] >> 27).max()) > 16
    s = orc.huffman_encode(v, bklen, fine=fine)
    assert int(s[4:8].view(np.int32)[0]) == sub
    np.testing.assert_array_equal(s, huffman_stream_def(v, bk4, rev, bklen, sub))


# ---------------------------------------------------------------------------
# Quantizer ABS / Bitpack (quantizer.cu:30-67; bitpack_stage.cu:30-124)
# ---------------------------------------------------------------------------
def test_quantizer_closed_form(orc):
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5, 3.49, 1e6, -7.0], np.float32)
    codes, vals, idx, cnt = orc.quant_abs_fwd(x, np.float32(1.0), 512, False, 2)
    q = np.rint(x.astype(np.float64)).astype(np.int64)  # half to even
    ok = np.abs(q) < 512
    np.testing.assert_array_equal(codes, np.where(ok, q & 0xFFFF, 0).astype(np.uint16))
    np.testing.assert_array_equal(idx, np.nonzero(~ok)[0])
    np.testing.assert_array_equal(vals, x[~ok])


def test_bitpack_known_answers(orc):
    for m, nb in ((15, 4), (200, 8), (255, 8), (65535, 16), (0, 1), (1, 1), (2, 2), (3, 2), (4, 4)):
        assert orc.bitpack_nbits_for_max(m, 2) == nb, m
    v = np.arange(1024, dtype=np.uint16) & 0xF
    for nb in (4, 8, 16):
        b = orc.bitpack_enc(v, nb)
        assert b.size == 1024 * nb // 8
        np.testing.assert_array_equal(orc.bitpack_dec(b, v.size, np.uint16, nb), v)
    b = orc.bitpack_enc(np.array([1, 2, 3, 0], np.uint16), 2)  # element k of a byte at bit 2k
    assert b.tolist() == [0b00111001]


# ---------------------------------------------------------------------------
# Concat layout of the pipeline test helper (src/pipeline/pipeline_utils.h:36-46)
# ---------------------------------------------------------------------------
def test_concat_layout_helper():
    import chains

    parts = [np.zeros(3, np.float32), np.array([5], np.uint32), np.arange(5, dtype=np.uint8)]
    b = chains.concat(parts)
    assert b[:4].view(np.uint32)[0] == 3
    assert b[4:28].view(np.uint64).tolist() == [12, 4, 5]
    assert b.size == 32 + 16 + 16 + 16
    assert [p.size for p in chains.split(b)] == [12, 4, 5]
    assert not b[28:32].any() and not b[32 + 12:48].any()


def test_huffman_fine_mode_layout(orc):
    """Fine mode (huffman_stage.cu:41, hf_hl.cc:69-80, hf_buf.cc:48-49):
    1024-symbol partitions when every codeword of the book is at most 8 bits
    -- unused symbols keep the all-ones word (hf_bk.cc:96-109), so only a
    book using every symbol qualifies -- else the coarse 768; the stream
    decodes back and its partition bit counts equal the codeword lengths."""
    rng = np.random.default_rng(21)
    v = rng.integers(0, 256, 1024 * 50 + 77).astype(np.uint8)
    s = orc.huffman_encode(v, 256, fine=True)
    hdr = s[:64]
    sublen, pardeg = int(hdr[4:8].view(np.int32)[0]), int(hdr[8:12].view(np.int32)[0])
    assert (sublen, pardeg) == (1024, -(-v.size // 1024))
    assert np.array_equal(orc.huffman_decode(s, np.uint8, v.size), v)
    bk4, _ = orc.huffman_book(np.bincount(v, minlength=256).astype(np.uint32), 1)
    lens = (bk4 >> 27)[v].astype(np.int64)
    ent = hdr[40:64].view(np.uint32)
    nbit = s[ent[2]:ent[3]].view(np.uint32)
    assert np.array_equal(nbit, np.add.reduceat(lens, np.arange(0, v.size, 1024)))
    # one unused symbol: the book's all-ones word forces the coarse stream
    w = np.where(v == 7, 8, v).astype(np.uint8)
    assert int(orc.huffman_encode(w, 256, fine=True)[4:8].view(np.int32)[0]) == 768
