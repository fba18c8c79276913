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
