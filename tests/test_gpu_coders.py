"""Parity of the HIP RLE, histogram, Huffman, Quantizer and Bitpack kernels
against the oracle (bit-exact streams) plus the reference's known answers.

Reference: modules/coders/rle/rle.cu, modules/util/histogram/histogram.cu,
modules/coders/huffman/phf/*, modules/quantizers/quantizer/quantizer.cu,
modules/coders/bitpack/bitpack_stage.cu; known answer tests/stages/test_rle.cpp:120-145.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rle_inputs():
    rng = np.random.default_rng(2)
    return {
        "known": np.array([7, 7, 7, 42, 42, 1], np.uint16),
        "single": np.array([5], np.uint16),
        "constant": np.full(100_000, 9, np.uint16),
        "runs": np.repeat(rng.integers(0, 5, 30_000).astype(np.uint16), rng.integers(1, 40, 30_000)),
        "noise": rng.integers(0, 3, 200_001).astype(np.uint16),
        "u8": np.repeat(rng.integers(0, 3, 5000).astype(np.uint8), 7),
        "u32": np.repeat(rng.integers(0, 1 << 30, 3000).astype(np.uint32), 3),
        "long_run_tail": np.concatenate([np.zeros(50_000, np.uint16), np.arange(10, dtype=np.uint16)]),
    }


@pytest.mark.parametrize("name", list(_rle_inputs()))
def test_rle_matches_oracle(gpu, orc, name):
    v = _rle_inputs()[name]
    s = gpu.rle_enc(v)
    np.testing.assert_array_equal(s, orc.rle_enc(v))
    np.testing.assert_array_equal(gpu.rle_dec(s, v.dtype, v.size), v)


def test_rle_known_answer(gpu):
    s = gpu.rle_enc(np.array([7, 7, 7, 42, 42, 1], np.uint16))
    assert int(s[:4].view(np.uint32)[0]) == 3


def test_rle_empty(gpu):
    s = gpu.rle_enc(np.zeros(0, np.uint16))
    assert s.size == 4 and int(s.view(np.uint32)[0]) == 0


def test_rle_decode_rejects_wrong_length(gpu, orc):
    v = _rle_inputs()["runs"]
    s = orc.rle_enc(v)
    with pytest.raises(ValueError):
        gpu.rle_dec(s, v.dtype, v.size + 5)


@pytest.mark.parametrize("bklen", [256, 1024, 4096])
def test_histogram_matches_oracle(gpu, orc, bklen):
    rng = np.random.default_rng(bklen)
    v = np.clip(rng.normal(bklen // 2, bklen / 20, 1_000_003), 0, bklen + 100).astype(np.uint16)
    np.testing.assert_array_equal(gpu.histogram(v, bklen), orc.histogram(v, bklen))


def test_huffman_book_matches_oracle_and_reference(gpu, orc):
    rng = np.random.default_rng(8)
    for trial in range(40):
        bklen = [256, 1024][trial % 2]
        f = (rng.geometric(0.05, bklen) * (rng.random(bklen) < 0.6)).astype(np.uint32)
        f[0] += 1
        g = gpu.huffman_book(f, 2)
        o = orc.huffman_book(f, 2)
        np.testing.assert_array_equal(g[0], o[0])
        np.testing.assert_array_equal(g[1], o[1])


def _huffman_inputs():
    rng = np.random.default_rng(12)
    return {
        "gauss": np.clip(rng.normal(512, 6, 300_000), 0, 1023).astype(np.uint16),
        "skewed": np.where(rng.random(250_001) < 0.97, 512, rng.integers(0, 1024, 250_001)).astype(np.uint16),
        "uniform": rng.integers(0, 1024, 100_000).astype(np.uint16),
        "tiny": np.array([1, 2, 3, 3, 3], np.uint16),
        "constant": np.full(5000, 17, np.uint16),
        "partial_partition": np.clip(rng.normal(100, 30, 768 * 10 + 13), 0, 1023).astype(np.uint16),
        "u8": np.clip(rng.normal(128, 10, 77_777), 0, 255).astype(np.uint8),
    }


@pytest.mark.parametrize("name", list(_huffman_inputs()))
def test_huffman_matches_oracle(gpu, orc, name):
    v = _huffman_inputs()[name]
    bklen = 256 if v.dtype == np.uint8 else 1024
    s = gpu.huffman_encode(v, bklen)
    o = orc.huffman_encode(v, bklen)
    assert s.size == o.size
    np.testing.assert_array_equal(s, o)
    np.testing.assert_array_equal(gpu.huffman_decode(o, v.dtype, v.size), v)


def test_huffman_long_codes(gpu, orc):
    """Fibonacci-like frequencies force codes longer than the 10-bit lookup."""
    fib = [1, 1]
    while len(fib) < 24:
        fib.append(fib[-1] + fib[-2])
    v = np.concatenate([np.full(c, i, np.uint16) for i, c in enumerate(fib)])
    v = np.random.default_rng(0).permutation(v)
    s = gpu.huffman_encode(v, 1024)
    np.testing.assert_array_equal(s, orc.huffman_encode(v, 1024))
    np.testing.assert_array_equal(gpu.huffman_decode(s, np.uint16, v.size), v)


@pytest.mark.parametrize("name", ["gauss", "skewed", "tiny", "constant", "partial_partition", "u8", "small_alpha",
                                  "u8_all_symbols", "u8_all_symbols_partial"])
def test_huffman_fine_matches_oracle(gpu, orc, name):
    """Fine mode: 1024-symbol partitions when every codeword of the book fits
    8 bits (hf_hl.cc:69-80), else the coarse stream.  Unused symbols keep
    the all-ones word (hf_bk.cc:96-109), whose length field reads 31, so the
    fine stream needs every symbol of [0, bklen) to occur."""
    rng = np.random.default_rng(3)
    inputs = _huffman_inputs()
    inputs["small_alpha"] = rng.integers(0, 40, 1024 * 37 + 5).astype(np.uint16)
    inputs["u8_all_symbols"] = rng.integers(0, 256, 1024 * 300).astype(np.uint8)
    inputs["u8_all_symbols_partial"] = rng.integers(0, 256, 1024 * 77 + 333).astype(np.uint8)
    v = inputs[name]
    bklen = 256 if v.dtype == np.uint8 else 1024
    s = gpu.huffman_encode(v, bklen, fine=True)
    o = orc.huffman_encode(v, bklen, fine=True)
    np.testing.assert_array_equal(s, o)
    np.testing.assert_array_equal(gpu.huffman_decode(s, v.dtype, v.size), v)
    sublen = int(o[4:8].view(np.int32)[0])
    bk4, _ = orc.huffman_book(np.bincount(v, minlength=bklen).astype(np.uint32), v.dtype.itemsize)
    assert sublen == (1024 if int((bk4 >> 27).max()) <= 8 else 768)


@pytest.mark.parametrize("fine", [False, True])
def test_huffman_fine_in_pipeline(gpu, orc, fine):
    from fzgpumodules_amd.pipeline import Pipeline
    v = np.random.default_rng(4).integers(0, 256, 200_000).astype(np.uint8)
    with Pipeline(v.nbytes) as p:
        h = p.add_huffman(1, 256)
        p.set_huffman_fine(h, fine)
        p.finalize()
        blob = p.compress_array(v)
        np.testing.assert_array_equal(p.decompress_array(blob, np.uint8)[: v.size], v)
    np.testing.assert_array_equal(blob, orc.huffman_encode(v, 256, fine=fine))
    assert int(blob[4:8].view(np.int32)[0]) == (1024 if fine else 768)  # phf_header.sublen


def test_huffman_rejects_out_of_range(gpu):
    with pytest.raises(ValueError):
        gpu.huffman_encode(np.array([1, 2, 2000], np.uint16), 1024)


# ---------------------------------------------------------------------------
# Huffman decoder on streams the encoders here never write, built with the
# numpy stream definition that pins the oracle's layout: partition lengths
# (phf_header.sublen) other than 768 and 1024, partitions denser than the
# decoder's LDS staging, and corrupt partition metadata.  The reference
# decoder takes any sublen (hf_kernels.cu:258-312).
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32])
@pytest.mark.parametrize("sublen", [1, 64, 333, 768, 1024, 2048, 5000])
def test_huffman_decode_any_sublen(gpu, orc, dtype, sublen):
    from test_cpu_oracle import huffman_stream_def
    rng = np.random.default_rng(sublen)
    bklen = 256 if dtype == np.uint8 else 1024
    n = 3001 if sublen < 100 else 40_011
    v = np.minimum(np.abs(rng.laplace(0, 8, n)), bklen - 1).astype(dtype)
    f = orc.histogram(v, bklen)
    bk4, rev = orc.huffman_book(f, np.dtype(dtype).itemsize)
    s = huffman_stream_def(v, bk4, rev, bklen, sublen)
    np.testing.assert_array_equal(orc.huffman_decode(s, dtype, v.size), v)
    np.testing.assert_array_equal(gpu.huffman_decode(s, dtype, v.size), v)


def _fib_book(orc):
    """Fibonacci frequencies on 24 of 1024 symbols: codewords of 1 to 23 bits."""
    fib = [1, 1]
    while len(fib) < 24:
        fib.append(fib[-1] + fib[-2])
    f = np.zeros(1024, np.uint32)
    f[:24] = fib[::-1]
    bk4, rev = orc.huffman_book(f, 2)
    lens = (bk4[:24] >> 27).astype(np.int64)
    return bk4, rev, lens


def test_huffman_decode_dense_partitions(gpu, orc):
    """Partitions of mostly 20+-bit codewords (over 448 cells per 768-symbol
    partition, past the decoder's LDS staging) next to ordinary ones."""
    from test_cpu_oracle import huffman_stream_def
    bk4, rev, lens = _fib_book(orc)
    rng = np.random.default_rng(5)
    rare = np.nonzero(lens >= 20)[0]
    common = np.nonzero(lens <= 6)[0]
    v = np.concatenate([rng.choice(common, 768 * 3), rng.choice(rare, 768 * 4 + 100),
                        rng.choice(common, 500)]).astype(np.uint16)
    s = huffman_stream_def(v, bk4, rev, 1024, 768)
    ent = s[40:64].view(np.uint32)
    nbit = s[ent[2]:ent[3]].view(np.uint32)
    assert (nbit[3:7] > 448 * 32).all() and (nbit[:3] < 448 * 32).all()
    np.testing.assert_array_equal(orc.huffman_decode(s, np.uint16, v.size), v)
    np.testing.assert_array_equal(gpu.huffman_decode(s, np.uint16, v.size), v)


@pytest.mark.parametrize("part", [0, 3, 389])
def test_huffman_decode_rejects_wrong_bit_count(gpu, orc, part):
    """A partition whose par_nbit is one more than its codewords' bits does
    not end on a codeword boundary with its symbol count: status set."""
    v = _huffman_inputs()["gauss"]
    s = orc.huffman_encode(v, 1024)
    ent = s[40:64].view(np.uint32)
    nb = s[ent[2]:ent[3]].view(np.uint32)
    assert nb.size == 391
    bad = s.copy()
    bad[ent[2]:ent[3]].view(np.uint32)[part] = int(nb[part]) + (1 if nb[part] % 32 else -1)
    with pytest.raises(ValueError):
        gpu.huffman_decode(bad, np.uint16, v.size)
    np.testing.assert_array_equal(gpu.huffman_decode(s, np.uint16, v.size), v)


@pytest.mark.parametrize("zigzag", [False, True])
@pytest.mark.parametrize("code_bytes,radius", [(2, 32768), (1, 64), (4, 1 << 20)])
def test_quantizer_matches_oracle(gpu, orc, zigzag, code_bytes, radius):
    rng = np.random.default_rng(radius)
    x = rng.normal(0, 1, 300_007).astype(np.float32)
    x[::997] *= 1e4
    eb = np.float32(1e-3)
    r = np.float32(1) / (np.float32(2) * eb)
    gc, gv, gi, gn = gpu.quant_abs_fwd(x, r, radius, zigzag, code_bytes, max_outliers=x.size)
    oc, ov, oi, on = orc.quant_abs_fwd(x, r, radius, zigzag, code_bytes, max_outliers=x.size)
    assert gn == on
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gv.view(np.uint32), ov.view(np.uint32))
    y = gpu.quant_abs_inv(gc, np.float32(2) * eb, zigzag, gv, gi, gn)
    np.testing.assert_array_equal(y.view(np.uint32), orc.quant_abs_inv(oc, np.float32(2) * eb, zigzag, ov, oi, on).view(np.uint32))


@pytest.mark.parametrize("nbits", [1, 2, 4, 8, 16])
def test_bitpack_matches_oracle(gpu, orc, nbits):
    rng = np.random.default_rng(nbits)
    v = rng.integers(0, 1 << nbits, 100_003).astype(np.uint16)
    b = gpu.bitpack_enc(v, nbits)
    np.testing.assert_array_equal(b, orc.bitpack_enc(v, nbits))
    np.testing.assert_array_equal(gpu.bitpack_dec(b, v.size, np.uint16, nbits), v)
    assert gpu.max_u32(v) == int(v.max())
    assert orc.bitpack_nbits_for_max(int(v.max()), 2) <= 16
