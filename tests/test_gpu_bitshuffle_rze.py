"""Parity of the HIP Bitshuffle and RZE kernels against the oracle (bit-exact
byte streams) plus the reference's known answers.

Reference: modules/shufflers/bitshuffle/bitshuffle_stage.cu,
modules/coders/rze/rze_stage.cu; known answers from
tests/stages/test_bitshuffle_stage.cpp:264-370 and test_rze_stage.cpp:197-214.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ew", [1, 2, 4, 8])
@pytest.mark.parametrize("nbytes", [16384, 16384 * 3 + 100, 16384 * 5])
def test_bitshuffle_matches_oracle(gpu, orc, ew, nbytes):
    rng = np.random.default_rng(ew * 1000 + nbytes)
    b = rng.integers(0, 256, nbytes).astype(np.uint8)
    enc = gpu.bitshuffle(b, 16384, ew)
    np.testing.assert_array_equal(enc, orc.bitshuffle_enc(b, 16384, ew))
    np.testing.assert_array_equal(gpu.bitshuffle(enc, 16384, ew, inverse=True), b)


@pytest.mark.parametrize("ew,bs", [(2, 2048), (4, 4096), (8, 8192), (1, 1024)])
def test_bitshuffle_small_blocks(gpu, orc, ew, bs):
    rng = np.random.default_rng(bs)
    b = rng.integers(0, 256, bs * 3).astype(np.uint8)
    np.testing.assert_array_equal(gpu.bitshuffle(b, bs, ew), orc.bitshuffle_enc(b, bs, ew))


def test_bitshuffle_known_answers(gpu):
    ones = np.full(16384, 0xFF, np.uint8)
    np.testing.assert_array_equal(gpu.bitshuffle(ones, 16384, 4), ones)
    msb = np.full(1024, 0x80000000, np.uint32)
    w = gpu.bitshuffle(msb.view(np.uint8), 4096, 4).view(np.uint32)
    assert (w[:32] == 0xFFFFFFFF).all() and (w[32:] == 0).all()


def _rze_cases():
    rng = np.random.default_rng(42)
    n = 16384 * 6
    sparse = (rng.random(n) < 0.05).astype(np.uint8) * rng.integers(1, 255, n).astype(np.uint8)
    runs = np.repeat(rng.integers(0, 4, n // 64).astype(np.uint8), 64)
    checker = np.tile(np.array([0, 0xAA], np.uint8), n // 2)
    return {
        "zeros": np.zeros(16384, np.uint8),
        "zeros_multi": np.zeros(16384 * 4, np.uint8),
        "sparse": sparse,
        "runs": runs,
        "checker": checker,
        "random_raw": rng.integers(0, 256, 16384 * 2).astype(np.uint8),
        "partial_chunk": sparse[: 16384 * 2 + 777],
        "tiny": np.array([0, 3, 0, 0, 5], np.uint8),
        "random_partial": rng.integers(0, 256, 50000).astype(np.uint8),
        "mixed": np.concatenate([np.zeros(16384, np.uint8), rng.integers(0, 256, 16384).astype(np.uint8),
                                 sparse[:16384]]),
    }


@pytest.mark.parametrize("name", list(_rze_cases()))
def test_rze_matches_oracle(gpu, orc, name):
    data = _rze_cases()[name]
    s_gpu = gpu.rze_enc(data)
    s_orc = orc.rze_enc(data)
    np.testing.assert_array_equal(s_gpu, s_orc)
    np.testing.assert_array_equal(gpu.rze_dec(s_orc, data.size), data)


def _many_chunks(nbytes, seed):
    """Sparse chunks with every fifth chunk random (stored raw) and every
    seventh all zero."""
    rng = np.random.default_rng(seed)
    b = rng.integers(0, 256, nbytes, dtype=np.uint8)
    r = rng.integers(0, 256, nbytes, dtype=np.uint8)
    b[(r & 0xF0) != 0] = 0
    full = nbytes // 16384
    rows = b[: full * 16384].reshape(full, 16384)
    rows[::5] = rng.integers(0, 256, (rows[::5].shape[0], 16384), dtype=np.uint8)
    rows[::7] = 0
    return b


@pytest.mark.parametrize("nbytes", [16384 * 4096, 16384 * 4096 + 1000, 16384 * 4200])
def test_rze_chunk_offsets_both_ways(gpu, orc, nbytes):
    """Up to 4096 chunks every workgroup sums the sizes before its chunk; from
    4097 on the offsets come from a separate scan.  Same bytes as the oracle
    either way, plain and bit-shuffled."""
    data = _many_chunks(nbytes, nbytes)
    s_gpu = gpu.rze_enc(data)
    np.testing.assert_array_equal(s_gpu, orc.rze_enc(data))
    np.testing.assert_array_equal(gpu.rze_dec(s_gpu, data.size), data)
    got = gpu.rze_enc(data, shuffle_width=2)
    np.testing.assert_array_equal(got, orc.rze_enc(orc.bitshuffle_enc(data, 16384, 2)))
    np.testing.assert_array_equal(gpu.rze_dec(got, data.size, shuffle_width=2), data)


def test_rze_all_zero_chunk_known_answer(gpu):
    s = gpu.rze_enc(np.zeros(16384, np.uint8))
    assert s.size <= 17  # test_rze_stage.cpp:197-214
    assert s.size == 16


def test_rze_decode_rejects_corrupt_header(gpu, orc):
    data = _rze_cases()["sparse"]
    s = orc.rze_enc(data).copy()
    s[0:4] = np.frombuffer(np.uint32(12345).tobytes(), np.uint8)
    with pytest.raises(ValueError):
        gpu.rze_dec(s, data.size)


def test_bitshuffle_rze_chain_on_codes(gpu, orc):
    """Lorenzo codes -> Bitshuffle -> RZE, the config-2 coder chain."""
    rng = np.random.default_rng(1)
    codes = np.clip(rng.normal(0, 3, 16384 * 8 // 2), -50, 50).astype(np.int16)
    z = ((codes.astype(np.int32) << 1) ^ (codes.astype(np.int32) >> 15)).astype(np.uint16)
    sh = gpu.bitshuffle(z.view(np.uint8), 16384, 2)
    s = gpu.rze_enc(sh)
    np.testing.assert_array_equal(s, orc.rze_enc(orc.bitshuffle_enc(z.view(np.uint8), 16384, 2)))
    back = gpu.bitshuffle(gpu.rze_dec(s, sh.size), 16384, 2, inverse=True)
    np.testing.assert_array_equal(back.view(np.uint16), z)


@pytest.mark.parametrize("ew", [1, 2])
@pytest.mark.parametrize("name", ["zeros_multi", "sparse", "runs", "random_raw", "partial_chunk", "mixed"])
def test_fused_bitshuffle_rze_matches_two_stages(gpu, orc, ew, name):
    """fzh_rze_enc_shuffled == RZE(Bitshuffle(x)) byte for byte, and the fused
    decoder returns x (bitshuffle_stage.cu + rze_stage.cu composed)."""
    b = _rze_cases()[name]
    want = orc.rze_enc(orc.bitshuffle_enc(b, 16384, ew))
    got = gpu.rze_enc(b, shuffle_width=ew)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(gpu.rze_dec(got, b.size, shuffle_width=ew), b)
