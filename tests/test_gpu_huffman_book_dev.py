"""The Huffman codebook built on the device (fzh_huffman_build_book_device)
against the host builder (fzh_huffman_build_book, which restates
modules/coders/huffman/phf/hf_bk_impl1.cc:55-163, hf_bk_internal.cc:58-101
and hf_canon.cc:86-131 and is pinned by the CPU oracle tests): bk4 and the
reverse book byte for byte over histograms chosen for the heap's tie order
(all counts equal, small integer counts, repeated values), degenerate trees
(no symbol, one symbol, two symbols) and deep trees whose codes are
truncated to 28 bits; then HuffmanStage with the device codebook, whose
stream must equal the host-codebook stream, eagerly and under graph replay,
and its out-of-range report."""
import zlib

import numpy as np
import pytest

import chains
from fzgpumodules_amd import kernels as K
from fzgpumodules_amd._fzh import DeviceBuffer, call
from fzgpumodules_amd.pipeline import Pipeline

pytestmark = pytest.mark.gpu


def fib(k):
    a, b, out = 1, 1, []
    for _ in range(k):
        out.append(a)
        a, b = b, a + b
    return out


def histograms(bklen, rng):
    z = np.zeros(bklen, np.uint32)
    yield "zeros", z.copy()
    h = z.copy(); h[0] = 5
    yield "single_first", h
    h = z.copy(); h[-1] = 7
    yield "single_last", h
    if bklen >= 2:
        h = z.copy(); h[bklen // 3] = 1; h[-1] = 1
        yield "two_equal", h
        h = z.copy(); h[1] = 1000; h[0] = 3
        yield "two", h
    yield "all_ones", np.ones(bklen, np.uint32)
    yield "all_equal_big", np.full(bklen, 123456, np.uint32)
    yield "ties_small", rng.integers(0, 4, bklen).astype(np.uint32)
    yield "ties_1_2", rng.integers(1, 3, bklen).astype(np.uint32)
    # powers of two: a chain of depth up to 31, truncated to 28-bit codes
    h = z.copy()
    k = min(bklen, 32)
    h[rng.permutation(bklen)[:k]] = (1 << np.arange(k)).astype(np.uint32)
    yield "geometric", h
    # Fibonacci counts: the deepest tree for its total
    h = z.copy()
    k = min(bklen, 46)
    h[np.sort(rng.permutation(bklen)[:k])] = np.array(fib(k), np.uint32)
    yield "fibonacci", h
    h = z.copy()
    m = max(1, bklen // 20)
    h[rng.permutation(bklen)[:m]] = rng.integers(1, 1 << 24, m).astype(np.uint32)
    yield "sparse", h
    c = np.arange(bklen) - bklen // 2
    yield "laplace", np.floor(1e6 * np.exp(-np.abs(c) / 3.0)).astype(np.uint32)
    yield "dense", rng.integers(1, 1 << 30, bklen).astype(np.uint32)


CASES = [(1, 256), (1, 1), (1, 7), (1, 64), (2, 1024), (2, 768), (2, 3), (2, 513), (4, 1024), (4, 100), (4, 2)]


@pytest.mark.parametrize("eb,bklen", CASES)
def test_device_book_equals_host_book(gpu, eb, bklen):
    rng = np.random.default_rng(zlib.crc32(repr((eb, bklen)).encode()))
    for name, h in histograms(bklen, rng):
        bk, rev = K.huffman_book(h, eb)
        dbk, drev, st = K.huffman_book_device(h, eb)
        assert st == 0, name
        np.testing.assert_array_equal(dbk, bk, err_msg=name)
        np.testing.assert_array_equal(drev, rev, err_msg=name)


def test_deep_tree_is_truncated(gpu):
    """Powers of two up to 2^31: code lengths 1..31, the ones above 27 stored
    as length 28, exactly as the host builder does."""
    h = (1 << np.arange(32)).astype(np.uint32)
    bk, _ = K.huffman_book(h, 4)
    assert int((bk >> 27).max()) == 28
    dbk, _, st = K.huffman_book_device(h, 4)
    assert st == 0
    np.testing.assert_array_equal(dbk, bk)


def test_histogram_sum_mismatch_is_reported(gpu):
    h = np.random.default_rng(3).integers(0, 50, 1024).astype(np.uint32)
    n = int(h.sum())
    _, _, st = K.huffman_book_device(h, 2, n + 1)
    assert st == 1
    dbk, drev, st = K.huffman_book_device(h, 2, n)
    assert st == 0
    bk, rev = K.huffman_book(h, 2)
    np.testing.assert_array_equal(dbk, bk)
    np.testing.assert_array_equal(drev, rev)


def test_invalid_arguments_rejected(gpu):
    from fzgpumodules_amd._fzh import FzhError
    with pytest.raises(FzhError, match="bklen"):
        K.huffman_book_device(np.ones(1025, np.uint32), 2)
    with pytest.raises(FzhError, match="1-byte"):
        K.huffman_book_device(np.ones(300, np.uint32), 1)


def _huffman_pipeline(eb, bklen, n, device):
    p = Pipeline(n * eb)
    h = p.add_huffman(eb, bklen)
    if device:
        p.set_huffman_device_book(h)
    p.finalize()
    return p


SYM_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}


@pytest.mark.parametrize("eb,bklen", [(1, 256), (2, 1024), (2, 64), (4, 1024), (4, 513)])
@pytest.mark.parametrize("n", [1, 767, 768, 100003, 2000000])
def test_stage_stream_equals_host_book_stream(gpu, eb, bklen, n):
    rng = np.random.default_rng(zlib.crc32(repr((eb, bklen, n)).encode()))
    c = np.clip(np.rint(rng.laplace(bklen // 2, 4.0, n)), 0, bklen - 1).astype(SYM_DT[eb])
    with _huffman_pipeline(eb, bklen, n, False) as ph, _huffman_pipeline(eb, bklen, n, True) as pd:
        want = ph.compress_array(c)
        got = pd.compress_array(c)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(pd.decompress_array(got, SYM_DT[eb])[:n], c)


def test_lorenzo_huffman_chain_device_book(gpu):
    """Lorenzo -> Huffman<u16> bklen 1024 (the 3D turbulence chain): the
    device-book stage gives the oracle chain's bytes, eagerly and in graph
    mode."""
    from fzgpumodules_amd import fields
    shape = (24, 40, 56)  # z, y, x
    x = fields.power_law_field(shape, -11.0 / 3.0, 1234).astype(np.float32)
    rng = np.random.default_rng(3)
    x.ravel()[rng.choice(x.size, x.size // 100, replace=False)] *= 100
    dims = (shape[2], shape[1], shape[0])
    want, info = chains.lorenzo_huffman(x, dims)
    for graph in (False, True):
        with Pipeline(x.nbytes) as p:
            lz = p.add_lorenzo(1e-4, mode="REL", quant_radius=512, outlier_capacity=0.2, zigzag=True)
            hf = p.add_huffman(2, 1024)
            p.connect(hf, lz, "codes")
            p.set_dims(*dims)
            p.set_huffman_device_book(hf)
            p.finalize()
            if graph:
                p.enable_graph_mode(True)
            d = DeviceBuffer.from_array(x)
            for it in range(3 if graph else 1):
                ptr, sz = p.compress(d, x.nbytes)
                got = np.empty(sz, np.uint8)
                call("fzh_memcpy_dtoh", got.ctypes.data, ptr, sz)
                np.testing.assert_array_equal(got, want, err_msg=f"graph={graph} call {it}")
            d.free()
            np.testing.assert_array_equal(p.decompress_array(want), chains.reconstruct(info, dims))


def test_graph_replay_with_device_book(gpu):
    """Histogram, codebook and encode captured in one graph: every replay,
    with other eager work between, equals the host-codebook stream."""
    n = 1 << 20
    rng = np.random.default_rng(21)
    a = np.clip(np.rint(rng.laplace(512, 3.0, n)), 0, 1023).astype(np.uint16)
    b = np.clip(np.rint(rng.laplace(300, 30.0, n)), 0, 1023).astype(np.uint16)
    with _huffman_pipeline(2, 1024, n, False) as ph:
        want_a = ph.compress_array(a)
        want_b = ph.compress_array(b)
    p = _huffman_pipeline(2, 1024, n, True)
    p.enable_graph_mode(True)
    with p:
        d = DeviceBuffer.from_array(a)
        for it in range(5):  # eager, capture+launch, replays
            for src, want in ((a, want_a), (b, want_b)):
                d.upload(src)
                ptr, sz = p.compress(d, a.nbytes)
                got = np.empty(sz, np.uint8)
                call("fzh_memcpy_dtoh", got.ctypes.data, ptr, sz)
                np.testing.assert_array_equal(got, want, err_msg=f"call {it}")
        d.free()


@pytest.mark.parametrize("eb", [2, 4])
def test_out_of_range_symbol_reported(gpu, eb):
    """The host path rejects symbols >= bklen before encoding; the device
    path encodes them with no bits and reports them after the sync."""
    n = 50000
    c = (np.arange(n) % 200).astype(SYM_DT[eb])
    c[777] = 300
    with _huffman_pipeline(eb, 256, n, False) as ph:
        with pytest.raises(Exception, match="out-of-range"):
            ph.compress_array(c)
    with _huffman_pipeline(eb, 256, n, True) as pd:
        with pytest.raises(Exception, match="out-of-range"):
            pd.compress_array(c)
    c[777] = 5
    with _huffman_pipeline(eb, 256, n, False) as ph, _huffman_pipeline(eb, 256, n, True) as pd:
        np.testing.assert_array_equal(pd.compress_array(c), ph.compress_array(c))


def test_device_book_with_fine_mode_rejected(gpu):
    p = Pipeline(4096)
    h = p.add_huffman(2, 1024)
    p.set_huffman_device_book(h)
    p.set_huffman_fine(h)
    p.finalize()
    with p:
        with pytest.raises(Exception, match="coarse"):
            p.compress_array(np.zeros(2048, np.uint16))
