"""Parity of the Quantizer REL kernels (log2-space bins,
modules/quantizers/quantizer/quantizer.cu:171-352) with the oracle: codes,
outlier lists and the inverse bit-exact; then the QuantizerStage in REL mode
through pipelines and FZM files, and the reference's own REL cases
(tests/stages/test_quantizer.cpp:354-466, :1164-1200)."""
import os
import tempfile
import zlib

import numpy as np
import pytest

import chains
from fzgpumodules_amd.pipeline import Pipeline

pytestmark = pytest.mark.gpu


def mixed(rng, n):
    x = (rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))).astype(np.float32)
    k = max(1, n // 50)
    for v in (0.0, -0.0, 1e-40, 1.0, -2.0):
        x[rng.integers(0, n, k)] = np.float32(v)
    x[rng.integers(0, n, 3)] = np.inf
    x[rng.integers(0, n, 3)] = np.nan
    return x


@pytest.mark.parametrize("n", [1, 5, 1000, 1024, 300001])
@pytest.mark.parametrize("eps", [1e-2, 1e-3, 1e-4])
@pytest.mark.parametrize("cb,radius", [(4, 512), (4, 1 << 20), (2, 32768), (1, 64)])
@pytest.mark.parametrize("threshold", [np.inf, 1e3])
def test_forward_and_inverse_bit_exact(gpu, orc, n, eps, cb, radius, threshold):
    rng = np.random.default_rng(zlib.crc32(repr((n, eps, cb, radius, threshold)).encode()))
    x = mixed(rng, n)
    gc, gv, gi, gn = gpu.quant_rel_fwd(x, eps, radius, cb, threshold)
    oc, ov, oi, on = orc.quant_rel_fwd(x, eps, radius, cb, threshold)
    assert gn == on
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gv.view(np.uint32), ov.view(np.uint32))
    gy = gpu.quant_rel_inv(gc, eps, gv, gi, gn)
    oy = orc.quant_rel_inv(oc, eps, ov, oi, on)
    np.testing.assert_array_equal(gy.view(np.uint32), oy.view(np.uint32))


def test_capacity_overflow_counts_all(gpu, orc):
    x = mixed(np.random.default_rng(5), 50000)
    gc, gv, gi, gn = gpu.quant_rel_fwd(x, 1e-3, 64, 1, max_outliers=100)
    oc, ov, oi, on = orc.quant_rel_fwd(x, 1e-3, 64, 1, max_outliers=100)
    assert gn == on > 100
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)


def _rel_pipeline(cb, radius, cap, threshold=None, bitpack=False):
    p = Pipeline(1 << 20)
    q = p.add_quantizer(0.01, mode="REL", quant_radius=radius, outlier_capacity=cap, code_bytes=cb)
    if threshold is not None:
        p.set_quantizer_outliers(q, threshold=threshold)
    if bitpack:
        b = p.add_bitpack(cb, 0)
        p.connect(b, q, "codes")
    p.finalize()
    return p


def test_reference_smooth_round_trip(gpu, orc):
    """test_quantizer.cpp:354-418 through a pipeline and an FZM file."""
    i = np.arange(1024, dtype=np.float32)
    x = ((np.abs(np.sin(i * np.float32(0.05))) + np.float32(0.1)) * np.float32(10)).astype(np.float32)
    c, v, idx, cnt = orc.quant_rel_fwd(x, 0.01, 512, 4, max_outliers=103)
    want = chains.concat([c, v, idx, np.array([cnt], np.uint32)])
    with _rel_pipeline(4, 512, 0.1) as p:
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)
        np.testing.assert_array_equal(y.view(np.uint32), orc.quant_rel_inv(c, 0.01, v, idx, cnt).view(np.uint32))
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "rel.fzm")
            p.write_file(path)
            yf = Pipeline.decompress_file(path)
        np.testing.assert_array_equal(yf.view(np.uint32), y.view(np.uint32))
    out = np.zeros(x.size, bool)
    out[idx] = True
    assert np.all(y[out] == x[out])
    assert np.max(np.abs(y[~out] - x[~out]) / np.abs(x[~out])) <= 0.01 * 1.5


def test_reference_zeros_go_to_outliers(gpu, orc):
    """test_quantizer.cpp:420-466."""
    i = np.arange(256)
    s = (np.sin(i.astype(np.float32) * np.float32(0.1) + np.float32(0.5)) * np.float32(5)).astype(np.float32)
    x = np.where(i % 4 == 0, np.float32(0), np.where(i % 4 == 1, -s, s)).astype(np.float32)
    with _rel_pipeline(4, 512, 0.5) as p:
        blob = p.compress_array(x)
        parts = chains.split(blob)
        cnt = int(parts[3].view(np.uint32)[0])
        assert cnt >= 64 and set(range(0, 256, 4)) <= set(parts[2].view(np.uint32).tolist())
        y = p.decompress_array(blob)
    assert np.all(y[::4] == 0.0)


def test_reference_threshold_forces_outliers(gpu, orc):
    """test_quantizer.cpp:1164-1200."""
    x = np.linspace(0.5, 8.0, 512).astype(np.float32)
    c, v, idx, cnt = orc.quant_rel_fwd(x, 0.01, 512, 4, threshold=4.0, max_outliers=512)
    with _rel_pipeline(4, 512, 1.0, threshold=4.0) as p:
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, chains.concat([c, v, idx, np.array([cnt], np.uint32)]))
        y = p.decompress_array(blob)
    big = np.abs(x) >= 4.0
    assert np.all(y[big] == x[big])


def test_rel_bitpack_chain(gpu, orc):
    rng = np.random.default_rng(9)
    x = (rng.uniform(0.5, 2.0, 200000) * np.sign(rng.standard_normal(200000))).astype(np.float32)
    with _rel_pipeline(2, 32768, 0.05, bitpack=True) as p:
        blob = p.compress_array(x)
        parts = chains.split(blob)
        codes, vals, idx, cnt = orc.quant_rel_fwd(x, 0.01, 32768, 2, max_outliers=10000)
        assert int(parts[3].view(np.uint32)[0]) == cnt
        nb = orc.bitpack_nbits_for_max(int(codes.max()), 2)
        np.testing.assert_array_equal(parts[0], orc.bitpack_enc(codes, nb))
        y = p.decompress_array(blob)
        np.testing.assert_array_equal(y.view(np.uint32), orc.quant_rel_inv(codes, 0.01, vals, idx, cnt).view(np.uint32))
