"""The reference's edge-case tests, through fz::Pipeline (C ABI), each blob
compared byte for byte with the oracle chain and each reconstruction bit for
bit with the oracle's inverse, plus the reference test's own check.

  DataPatterns (tests/pipeline/test_data_patterns.cpp:75-300): Lorenzo-only
    pipelines over all-zero, constant, ramp, white-noise, 1..4-element,
    4 Mi-element, tight- and loose-bound inputs, and outlier capacity 0
    (outliers dropped: no failure, the decoded size is the input size).
    White noise is drawn with numpy's generator rather than std::mt19937_64
    (same distribution U[-1, 1]; both sides of each comparison see the same
    values).
  RZEStage (tests/stages/test_rze_stage.cpp:258-330): chunk sizes other than
    16384 and level counts other than 4 are rejected when the stage runs,
    forward and inverse.
  Golden configuration (tests/golden/regenerate.cpp:36-95): 4096-element
    two-sinusoid input, ABS 1e-2, radius 512, capacity 0.2,
    Lorenzo<float,u16> -> Bitshuffle (default 16 KiB blocks, element width
    4), written to an FZM file and read back.  The reference's golden .fzm
    file itself is not in the repository, so the oracle stands in for it.
"""
import zlib

import numpy as np
import pytest

import chains
import oracle as orc
from fzgpumodules_amd._fzh import FzhError
from fzgpumodules_amd.pipeline import Pipeline

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[edge]", *a, flush=True)


def lorenzo_only_expected(x, eb, radius, cap):
    """Oracle outputs of a one-stage LorenzoQuant<float,u16> pipeline (ABS,
    no zigzag): [codes][outlier_errors][outlier_indices][outlier_count]."""
    xp, codes, err, idx, cnt, abs_eb, ebx2 = chains.lorenzo_part(x, (0, 1, 1), 0, eb, radius, cap, False, 2, 1)
    blob = chains.concat([codes, err, idx, np.array([cnt], np.uint32)])
    return blob, dict(xp=xp, codes=codes, err=err, idx=idx, cnt=cnt, abs_eb=abs_eb, ebx2=ebx2)


def run_pattern(x, eb, radius=32768, cap=0.25):
    """Compress + decompress through a Lorenzo-only pipeline; checks the blob
    and the reconstruction against the oracle; returns (max|err|, blob size,
    outlier count)."""
    x = np.ascontiguousarray(x, np.float32)
    want, info = lorenzo_only_expected(x, eb, radius, cap)
    with Pipeline(x.nbytes) as p:
        p.add_lorenzo(eb, mode="ABS", quant_radius=radius, outlier_capacity=cap)
        p.finalize()
        blob = p.compress_array(x)
        assert blob.size == want.size
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)
    assert y.size == x.size
    recon = chains.reconstruct(info, (0, 1, 1), zigzag=False)
    np.testing.assert_array_equal(y, recon[: x.size])
    err = float(np.max(np.abs(y.astype(np.float64) - x.astype(np.float64)))) if x.size else 0.0
    return err, blob.size, info["cnt"]


def test_all_zeros():
    x = np.zeros(1 << 14, np.float32)
    err, size, _ = run_pattern(x, 1e-2, 32768, 0.25)
    assert err == 0.0
    assert size < x.nbytes


def test_all_same_constant():
    x = np.full(1 << 14, 3.14159, np.float32)
    err, _, _ = run_pattern(x, 1e-2)
    assert err <= np.float32(1e-2)


def test_linear_ramp():
    x = (np.arange(1 << 14, dtype=np.float32) * np.float32(0.001)).astype(np.float32)
    err, _, _ = run_pattern(x, 1e-2)
    assert err <= 1e-2 * 1.01


def test_white_noise_all_outliers():
    x = np.random.default_rng(42).uniform(-1.0, 1.0, 1 << 14).astype(np.float32)
    err, _, cnt = run_pattern(x, 1e-2, 32768, 1.0)
    _say(f"white noise: {cnt} outliers of {x.size}")
    assert err <= 1e-2 * 1.01


@pytest.mark.parametrize("vals", [[42.0], [1.0, 2.0], [1.0, 2.0, 3.0], [1.0, 2.0, 4.0, 8.0]],
                         ids=["n1", "n2", "n3", "n4"])
def test_tiny_inputs(vals):
    err, _, _ = run_pattern(np.array(vals, np.float32), 1e-2, 512, 1.0)
    assert err <= 1e-2 * 1.01


def test_four_million_floats():
    i = np.arange(1 << 22, dtype=np.float32)
    x = (np.float32(50.0) * np.sin(i * np.float32(0.001))).astype(np.float32)
    err, _, _ = run_pattern(x, 1e-2)
    assert err <= 1e-2 * 1.01


def test_tight_error_bound():
    i = np.arange(1 << 14, dtype=np.float32)
    x = np.sin(i * np.float32(0.05)).astype(np.float32)
    err, _, _ = run_pattern(x, 1e-5, 32768, 0.5)
    assert err <= 1e-5 * 1.01


def test_loose_error_bound_compresses_no_worse():
    i = np.arange(1 << 14, dtype=np.float32)
    x = (np.float32(50.0) * np.sin(i * np.float32(0.01))).astype(np.float32)
    err_loose, size_loose, _ = run_pattern(x, 10.0, 512, 0.2)
    _, size_tight, _ = run_pattern(x, 1e-2, 512, 0.2)
    assert err_loose <= 10.0 * 1.01
    assert size_loose <= size_tight


def test_zero_outlier_capacity_drops_outliers():
    # radius 512 instead of the reference test's 32768: with 32768 codes the
    # quantizer spans +-6.5 at eb 1e-4 and U[-1, 1] noise makes no outlier,
    # so nothing would be dropped
    x = np.random.default_rng(42).uniform(-1.0, 1.0, 4096).astype(np.float32)
    want, info = lorenzo_only_expected(x, 1e-4, 512, 0.0)
    assert info["cnt"] > 0 and info["err"].size == 0
    with Pipeline(x.nbytes) as p:
        p.add_lorenzo(1e-4, mode="ABS", quant_radius=512, outlier_capacity=0.0)
        p.finalize()
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)
    assert y.nbytes == x.nbytes
    np.testing.assert_array_equal(y, chains.reconstruct(info, (0, 1, 1), zigzag=False)[: x.size])
    _say(f"capacity 0: {info['cnt']} outliers dropped")


# ---------------------------------------------------------------------------
# RZE configuration checks
# ---------------------------------------------------------------------------
def _lorenzo_rze(x, chunk, levels):
    p = Pipeline(x.nbytes)
    lz = p.add_lorenzo(1e-3, mode="ABS", quant_radius=32768, outlier_capacity=0.2)
    rz = p.add_rze(chunk, levels)
    p.connect(rz, lz, "codes")
    return p


def _sine(n):
    return np.sin(np.arange(n, dtype=np.float32) * np.float32(0.01)).astype(np.float32)


@pytest.mark.parametrize("chunk,levels", [(8192, 4), (16384, 1), (16384, 2), (16384, 3)])
def test_rze_unsupported_config_rejected_forward(chunk, levels):
    x = _sine(8192)
    with _lorenzo_rze(x, chunk, levels) as p:
        p.finalize()
        with pytest.raises(FzhError, match="chunk_size|levels"):
            p.compress_array(x)


def _rewrite_stage_config(raw, stage_type, offset, value_bytes):
    """Patch one stage's config bytes in an FZM file and refresh the header
    checksum (CRC32 of the core with the checksum field zeroed, the stage
    records and the buffer records)."""
    raw = raw.copy()
    num_buffers = int(raw[6:8].view(np.uint16)[0])
    num_stages = int(raw[32:36].view(np.uint32)[0])
    hit = 0
    for s in range(num_stages):
        rec = 80 + 256 * s
        if int(raw[rec:rec + 2].view(np.uint16)[0]) == stage_type:
            raw[rec + 40 + offset: rec + 40 + offset + len(value_bytes)] = np.frombuffer(value_bytes, np.uint8)
            hit += 1
    assert hit == 1
    end = 80 + 256 * (num_stages + num_buffers)
    core = raw[:80].copy()
    core[76:80] = 0
    crc = zlib.crc32(raw[80:end].tobytes(), zlib.crc32(core.tobytes()))
    raw[76:80] = np.frombuffer(np.uint32(crc).tobytes(), np.uint8)
    return raw


@pytest.mark.parametrize("field,value", [("chunk", 8192), ("levels", 3)])
def test_rze_unsupported_config_rejected_inverse(tmp_path, field, value):
    x = _sine(8192)
    with _lorenzo_rze(x, 16384, 4) as p:
        p.finalize()
        p.compress_array(x)
        path = str(tmp_path / "ok.fzm")
        p.write_file(path)
    raw = np.fromfile(path, np.uint8)
    y = Pipeline.decompress_file(path)  # the untouched file decodes
    assert y.size >= x.size
    RZE = 18
    patch = np.uint32(value).tobytes() if field == "chunk" else bytes([value])
    bad = _rewrite_stage_config(raw, RZE, 0 if field == "chunk" else 4, patch)
    bad_path = str(tmp_path / "bad.fzm")
    bad.tofile(bad_path)
    with pytest.raises(FzhError, match="chunk_size|levels"):
        Pipeline.decompress_file(bad_path)


# ---------------------------------------------------------------------------
# Golden configuration
# ---------------------------------------------------------------------------
def golden_input():
    i = np.arange(4096, dtype=np.float64)
    return (50.0 * np.sin(i * 0.01) + 20.0 * np.cos(i * 0.003)).astype(np.float32)


def test_golden_configuration_lorenzo_bitshuffle(tmp_path):
    x = golden_input()
    # Bitshuffle's alignment (its 16 KiB block) pads the input; 4096 floats are
    # exactly one block.
    xp, codes, err, idx, cnt, abs_eb, ebx2 = chains.lorenzo_part(x, (0, 1, 1), 0, 1e-2, 512, 0.2, False, 2, 16384)
    bs = orc.bitshuffle_enc(codes.view(np.uint8), 16384, 4)
    want = chains.concat([bs, err, idx, np.array([cnt], np.uint32)])
    with Pipeline(x.nbytes) as p:
        lz = p.add_lorenzo(1e-2, mode="ABS", quant_radius=512, outlier_capacity=0.2)
        b = p.add_bitshuffle(16384, 4)
        p.connect(b, lz, "codes")
        p.finalize()
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        path = str(tmp_path / "ref_lorenzo_bitshuffle.fzm")
        p.write_file(path)
    raw = np.fromfile(path, np.uint8)
    assert raw[:4].view(np.uint32)[0] == 0x464D5A32
    num_stages = int(raw[32:36].view(np.uint32)[0])
    num_buffers = int(raw[6:8].view(np.uint16)[0])
    assert (num_stages, num_buffers) == (2, 4)
    y = Pipeline.decompress_file(path)
    info = dict(codes=codes, err=err, idx=idx, cnt=cnt, ebx2=ebx2)
    recon = chains.reconstruct(info, (0, 1, 1), zigzag=False)
    np.testing.assert_array_equal(y[: x.size], recon[: x.size])
    assert float(np.max(np.abs(y[: x.size].astype(np.float64) - x.astype(np.float64)))) <= 1e-2 * 1.01
    _say(f"golden: {x.nbytes} B -> {blob.size} B, {cnt} outliers")
