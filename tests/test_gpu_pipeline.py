"""Whole-pipeline parity: fz::Pipeline (through the C ABI) against the oracle
chains, byte for byte, plus round trips, graph replay and FZM files.

Reference behaviour: src/pipeline/compressor.cpp (finalize, concat),
compressor_exec.cpp (padding, compress/decompress), compressor_io.cpp (FZM).
"""
import os
import zlib

import numpy as np
import pytest

import chains
from fzgpumodules_amd import fields
from fzgpumodules_amd._fzh import DeviceBuffer, FzhError, call, lib
from fzgpumodules_amd.pipeline import Pipeline

pytestmark = pytest.mark.gpu


def _seed(name):
    return zlib.crc32(name.encode())


def cloud(ny, nx, seed):
    g = fields.power_law_field((ny, nx), -3.0, seed)
    q = np.quantile(g, 0.3)
    x = (g - q) / (g.max() - q)
    return np.clip(x, 0.0, 1.0).astype(np.float32)


def build_cfg2(nx, ny, graph=False, mode="REL", eb=1e-3, cap=0.1):
    p = Pipeline(nx * ny * 4)
    lz = p.add_lorenzo(eb, mode=mode, quant_radius=32768, outlier_capacity=cap, zigzag=True)
    bs = p.add_bitshuffle(16384, 2)
    rz = p.add_rze(16384, 4)
    p.connect(bs, lz, "codes")
    p.connect(rz, bs)
    p.set_dims(nx, ny, 1)
    p.finalize()
    if graph:
        p.enable_graph_mode(True)
    return p, lz


def check_bound(x, y, abs_eb, slack=1.0001):
    err = np.abs(y[: x.size].astype(np.float64) - x.astype(np.float64))
    assert err.max() <= abs_eb * slack, (err.max(), abs_eb)


@pytest.mark.parametrize("shape", [(300, 700), (64, 512), (33, 97)])
def test_cfg2_chain_bit_exact(shape):
    ny, nx = shape
    x = cloud(ny, nx, _seed(f"cfg2{shape}"))
    want, info = chains.lorenzo_bitshuffle_rze(x, (nx, ny, 1))
    with build_cfg2(nx, ny)[0] as p:
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        names = [o[0] for o in p.outputs()]
        assert names == ["output", "outlier_errors", "outlier_indices", "outlier_count"]
        y = p.decompress_array(blob)
        assert y.size == x.size
        ref = chains.reconstruct(info, (nx, ny, 1))
        np.testing.assert_array_equal(y, ref[: x.size])
        check_bound(x.ravel(), y, float(info["abs_eb"]), slack=1.001)


def test_cfg1_chain_rle_1d():
    x = fields.sinusoid_1d(200003, seed=_seed("cfg1"))
    want, info = chains.lorenzo_rle(x)
    with Pipeline(x.nbytes) as p:
        lz = p.add_lorenzo(1e-4, mode="ABS", outlier_capacity=0.2)
        rl = p.add_rle(2)
        p.connect(rl, lz, "codes")
        p.finalize()
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)
        np.testing.assert_array_equal(y, chains.reconstruct(info, (0, 1, 1), zigzag=False))
        check_bound(x, y, 1e-4, slack=1.01)


def test_cfg3_chain_huffman_3d():
    shape = (24, 40, 56)  # z, y, x
    # smoother than the 512^3 workload's spectrum so a grid this small stays
    # within the 20% outlier capacity
    g = fields.power_law_field(shape, -11.0 / 3.0, _seed("cfg3"))
    x = g.astype(np.float32)
    rng = np.random.default_rng(3)
    spikes = rng.choice(x.size, x.size // 100, replace=False)
    x.ravel()[spikes] *= 100
    dims = (shape[2], shape[1], shape[0])
    want, info = chains.lorenzo_huffman(x, dims)
    with Pipeline(x.nbytes) as p:
        lz = p.add_lorenzo(1e-4, mode="REL", quant_radius=512, outlier_capacity=0.2, zigzag=True)
        hf = p.add_huffman(2, 1024)
        p.connect(hf, lz, "codes")
        p.set_dims(*dims)
        p.finalize()
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)
        np.testing.assert_array_equal(y, chains.reconstruct(info, dims))
        assert info["cnt"] <= chains.max_outliers(x.size, 0.2)
        check_bound(x.ravel(), y, float(info["abs_eb"]), slack=1.001)


def test_1d_padded_view_matches_zero_padded_copy():
    """1D input not a multiple of 16 KiB: the Lorenzo stage reads the caller's
    buffer and treats the tail as zeros; results equal the reference's padded
    copy (src/pipeline/compressor_exec.cpp:35-54)."""
    x = fields.sinusoid_1d(50001, seed=_seed("pad1d")) * np.float32(-3.0) - np.float32(2.0)
    for mode in ("REL", "NOA"):
        want, info = chains.lorenzo_bitshuffle_rze(x, (0, 1, 1), mode=1 if mode == "REL" else 2, cap=0.2)
        with Pipeline(x.nbytes) as p:
            lz = p.add_lorenzo(1e-3, mode=mode, outlier_capacity=0.2, zigzag=True)
            bs = p.add_bitshuffle(16384, 2)
            rz = p.add_rze()
            p.connect(bs, lz, "codes")
            p.connect(rz, bs)
            p.finalize()
            blob = p.compress_array(x)
            np.testing.assert_array_equal(blob, want)
            assert abs(p.abs_error_bound(lz) - float(info["abs_eb"])) == 0.0
            y = p.decompress_array(blob)
            np.testing.assert_array_equal(y, chains.reconstruct(info, (0, 1, 1))[: x.size])


def test_graph_replay_matches_eager():
    ny, nx = 200, 600
    x = cloud(ny, nx, _seed("graph"))
    x2 = cloud(ny, nx, _seed("graph2"))
    want, _ = chains.lorenzo_bitshuffle_rze(x, (nx, ny, 1))
    want2, _ = chains.lorenzo_bitshuffle_rze(x2, (nx, ny, 1))
    p, _ = build_cfg2(nx, ny, graph=True)
    with p:
        d = DeviceBuffer.from_array(x)
        for it in range(4):  # eager, capture+launch, replay, replay
            ptr, n = p.compress(d, x.nbytes)
            got = np.empty(n, np.uint8)
            call("fzh_memcpy_dtoh", got.ctypes.data, ptr, n)
            np.testing.assert_array_equal(got, want, err_msg=f"call {it}")
            # eager work between replays: overwrite with another field and back
            d.upload(x2)
            ptr, n = p.compress(d, x.nbytes)
            got = np.empty(n, np.uint8)
            call("fzh_memcpy_dtoh", got.ctypes.data, ptr, n)
            np.testing.assert_array_equal(got, want2, err_msg=f"call {it} (second field)")
            d.upload(x)
        d.free()


def test_file_round_trip(tmp_path):
    ny, nx = 120, 500
    x = cloud(ny, nx, _seed("file"))
    want, info = chains.lorenzo_bitshuffle_rze(x, (nx, ny, 1))
    p, _ = build_cfg2(nx, ny)
    with p:
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        path = str(tmp_path / "f.fzm")
        p.write_file(path)
    raw = np.fromfile(path, np.uint8)
    core = raw[:80]
    assert core[:4].view(np.uint32)[0] == 0x464D5A32
    num_buffers = int(core[6:8].view(np.uint16)[0])
    num_stages = int(core[32:36].view(np.uint32)[0])
    assert (num_stages, num_buffers) == (3, 4)
    hdr = 80 + 256 * (num_stages + num_buffers)
    payload = raw[hdr:]
    # payload = outputs back to back, no padding (compressor_io.cpp:156-165)
    np.testing.assert_array_equal(payload, np.concatenate([np.ascontiguousarray(s) for s in chains.split(want)]))
    assert core[72:76].view(np.uint32)[0] == zlib.crc32(payload.tobytes())
    y = Pipeline.decompress_file(path)
    ref = chains.reconstruct(info, (nx, ny, 1))
    assert y.size == ref.size  # padded length, as the reference returns it
    np.testing.assert_array_equal(y[: x.size], ref[: x.size])
    # a flipped payload byte fails the data checksum
    bad = raw.copy()
    bad[hdr + 5] ^= 0x40
    bad_path = str(tmp_path / "bad.fzm")
    bad.tofile(bad_path)
    with pytest.raises(FzhError, match="checksum"):
        Pipeline.decompress_file(bad_path)


def test_user_owned_buffers():
    ny, nx = 64, 300
    x = cloud(ny, nx, _seed("user"))
    want, _ = chains.lorenzo_bitshuffle_rze(x, (nx, ny, 1))
    p, _ = build_cfg2(nx, ny)
    with p:
        d = DeviceBuffer.from_array(x)
        cap = p.max_compressed_size(x.nbytes)
        out = DeviceBuffer(cap)
        n = p.compress_into(d, x.nbytes, out, cap)
        np.testing.assert_array_equal(out.download(np.uint8, n), want)
        y = DeviceBuffer(x.nbytes)
        m = p.decompress_into(out, n, y, x.nbytes)
        assert m == x.nbytes
        with pytest.raises(FzhError, match="too small"):
            p.decompress_into(out, n, y, x.nbytes - 4)
        for b in (d, out, y):
            b.free()


def test_quantizer_bitpack_chain(orc):
    rng = np.random.default_rng(_seed("qbp"))
    x = rng.integers(0, 20, 100000).astype(np.float32)
    with Pipeline(x.nbytes) as p:
        q = p.add_quantizer(0.5, mode="ABS", quant_radius=32768, outlier_capacity=0.05, zigzag=True)
        b = p.add_bitpack(2, 0)
        p.connect(b, q, "codes")
        p.finalize()
        blob = p.compress_array(x)
        parts = chains.split(blob)
        r, ebx2 = np.float32(1) / np.float32(1.0), np.float32(1.0)
        codes, vals, idx, cnt = orc.quant_abs_fwd(x, r, 32768, True, 2, max_outliers=5000)
        assert int(parts[3].view(np.uint32)[0]) == cnt
        nb = orc.bitpack_nbits_for_max(int(codes.max()), 2)
        np.testing.assert_array_equal(parts[0], orc.bitpack_enc(codes, nb))
        y = p.decompress_array(blob)
        np.testing.assert_array_equal(y, orc.quant_abs_inv(codes, ebx2, True, vals, idx, cnt))


def test_empty_and_errors():
    with pytest.raises(FzhError):
        p = Pipeline()
        p.finalize()  # no stages
    with Pipeline() as p:
        a = p.add_lorenzo(1e-3)
        b = p.add_rle(2)
        with pytest.raises(FzhError, match="no output named"):
            p.connect(b, a, "nope")
        p.connect(b, a, "codes")
        p.finalize()
        with pytest.raises(FzhError, match="compress"):
            p.decompress(0x1000, 64)


@pytest.mark.parametrize("n", [1 << 20, 34_000_000])
def test_rze_pack_in_the_concat_launch(n):
    """The RZE stage writing the concat's first slot leaves its pack step to
    the final concat + readback launch; the stream stays the oracle's byte
    for byte, below and above 4096 chunks (the offset-scan path), and
    decompresses."""
    x = fields.sinusoid_1d(n, seed=_seed(f"packfuse{n}"))
    want, info = chains.lorenzo_bitshuffle_rze(x, (0, 1, 1), mode=0, user_eb=1e-4)
    with Pipeline(x.nbytes) as p:
        lz = p.add_lorenzo(1e-4, mode="ABS", outlier_capacity=0.1, zigzag=True)
        bs = p.add_bitshuffle(16384, 2)
        co = p.add_rze()
        p.connect(bs, lz, "codes")
        p.connect(co, bs)
        p.finalize()
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)[: x.size]
        assert np.abs(y.astype(np.float64) - x).max() <= 1e-4 * 1.01


def _boundary_field(ny, nx, eb, seed):
    """Values within a few ulps of eb from a quantisation point: the strict
    test |fl(s*ebx2) - x| <= eb is decided near its boundary everywhere."""
    rng = np.random.default_rng(seed)
    eb = np.float32(eb)
    ebx2 = np.float32(2) * eb
    yy, xx = np.mgrid[0:ny, 0:nx]
    k = np.round(200 * np.sin(xx / 37.0) * np.cos(yy / 23.0)).astype(np.float32)
    xh = k * ebx2
    u = rng.integers(-4, 5, (ny, nx)).astype(np.float32) * np.float32(2.0 ** -23)
    delta = eb * (np.float32(1) + u)
    sign = np.where(rng.random((ny, nx)) < 0.5, np.float32(-1), np.float32(1))
    return (xh + sign * delta).astype(np.float32)


@pytest.mark.parametrize("case", ["1d_rle", "1d_rle_many_tiles", "2d_rze", "2d_boundary", "3d_rze"])
def test_strict_error_bound(case):
    """Strict mode: codes/outliers/coder streams identical to the plain chain,
    plus the patch list = exactly the elements the reference arithmetic
    reconstructs outside eb; every decompressed value is within eb.  The
    many-tiles case (16400 tiles) takes the scanned compaction path."""
    if case.startswith("1d_rle"):
        x = fields.sinusoid_1d(1 << 20 if case == "1d_rle" else 1024 * 16400, seed=_seed("strict1"))
        dims = (x.size, 1, 1)
        want, info = chains.lorenzo_rle(x)
    elif case == "2d_rze":
        ny, nx = 300, 700
        x = cloud(ny, nx, _seed("strict2")).ravel()
        dims = (nx, ny, 1)
        want, info = chains.lorenzo_bitshuffle_rze(x, dims)
    elif case == "2d_boundary":
        ny, nx = 300, 700
        x = _boundary_field(ny, nx, 1e-3, _seed("strict4")).ravel()
        dims = (nx, ny, 1)
        want, info = chains.lorenzo_bitshuffle_rze(x, dims, mode=0, user_eb=1e-3)
    else:
        shape = (24, 40, 56)
        x = fields.power_law_field(shape, -11.0 / 3.0, _seed("strict3")).astype(np.float32).ravel()
        dims = (56, 40, 24)
        want, info = chains.lorenzo_bitshuffle_rze(x, dims, user_eb=1e-5, cap=0.2)
    is_1d = case.startswith("1d_rle")
    recon = chains.reconstruct(info, dims if not is_1d else (0, 1, 1), zigzag=not is_1d)
    pidx, pval = chains.strict_patches(x, recon, info["abs_eb"], dims)
    with Pipeline(x.nbytes) as p:
        if is_1d:
            lz = p.add_lorenzo(1e-4, mode="ABS", outlier_capacity=0.2)
            co = p.add_rle(2)
        elif case == "2d_boundary":
            lz = p.add_lorenzo(1e-3, mode="ABS", outlier_capacity=0.1, zigzag=True)
            bs = p.add_bitshuffle(16384, 2)
            co = p.add_rze()
        else:
            lz = p.add_lorenzo(1e-3 if case == "2d_rze" else 1e-5, mode="REL",
                               outlier_capacity=0.1 if case == "2d_rze" else 0.2, zigzag=True)
            bs = p.add_bitshuffle(16384, 2)
            co = p.add_rze()
        if case == "2d_boundary":
            p.set_strict(lz, True, 1.0)  # about half the elements are patched
        else:
            p.set_strict(lz, True)
        if not is_1d:
            p.connect(bs, lz, "codes")
            p.connect(co, bs)
        else:
            p.connect(co, lz, "codes")
        if not is_1d:
            p.set_dims(*dims)
        p.finalize()
        blob = p.compress_array(x)
        got = chains.split(blob)
        ref = chains.split(want)
        assert len(got) == 7
        assert pidx.size > 0  # the data exercise the patch path
        for i in range(4):
            np.testing.assert_array_equal(got[i], ref[i], err_msg=f"part {i}")
        np.testing.assert_array_equal(got[5].view(np.uint32), pidx)
        np.testing.assert_array_equal(got[4].view(np.float32), pval)
        assert got[6].view(np.uint32)[0] == pidx.size
        y = p.decompress_array(blob)[: x.size]
        err = np.abs(y.astype(np.float64) - x.astype(np.float64))
        assert err.max() <= float(info["abs_eb"])  # exhaustive, no slack
        expect = recon[: x.size].copy()
        expect[pidx] = pval
        np.testing.assert_array_equal(y, expect)


@pytest.mark.parametrize("fusion", [True, False])
@pytest.mark.parametrize("ew", [1, 2])
def test_stage_fusion_same_bytes(fusion, ew):
    """Bitshuffle folded into RZE (default) and the two-stage path give the
    oracle's bytes; element width 1 runs on u8 codes (radius 64)."""
    ny, nx = 150, 260
    x = cloud(ny, nx, _seed(f"fuse{ew}"))
    code_bytes, radius = (2, 32768) if ew == 2 else (1, 64)
    xp, codes, err, idx, cnt, abs_eb, ebx2 = chains.lorenzo_part(x, (nx, ny, 1), 1, 1e-3, radius, 0.1, True,
                                                                 code_bytes, 16384)
    rze = chains.orc.rze_enc(chains.orc.bitshuffle_enc(codes.view(np.uint8), 16384, ew))
    want = chains.concat([rze, err, idx, np.array([cnt], np.uint32)])
    with Pipeline(x.nbytes) as p:
        lz = p.add_lorenzo(1e-3, mode="REL", quant_radius=radius, outlier_capacity=0.1, zigzag=True,
                           code_bytes=code_bytes)
        bs = p.add_bitshuffle(16384, ew)
        rz = p.add_rze()
        p.connect(bs, lz, "codes")
        p.connect(rz, bs)
        p.set_dims(nx, ny, 1)
        p.set_stage_fusion(fusion)
        p.finalize()
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)
        ref = chains.orc.lorenzo_inv(codes, (nx, ny, 1), err, idx, cnt, ebx2, True)
        np.testing.assert_array_equal(y, ref[: x.size])
