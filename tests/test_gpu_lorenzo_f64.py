"""Parity of the FP64 Lorenzo kernels (LorenzoQuantStage<double, u16|u32>,
modules/fused/lorenzo_quant/lorenzo_quant.h:335-339) with the oracle.

Codes, outlier lists (double errors) and counts bit-exact, the inverse
bit-exact (the reference's addition trees in double), the value base and the
bound resolved in double; then whole pipelines with an FP64 Lorenzo stage,
their blobs byte for byte against chains composed from the oracle, their
FZM files read back with the header's float bound (lorenzo_quant.h:255).
"""
import os
import tempfile
import zlib

import numpy as np
import pytest

import chains
from fzgpumodules_amd.pipeline import Pipeline

pytestmark = pytest.mark.gpu

CASES = [
    ("1d_ragged", (5000, 1, 1), None),
    ("1d_tiles", (4096, 1, 1), None),
    ("1d_tiny", (3, 1, 1), None),
    ("1d_padded", (3000, 1, 1), 4100),
    ("2d_partial", (100, 70, 1), None),
    ("2d_exact", (64, 64, 1), None),
    ("2d_padded", (37, 33, 1), 37 * 33 + 333),
    ("3d_partial", (37, 29, 21), None),
    ("3d_exact", (16, 16, 16), None),
    ("3d_padded", (9, 9, 9), 729 + 100),
]


def field(rng, cover, n, spikes=True):
    x = np.cumsum(rng.normal(0, 1, cover)) * 0.01
    if spikes and cover > 10:
        x[rng.integers(0, cover, max(1, cover // 100))] *= 1000
    out = np.zeros(n)
    out[:cover] = x
    return out


def _seed(*k):
    return zlib.crc32(repr(k).encode())


@pytest.mark.parametrize("name,dims,n", CASES)
@pytest.mark.parametrize("zigzag", [False, True])
@pytest.mark.parametrize("code_bytes,radius", [(2, 32768), (2, 512), (4, 512), (4, 1 << 20)])
def test_forward_and_inverse_bit_exact(gpu, orc, name, dims, n, zigzag, code_bytes, radius):
    rng = np.random.default_rng(_seed(name, zigzag, code_bytes, radius))
    nx, ny, nz = dims
    cover = nx * ny * nz
    n = cover if n is None else n
    x = field(rng, cover, n)
    if nz == 1 and ny == 1:
        x[nx:] = 0.0  # 1D reads [nx, n) as zeros
    eb = 1e-3
    r, ebx2 = 1.0 / (2.0 * eb), 2.0 * eb
    cap = max(1, int(np.ceil(n * 0.2)))
    oc, oe, oi, ocnt = orc.lorenzo_fwd(x, (n if ny == nz == 1 else nx, ny, nz), r, radius, zigzag, code_bytes, cap, n)
    gc, ge, gi, gcnt = gpu.lorenzo_fwd64(x[:nx] if ny == nz == 1 else x, dims, r, radius, zigzag, code_bytes, cap, n)
    assert gcnt == ocnt
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(ge.view(np.uint64), oe.view(np.uint64))
    oy = orc.lorenzo_inv(oc, (n if ny == nz == 1 else nx, ny, nz), oe, oi, ocnt, ebx2, zigzag, cap)
    gy = gpu.lorenzo_inv64(gc, (n if ny == nz == 1 else nx, ny, nz), ge, gi, gcnt, ebx2, zigzag, cap)
    m = n if ny == nz == 1 else cover
    np.testing.assert_array_equal(gy[:m].view(np.uint64), oy[:m].view(np.uint64))
    assert np.max(np.abs(gy[:cover] - x[:cover])) <= eb * (1 + 1e-9)


@pytest.mark.parametrize("dims", [(5000, 1, 1), (100, 70, 1), (37, 29, 21)])
@pytest.mark.parametrize("many", [False, True])
def test_inverse_any_outlier_order(gpu, orc, dims, many):
    """Binned (more than 256 slots) and direct (at most 256) outlier paths,
    with the list permuted."""
    rng = np.random.default_rng(_seed(dims, many))
    n = int(np.prod(dims))
    x = field(rng, n, n)
    r, ebx2 = 1.0 / 2e-3, 2e-3
    c, e, i, cnt = orc.lorenzo_fwd(x, dims, r, 512, True, 2, n)
    if not many:
        c, e, i, cnt = orc.lorenzo_fwd(x, dims, r, 512, True, 2, 256)
        cnt = min(cnt, 256)
    assert cnt > 0
    perm = rng.permutation(cnt)
    cap = n if many else 256
    y_ref = orc.lorenzo_inv(c, dims, e, i, cnt, ebx2, True, cap)
    y_gpu = gpu.lorenzo_inv64(c, dims, e[perm], i[perm], cnt, ebx2, True, cap)
    np.testing.assert_array_equal(y_gpu.view(np.uint64), y_ref.view(np.uint64))


def test_outlier_capacity_overflow_counts_all(gpu, orc):
    rng = np.random.default_rng(3)
    x = rng.normal(0, 100, 20000)
    r = 1.0 / 2e-3
    oc, oe, oi, ocnt = orc.lorenzo_fwd(x, (20000, 1, 1), r, 16, False, 2, 100)
    gc, ge, gi, gcnt = gpu.lorenzo_fwd64(x, (20000, 1, 1), r, 16, False, 2, 100)
    assert gcnt == ocnt > 100
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(ge, oe)


def test_known_answer_and_u32_saturation(gpu):
    x = np.array([0.0, 0.25, 0.5, 1.0, -3.7])
    c, e, i, cnt = gpu.lorenzo_fwd64(x, (5, 1, 1), 1.0 / (2.0 * 0.1), 16, False, 2)
    np.testing.assert_array_equal(c, [0, 1, 2, 2, 0])
    assert cnt == 1 and list(i) == [4] and list(e) == [-24.0]
    x = np.array([3e9, 0.0, -3e9, 5.0])
    c, _, _, cnt = gpu.lorenzo_fwd64(x, (4, 1, 1), 1.0, 0xFFFFFFFF, True, 4)
    assert cnt == 0
    np.testing.assert_array_equal(c, [0xFFFFFFFE, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFE])
    x = np.array([2.5 - 1e-10, 0.0])
    c, _, _, _ = gpu.lorenzo_fwd64(x, (2, 1, 1), 1.0, 512, False, 2)
    np.testing.assert_array_equal(c, [2, 0xFFFE])


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("n", [1, 7, 4096, 1_000_003])
def test_value_base_and_bound_in_double(gpu, orc, mode, n):
    rng = np.random.default_rng(_seed(mode, n))
    x = rng.normal(0, 1, n) * 1e5 - 7.0
    user_eb = 1e-4
    p = gpu.value_base_params64(x, mode, user_eb)
    vb = orc.value_base(x, mode)
    abs_eb = float(np.float32(user_eb)) * vb if vb > 0 else float(np.float32(user_eb))
    assert p[0] == vb
    assert p[1] == abs_eb
    assert p[2] == 1.0 / (2.0 * abs_eb)
    assert p[3] == 2.0 * abs_eb


def test_device_resolved_bound_drives_the_forward(gpu, orc):
    """d_ebx2_r read from fzh_value_base_f64's params equals the host value."""
    from fzgpumodules_amd._fzh import DeviceBuffer, call, lib
    from fzgpumodules_amd.kernels import workspace
    rng = np.random.default_rng(11)
    x = field(rng, 30000, 30000)
    din = DeviceBuffer.from_array(x)
    par = DeviceBuffer(32)
    ws = workspace(lib().fzh_value_base_f64_ws_bytes(x.size))
    call("fzh_value_base_f64", None, din.ptr, x.size, 1, np.float32(1e-3), par.ptr, ws.ptr, ws.nbytes)
    p = par.download(np.float64, 4)
    gc, ge, gi, gcnt = gpu.lorenzo_fwd64(x, (30000, 1, 1), 0.0, 512, True, 2, d_params=par.ptr + 16)
    oc, oe, oi, ocnt = orc.lorenzo_fwd(x, (30000, 1, 1), p[2], 512, True, 2)
    assert gcnt == ocnt
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(ge, oe)


def test_rejects_bad_code_width_and_alignment(gpu):
    from fzgpumodules_amd._fzh import FzhError
    x = np.zeros(64)
    with pytest.raises(FzhError, match="code_bytes"):
        gpu.lorenzo_fwd64(x, (64, 1, 1), 1.0, 512, False, 1)
    with pytest.raises(FzhError, match="code_bytes"):
        gpu.lorenzo_inv64(np.zeros(64, np.uint8), (64, 1, 1), np.zeros(0), np.zeros(0, np.uint32), 0, 1.0, False)


# ---------------------------------------------------------------------------
# Pipelines with an FP64 Lorenzo stage
# ---------------------------------------------------------------------------
PIPE_CASES = [
    # coder, dims, mode, code_bytes, radius
    ("rze", (1800, 900, 1), 1, 2, 32768),
    ("rze", (0, 1, 1), 2, 2, 512),
    ("huffman", (64, 48, 40), 1, 2, 512),
    ("huffman", (1000, 700, 1), 0, 2, 512),
    ("rle", (0, 1, 1), 0, 4, 1 << 16),
    ("rle", (300, 200, 1), 1, 2, 512),
]


def _pipeline(coder, dims, mode, code_bytes, radius, nbytes, user_eb):
    p = Pipeline(nbytes)
    lz = p.add_lorenzo(user_eb, mode=("ABS", "REL", "NOA")[mode], quant_radius=radius, outlier_capacity=0.2,
                       zigzag=True, code_bytes=code_bytes, dtype="f64")
    if coder == "rze":
        bs = p.add_bitshuffle(16384, 2)
        rz = p.add_rze()
        p.connect(bs, lz, "codes")
        p.connect(rz, bs)
    elif coder == "huffman":
        hf = p.add_huffman(2, 2 * radius)
        p.connect(hf, lz, "codes")
    else:
        rl = p.add_rle(code_bytes)
        p.connect(rl, lz, "codes")
    if dims[1] > 1 or dims[2] > 1:
        p.set_dims(*dims)
    p.finalize()
    return p, lz


@pytest.mark.parametrize("coder,dims,mode,code_bytes,radius", PIPE_CASES)
def test_pipeline_blob_and_round_trip(gpu, coder, dims, mode, code_bytes, radius):
    rng = np.random.default_rng(_seed(coder, dims, mode))
    nx, ny, nz = dims
    n = nx * ny * nz if nx else 250001
    x = field(rng, n, n)
    user_eb = 1e-3 if mode else 1e-2
    want, info = chains.lorenzo64_chain(x, dims, coder, mode=mode, user_eb=user_eb, radius=radius,
                                        code_bytes=code_bytes, bklen=2 * radius)
    p, lz = _pipeline(coder, dims, mode, code_bytes, radius, x.nbytes, user_eb)
    try:
        blob = p.compress_array(x)
        np.testing.assert_array_equal(blob, want)
        assert p.abs_error_bound(lz) == info["abs_eb"]
        y = p.decompress_array(blob, dtype=np.float64)
        full = chains.reconstruct(info, dims)
        np.testing.assert_array_equal(y.view(np.uint64), full[: y.size].view(np.uint64))
        assert np.max(np.abs(y[:n] - x)) <= info["abs_eb"] * (1 + 1e-9)
        # FZM file: the header carries the bound as a float, and a reader
        # decodes with (double)(float)abs_eb (lorenzo_quant.h:218, :255)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "f64.fzm")
            p.write_file(path)
            yf = Pipeline.decompress_file(path, dtype=np.float64)
        info_f = dict(info, ebx2=2.0 * float(np.float32(info["abs_eb"])))
        full_f = chains.reconstruct(info_f, dims)
        np.testing.assert_array_equal(yf.view(np.uint64), full_f[: yf.size].view(np.uint64))
    finally:
        p.close()


def test_strict_bound_rejected_for_f64(gpu):
    from fzgpumodules_amd._fzh import FzhError
    p = Pipeline(8 * 1024)
    try:
        lz = p.add_lorenzo(1e-3, dtype="f64")
        with pytest.raises(FzhError, match="float input only"):
            p.set_strict(lz, True)
    finally:
        p.close()
