"""Parity of the HIP transform kernels (Zigzag, Negabinary, Difference,
integer Lorenzo) with the oracle, bit-exact, and their stages in pipelines
and FZM files.  Reference behaviour: modules/transforms/*,
modules/predictors/diff/diff.cu, modules/predictors/lorenzo/lorenzo_stage.cu."""
import numpy as np
import pytest

from oracle import transforms_ref as T

pytestmark = pytest.mark.gpu

_CODE = {"u8": 0, "u16": 1, "u32": 2, "u64": 3, "i8": 4, "i16": 5, "i32": 6, "i64": 7, "f32": 8, "f64": 9}
_NP = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64, "i8": np.int8, "i16": np.int16,
       "i32": np.int32, "i64": np.int64, "f32": np.float32, "f64": np.float64}


def _dev(a):
    from fzgpumodules_amd._fzh import DeviceBuffer
    return DeviceBuffer.from_array(a if a.size else np.zeros(16, np.uint8))


def _rand(rng, dt, n):
    dt = np.dtype(dt)
    if dt.kind == "f":
        return (rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)).astype(dt)
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, n, dtype=np.int64, endpoint=True).astype(dt)


@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.int64])
@pytest.mark.parametrize("n", [1, 15, 4096, 100003])
@pytest.mark.parametrize("kind", ["zigzag", "negabinary"])
def test_codes(gpu, dt, n, kind):
    from fzgpumodules_amd._fzh import DeviceBuffer, call
    rng = np.random.default_rng(n)
    x = _rand(rng, dt, n)
    enc, dec = (T.zigzag_enc, T.zigzag_dec) if kind == "zigzag" else (T.negabinary_enc, T.negabinary_dec)
    want = enc(x)
    dx = _dev(x)
    out = DeviceBuffer(max(x.nbytes, 16))
    fn = "fzh_zigzag" if kind == "zigzag" else "fzh_negabinary"
    call(fn, None, dx.ptr, out.ptr, n, x.itemsize, 0)
    got = out.download(want.dtype, n)
    assert np.array_equal(got, want)
    call(fn, None, out.ptr, out.ptr, n, x.itemsize, 1)  # in place
    assert np.array_equal(out.download(dt, n), x)


CASES = [("u8", False), ("u16", False), ("u32", False), ("i32", False), ("i64", False), ("f32", False),
         ("f64", False), ("i8", True), ("i16", True), ("i32", True), ("i64", True)]


@pytest.mark.parametrize("dname,nb", CASES)
@pytest.mark.parametrize("chunk", [0, 4, 512, 4096])
@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_difference(gpu, dname, nb, chunk, n):
    from fzgpumodules_amd._fzh import DeviceBuffer, call, lib, workspace
    dt = _NP[dname]
    rng = np.random.default_rng(n + chunk)
    x = _rand(rng, dt, n)
    want = T.diff_fwd(x, chunk, nb)
    dx = _dev(x)
    d = DeviceBuffer(max(x.nbytes, 16))
    call("fzh_diff_fwd", None, dx.ptr, d.ptr, n, _CODE[dname], chunk, int(nb))
    got = d.download(want.dtype, n)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    # inverse of the reference's differences, in the reference's association
    back = T.diff_inv(want, dt, chunk, nb)
    y = DeviceBuffer(max(x.nbytes, 16))
    ws = workspace(lib().fzh_diff_inv_ws_bytes(n, _CODE[dname], chunk))
    call("fzh_diff_inv", None, d.ptr, y.ptr, n, _CODE[dname], chunk, int(nb), ws.ptr, ws.nbytes)
    gy = y.download(dt, n)
    assert np.array_equal(gy.view(np.uint8), back.view(np.uint8))
    if np.dtype(dt).kind != "f":
        assert np.array_equal(gy, x)


def test_difference_rejects(gpu):
    from fzgpumodules_amd._fzh import DeviceBuffer, FzhError, call
    b = DeviceBuffer(64)
    with pytest.raises(FzhError):
        call("fzh_diff_fwd", None, b.ptr, b.ptr, 8, 6, 3, 0)  # chunk not a power of two
    with pytest.raises(FzhError):
        call("fzh_diff_fwd", None, b.ptr, b.ptr, 8, 8, 0, 1)  # negabinary on float


@pytest.mark.parametrize("dims", [(1000, 1, 1), (256, 1, 1), (64, 64, 1), (37, 23, 1), (16, 16, 16), (9, 7, 5),
                                  (300, 200, 1), (70, 40, 30)])
@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.int64])
def test_integer_lorenzo(gpu, dims, dt):
    from fzgpumodules_amd._fzh import DeviceBuffer, call, lib, workspace
    rng = np.random.default_rng(int(np.prod(dims)))
    n = int(np.prod(dims))
    x = _rand(rng, dt, n)
    want = T.lorenzo_int_fwd(x, dims)
    dx = _dev(x)
    d = DeviceBuffer(max(x.nbytes, 16))
    call("fzh_lorenzo_int_fwd", None, dx.ptr, d.ptr, *dims, x.itemsize)
    got = d.download(dt, n)
    assert np.array_equal(got, want)
    y = DeviceBuffer(max(x.nbytes, 16))
    ws = workspace(lib().fzh_lorenzo_int_inv_ws_bytes(*dims))
    call("fzh_lorenzo_int_inv", None, d.ptr, y.ptr, *dims, x.itemsize, ws.ptr, ws.nbytes)
    assert np.array_equal(y.download(dt, n), x)


def test_pfpl_style_chain_round_trip(gpu, tmp_path):
    """Difference(int32 -> negabinary, 16 KiB chunks) -> Bitshuffle(ew 4) ->
    RZE, through the pipeline, an FZM file and createStage."""
    from fzgpumodules_amd.pipeline import Pipeline
    rng = np.random.default_rng(9)
    x = np.cumsum(rng.integers(-3, 4, 300000)).astype(np.int32)
    want_codes = T.diff_fwd(x, 4096, True)
    with Pipeline(x.nbytes) as p:
        df = p.add_difference("i32", negabinary=True, chunk_bytes=16384)
        bs = p.add_bitshuffle(16384, 4)
        rz = p.add_rze(16384, 4)
        p.connect(bs, df)
        p.connect(rz, bs)
        p.finalize()
        blob = p.compress_array(x)
        y = p.decompress_array(blob, np.int32)
        assert np.array_equal(y[: x.size], x)
        assert blob.nbytes < x.nbytes / 2
        path = str(tmp_path / "d.fzm")
        p.write_file(path)
    z = Pipeline.decompress_file(path, np.int32)
    assert np.array_equal(z[: x.size], x)
    assert want_codes.dtype == np.uint32


@pytest.mark.parametrize("kind", ["zigzag", "negabinary", "lorenzo_int"])
def test_stage_chains(gpu, kind, tmp_path):
    from fzgpumodules_amd.pipeline import Pipeline
    rng = np.random.default_rng(10)
    nx, ny = 257, 129
    x = (np.cumsum(rng.integers(-50, 51, nx * ny)) % 5000 - 2500).astype(np.int32)
    with Pipeline(x.nbytes) as p:
        if kind == "zigzag":
            s = p.add_zigzag(4)
        elif kind == "negabinary":
            s = p.add_negabinary(4)
        else:  # integer Lorenzo -> Zigzag (int32 -> uint32), as RLE takes unsigned codes
            lz = p.add_lorenzo_int(4)
            s = p.add_zigzag(4)
            p.connect(s, lz)
        rz = p.add_rle(4)
        p.connect(rz, s)
        p.set_dims(nx, ny, 1)
        p.finalize()
        blob = p.compress_array(x)
        y = p.decompress_array(blob, np.int32)
        assert np.array_equal(y[: x.size], x)
        path = str(tmp_path / f"{kind}.fzm")
        p.write_file(path)
    z = Pipeline.decompress_file(path, np.int32)
    assert np.array_equal(z[: x.size], x)
