"""Quantizer with in-place outliers (QuantizerStage::setInplaceOutliers,
modules/quantizers/quantizer/quantizer.cu:83-168): codes and
reconstructions bit-exact against the oracle, the one-port stage in a
pipeline and an FZM file, and the reference's argument errors."""
import numpy as np
import pytest

from oracle import quant_inplace_ref as Q

pytestmark = pytest.mark.gpu


def _field(rng, n):
    x = (np.cumsum(rng.normal(0, 1, n)) * 0.05).astype(np.float32)
    x[rng.integers(0, n, max(1, n // 100))] *= 1e4  # outliers
    x[:4] = [np.inf, -np.inf, np.nan, 0.0]
    return x


@pytest.mark.parametrize("n", [1, 4, 1000, 300001])
@pytest.mark.parametrize("radius,threshold", [(32768, np.inf), (8, np.inf), (1 << 22, 50.0)])
def test_inplace_kernels_match_oracle(gpu, n, radius, threshold):
    from fzgpumodules_amd._fzh import DeviceBuffer, call
    rng = np.random.default_rng(n + radius)
    x = _field(rng, max(n, 4))[:n]
    eb = np.float32(1e-3)
    ebx2 = np.float32(2) * eb
    r = np.float32(1) / ebx2
    want = Q.fwd(x, r, radius, threshold)
    dx = DeviceBuffer.from_array(x)
    dc = DeviceBuffer(4 * n)
    call("fzh_quant_abs_fwd_inplace_f32", None, dx.ptr, n, r, np.float32(threshold), radius, dc.ptr)
    got = dc.download(np.uint32, n)
    np.testing.assert_array_equal(got, want)
    dy = DeviceBuffer(4 * n)
    call("fzh_quant_abs_inv_inplace_f32", None, dc.ptr, n, ebx2, radius, dy.ptr)
    y = dy.download(np.float32, n)
    np.testing.assert_array_equal(y.view(np.uint32), Q.inv(want, ebx2, radius).view(np.uint32))
    fin = np.isfinite(x)
    if fin.any():
        assert np.abs(y[fin].astype(np.float64) - x[fin]).max() <= float(eb) * 1.01


@pytest.mark.parametrize("mode", ["ABS", "NOA"])
def test_inplace_stage_in_pipeline_and_file(gpu, tmp_path, mode):
    from fzgpumodules_amd.pipeline import Pipeline
    rng = np.random.default_rng(5)
    x = _field(rng, 200000)[4:]  # finite values only for the value range
    with Pipeline(x.nbytes) as p:
        q = p.add_quantizer(1e-3, mode=mode, quant_radius=32768, zigzag=True, code_bytes=4)
        p.set_quantizer_outliers(q, inplace=True)
        p.finalize()
        blob = p.compress_array(x)
        eb = np.float32(1e-3) if mode == "ABS" else np.float32(1e-3) * (x.max() - x.min())
        ebx2 = np.float32(2) * eb
        want = Q.fwd(x, np.float32(1) / ebx2, 32768)
        assert blob.nbytes == x.nbytes  # the single port: u32 codes, no outlier lists
        np.testing.assert_array_equal(blob.view(np.uint32)[: x.size], want)
        y = p.decompress_array(blob, np.float32)[: x.size]
        np.testing.assert_array_equal(y.view(np.uint32), Q.inv(want, ebx2, 32768).view(np.uint32))
        path = str(tmp_path / "q.fzm")
        p.write_file(path)
    z = Pipeline.decompress_file(path, np.float32)
    np.testing.assert_array_equal(z[: x.size].view(np.uint32), y.view(np.uint32))


def test_inplace_argument_errors(gpu):
    from fzgpumodules_amd._fzh import FzhError
    from fzgpumodules_amd.pipeline import Pipeline
    x = np.linspace(0, 1, 4096, dtype=np.float32)
    for zigzag, cb in ((False, 4), (True, 2)):
        with Pipeline(x.nbytes) as p:
            q = p.add_quantizer(1e-3, zigzag=zigzag, code_bytes=cb)
            p.set_quantizer_outliers(q, inplace=True)
            p.finalize()
            with pytest.raises(FzhError, match="setInplaceOutliers"):
                p.compress_array(x)
