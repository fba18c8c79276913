"""The benchmark workloads at their full sizes, through fz::Pipeline (C ABI),
byte for byte against the oracle chains.

  1D  sinusoid_1d(2^24)           Lorenzo ABS 1e-4, radius 32768, cap 0.2 -> RLE
                                  (README.md:51-58 quick start)
  2D  cloud_2d(1800 x 3600)       Lorenzo REL 1e-3, radius 32768, zigzag, cap 0.1
                                  -> Bitshuffle(16 KiB, ew 2) -> RZE(16 KiB, 4)
                                  (src/utils/cli/cli.cpp:435-462)
  3D  turbulence_3d(512^3)        Lorenzo REL 1e-4, radius 512, zigzag, cap 0.2
                                  -> Huffman(bklen 1024, coarse)
                                  (src/utils/cli/cli.cpp:470-479)

Each case checks the plain chain's blob against the oracle, the decompressed
field against the oracle's inverse bit for bit, then the strict-bound chain:
the same first four parts, the patch list equal to the elements the oracle's
reconstruction puts outside eb, and |x - x_hat| <= eb for every element.
"""
import numpy as np
import pytest

import chains
from fzgpumodules_amd import fields
from fzgpumodules_amd.pipeline import Pipeline

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[workload]", *a, flush=True)


@pytest.fixture(scope="module")
def turbulence():
    return fields.turbulence_3d(512)


def _pipeline(kind, nbytes, dims, strict):
    p = Pipeline(nbytes)
    if kind == "1d":
        lz = p.add_lorenzo(1e-4, mode="ABS", quant_radius=32768, outlier_capacity=0.2)
        co = p.add_rle(2)
        p.connect(co, lz, "codes")
    elif kind == "2d":
        lz = p.add_lorenzo(1e-3, mode="REL", quant_radius=32768, outlier_capacity=0.1, zigzag=True)
        bs = p.add_bitshuffle(16384, 2)
        co = p.add_rze(16384, 4)
        p.connect(bs, lz, "codes")
        p.connect(co, bs)
    else:
        lz = p.add_lorenzo(1e-4, mode="REL", quant_radius=512, outlier_capacity=0.2, zigzag=True)
        co = p.add_huffman(2, 1024)
        p.connect(co, lz, "codes")
    if strict:
        p.set_strict(lz, True)
    if kind != "1d":
        p.set_dims(*dims)
    p.finalize()
    return p, lz


def _check(kind, x, dims, want, info):
    zig = kind != "1d"
    odims = (0, 1, 1) if kind == "1d" else dims
    cap = chains.max_outliers(info["xp"].size, 0.2 if kind != "2d" else 0.1)
    _say(kind, f"n={x.size} outliers={info['cnt']} capacity={cap} oracle blob={want.size} B")
    assert info["cnt"] <= cap
    recon = chains.reconstruct(info, odims, zigzag=zig)
    with _pipeline(kind, x.nbytes, dims, strict=False)[0] as p:
        blob = p.compress_array(x)
        assert blob.size == want.size
        np.testing.assert_array_equal(blob, want)
        y = p.decompress_array(blob)
        np.testing.assert_array_equal(y[: x.size], recon[: x.size])
    _say(kind, "plain chain: blob and reconstruction equal the oracle")
    pidx, pval = chains.strict_patches(x.ravel(), recon, info["abs_eb"], dims if kind != "1d" else (x.size, 1, 1))
    with _pipeline(kind, x.nbytes, dims, strict=True)[0] as p:
        blob = p.compress_array(x)
        got, ref = chains.split(blob), chains.split(want)
        assert len(got) == 7
        for i in range(4):
            np.testing.assert_array_equal(got[i], ref[i], err_msg=f"part {i}")
        np.testing.assert_array_equal(got[5].view(np.uint32), pidx)
        np.testing.assert_array_equal(got[4].view(np.float32), pval)
        assert int(got[6].view(np.uint32)[0]) == pidx.size
        y = p.decompress_array(blob)
    err = np.abs(y[: x.size].astype(np.float64) - x.ravel().astype(np.float64))
    worst = float(err.max())
    _say(kind, f"strict chain: {pidx.size} patches, max|err| {worst:.6e} <= eb {float(info['abs_eb']):.6e}")
    assert worst <= float(info["abs_eb"])


def test_sinusoid_1d_2p24_rle():
    x = fields.sinusoid_1d()
    want, info = chains.lorenzo_rle(x)
    _check("1d", x, (x.size, 1, 1), want, info)


def test_cloud_2d_1800x3600_bitshuffle_rze():
    x = fields.cloud_2d()
    dims = (3600, 1800, 1)
    want, info = chains.lorenzo_bitshuffle_rze(x, dims)
    _check("2d", x, dims, want, info)


def test_turbulence_3d_512_huffman(turbulence):
    x = turbulence
    dims = (512, 512, 512)
    want, info = chains.lorenzo_huffman(x, dims)
    _check("3d", x, dims, want, info)
