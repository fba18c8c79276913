"""Batched compression (one persistent launch per stage over many fields):
every field's bytes must equal fz::Pipeline::compress of that field alone
with the same settings, whatever the launch geometry, in eager and graph
mode; a subset is also checked against the oracle chain directly, and every
field decompresses through fz::Pipeline within eb."""

import numpy as np
import pytest

import chains
from fzgpumodules_amd import fields
from fzgpumodules_amd._fzh import DeviceBuffer, FzhError
from fzgpumodules_amd.batch import Batch
from fzgpumodules_amd.pipeline import Pipeline

pytestmark = pytest.mark.gpu


def _field(shape, expo, seed, zero_frac=0.0):
    g = fields.power_law_field(shape, expo, seed)
    if zero_frac:
        q = np.quantile(g, zero_frac)
        g = np.clip((g - q) / (g.max() - q), 0.0, None)
    return g.astype(np.float32)


def _mixed_fields(count, seed, max_elems=1 << 18):
    """Fields of mixed dimensionality and uneven sizes (64 KiB .. 1 MiB)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nd = 1 + i % 3
        n = int(np.exp(rng.uniform(np.log(16384), np.log(max_elems))))
        expo = float(rng.uniform(-5.0, -1.0))
        if nd == 1:
            shape = (n,)
        elif nd == 2:
            a = float(np.exp(rng.uniform(-1.0, 1.0)))
            nx = max(1, int(round(np.sqrt(n * a))))
            shape = (max(2, n // nx), nx)
        else:
            a, b = np.exp(rng.uniform(-0.7, 0.7, 2))
            side = n ** (1 / 3)
            nx, ny = max(1, int(round(side * a))), max(2, int(round(side * b)))
            shape = (max(2, n // (nx * ny)), ny, nx)
        out.append(_field(shape, expo, seed * 1000 + i, zero_frac=0.3 if i % 4 == 0 else 0.0))
    return out


def _dims(x):
    s = x.shape
    return (s[-1], s[-2] if len(s) > 1 else 1, s[-3] if len(s) > 2 else 1)


def _single(x, eb=1e-3, mode="REL", radius=32768, cap=0.1, strict=False, patch_cap=0.0):
    nx, ny, nz = _dims(x)
    with Pipeline(x.nbytes) as p:
        lz = p.add_lorenzo(eb, mode=mode, quant_radius=radius, outlier_capacity=cap, zigzag=True)
        if strict:
            p.set_strict(lz, True, patch_cap)
        bs = p.add_bitshuffle(16384, 2)
        rz = p.add_rze(16384, 4)
        p.connect(bs, lz, "codes")
        p.connect(rz, bs)
        p.set_dims(nx, ny, nz)
        p.finalize()
        blob = p.compress_array(x)
        y = p.decompress_array(blob)
        abs_eb = p.abs_error_bound(lz)
    return blob, y, abs_eb


class _Batch:
    def __init__(self, xs, **cfg):
        self.bufs = [DeviceBuffer.from_array(np.ascontiguousarray(x)) for x in xs]
        self.b = Batch(**cfg)
        self.b.set_fields([(d.ptr, _dims(x)) for d, x in zip(self.bufs, xs)])

    def run(self):
        self.b.compress()
        return [self.b.field_bytes(i) for i in range(len(self.bufs))]

    def close(self):
        self.b.close()
        for d in self.bufs:
            d.free()


@pytest.fixture(scope="module")
def mixed():
    return _mixed_fields(18, 7)


@pytest.fixture(scope="module")
def mixed_reference(mixed):
    return [_single(x) for x in mixed]


def test_batch_equals_single_field_pipeline(mixed, mixed_reference):
    bt = _Batch(mixed)
    try:
        got = bt.run()
        for i, (x, (want, y, abs_eb)) in enumerate(zip(mixed, mixed_reference)):
            np.testing.assert_array_equal(got[i], want, err_msg=f"field {i} shape {x.shape}")
            err = np.max(np.abs(y[: x.size].astype(np.float64) - x.ravel().astype(np.float64)))
            assert err <= abs_eb * 1.01, (i, err, abs_eb)  # the reference's own tolerance (test_lorenzo.cpp:189)
        tin, tout = bt.b.totals()
        assert tin == sum(x.nbytes for x in mixed)
        assert tout == sum(g.size for g in got)
    finally:
        bt.close()


def test_batch_fields_match_the_oracle_chain(mixed, mixed_reference):
    # the single-field pipeline is itself pinned against the oracle; check a
    # few batch fields against the oracle chain directly as well
    bt = _Batch(mixed[:6])
    try:
        got = bt.run()
        for i, x in enumerate(mixed[:6]):
            nx, ny, nz = _dims(x)
            dims = (0, 1, 1) if x.ndim == 1 else (nx, ny, nz)
            want, _ = chains.lorenzo_bitshuffle_rze(x, dims)
            np.testing.assert_array_equal(got[i], want, err_msg=f"field {i} shape {x.shape}")
    finally:
        bt.close()


@pytest.mark.parametrize("launch", [
    dict(vb_wg_per_cu=1, lz_wg_per_cu=1, lz_tiles_per_unit=1, pack_wg_per_cu=1, scatter_wg_per_cu=1),
    dict(vb_wg_per_cu=8, lz_wg_per_cu=16, lz_tiles_per_unit=16, pack_wg_per_cu=8, scatter_wg_per_cu=8),
    dict(lz_tiles_per_unit=3, pack_wg_per_cu=3),
    dict(lz_wg_per_cu=1 << 20, lz_tiles_per_unit=1),  # one tile per wave, no loop
])
def test_batch_bytes_do_not_depend_on_launch_geometry(mixed, mixed_reference, launch):
    bt = _Batch(mixed)
    try:
        bt.b.set_launch(**launch)
        got = bt.run()
        for i, (want, _, _) in enumerate(mixed_reference):
            np.testing.assert_array_equal(got[i], want, err_msg=f"field {i} {launch}")
    finally:
        bt.close()


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_batch_of_one_dimensionality(nd):
    """Only one tile shape present: the other shapes' launches are skipped."""
    xs = [x for x in _mixed_fields(9, 11, max_elems=1 << 16) if x.ndim == nd]
    want = [_single(x)[0] for x in xs]
    for launch in (dict(lz_wg_per_cu=1 << 20, lz_tiles_per_unit=1), dict(lz_wg_per_cu=2, lz_tiles_per_unit=4)):
        bt = _Batch(xs)
        try:
            bt.b.set_launch(**launch)
            got = bt.run()
            for i in range(len(xs)):
                np.testing.assert_array_equal(got[i], want[i], err_msg=f"field {i} {launch}")
        finally:
            bt.close()


def test_batch_graph_replay_and_autotune(mixed, mixed_reference):
    bt = _Batch(mixed)
    try:
        for rep in range(3):  # eager, capture, replay
            got = bt.run()
            for i, (want, _, _) in enumerate(mixed_reference):
                np.testing.assert_array_equal(got[i], want, err_msg=f"field {i} call {rep}")
        best = bt.b.autotune(reps=2)
        assert best == bt.b.get_launch()
        for rep in range(2):
            got = bt.run()
            for i, (want, _, _) in enumerate(mixed_reference):
                np.testing.assert_array_equal(got[i], want, err_msg=f"field {i} after autotune")
        bt.b.enable_profiling(True)
        bt.run()
        names = [n for n, _ in bt.b.stage_ms()]
        assert names == ["value_base", "lorenzo", "bitshuffle_rze", "scan", "finalize", "pack", "scatter"]
    finally:
        bt.close()


def test_batch_dense_outliers_recompute_path():
    # a small radius makes most tiles overflow the per-tile list slots, so the
    # scatter kernel recomputes their lists; ABS mode, outlier overflow in one field
    xs = [_field((40, 300), -1.5, 11), _field((70000,), -2.0, 12), _field((12, 20, 36), -1.0, 13)]
    xs[1] *= 50.0
    cfg = dict(error_bound=1e-3, mode="ABS", quant_radius=64, outlier_capacity=0.3)
    bt = _Batch(xs, **cfg)
    try:
        got = bt.run()
        for i, x in enumerate(xs):
            want, y, abs_eb = _single(x, eb=1e-3, mode="ABS", radius=64, cap=0.3)
            np.testing.assert_array_equal(got[i], want, err_msg=f"field {i}")
            n_out = int(chains.split(want)[3].view(np.uint32)[0])
            assert n_out > 16 * 4, "the data must overflow the per-tile slots"
    finally:
        bt.close()


def test_batch_strict_mode(mixed):
    xs = mixed
    bt = _Batch(xs, error_bound=1e-4, strict=True, patch_capacity=0.05)
    try:
        got = bt.run()
        patched = 0
        for i, x in enumerate(xs):
            want, y, abs_eb = _single(x, eb=1e-4, strict=True, patch_cap=0.05)
            np.testing.assert_array_equal(got[i], want, err_msg=f"field {i}")
            err = np.max(np.abs(y[: x.size].astype(np.float64) - x.ravel().astype(np.float64)))
            assert err <= abs_eb, (i, err, abs_eb)
            patched += int(chains.split(want)[6].view(np.uint32)[0])
        assert patched > 0
    finally:
        bt.close()


def test_batch_rejects_bad_fields():
    b = Batch()
    try:
        with pytest.raises(FzhError):
            b.set_fields([])
        d = DeviceBuffer(4096)
        with pytest.raises(FzhError, match="aligned"):
            b.set_fields([(d.ptr + 4, (100, 1, 1))])
        with pytest.raises(FzhError, match="empty"):
            b.set_fields([(d.ptr, (0, 1, 1))])
        with pytest.raises(FzhError, match="setFields"):
            b.compress()
        d.free()
    finally:
        b.close()
