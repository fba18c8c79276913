"""Learned chain selector on the GPU: features and MLP kernels against the
CPU reference (oracle/selector_ref.py), the exhaustive labeller against the
real fz::Pipeline compress of every chunk under every chain, and a small
training run that must beat the selection rules without the network."""
import numpy as np
import pytest

from fzgpumodules_amd import fields
from fzgpumodules_amd._fzh import DeviceBuffer, FzhError, call
from fzgpumodules_amd import selector as S
from fzgpumodules_amd.pipeline import Pipeline
from oracle import selector_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_set():
    return fields.selector_set(5, 101)  # 5 parent fields per dimensionality, one per regime


def _dev(a):
    return DeviceBuffer.from_array(np.ascontiguousarray(a))


@pytest.mark.parametrize("nd", [2, 3])
def test_features_kernel_matches_reference(small_set, nd):
    chunks, eb, regime = small_set[nd]
    pick = np.concatenate([np.where(regime == r)[0][:2] for r in range(5)])
    c = chunks[pick]
    d_c, d_eb = _dev(c), _dev(eb[pick])
    d_f = S.features(d_c.ptr, len(pick), c.shape[1:], d_eb.ptr)
    got = d_f.download(np.float32, 32 * len(pick)).reshape(-1, 32)
    for i, j in enumerate(pick):
        want = R.features(chunks[j], eb[j])
        np.testing.assert_allclose(got[i], want, rtol=2e-6, atol=2e-6, err_msg=f"chunk {j} regime {regime[j]}")


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(R.NP) * 0.25).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(32) * 0.2, 0.5 + rng.random(32)]).astype(np.float32)
    return p, norm


def test_mlp_forward_is_bit_exact_with_the_bf16_reference():
    p, norm = _rand_params(1)
    rng = np.random.default_rng(2)
    feat = (rng.standard_normal((333, 32)) * 3).astype(np.float32)
    sel = S.Selector()
    try:
        sel.set_params(p)
        sel.norm.upload(norm)
        d_f = _dev(feat)
        z = sel.forward(d_f.ptr, feat.shape[0])
        want, _ = R.forward(p, norm, feat, "exact32")
        np.testing.assert_array_equal(z, want)
        ch = sel.select(d_f.ptr, feat.shape[0])
        np.testing.assert_array_equal(ch, np.argmax(want, axis=1))
    finally:
        sel.close()


def test_mlp_gradients_match_the_reference():
    p, norm = _rand_params(3)
    rng = np.random.default_rng(4)
    n = 200
    feat = (rng.standard_normal((n, 32)) * 2).astype(np.float32)
    cost = (rng.random((n, 4)) * 0.6).astype(np.float32)
    cost[np.arange(n), rng.integers(0, 4, n)] = 0
    sel = S.Selector()
    try:
        sel.set_params(p)
        sel.norm.upload(norm)
        d_f, d_c = _dev(feat), _dev(cost)
        loss = sel.train_step(d_f.ptr, d_c.ptr, n, 0.05)
        g = sel.grad.download(np.float32, R.NP)
        l32, g32 = R.loss_and_grad(p, norm, feat, cost, 0.05, "exact32")
        l64, g64 = R.loss_and_grad(p.astype(np.float64), norm.astype(np.float64), feat, cost, 0.05, "f64")
        assert abs(loss - l32) <= 1e-5 * abs(l32) + 1e-6
        # same rounding points: agreement to FP32 noise, except where exp()
        # differs by an ulp before a BF16 rounding of a delta
        assert np.max(np.abs(g - g32)) <= 2e-3 * np.max(np.abs(g32))
        assert np.linalg.norm(g - g32) <= 1e-3 * np.linalg.norm(g32)
        # and the BF16 network's gradient tracks the float64 one
        assert np.linalg.norm(g - g64) <= 0.05 * np.linalg.norm(g64)
        # Adam step
        m0, v0 = np.zeros(R.NP, np.float32), np.zeros(R.NP, np.float32)
        sel.adam(1e-3, wd=1e-4)
        p1 = sel.get_params()
        want, _, _ = R.adam(p, g, m0, v0, 1, 1e-3, wd=1e-4)
        np.testing.assert_allclose(p1, want, rtol=1e-6, atol=1e-7)
    finally:
        sel.close()


@pytest.mark.parametrize("nd", [2, 3])
def test_labeller_equals_pipeline_sizes(small_set, nd):
    chunks, eb, regime = small_set[nd]
    pick = np.concatenate([np.where(regime == r)[0][:2] for r in range(5)])
    c, e = chunks[pick], eb[pick]
    d_c, d_eb = _dev(c), _dev(e)
    sizes = S.label_sizes(d_c.ptr, len(pick), c.shape[1:], d_eb.ptr)
    for k in range(4):
        for i, j in enumerate(pick):
            with S.chain_pipeline(k, float(e[i]), c.shape[1:]) as p:
                blob = p.compress_array(chunks[j])
                assert sizes[i, k] == blob.size, (nd, S.CHAINS[k], int(regime[j]), int(sizes[i, k]), blob.size)
                if i % 3 == 0:  # decode from the recorded chain stays within eb
                    y = p.decompress_array(blob)
                    err = np.max(np.abs(y[: c[i].size].astype(np.float64) - c[i].ravel().astype(np.float64)))
                    assert err <= float(e[i]) * 1.01


def _featurize(data):
    feats, sizes = [], []
    for nd in (2, 3):
        ch, eb, _ = data[nd]
        d_c, d_eb = _dev(ch), _dev(eb)
        f = S.features(d_c.ptr, len(ch), ch.shape[1:], d_eb.ptr).download(np.float32, 32 * len(ch))
        feats.append(f.reshape(-1, 32))
        sizes.append(S.label_sizes(d_c.ptr, len(ch), ch.shape[1:], d_eb.ptr))
        d_c.free()
        d_eb.free()
    sz = np.concatenate(sizes)
    return np.concatenate(feats), S.regret_costs(sz), sz


@pytest.fixture(scope="module")
def trained():
    f_tr, c_tr, _ = _featurize(fields.selector_set(30, 202))
    sel = S.Selector(seed=5)
    try:
        hist = sel.fit(f_tr, c_tr, epochs=60, batch=64, lr=3e-3, tau=0.02)
        return sel.get_params(), sel.get_norm(), hist
    finally:
        sel.close()


def test_selector_training_beats_rules(small_set, trained):
    params, norm, hist = trained
    assert hist[-1] < hist[0]
    f_te, _, s_te = _featurize(small_set)
    sel = S.Selector()
    try:
        sel.set_params(params)
        sel.norm.upload(norm)
        choice = sel.select(_dev(f_te).ptr, f_te.shape[0])
    finally:
        sel.close()
    best = s_te.min(axis=1).sum()
    learned = s_te[np.arange(len(choice)), choice].sum()
    regret = learned / best - 1
    # the method without the network: the best single chain for every chunk
    single = min(s_te[:, k].sum() for k in range(4)) / best - 1
    assert regret <= 0.02, regret
    assert regret < single, (regret, single)


def _composite_field():
    """A 512 x 1024 field whose eight 256 x 256 chunks come from all regimes."""
    tiles = [fields.regime_field(r % 5, (256, 256), 900 + r) for r in range(8)]
    rows = [np.concatenate(tiles[4 * i:4 * i + 4], axis=1) for i in range(2)]
    return np.ascontiguousarray(np.concatenate(rows, axis=0))


def test_selecting_compressor_round_trip(trained):
    params, norm, _ = trained
    x = _composite_field()
    d = _dev(x)
    out = DeviceBuffer(x.nbytes)
    sc = S.SelectingCompressor(fields.CHUNK_2D)
    try:
        sc.set_model(params, norm)
        eb = float(np.float32(np.float32(1e-3) * np.float32(np.abs(x).max())))
        sizes = {}
        for mode in (S.LEARNED, S.EXHAUSTIVE, S.FIXED):
            blob = sc.compress(d.ptr, x.shape, 1e-3, mode=mode, chain=1)
            ch = sc.choices()
            assert ch.size == 8
            if mode == S.FIXED:
                assert np.all(ch == 1)
            sizes[mode] = blob.size
            out.upload(np.zeros(x.size, np.float32))
            sc.decompress(blob, out.ptr, x.size)
            y = out.download(np.float32, x.size).reshape(x.shape)
            err = np.max(np.abs(y.astype(np.float64) - x.astype(np.float64)))
            assert err <= eb * 1.01, (mode, err, eb)
        assert sizes[S.EXHAUSTIVE] <= sizes[S.LEARNED] <= sizes[S.FIXED] * 1.02
        blob = sc.compress(d.ptr, x.shape, 1e-3, mode=S.LEARNED)
        bad = blob.copy()
        bad[-5] ^= 0xFF
        with pytest.raises(FzhError, match="checksum"):
            sc.decompress(bad, out.ptr, x.size)
    finally:
        sc.close()
        d.free()
        out.free()


def test_chain1_chunks_equal_the_pipeline(trained):
    """Chunks on chain 1 (LorenzoQuant ABS -> Bitshuffle(16384, 2) ->
    RZE(16384, 4)) are compressed as one batch; every chunk's bytes in the
    container equal fz::Pipeline::compress of that chunk alone with the
    chain's settings (radius 512, outlier capacity 0.2, zigzag)."""
    x = _composite_field()
    d = _dev(x)
    sc = S.SelectingCompressor(fields.CHUNK_2D)
    try:
        for _ in range(2):  # the second call reuses the batch plan
            blob = sc.compress(d.ptr, x.shape, 1e-3, mode=S.FIXED, chain=1)
        assert np.all(sc.choices() == 1)
    finally:
        sc.close()
        d.free()
    abs_eb = float(blob[56:60].view(np.float32)[0])
    want = []
    for cy in range(x.shape[0] // 256):
        for cx in range(x.shape[1] // 256):
            tile = np.ascontiguousarray(x[256 * cy:256 * (cy + 1), 256 * cx:256 * (cx + 1)])
            with Pipeline(tile.nbytes) as p:
                lz = p.add_lorenzo(abs_eb, mode="ABS", quant_radius=512, outlier_capacity=0.2, zigzag=True)
                bs = p.add_bitshuffle(16384, 2)
                rz = p.add_rze(16384, 4)
                p.connect(bs, lz, "codes")
                p.connect(rz, bs, "output")
                p.set_dims(256, 256, 1)
                p.finalize()
                want.append(p.compress_array(tile))
    pad = [(w.size + 15) & ~15 for w in want]
    start = blob.size - sum(pad)
    sizes = blob[start - 8 * len(want):start].view(np.uint64)
    np.testing.assert_array_equal(sizes, [w.size for w in want])
    o = start
    for c, w in enumerate(want):
        np.testing.assert_array_equal(blob[o:o + w.size], w, err_msg=f"chunk {c}")
        o += pad[c]


# ---------------------------------------------------------------------------
# Device-generated selector chunks (fzs_synth_chunks)
# ---------------------------------------------------------------------------
def _synth(nd, n, seed):
    E = 65536
    base = S.synth_base(nd, 99)
    d_c, d_eb, d_r = S.synth_chunks(nd, n, seed, base, regime=True)
    base.free()
    shape = (n, 256, 256) if nd == 2 else (n, 16, 64, 64)
    out = (d_c.download(np.float32, E * n).reshape(shape), d_eb.download(np.float32, n),
           d_r.download(np.uint8, n))
    for b in (d_c, d_eb, d_r):
        b.free()
    return out


@pytest.mark.parametrize("nd", [2, 3])
def test_synth_chunks_regimes_and_bounds(nd):
    cpp = 4 if nd == 2 else 8
    n = 10 * cpp  # two parent fields of every regime
    c, eb, reg = _synth(nd, n, 7)
    np.testing.assert_array_equal(reg, (np.arange(n) // cpp) % 5)
    for f in range(n // cpp):
        ch = c[f * cpp:(f + 1) * cpp]
        vb = float(np.max(np.abs(ch)))
        want = np.float32(np.float32(1e-3) * np.float32(vb)) if vb > 0 else np.float32(1e-3)
        assert (eb[f * cpp:(f + 1) * cpp] == want).all()
        r = f % 5
        v = ch.astype(np.float64).ravel()
        if r == 0:  # smooth: neighbours along x nearly equal
            a = ch.reshape(-1, ch.shape[-1])
            cor = np.corrcoef(a[:, :-1].ravel(), a[:, 1:].ravel())[0, 1]
            assert cor > 0.9
        elif r == 1:
            assert abs(v.mean()) < 0.02 and abs(v.std() - 1) < 0.02
        elif r == 2:
            z = np.mean(v == 0.0)
            assert 0.78 < z < 0.82
            assert abs(v[v != 0].std() - 1) < 0.03
        elif r == 3:
            a = ch.reshape(-1, ch.shape[-1])
            assert np.mean(a[:, 1:] == a[:, :-1]) > 0.9
        else:
            assert np.array_equal(v, np.round(v)) and v.min() >= 0
            assert abs(v.mean() - 20) < 0.1 and abs(v.var() - 20) < 0.5


def test_synth_chunks_deterministic_and_seeded():
    a = _synth(2, 8, 3)[0]
    b = _synth(2, 8, 3)[0]
    c = _synth(2, 8, 4)[0]
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)


def test_synth_chunks_rejects_partial_parent():
    d = DeviceBuffer(16)
    with pytest.raises(FzhError, match="whole number of parent"):
        call("fzs_synth_chunks", None, 3, 4, 1, d.ptr, 4, 4, 4, d.ptr, d.ptr, None)  # 3D parents hold 8 chunks
    d.free()
