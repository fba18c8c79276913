"""Pins the CPU reference of the chain selector (oracle/selector_ref.py):
BF16 rounding against an independent construction, the backward formulas
against central finite differences in float64, and the vectorised features
against a direct per-element loop on tiny chunks."""
import math

import numpy as np

from oracle import selector_ref as R


def _bf16_slow(v: float) -> float:
    """Round a float32 value to 8 significant bits, ties to even (frexp)."""
    if v == 0 or not math.isfinite(v):
        return v
    m, e = math.frexp(v)  # v = m * 2^e, 0.5 <= |m| < 1
    scaled = m * 256.0  # 8 significant bits -> integer part
    r = float(np.rint(scaled))  # ties to even
    return math.ldexp(r / 256.0, e)


def test_bf16_rounding_known_answers_and_random():
    one = 1.0
    cases = {one: one, one + 2 ** -8: one, one + 3 * 2 ** -8: one + 2 ** -6, -(one + 2 ** -8): -one,
             one + 2 ** -7: one + 2 ** -7, 3.0e38: float(np.float32(_bf16_slow(3.0e38)))}
    for x, want in cases.items():
        got = float(R.bf16r(np.array([x], np.float32))[0])
        assert got == want, (x, got, want)
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    got = R.bf16r(x)
    want = np.array([_bf16_slow(float(v)) for v in x], np.float32)
    np.testing.assert_array_equal(got, want)
    assert np.isinf(R.bf16r(np.array([np.inf], np.float32))[0])
    assert np.isnan(R.bf16r(np.array([np.nan], np.float32))[0])


def _rand_net(seed, n=24):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(R.NP) * 0.3).astype(np.float64)
    norm = np.concatenate([rng.standard_normal(R.NF) * 0.1, 1.0 + rng.random(R.NF)])
    feat = rng.standard_normal((n, R.NF))
    cost = rng.random((n, R.NC)) * 0.5
    cost[np.arange(n), rng.integers(0, R.NC, n)] = 0.0
    return p, norm, feat, cost


def test_backward_matches_finite_differences_f64():
    p, norm, feat, cost = _rand_net(3)
    loss, g = R.loss_and_grad(p, norm, feat, cost, 0.05, mode="f64")
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(R.OB1, 12, replace=False), R.OB1 + rng.choice(64, 4, replace=False),
                          R.OW2 + rng.choice(4096, 12, replace=False), R.OB2 + rng.choice(64, 4, replace=False),
                          R.OW3 + rng.choice(256, 8, replace=False), R.OB3 + np.arange(4)])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (R.loss_and_grad(pp, norm, feat, cost, 0.05, mode="f64")[0] -
              R.loss_and_grad(pm, norm, feat, cost, 0.05, mode="f64")[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 + 1e-4 * abs(fd), (q, fd, g[q])


def test_bf16_mode_tracks_f64():
    p, norm, feat, cost = _rand_net(5, 64)
    l64, g64 = R.loss_and_grad(p, norm, feat, cost, 0.05, mode="f64")
    l32, g32 = R.loss_and_grad(p.astype(np.float32), norm.astype(np.float32), feat.astype(np.float32), cost, 0.05)
    assert abs(l32 - l64) < 0.05 * abs(l64) + 1e-3
    # BF16 inputs carry ~3 significant digits: the gradients agree to that level
    assert np.linalg.norm(g32 - g64) <= 0.05 * np.linalg.norm(g64)


def test_adam_reference_first_step_is_sign_times_lr():
    p = np.ones(R.NP, np.float32)
    g = np.linspace(-1, 1, R.NP).astype(np.float32)
    p2, m, v = R.adam(p, g, np.zeros(R.NP, np.float32), np.zeros(R.NP, np.float32), 1, 0.01, eps=1e-12)
    nz = g != 0
    np.testing.assert_allclose(p2[nz], (1 - 0.01 * np.sign(g))[nz], rtol=1e-5)


def _features_slow(x, eb):
    """Direct per-element loop restatement of the feature list (tiny chunks)."""
    x = np.asarray(x, np.float32)
    if x.ndim == 2:
        x = x[None]
    nz, ny, nx = x.shape
    f32 = np.float32
    inv2eb = f32(0.5) / f32(eb)
    E = x.size
    cnt = dict(zero=0, integ=0, rep=0, smooth=0, out=0, qout=0, pos=0)
    sdx = sdy = sdz = sl = sabs = 0.0
    hist = [0] * 15
    for z in range(nz):
        for y in range(ny):
            for i in range(nx):
                def at(dz, dy, dx):
                    zz, yy, xx = z - dz, y - dy, i - dx
                    return x[zz, yy, xx] if zz >= 0 and yy >= 0 and xx >= 0 else f32(0)
                v = x[z, y, i]
                w, n_, wn = at(0, 0, 1), at(0, 1, 0), at(0, 1, 1)
                if nz > 1:
                    r = f32(f32(f32(f32(f32(f32(v - w) - n_) - at(1, 0, 0)) + wn) + at(1, 0, 1)) + at(1, 1, 0))
                    r = f32(r - at(1, 1, 1))
                    if z:
                        sdz += abs(float(v) - float(at(1, 0, 0)))
                else:
                    r = f32(f32(f32(v - w) - n_) + wn)
                l = f32(abs(r) * inv2eb)
                sl += float(l)
                lp = f32(l + f32(0.5))
                b = 0 if lp < 1 else min(14, int(math.floor(math.log2(min(float(lp), 16383.99)))) + 1)
                hist[b] += 1
                cnt["zero"] += v == 0
                cnt["integ"] += v == np.rint(v)
                cnt["pos"] += v > 0
                cnt["out"] += l >= 511.5
                cnt["qout"] += f32(abs(v) * inv2eb) >= 32767.5
                sabs += abs(float(v))
                if i:
                    cnt["rep"] += v == w
                    d = abs(float(v) - float(w))
                    sdx += d
                    cnt["smooth"] += d * float(inv2eb) < 0.5
                if y:
                    sdy += abs(float(v) - float(n_))
    e = float(np.float32(eb))
    px, py, pz = (nx - 1) * ny * nz, nx * (ny - 1) * nz, nx * ny * (nz - 1)
    xd = x.astype(np.float64)
    f = np.zeros(32)
    f[0] = math.log2(1 + math.sqrt(max((xd ** 2).mean() - xd.mean() ** 2, 0)) / e)
    f[1] = math.log2(1 + (xd.max() - xd.min()) / e)
    f[2], f[3] = cnt["zero"] / E, cnt["integ"] / E
    f[4] = cnt["rep"] / px
    f[5], f[6] = math.log2(1 + sdx / px / e), math.log2(1 + sdy / py / e)
    f[7] = math.log2(1 + sdz / pz / e) if nz > 1 else 0
    f[8] = math.log2(1 + sl / E)
    for k in range(15):
        f[9 + k] = hist[k] / E
    f[24] = cnt["smooth"] / px
    f[25] = math.log2(1 + sabs / E / e)
    f[26] = math.log2(1 + float(np.abs(x).max()) / e)
    f[27], f[28] = cnt["out"] / E, cnt["qout"] / E
    f[29] = 1.0 if nz > 1 else 0.0
    f[30] = -sum(h / E * math.log2(h / E) for h in hist if h)
    f[31] = cnt["pos"] / E
    return f


def test_features_reference_matches_direct_loops():
    rng = np.random.default_rng(9)
    cases = [
        (rng.standard_normal((6, 7)).astype(np.float32), 1e-2),
        (np.round(rng.standard_normal((5, 4, 6)) * 3).astype(np.float32), 0.05),
        (np.zeros((4, 5), np.float32), 1e-3),
        ((np.arange(30, dtype=np.float32).reshape(5, 6) * 0.25), 0.1),
        ((rng.random((3, 4, 5)) * 1e4).astype(np.float32), 1e-3),
    ]
    for x, eb in cases:
        np.testing.assert_allclose(R.features(x, eb), _features_slow(x, eb), rtol=1e-9, atol=1e-12)


def test_features_closed_forms():
    # a 2D ramp: the Lorenzo residual vanishes away from the first row/column
    ny, nx = 8, 8
    x = (np.arange(ny)[:, None] * 2.0 + np.arange(nx)[None, :] * 3.0).astype(np.float32)
    f = R.features(x, 0.5)
    interior = (ny - 1) * (nx - 1)
    assert f[9] >= interior / x.size  # bin 0 holds at least the interior
    assert f[5] == np.log2(1 + 3.0 / 0.5) and f[6] == np.log2(1 + 2.0 / 0.5)
    assert f[3] == 1.0 and f[29] == 0.0
    # constant chunk: std 0, no west changes
    c = np.full((4, 4, 4), 7.0, np.float32)
    g = R.features(c, 1.0)
    assert g[0] == 0.0 and g[4] == 1.0 and g[24] == 1.0 and g[29] == 1.0 and g[31] == 1.0
