"""Pins the oracle's Quantizer REL mode (orc_quant_rel_*), which restates
modules/quantizers/quantizer/quantizer.cu:171-352 (kernels) and :426-436,
:563-580 (the stage's constants), with log2 and 2^t taken correctly rounded
to float where the reference calls __log2f / exp2f.

Pins: the host constants against their double-precision values; an
independent vectorised numpy statement of the forward and inverse compared
bit for bit; hand-derived codes; the reference's own REL cases
(tests/stages/test_quantizer.cpp:354-466, :1164-1200: smooth round trip
within eps, zeros/denormals/inf/NaN as outliers, the threshold); and the
pointwise bound |x_hat - x| <= eps |x| on every quantised element.
"""
import numpy as np
import pytest

CODE_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}
F = np.float32


def params_np(eps):
    e = F(eps)
    one = F(1) + e
    log2eb = F(2) * F(np.log2(np.float64(one)))
    return np.array([log2eb, F(1) / log2eb, one, F(1) / one], np.float32)


def rel_numpy(x, eps, radius, cb, threshold=np.inf):
    x = np.ascontiguousarray(x, np.float32)
    p = params_np(eps)
    log2eb, log2eb_r, opp, oopp = p
    rad = int(np.array(radius).astype(CODE_DT[cb]))
    cmax = np.iinfo(CODE_DT[cb]).max
    expo = (x.view(np.uint32) >> 23) & 0xFF
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        cand = ~(np.abs(x) >= F(threshold)) & (expo != 0) & (expo != 255)
        ax = np.where(cand, np.abs(x), F(1))
        lf = np.log2(ax.astype(np.float64)).astype(np.float32)
        lb = np.rint(lf * log2eb_r).astype(np.int64)
        alb = np.abs(lb)
        stored = ((alb << 2) | ((lb < 0).astype(np.int64) << 1) | (x < 0).astype(np.int64)) + 1
        fits = (alb < rad) & (stored <= cmax)
        rec = np.exp2((lb.astype(np.float32) * log2eb).astype(np.float64)).astype(np.float32)
        ok = cand & fits & (rec >= ax * oopp) & (rec <= ax * opp)
    codes = np.where(ok, stored, 0).astype(CODE_DT[cb])
    o = np.nonzero(~ok)[0]
    return codes, x[o], o.astype(np.uint32)


def rel_inv_numpy(codes, eps, vals, idx):
    log2eb = params_np(eps)[0]
    st = codes.astype(np.int64)
    cv = st - 1
    alb = cv >> 2
    lb = np.where((cv >> 1) & 1, -alb, alb)
    rec = np.exp2((lb.astype(np.float32) * log2eb).astype(np.float64)).astype(np.float32)
    out = np.where(st == 0, F(0), np.where(cv & 1, -rec, rec)).astype(np.float32)
    out[idx] = vals
    return out


@pytest.mark.parametrize("eps", [1e-2, 5e-3, 1e-3, 1e-4])
def test_constants(orc, eps):
    np.testing.assert_array_equal(orc.quant_rel_params(eps), params_np(eps))


def mixed(rng, n):
    x = (rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))).astype(np.float32)
    k = n // 50
    x[rng.integers(0, n, k)] = 0.0
    x[rng.integers(0, n, k)] = -0.0
    x[rng.integers(0, n, k)] = np.float32(1e-40)  # denormal
    x[rng.integers(0, n, 3)] = np.inf
    x[rng.integers(0, n, 3)] = np.nan
    x[rng.integers(0, n, k)] = 1.0
    x[rng.integers(0, n, k)] = -2.0
    return x


@pytest.mark.parametrize("eps", [1e-2, 1e-3])
@pytest.mark.parametrize("cb,radius", [(4, 512), (4, 1 << 20), (2, 32768), (1, 64)])
@pytest.mark.parametrize("threshold", [np.inf, 1e3])
def test_forward_and_inverse_match_numpy(orc, eps, cb, radius, threshold):
    rng = np.random.default_rng(int(eps * 1e4) + cb)
    x = mixed(rng, 200003)
    c, v, i, cnt = orc.quant_rel_fwd(x, eps, radius, cb, threshold)
    nc, nv, ni = rel_numpy(x, eps, radius, cb, threshold)
    np.testing.assert_array_equal(c, nc)
    np.testing.assert_array_equal(i, ni)
    np.testing.assert_array_equal(v.view(np.uint32), nv.view(np.uint32))
    assert cnt == ni.size
    y = orc.quant_rel_inv(c, eps, v, i, cnt)
    np.testing.assert_array_equal(y.view(np.uint32), rel_inv_numpy(nc, eps, nv, ni).view(np.uint32))
    q = c != 0
    assert q.any()
    # the bound test runs on float products |x| (1 + eps), |x| / (1 + eps):
    # exact up to their rounding, 2^-24 |x| each
    err = np.abs(y[q].astype(np.float64) - x[q].astype(np.float64))
    assert np.all(err <= np.abs(x[q].astype(np.float64)) * (float(np.float32(eps)) + 2.0 ** -23))


def test_known_codes():
    import oracle
    eps = 1e-3
    log2eb = float(params_np(eps)[0])
    lb2 = int(np.rint(np.float32(1.0) * np.float32(1 / np.float32(log2eb))))
    assert lb2 == 347  # 1 / (2 log2(1.001)) = 346.92...
    x = np.array([1.0, -1.0, 2.0, 0.5, -0.5, 0.0, np.inf, 1e-40], np.float32)
    c, v, i, cnt = oracle.quant_rel_fwd(x, eps, 512, 4)
    # (|lb| << 2 | (lb<0) << 1 | (x<0)) + 1
    np.testing.assert_array_equal(c, [1, 2, (347 << 2) + 1, ((347 << 2) | 2) + 1, ((347 << 2) | 3) + 1, 0, 0, 0])
    assert list(i) == [5, 6, 7] and cnt == 3
    # u8 codes: stored must fit in 255 -> |lb| <= 63, so 2.0 becomes an outlier
    c8, _, i8, _ = oracle.quant_rel_fwd(x, eps, 64, 1)
    assert c8[2] == 0 and 2 in list(i8) and c8[0] == 1
    # radius bounds |lb|: |lb| = 347 < radius 347 is false
    cr, _, ir, _ = oracle.quant_rel_fwd(x[:3], eps, 347, 4)
    assert cr[2] == 0 and list(ir) == [2]


def test_reference_smooth_round_trip(orc):
    """test_quantizer.cpp:354-418: (|sin(0.05 i)| + 0.1) * 10, eps 1e-2,
    radius 512, u32 codes, capacity 0.1."""
    i = np.arange(1024, dtype=np.float32)
    x = ((np.abs(np.sin(i * np.float32(0.05))) + np.float32(0.1)) * np.float32(10)).astype(np.float32)
    c, v, idx, cnt = orc.quant_rel_fwd(x, 0.01, 512, 4, max_outliers=103)
    y = orc.quant_rel_inv(c, 0.01, v, idx, cnt)
    out = np.zeros(x.size, bool)
    out[idx] = True
    np.testing.assert_array_equal(y[out], x[out])
    rel = np.abs(y[~out] - x[~out]) / np.abs(x[~out])
    assert rel.max() <= 0.01 * 1.5


def test_reference_zeros_go_to_outliers(orc):
    """test_quantizer.cpp:420-466: every fourth element 0, the rest +-5 sin."""
    i = np.arange(256)
    s = (np.sin(i.astype(np.float32) * np.float32(0.1) + np.float32(0.5)) * np.float32(5)).astype(np.float32)
    x = np.where(i % 4 == 0, np.float32(0), np.where(i % 4 == 1, -s, s)).astype(np.float32)
    c, v, idx, cnt = orc.quant_rel_fwd(x, 0.01, 512, 4, max_outliers=128)
    assert cnt >= 64 and set(range(0, 256, 4)) <= set(idx.tolist())
    y = orc.quant_rel_inv(c, 0.01, v, idx, cnt)
    assert np.all(y[::4] == 0.0)


def test_reference_threshold_forces_outliers(orc):
    """test_quantizer.cpp:1164-1200: |x| >= threshold is an outlier in REL."""
    x = np.linspace(0.5, 8.0, 512).astype(np.float32)
    c, v, idx, cnt = orc.quant_rel_fwd(x, 0.01, 512, 4, threshold=4.0)
    big = np.nonzero(np.abs(x) >= 4.0)[0]
    assert set(big.tolist()) <= set(idx.tolist())
    assert np.all(c[big] == 0)
