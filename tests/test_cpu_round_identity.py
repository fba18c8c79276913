"""The Lorenzo quantiser's rounding: round(v) (halves away from zero, C
round(), as lorenzo_quant.cu:120-135 uses it) computed as
trunc(v + copysign(c, v)) with c = 0.49999997 (the largest float below 1/2)
in float32 round-to-nearest-even arithmetic.  The identity holds for all 2^32
floats (checked once exhaustively); here every float within 8 ulps of a
half-integer in every binade, the binade edges, specials and 2^20 random bit
patterns are compared with a float64 reference."""
import numpy as np

C = np.float32(0.49999997)


def _fast(v):
    v = np.asarray(v, np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        return np.trunc(v + np.copysign(C, v)).astype(np.float32)


def _ref(v):
    with np.errstate(invalid="ignore"):
        d = np.asarray(v, np.float32).astype(np.float64)
        return np.copysign(np.floor(np.abs(d) + 0.5), d).astype(np.float32)


def _same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint32) == b.view(np.uint32)) | both_nan


def test_constant_is_largest_float_below_half():
    assert C == np.nextafter(np.float32(0.5), np.float32(0.0))


def test_near_half_integers_every_binade():
    vals = []
    for e in range(-3, 25):
        base = np.float32(2.0) ** e
        k = np.arange(0, 64, dtype=np.float64)
        halves = (np.floor(base) + k + 0.5).astype(np.float32)
        for h in halves:
            u = h.view(np.uint32)
            vals.append((u + np.arange(-8, 9, dtype=np.int64)).astype(np.uint32).view(np.float32))
    v = np.concatenate(vals)
    v = np.concatenate([v, -v])
    assert _same(_fast(v), _ref(v)).all()


def test_binade_edges_and_specials():
    e = np.arange(0, 255, dtype=np.uint32) << np.uint32(23)
    u = np.concatenate([e + d for d in (0, 1, 2)] + [e - 1, e - 2])
    v = u.astype(np.uint32).view(np.float32)
    v = np.concatenate([v, -v, np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 0.5, -0.5, 1.5, 2.5, -2.5,
                                         8388607.5, 8388608.0, 16777217.0], np.float32)])
    assert _same(_fast(v), _ref(v)).all()


def test_random_bit_patterns():
    u = np.random.default_rng(0).integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    v = u.view(np.float32)
    assert _same(_fast(v), _ref(v)).all()


def test_a_plain_half_constant_would_fail():
    """The pin: with c = 1/2 the sum rounds up just below a half (the
    classic mistake), so the check above would catch that formula."""
    v = np.float32(0.49999997)
    with np.errstate(invalid="ignore"):
        naive = np.trunc(v + np.float32(0.5))
    assert naive == 1.0 and _ref(v) == 0.0
