"""Pins of the in-place outlier quantizer oracle (oracle/quant_inplace_ref.py):
the reference's own cases (tests/stages/test_quantizer.cpp:875-1010: ABS and
NOA round trips within eb, all-outlier field decoded bit for bit), the code
layout in closed form, and the threshold rule."""
import numpy as np

from oracle import quant_inplace_ref as Q


def _sine(n, freq, amp):
    return (amp * np.sin(freq * np.arange(n))).astype(np.float32)


def test_abs_round_trip_within_eb():
    x = _sine(1024, 0.05, 20.0)
    eb = np.float32(1e-2)
    ebx2 = np.float32(2) * eb
    c = Q.fwd(x, np.float32(1) / ebx2, 1 << 22)
    y = Q.inv(c, ebx2, 1 << 22)
    assert np.abs(y.astype(np.float64) - x).max() <= float(eb) * 1.01


def test_noa_round_trip_within_eb():
    x = _sine(1024, 0.04, 8.0)
    eb = np.float32(0.01) * (x.max() - x.min())
    ebx2 = np.float32(2) * eb
    c = Q.fwd(x, np.float32(1) / ebx2, 32768)
    assert np.abs(Q.inv(c, ebx2, 32768).astype(np.float64) - x).max() <= float(eb) * 1.02


def test_all_outliers_are_raw_bits():
    x = np.full(512, 5.0, np.float32)
    c = Q.fwd(x, np.float32(1) / np.float32(2e-3), 8)
    assert (c == np.float32(5.0).view(np.uint32)).all()
    assert np.array_equal(Q.inv(c, np.float32(2e-3), 8), x)


def test_code_layout_closed_form():
    q = np.arange(-1000, 1000)
    x = q.astype(np.float32)  # ebx2 = 1: x * r is exactly q
    c = Q.fwd(x, 1.0, 1001)
    want = np.where(q >= 0, 2 * q, -2 * q - 1)
    assert np.array_equal(c.astype(np.int64), want)
    assert np.array_equal(Q.inv(c, np.float32(1), 1001), x)
    # |q| = radius is an outlier: raw bits
    c2 = Q.fwd(np.array([1001.0, -1001.0, 1000.0], np.float32), 1.0, 1001)
    assert c2[0] == np.float32(1001.0).view(np.uint32) and c2[1] == np.float32(-1001.0).view(np.uint32)
    assert c2[2] == 2000


def test_ties_round_half_to_even_and_threshold():
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5], np.float32)
    assert np.array_equal(Q.fwd(x, 1.0, 100), np.array([0, 4, 4, 0, 3], np.uint32))
    t = Q.fwd(np.array([3.0, 4.0, -4.0], np.float32), 1.0, 100, threshold=4.0)
    assert t[0] == 6 and t[1] == np.float32(4.0).view(np.uint32) and t[2] == np.float32(-4.0).view(np.uint32)


def test_nan_and_inf_round_trip_as_raw_bits():
    x = np.array([np.inf, -np.inf, 1.0], np.float32)
    c = Q.fwd(x, 1.0, 100)
    y = Q.inv(c, np.float32(1), 100)
    assert np.isposinf(y[0]) and np.isneginf(y[1]) and y[2] == 1.0
