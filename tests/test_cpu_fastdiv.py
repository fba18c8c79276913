"""The division-by-invariant method the tile mapping uses on the device
(Granlund and Montgomery, PLDI 1994, fig. 4.1; FastDiv in
fzgpumodules_amd/csrc/fzh_lorenzo_common.h): q = (t + ((n - t) >> s1)) >> s2
with t = mulhi(m, n), m = floor(2^32 (2^l - d) / d) + 1, l = ceil(log2 d),
checked against integer division for every divisor up to 4096 and random
large ones, over the 32-bit range's edge values and random numerators."""
import numpy as np


def params(d):
    l = 0
    while l < 32 and (1 << l) < d:
        l += 1
    m = ((((1 << l) - d) << 32) // d + 1) & 0xFFFFFFFF
    return m, (1 if l else 0), (l - 1 if l else 0)


def fastdiv(n, d):
    m, s1, s2 = params(d)
    n = n.astype(np.uint64)
    t = (n * np.uint64(m)) >> np.uint64(32)
    return ((t + ((n - t) >> np.uint64(s1))) >> np.uint64(s2)).astype(np.uint64)


def test_fastdiv_matches_integer_division():
    rng = np.random.default_rng(0)
    edge = np.array([0, 1, 2, 3, 4095, 4096, 2**24 - 1, 2**24, 2**31 - 1, 2**31, 2**32 - 2, 2**32 - 1], np.uint64)
    divisors = list(range(1, 4097)) + [int(v) for v in rng.integers(4097, 2**32, 200)] + [2**31, 2**31 + 1, 2**32 - 1]
    for d in divisors:
        n = np.concatenate([edge, rng.integers(0, 2**32, 500, dtype=np.uint64),
                            (np.arange(1, 6, dtype=np.uint64) * np.uint64(d)) % np.uint64(2**32),
                            (np.arange(1, 6, dtype=np.uint64) * np.uint64(d) - np.uint64(1)) % np.uint64(2**32)])
        np.testing.assert_array_equal(fastdiv(n, d), n // np.uint64(d), err_msg=f"d={d}")
