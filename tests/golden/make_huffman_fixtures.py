"""Writes tests/golden/huffman_books.npz: codebooks produced by the
reference's own host-side builder (phf_CPU_build_canonized_codebook_v2,
modules/coders/huffman/phf/hf_bk.cc:84-140, compiled from its sources into
oracle/_ref by oracle/Makefile) for a fixed set of seeded histograms.

Run where the reference sources are present:
    make -C oracle && python tests/golden/make_huffman_fixtures.py
The fixture is data only (histograms in, bk4 words and revbook bytes out).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def histograms():
    """(name, freq u32[bklen], elem_bytes) cases covering ties, skew, sparse
    and near-degenerate books."""
    rng = np.random.default_rng(20261202)
    cases = []
    for bklen in (256, 1024):
        g = np.rint(np.abs(rng.normal(0, 30, 200000))).astype(np.int64)
        cases.append((f"gauss{bklen}", np.bincount(np.minimum(g, bklen - 1), minlength=bklen), 2))
        cases.append((f"uniform{bklen}", np.full(bklen, 7), 2))
        z = rng.zipf(1.5, 100000)
        cases.append((f"zipf{bklen}", np.bincount(np.minimum(z, bklen - 1), minlength=bklen), 2))
        t = np.zeros(bklen, np.int64)
        t[rng.choice(bklen, 40, replace=False)] = rng.integers(1, 4, 40)
        cases.append((f"ties{bklen}", t, 2))
    fib = np.zeros(1024, np.int64)
    a, b = 1, 1
    for i in range(30):
        fib[i] = a
        a, b = b, a + b
    cases.append(("fibonacci30", fib, 2))  # codes longer than 27 bits
    one = np.zeros(1024, np.int64)
    one[5] = 1000
    cases.append(("single", one, 2))
    two = np.zeros(1024, np.int64)
    two[[0, 1023]] = [3, 5]
    cases.append(("two", two, 2))
    cases.append(("u8_gauss", np.bincount(np.minimum(np.rint(np.abs(rng.normal(0, 10, 50000))).astype(np.int64), 255),
                                          minlength=256), 1))
    cases.append(("u32_small", np.bincount(rng.integers(0, 64, 5000), minlength=128), 4))
    for k in range(8):
        bk = int(rng.choice([256, 512, 1024, 2048]))
        f = rng.integers(0, 50, bk) * (rng.random(bk) < 0.3)
        cases.append((f"random{k}", f, 2))
    return [(n, np.asarray(f, np.uint32), eb) for n, f, eb in cases]


def main():
    oracle.lib()
    if oracle.ref_phf() is None:
        sys.exit("oracle/_ref/libref_phf.so is not built (needs the reference sources)")
    out = {}
    names = []
    for name, f, eb in histograms():
        bk4, rev = oracle.ref_huffman_book(f, eb)
        out[f"{name}.freq"] = f
        out[f"{name}.bk4"] = bk4
        out[f"{name}.rev"] = rev
        out[f"{name}.eb"] = np.array([eb], np.uint8)
        names.append(name)
    out["names"] = np.array(names)
    path = os.path.join(HERE, "huffman_books.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(names)} codebooks")


if __name__ == "__main__":
    main()
