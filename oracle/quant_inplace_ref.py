"""CPU restatement of the quantizer's in-place outlier mode (numpy).

TEST INFRASTRUCTURE: this module is a checker for the HIP path.  Only
tests/ may import it.

  fwd  modules/quantizers/quantizer/quantizer.cu:83-104
       q = __float2int_rn(x * ebx2_r) (round half to even; saturating; NaN
       -> 0); |q| < radius and |x| < threshold -> code = (q << 1) ^ (q >> 31)
       as u32; otherwise code = the raw IEEE-754 bits of x.
  inv  quantizer.cu:147-168
       (code >> 1) >= radius -> the raw bits as float; otherwise
       q = c odd ? -((c + 1) >> 1) : c >> 1 and x = (float)q * ebx2 (FP32).
"""
from __future__ import annotations

import numpy as np


def _f2i_rn(p: np.ndarray) -> np.ndarray:
    p = p.astype(np.float64)
    q = np.rint(np.where(np.isnan(p), 0.0, p))
    return np.clip(q, -2147483648.0, 2147483647.0).astype(np.int64)


def fwd(x: np.ndarray, ebx2_r: float, radius: int, threshold: float = np.inf) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    prod = x * np.float32(ebx2_r)  # FP32 product, as the kernel forms it
    q = _f2i_rn(prod)
    ok = (q > -radius) & (q < radius) & (np.abs(x) < np.float32(threshold))
    zz = ((q << 1) ^ (q >> 63)) & 0xFFFFFFFF
    return np.where(ok, zz, x.view(np.uint32).astype(np.int64)).astype(np.uint32)


def inv(codes: np.ndarray, ebx2: float, radius: int) -> np.ndarray:
    c = np.ascontiguousarray(codes, np.uint32)
    out_raw = c.view(np.float32)
    ci = c.astype(np.int64)
    q = np.where(ci & 1, -((ci + 1) >> 1), ci >> 1)
    val = q.astype(np.float32) * np.float32(ebx2)
    return np.where((c >> 1) >= radius, out_raw, val).astype(np.float32)
