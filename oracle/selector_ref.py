"""CPU reference of the chain selector's features and MLP (numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py as the checker of the
selector kernels (fzgpumodules_amd/csrc/select.hip); the product never
imports it.  The selector has no counterpart in the reference, so this file
restates the definitions the kernels implement (feature list and MLP in
include/fzs.h / select.hip) in plain numpy.

MLP arithmetic modes:
  * ``exact32`` -- what the kernels do: every layer input rounded to BF16
    (round to nearest even), products of two BF16 values (exact in FP32),
    FP32 sums in index order, propagated deltas rounded to BF16;
  * ``f64`` -- the same network in float64 without rounding, for the
    finite-difference check of the backward formulas.
"""
from __future__ import annotations

import numpy as np

NF, H1, H2, NC = 32, 64, 64, 4
OW1 = 0
OB1 = OW1 + H1 * NF
OW2 = OB1 + H1
OB2 = OW2 + H2 * H1
OW3 = OB2 + H2
OB3 = OW3 + NC * H2
NP = OB3 + NC


def bf16r(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    is_nan = np.isnan(x)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    out[np.isinf(x)] = x[np.isinf(x)]
    out[is_nan] = np.float32("nan")
    return out


def unpack(params):
    p = np.asarray(params)
    return (p[OW1:OB1].reshape(H1, NF), p[OB1:OW2], p[OW2:OB2].reshape(H2, H1), p[OB2:OW3],
            p[OW3:OB3].reshape(NC, H2), p[OB3:NP])


def _layer32(W, b, x):
    """FP32 sums in index order: acc = b; acc = fl(acc + W[:, i] * x[:, i])."""
    acc = np.broadcast_to(b.astype(np.float32), (x.shape[0], W.shape[0])).copy()
    for i in range(W.shape[1]):
        acc = (acc + W[:, i][None, :] * x[:, i][:, None]).astype(np.float32)
    return acc


def forward(params, norm, feat, mode="exact32"):
    """Returns (logits, cache)."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    shift, scale = np.asarray(norm[:NF]), np.asarray(norm[NF:])
    if mode == "f64":
        x = (feat.astype(np.float64) - shift) * scale
        a1 = x @ W1.T.astype(np.float64) + b1
        h1 = np.maximum(a1, 0.0)
        a2 = h1 @ W2.T.astype(np.float64) + b2
        h2 = np.maximum(a2, 0.0)
        z = h2 @ W3.T.astype(np.float64) + b3
        return z, dict(x=x, h1=h1, h2=h2, W2=W2.astype(np.float64), W3=W3.astype(np.float64))
    f32 = np.float32
    x = bf16r(((feat.astype(f32) - shift.astype(f32)).astype(f32) * scale.astype(f32)).astype(f32))
    W1b, W2b, W3b = bf16r(W1.astype(f32)), bf16r(W2.astype(f32)), bf16r(W3.astype(f32))
    a1 = _layer32(W1b, b1.astype(f32), x)
    h1 = bf16r(np.maximum(a1, f32(0)))
    a2 = _layer32(W2b, b2.astype(f32), h1)
    h2 = bf16r(np.maximum(a2, f32(0)))
    z = _layer32(W3b, b3.astype(f32), h2)
    return z, dict(x=x, h1=h1, h2=h2, W2=W2b, W3=W3b)


def soft_target(cost, tau, dtype):
    nc = -np.asarray(cost, dtype) / dtype(tau)
    nc = nc - nc.max(axis=1, keepdims=True)
    e = np.exp(nc)
    return e / e.sum(axis=1, keepdims=True)


def loss_and_grad(params, norm, feat, cost, tau, mode="exact32"):
    """Mean soft-target cross entropy and its gradient w.r.t. the parameters."""
    dt = np.float64 if mode == "f64" else np.float32
    z, c = forward(params, norm, feat, mode)
    z = z.astype(dt)
    zm = z - z.max(axis=1, keepdims=True)
    e = np.exp(zm)
    ps = e.sum(axis=1, keepdims=True)
    p = e / ps
    t = soft_target(cost, tau, dt)
    logp = zm - np.log(ps)
    loss = float(np.mean(-(t * logp).sum(axis=1)))
    n = feat.shape[0]
    d3 = (p - t).astype(dt)
    rnd = (lambda a: a) if mode == "f64" else bf16r
    d3 = rnd(d3)
    g2 = d3 @ c["W3"].astype(dt) if mode == "f64" else _seq_mat(d3, c["W3"])
    d2 = rnd(np.where(c["h2"] > 0, g2, 0).astype(dt))
    g1 = d2 @ c["W2"].astype(dt) if mode == "f64" else _seq_mat(d2, c["W2"])
    d1 = rnd(np.where(c["h1"] > 0, g1, 0).astype(dt))
    if mode == "f64":
        gW1, gb1 = d1.T @ c["x"] / n, d1.sum(0) / n
        gW2, gb2 = d2.T @ c["h1"] / n, d2.sum(0) / n
        gW3, gb3 = d3.T @ c["h2"] / n, d3.sum(0) / n
    else:
        gW1, gb1 = _seq_outer(d1, c["x"], n), _seq_sum(d1, n)
        gW2, gb2 = _seq_outer(d2, c["h1"], n), _seq_sum(d2, n)
        gW3, gb3 = _seq_outer(d3, c["h2"], n), _seq_sum(d3, n)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3]).astype(dt)
    return loss, g


def _seq_mat(d, W):
    """g[:, j] = sum_k W[k, j] * d[:, k] in FP32, k in order."""
    g = np.zeros((d.shape[0], W.shape[1]), np.float32)
    for k in range(W.shape[0]):
        g = (g + W[k][None, :] * d[:, k][:, None]).astype(np.float32)
    return g


def _seq_outer(d, h, n):
    """sum over samples (in order) of d[s, o] * h[s, k], FP32, then / n."""
    acc = np.zeros((d.shape[1], h.shape[1]), np.float32)
    for s in range(d.shape[0]):
        acc = (acc + d[s][:, None] * h[s][None, :]).astype(np.float32)
    return (acc / np.float32(n)).astype(np.float32)


def _seq_sum(d, n):
    acc = np.zeros(d.shape[1], np.float32)
    for s in range(d.shape[0]):
        acc = (acc + d[s]).astype(np.float32)
    return (acc / np.float32(n)).astype(np.float32)


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    f = np.float32
    p, g = params.astype(f), grad.astype(f)
    m = (f(b1) * m + (f(1) - f(b1)) * g).astype(f)
    v = (f(b2) * v + (f(1) - f(b2)) * g * g).astype(f)
    c1 = f(1) - f(b1) ** f(step)
    c2 = f(1) - f(b2) ** f(step)
    p = (p - f(lr) * ((m / c1) / (np.sqrt(v / c2) + f(eps)) + f(wd) * p)).astype(f)
    return p, m, v


def features(chunk, abs_eb):
    """The 32 features of one chunk (nz, ny, nx) or (ny, nx), in float64."""
    x = np.asarray(chunk, np.float32)
    if x.ndim == 2:
        x = x[None]
    nz, ny, nx = x.shape
    e = float(np.float32(abs_eb))
    inv2eb = float(np.float32(0.5) / np.float32(abs_eb))
    xd = x.astype(np.float64)
    E = x.size

    def sh(a, dz, dy, dx):
        out = np.zeros_like(a)
        out[dz:, dy:, dx:] = a[: nz - dz, : ny - dy, : nx - dx]
        return out

    x32 = x
    w, n_, wn = sh(x32, 0, 0, 1), sh(x32, 0, 1, 0), sh(x32, 0, 1, 1)
    f32 = np.float32
    if nz > 1:
        b, wb, nb, wnb = sh(x32, 1, 0, 0), sh(x32, 1, 0, 1), sh(x32, 1, 1, 0), sh(x32, 1, 1, 1)
        res = ((((((x32 - w).astype(f32) - n_).astype(f32) - b).astype(f32) + wn).astype(f32) + wb).astype(f32)
               + nb).astype(f32)
        res = (res - wnb).astype(f32)
    else:
        res = ((((x32 - w).astype(f32) - n_).astype(f32)) + wn).astype(f32)
    l = (np.abs(res) * f32(inv2eb)).astype(f32)
    lp = (l + f32(0.5)).astype(f32)
    bins = np.zeros(l.shape, np.int64)
    big = lp >= 1.0
    bins[big] = np.minimum(14, np.floor(np.log2(np.minimum(lp[big], 16383.99).astype(np.float64))).astype(np.int64) + 1)
    bins[lp >= 16384.0] = 14
    hist = np.bincount(bins.ravel(), minlength=15)[:15].astype(np.float64)
    px, py, pz = (nx - 1) * ny * nz, nx * (ny - 1) * nz, nx * ny * (nz - 1)
    dx = np.abs(xd[:, :, 1:] - xd[:, :, :-1])
    dy = np.abs(xd[:, 1:, :] - xd[:, :-1, :])
    dz = np.abs(xd[1:] - xd[:-1]) if nz > 1 else None
    mean = xd.mean()
    var = max((xd * xd).mean() - mean * mean, 0.0)
    f = np.zeros(32)
    f[0] = np.log2(1 + np.sqrt(var) / e)
    f[1] = np.log2(1 + (xd.max() - xd.min()) / e)
    f[2] = np.mean(x == 0)
    f[3] = np.mean(x == np.rint(x))
    f[4] = np.sum(x[:, :, 1:] == x[:, :, :-1]) / px if px else 0
    f[5] = np.log2(1 + dx.sum() / px / e) if px else 0
    f[6] = np.log2(1 + dy.sum() / py / e) if py else 0
    f[7] = np.log2(1 + dz.sum() / pz / e) if (nz > 1 and pz) else 0
    f[8] = np.log2(1 + l.astype(np.float64).sum() / E)
    p = hist / E
    f[9:24] = p
    f[24] = np.sum(dx * inv2eb < 0.5) / px if px else 0
    f[25] = np.log2(1 + np.abs(xd).sum() / E / e)
    f[26] = np.log2(1 + float(np.abs(x).max()) / e)
    f[27] = np.mean(l >= 511.5)
    f[28] = np.mean((np.abs(x) * f32(inv2eb)).astype(f32) >= 32767.5)
    f[29] = 1.0 if nz > 1 else 0.0
    nzp = p[p > 0]
    f[30] = float(-(nzp * np.log2(nzp)).sum())
    f[31] = np.mean(x > 0)
    return f
