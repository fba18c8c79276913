"""Python view of the learned chain selector (include/fzs.h).

``features``/``label_sizes`` wrap the device kernels; ``Selector`` holds the
MLP (FP32 master weights on the device, BF16 where a layer reads them) and
trains it with the device forward/backward/Adam kernels; ``chain_pipeline``
builds the fz::Pipeline of a candidate chain with the labeller's settings,
which is how a chunk compressed under a recorded choice is decoded.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._fzh import DeviceBuffer, call, lib
from .pipeline import Pipeline

NFEAT, NCHAINS = 32, 4
CHAINS = ("lorenzo-rle", "lorenzo-bitshuffle-rze", "lorenzo-huffman", "quantizer-bitpack")
LZ_RADIUS, HUF_BKLEN, Q_RADIUS, OUTLIER_CAP = 512, 1024, 32768, 0.2


def _dims(shape):
    """(nx, ny, nz) of a chunk shape given as (ny, nx) or (nz, ny, nx)."""
    s = tuple(shape)
    return (s[-1], s[-2] if len(s) > 1 else 1, s[-3] if len(s) > 2 else 1)


def chain_pipeline(k: int, abs_eb: float, shape) -> Pipeline:
    """fz::Pipeline of candidate chain k for chunks of `shape` (fzs.h)."""
    nx, ny, nz = _dims(shape)
    nbytes = 4 * nx * ny * nz
    p = Pipeline(nbytes)
    if k == 3:
        q = p.add_quantizer(abs_eb, mode="ABS", quant_radius=Q_RADIUS, outlier_capacity=OUTLIER_CAP, zigzag=True)
        bp = p.add_bitpack(2, 0)
        p.connect(bp, q, "codes")
    else:
        lz = p.add_lorenzo(abs_eb, mode="ABS", quant_radius=LZ_RADIUS, outlier_capacity=OUTLIER_CAP, zigzag=True)
        if k == 0:
            c = p.add_rle(2)
            p.connect(c, lz, "codes")
        elif k == 1:
            bs = p.add_bitshuffle(16384, 2)
            rz = p.add_rze(16384, 4)
            p.connect(bs, lz, "codes")
            p.connect(rz, bs)
        elif k == 2:
            c = p.add_huffman(2, HUF_BKLEN)
            p.connect(c, lz, "codes")
        else:
            raise ValueError("chain must be 0..3")
    p.set_dims(nx, ny, nz)
    p.finalize()
    return p


def features(d_chunks, n: int, shape, d_eb, stream=None) -> DeviceBuffer:
    nx, ny, nz = _dims(shape)
    out = DeviceBuffer(max(4 * NFEAT * n, 16))
    call("fzs_features_f32", stream, int(d_chunks), n, nx, ny, nz, int(d_eb), out.ptr)
    return out


def label_sizes(d_chunks, n: int, shape, d_eb, stream=None, host_threads=0) -> np.ndarray:
    """Exact fz::Pipeline::compress sizes, shape (n, 4) uint64."""
    nx, ny, nz = _dims(shape)
    ws = DeviceBuffer(max(lib().fzs_label_ws_bytes(n, nx, ny, nz), 256))
    out = np.zeros((n, NCHAINS), np.uint64)
    call("fzs_label_sizes", stream, int(d_chunks), n, nx, ny, nz, int(d_eb), out.ctypes.data, ws.ptr, ws.nbytes,
         int(host_threads))
    ws.free()
    return out


# Base fields for the smooth regime of the device-generated chunks (parents
# of regime 0 are seeded windows of them): k^-4 Gaussian random fields.
SYNTH_BASE = {2: (2048, 2048, 1), 3: (256, 256, 128)}


def synth_base(nd: int, seed: int, stream=None) -> DeviceBuffer:
    bx, by, bz = SYNTH_BASE[nd]
    base = DeviceBuffer(4 * bx * by * bz)
    call("fzh_synth_power_law_f32", stream, base.ptr, bx, by, bz, np.float32(-4.0), seed)
    return base


def synth_chunks(nd: int, n: int, seed: int, base: DeviceBuffer, stream=None, regime=False):
    """n device chunks of nd dimensions from `seed` (fzs_synth_chunks):
    (chunks, bounds[, regimes]) as DeviceBuffers."""
    bx, by, bz = SYNTH_BASE[nd]
    d_c, d_eb = DeviceBuffer(4 * 65536 * n), DeviceBuffer(max(4 * n, 16))
    d_r = DeviceBuffer(max(n, 16)) if regime else None
    call("fzs_synth_chunks", stream, nd, n, seed, base.ptr, bx, by, bz, d_c.ptr, d_eb.ptr, d_r.ptr if d_r else None)
    return (d_c, d_eb, d_r) if regime else (d_c, d_eb)


def regret_costs(sizes: np.ndarray) -> np.ndarray:
    s = sizes.astype(np.float64)
    return (s / s.min(axis=1, keepdims=True) - 1.0).astype(np.float32)


class Selector:
    """The MLP; parameters live on the device (FP32 master copy)."""

    def __init__(self, seed: int = 0):
        self.P = int(lib().fzs_mlp_param_count())
        rng = np.random.default_rng(seed)
        p = np.zeros(self.P, np.float32)
        o = 0
        for fan_in, fan_out in ((NFEAT, 64), (64, 64), (64, NCHAINS)):
            lim = np.sqrt(6.0 / fan_in)
            p[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
            o += fan_in * fan_out + fan_out
        self.params = DeviceBuffer.from_array(p)
        self.norm = DeviceBuffer.from_array(np.concatenate([np.zeros(NFEAT, np.float32), np.ones(NFEAT, np.float32)]))
        self.m = DeviceBuffer.zeros(4 * self.P)
        self.v = DeviceBuffer.zeros(4 * self.P)
        self.grad = DeviceBuffer(4 * self.P)
        self.loss = DeviceBuffer(16)
        self.step = 0
        self._ws = None

    def set_params(self, p: np.ndarray) -> None:
        self.params.upload(np.ascontiguousarray(p, np.float32))

    def get_params(self) -> np.ndarray:
        return self.params.download(np.float32, self.P)

    def set_norm_from(self, feat: np.ndarray) -> None:
        mu = feat.mean(axis=0).astype(np.float32)
        sd = feat.std(axis=0).astype(np.float32)
        sc = np.where(sd > 1e-6, 1.0 / np.maximum(sd, 1e-6), 1.0).astype(np.float32)
        self.norm.upload(np.concatenate([mu, sc]).astype(np.float32))

    def get_norm(self) -> np.ndarray:
        return self.norm.download(np.float32, 2 * NFEAT)

    def forward(self, d_feat, n: int, stream=None) -> np.ndarray:
        out = DeviceBuffer(max(16 * n, 16))
        call("fzs_mlp_forward", stream, self.params.ptr, self.norm.ptr, int(d_feat), n, out.ptr)
        z = out.download(np.float32, n * NCHAINS).reshape(n, NCHAINS)
        out.free()
        return z

    def select(self, d_feat, n: int, stream=None) -> np.ndarray:
        out = DeviceBuffer(max(16 * n, 16))
        ch = DeviceBuffer(max(n, 16))
        call("fzs_mlp_forward", stream, self.params.ptr, self.norm.ptr, int(d_feat), n, out.ptr)
        call("fzs_select", stream, out.ptr, n, ch.ptr)
        r = ch.download(np.uint8, n)
        out.free()
        ch.free()
        return r

    def train_step(self, d_feat, d_cost, n: int, tau: float, stream=None) -> float:
        """Gradient of one batch (into self.grad); returns the mean loss."""
        need = lib().fzs_mlp_train_ws_bytes(n)
        if self._ws is None or self._ws.nbytes < need:
            if self._ws is not None:
                self._ws.free()
            self._ws = DeviceBuffer(need)
        call("fzs_mlp_train_step", stream, self.params.ptr, self.norm.ptr, int(d_feat), int(d_cost), n,
             np.float32(tau), self.grad.ptr, self.loss.ptr, self._ws.ptr, self._ws.nbytes)
        return float(self.loss.download(np.float32, 1)[0])

    def adam(self, lr: float, b1=0.9, b2=0.999, eps=1e-8, wd=0.0, stream=None) -> None:
        self.step += 1
        call("fzs_mlp_adam", stream, self.params.ptr, self.grad.ptr, self.m.ptr, self.v.ptr, self.step,
             np.float32(lr), np.float32(b1), np.float32(b2), np.float32(eps), np.float32(wd))

    def fit(self, feat: np.ndarray, cost: np.ndarray, epochs=20, batch=1024, lr=3e-3, tau=0.02, seed=0,
            wd=1e-4, stream=None) -> list:
        """Mini-batch training on the device; returns the mean loss per epoch.
        Each epoch uploads one shuffled copy of the set; batches are slices
        of it and each step's loss lands in a device array read once per
        epoch (no host round trip per step)."""
        n = feat.shape[0]
        self.set_norm_from(feat)
        rng = np.random.default_rng(seed)
        nb = max(1, n // batch) if n >= batch else 1
        bsz = batch if n >= batch else n
        d_feat = DeviceBuffer(4 * NFEAT * nb * bsz)
        d_cost = DeviceBuffer(4 * NCHAINS * nb * bsz)
        d_loss = DeviceBuffer(max(4 * nb, 16))
        need = lib().fzs_mlp_train_ws_bytes(bsz)
        if self._ws is None or self._ws.nbytes < need:
            if self._ws is not None:
                self._ws.free()
            self._ws = DeviceBuffer(need)
        hist = []
        for _ in range(epochs):
            perm = rng.permutation(n)[: nb * bsz]
            d_feat.upload(np.ascontiguousarray(feat[perm], np.float32))
            d_cost.upload(np.ascontiguousarray(cost[perm], np.float32))
            for b in range(nb):
                call("fzs_mlp_train_step", stream, self.params.ptr, self.norm.ptr, d_feat.ptr + 4 * NFEAT * bsz * b,
                     d_cost.ptr + 4 * NCHAINS * bsz * b, bsz, np.float32(tau), self.grad.ptr, d_loss.ptr + 4 * b,
                     self._ws.ptr, self._ws.nbytes)
                self.adam(lr, wd=wd, stream=stream)
            hist.append(float(d_loss.download(np.float32, nb).astype(np.float64).mean()))
        d_feat.free()
        d_cost.free()
        d_loss.free()
        return hist

    def close(self):
        for b in (self.params, self.norm, self.m, self.v, self.grad, self.loss, self._ws):
            if b is not None:
                b.free()


LEARNED, EXHAUSTIVE, FIXED = 0, 1, 2


class SelectingCompressor:
    """fzs_compress_field / fzs_decompress_field: a field cut into chunks of
    `chunk` ((ny, nx) or (nz, ny, nx)), each compressed with its chain."""

    def __init__(self, chunk):
        self.chunk = tuple(chunk)
        nx, ny, nz = _dims(chunk)
        h = C.c_void_p()
        call("fzs_create", nx, ny, nz, C.byref(h))
        self.h = h.value

    def set_model(self, params: np.ndarray, norm: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        q = np.ascontiguousarray(norm, np.float32)
        call("fzs_set_model", self.h, p.ctypes.data, q.ctypes.data)

    def compress(self, d_field, shape, rel_eb=1e-3, mode=LEARNED, chain=0, stream=None) -> np.ndarray:
        nx, ny, nz = _dims(shape)
        p, n = C.c_void_p(), C.c_size_t()
        call("fzs_compress_field", self.h, stream, int(d_field), nx, ny, nz, np.float32(rel_eb), int(mode),
             int(chain), C.byref(p), C.byref(n))
        return np.ctypeslib.as_array((C.c_uint8 * n.value).from_address(p.value)).copy() if n.value else \
            np.zeros(0, np.uint8)

    def decompress(self, blob: np.ndarray, d_out, n_elems: int, stream=None) -> None:
        b = np.ascontiguousarray(blob, np.uint8)
        call("fzs_decompress_field", self.h, stream, b.ctypes.data, b.size, int(d_out), int(n_elems))

    def choices(self) -> np.ndarray:
        p, n = C.c_void_p(), C.c_size_t()
        call("fzs_last_choices", self.h, C.byref(p), C.byref(n))
        if not n.value:
            return np.zeros(0, np.uint8)
        return np.ctypeslib.as_array((C.c_uint8 * n.value).from_address(p.value)).copy()

    def timing(self):
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        call("fzs_last_timing", self.h, C.byref(a), C.byref(b), C.byref(c))
        return {"select_ms": a.value, "compress_ms": b.value, "total_ms": c.value}

    def close(self):
        if getattr(self, "h", None):
            lib().fzs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Selector", "SelectingCompressor", "features", "label_sizes", "chain_pipeline", "regret_costs", "CHAINS",
           "LEARNED", "EXHAUSTIVE", "FIXED"]
