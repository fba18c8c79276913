"""Python view of the batched compressor C ABI (include/fzb.h).

A thin ctypes layer over ``fz::BatchCompressor``: many fields through one
persistent launch per stage; field i's bytes equal ``Pipeline.compress`` of
that field alone with the same settings.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._fzh import call, lib
from .pipeline import _mode


class FieldC(C.Structure):
    _fields_ = [("d_in", C.c_void_p), ("nx", C.c_uint64), ("ny", C.c_uint64), ("nz", C.c_uint64)]


class ConfigC(C.Structure):
    _fields_ = [
        ("error_bound", C.c_float),
        ("eb_mode", C.c_int),
        ("quant_radius", C.c_uint32),
        ("outlier_capacity", C.c_float),
        ("zigzag", C.c_int),
        ("strict", C.c_int),
        ("patch_capacity", C.c_float),
    ]


class LaunchC(C.Structure):
    _fields_ = [
        ("vb_wg_per_cu", C.c_uint32),
        ("lz_wg_per_cu", C.c_uint32),
        ("lz_tiles_per_unit", C.c_uint32),
        ("pack_wg_per_cu", C.c_uint32),
        ("scatter_wg_per_cu", C.c_uint32),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Batch:
    """fz::BatchCompressor through the C ABI; methods raise FzhError."""

    def __init__(self, error_bound=1e-3, mode="REL", quant_radius=32768, outlier_capacity=0.1, zigzag=True,
                 strict=False, patch_capacity=0.01):
        cfg = ConfigC(float(error_bound), _mode(mode), int(quant_radius), float(outlier_capacity),
                      int(bool(zigzag)), int(bool(strict)), float(patch_capacity))
        h = C.c_void_p()
        call("fzb_create", C.byref(cfg), C.byref(h))
        self.h = h.value
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            lib().fzb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_fields(self, fields):
        """fields: iterable of (device pointer, (nx, ny, nz))."""
        arr = (FieldC * len(fields))()
        for i, (ptr, dims) in enumerate(fields):
            nx, ny, nz = (list(dims) + [1, 1])[:3]
            arr[i] = FieldC(int(ptr), int(nx), int(ny), int(nz))
        call("fzb_set_fields", self.h, arr, len(fields))
        self.n = len(fields)

    def compress(self, stream=None):
        call("fzb_compress", self.h, stream)

    def output(self):
        p, cap = C.c_void_p(), C.c_uint64()
        call("fzb_output", self.h, C.byref(p), C.byref(cap))
        return p.value, cap.value

    def field_result(self, i):
        off, nb = C.c_uint64(), C.c_uint64()
        call("fzb_field_result", self.h, int(i), C.byref(off), C.byref(nb))
        return off.value, nb.value

    def field_bytes(self, i) -> np.ndarray:
        """Host copy of field i's compressed bytes."""
        base, _ = self.output()
        off, nb = self.field_result(i)
        out = np.empty(nb, np.uint8)
        if nb:
            call("fzh_memcpy_dtoh", out.ctypes.data, base + off, nb)
        return out

    def totals(self):
        a, b = C.c_uint64(), C.c_uint64()
        call("fzb_totals", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_launch(self, **kw):
        cur = self.get_launch()
        cur.update({k: int(v) for k, v in kw.items()})
        l = LaunchC(**cur)
        call("fzb_set_launch", self.h, C.byref(l))

    def get_launch(self):
        l = LaunchC()
        call("fzb_get_launch", self.h, C.byref(l))
        return l.as_dict()

    def autotune(self, reps=3, stream=None):
        l = LaunchC()
        call("fzb_autotune", self.h, stream, int(reps), C.byref(l))
        return l.as_dict()

    def enable_graph_mode(self, on=True):
        call("fzb_enable_graph_mode", self.h, int(bool(on)))

    def enable_profiling(self, on=True):
        call("fzb_enable_profiling", self.h, int(bool(on)))

    def stage_ms(self):
        res = []
        for i in range(16):
            name = C.create_string_buffer(32)
            ms = C.c_float()
            if lib().fzb_stage_ms(self.h, i, name, 32, C.byref(ms)) != 0:
                break
            res.append((name.value.decode(), ms.value))
        return res


__all__ = ["Batch"]
