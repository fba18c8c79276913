"""Python view of the pipeline C ABI (include/fzp.h).

A thin ctypes layer over ``fz::Pipeline`` for tests and the benchmark.  The
stage names, port names and defaults are the reference's
(include/pipeline/compressor.h, the stage headers under modules/); all work
runs in libfzh.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._fzh import DeviceBuffer, FzhError, call, lib

ABS, REL, NOA = 0, 1, 2
_MODES = {"ABS": ABS, "REL": REL, "NOA": NOA}


_DTYPES = {"u8": 0, "u16": 1, "u32": 2, "u64": 3, "i8": 4, "i16": 5, "i32": 6, "i64": 7, "f32": 8, "f64": 9}


def dtype_code(d) -> int:
    """FZM DataType code of a type name ('u8' .. 'f64') or numpy dtype."""
    if isinstance(d, int):
        return d
    if isinstance(d, str) and d in _DTYPES:
        return _DTYPES[d]
    k = np.dtype(d)
    return _DTYPES[{"b": "i", "i": "i", "u": "u", "f": "f"}[k.kind] + str(8 * k.itemsize)]


def _mode(m) -> int:
    return _MODES[m.upper()] if isinstance(m, str) else int(m)


def _ptr(x) -> int:
    if isinstance(x, DeviceBuffer):
        return x.ptr
    return int(x)


class Pipeline:
    """fz::Pipeline through the C ABI; methods raise FzhError on failure."""

    def __init__(self, input_size_hint: int = 0, pool_multiplier: float = 3.0):
        h = C.c_void_p()
        call("fzp_create", int(input_size_hint), float(pool_multiplier), C.byref(h))
        self.h = h.value

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().fzp_destroy(self.h)
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

    # -- stages -------------------------------------------------------------
    def _id(self, fn, *args) -> int:
        i = C.c_int(-1)
        call(fn, self.h, *args, C.byref(i))
        return i.value

    def add_lorenzo(self, error_bound, mode="ABS", quant_radius=32768, outlier_capacity=0.2, zigzag=False,
                    code_bytes=2, value_base=0.0, dtype="f32") -> int:
        """LorenzoQuantStage<float, u8|u16> (dtype "f32", code_bytes 1 or 2) or
        LorenzoQuantStage<double, u16|u32> (dtype "f64", code_bytes 2 or 4)."""
        fn = {"f32": "fzp_add_lorenzo", "f64": "fzp_add_lorenzo_f64"}[dtype]
        return self._id(fn, float(error_bound), _mode(mode), int(quant_radius),
                        float(outlier_capacity), int(bool(zigzag)), int(code_bytes), float(value_base))

    def set_strict(self, stage_id: int, on: bool = True, patch_capacity: float = 0.0) -> None:
        """Strict error bound on a LorenzoQuant stage (extension)."""
        call("fzp_set_lorenzo_strict", self.h, stage_id, int(bool(on)), float(patch_capacity))

    def add_quantizer(self, error_bound, mode="ABS", quant_radius=32768, outlier_capacity=0.05, zigzag=False,
                      code_bytes=2, value_base=0.0) -> int:
        return self._id("fzp_add_quantizer", float(error_bound), _mode(mode), int(quant_radius),
                        float(outlier_capacity), int(bool(zigzag)), int(code_bytes), float(value_base))

    def set_quantizer_outliers(self, stage_id: int, threshold: float = float("inf"), inplace: bool = False) -> None:
        """QuantizerStage::setOutlierThreshold / setInplaceOutliers."""
        call("fzp_set_quantizer_outliers", self.h, stage_id, float(threshold), int(bool(inplace)))

    def add_bitshuffle(self, block_size=16384, element_width=2) -> int:
        return self._id("fzp_add_bitshuffle", int(block_size), int(element_width))

    def add_rze(self, chunk_size=16384, levels=4) -> int:
        return self._id("fzp_add_rze", int(chunk_size), int(levels))

    def add_rle(self, elem_bytes=2) -> int:
        return self._id("fzp_add_rle", int(elem_bytes))

    def add_huffman(self, elem_bytes=2, bklen=0) -> int:
        return self._id("fzp_add_huffman", int(elem_bytes), int(bklen))

    def add_bitpack(self, elem_bytes=2, nbits=0) -> int:
        return self._id("fzp_add_bitpack", int(elem_bytes), int(nbits))

    def set_huffman_fine(self, stage_id: int, on: bool = True) -> None:
        call("fzp_set_huffman_fine", self.h, int(stage_id), int(bool(on)))

    def set_huffman_device_book(self, stage_id: int, on: bool = True) -> None:
        """Build the Huffman codebook on the device (coarse encoding)."""
        call("fzp_set_huffman_device_book", self.h, int(stage_id), int(bool(on)))

    def add_difference(self, dtype="f32", negabinary=False, chunk_bytes=0) -> int:
        return self._id("fzp_add_difference", dtype_code(dtype), int(bool(negabinary)), int(chunk_bytes))

    def add_zigzag(self, elem_bytes=4) -> int:
        return self._id("fzp_add_zigzag", int(elem_bytes))

    def add_negabinary(self, elem_bytes=4) -> int:
        return self._id("fzp_add_negabinary", int(elem_bytes))

    def add_lorenzo_int(self, elem_bytes=4) -> int:
        return self._id("fzp_add_lorenzo_int", int(elem_bytes))

    def connect(self, dependent: int, producer: int, output_name: str | None = None) -> None:
        call("fzp_connect", self.h, dependent, producer, output_name.encode() if output_name else None)

    def set_dims(self, x, y=1, z=1) -> None:
        call("fzp_set_dims", self.h, int(x), int(y), int(z))

    def finalize(self) -> None:
        call("fzp_finalize", self.h)

    def enable_graph_mode(self, on=True) -> None:
        call("fzp_enable_graph_mode", self.h, int(bool(on)))

    def set_stage_fusion(self, on=True) -> None:
        call("fzp_set_stage_fusion", self.h, int(bool(on)))

    def enable_profiling(self, on=True) -> None:
        call("fzp_enable_profiling", self.h, int(bool(on)))

    def set_pool_managed_decomp_output(self, on=True) -> None:
        call("fzp_set_pool_managed_decomp_output", self.h, int(bool(on)))

    def max_compressed_size(self, nbytes: int) -> int:
        r = C.c_size_t()
        call("fzp_max_compressed_size", self.h, int(nbytes), C.byref(r))
        return r.value

    # -- execution ----------------------------------------------------------
    def compress(self, d_in, nbytes: int, stream=None):
        """(device pointer, size) of the pool-owned result."""
        p, n = C.c_void_p(), C.c_size_t()
        call("fzp_compress", self.h, stream, _ptr(d_in), int(nbytes), C.byref(p), C.byref(n))
        return p.value, n.value

    def decompress(self, d_in, nbytes: int, stream=None):
        """(device pointer, size); caller-owned unless pool-managed output is on."""
        p, n = C.c_void_p(), C.c_size_t()
        call("fzp_decompress", self.h, stream, _ptr(d_in), int(nbytes), C.byref(p), C.byref(n))
        return p.value, n.value

    def compress_into(self, d_in, nbytes, d_out, cap, stream=None) -> int:
        n = C.c_size_t()
        call("fzp_compress_into", self.h, stream, _ptr(d_in), int(nbytes), _ptr(d_out), int(cap), C.byref(n))
        return n.value

    def decompress_into(self, d_in, nbytes, d_out, cap, stream=None) -> int:
        n = C.c_size_t()
        call("fzp_decompress_into", self.h, stream, _ptr(d_in), int(nbytes), _ptr(d_out), int(cap), C.byref(n))
        return n.value

    def compress_array(self, x: np.ndarray) -> np.ndarray:
        """Host array in, compressed bytes (host copy) out."""
        d = DeviceBuffer.from_array(x)
        p, n = self.compress(d, x.nbytes)
        out = np.empty(n, np.uint8)
        if n:
            call("fzh_memcpy_dtoh", out.ctypes.data, p, n)
        d.free()
        return out

    def decompress_array(self, blob: np.ndarray, dtype=np.float32) -> np.ndarray:
        d = DeviceBuffer.from_array(blob)
        p, n = self.decompress(d, blob.nbytes)
        out = np.empty(n // np.dtype(dtype).itemsize, dtype)
        if n:
            call("fzh_memcpy_dtoh", out.ctypes.data, p, out.nbytes)
        lib().fzh_device_free(p)
        d.free()
        return out

    def write_file(self, path: str) -> None:
        call("fzp_write_file", self.h, str(path).encode())

    @staticmethod
    def decompress_file(path: str, dtype=np.float32) -> np.ndarray:
        p, n = C.c_void_p(), C.c_size_t()
        call("fzp_decompress_file", str(path).encode(), None, C.byref(p), C.byref(n))
        out = np.empty(n.value // np.dtype(dtype).itemsize, dtype)
        try:
            if out.nbytes:
                call("fzh_memcpy_dtoh", out.ctypes.data, p.value, out.nbytes)
        finally:
            lib().fzh_device_free(p.value)
        return out

    # -- introspection ------------------------------------------------------
    def outputs(self):
        n = C.c_int()
        call("fzp_num_outputs", self.h, C.byref(n))
        res = []
        for i in range(n.value):
            name = C.create_string_buffer(64)
            sz = C.c_size_t()
            call("fzp_output_info", self.h, i, name, 64, C.byref(sz))
            res.append((name.value.decode(), sz.value))
        return res

    def abs_error_bound(self, stage_id: int) -> float:
        """The stage's resolved absolute bound in its input precision (a
        float's value for FP32 stages, the double for FP64 stages)."""
        v = C.c_double()
        call("fzp_stage_abs_error_bound_f64", self.h, stage_id, C.byref(v))
        return v.value

    def last_perf(self):
        d, h = C.c_float(), C.c_float()
        i, o = C.c_size_t(), C.c_size_t()
        call("fzp_last_perf", self.h, C.byref(d), C.byref(h), C.byref(i), C.byref(o))
        return {"dag_ms": d.value, "host_ms": h.value, "input_bytes": i.value, "output_bytes": o.value}


__all__ = ["Pipeline", "ABS", "REL", "NOA", "FzhError"]
