"""ctypes binding of libfzh.so (the C ABI declared in include/fzh.h).

This is the Python-side view of the drop-in boundary: every wrapper calls one
``fzh_*`` entry point with plain device pointers.  There is no CPU fallback:
if libfzh.so is missing or no GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libfzh.so")

_lib = None

_P = C.c_void_p
_Z = C.c_size_t
_I = C.c_int
_U32 = C.c_uint32
_F = C.c_float
_D = C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "fzh_last_error": (C.c_char_p, []),
    "fzh_version": (C.c_char_p, []),
    "fzh_workspace_init": (_I, [_P, _P, _Z]),
    "fzh_lorenzo_quant_fwd_ws_bytes": (_Z, [_Z, _Z, _Z, _Z]),
    "fzh_lorenzo_quant_fwd_f32": (_I, [_P, _P, _Z, _Z, _Z, _Z, _F, _P, _U32, _I, _I, _P, _P, _P, _P, _Z, _P, _Z]),
    "fzh_lorenzo_quant_fwd_strict_ws_bytes": (_Z, [_Z, _Z, _Z, _Z]),
    "fzh_lorenzo_quant_fwd_strict_f32": (_I, [_P, _P, _Z, _Z, _Z, _Z, _F, _P, _U32, _I, _I, _P, _P, _P, _P, _Z,
                                             _F, _F, _P, _P, _P, _P, _Z, _P, _Z]),
    "fzh_lorenzo_apply_patches": (_I, [_P, _P, _P, _P, _P, _Z, _Z]),
    "fzh_lorenzo_quant_inv_ws_bytes": (_Z, [_Z, _Z, _Z, _Z, _Z]),
    "fzh_lorenzo_quant_inv_f32": (_I, [_P, _P, _Z, _Z, _Z, _Z, _P, _P, _P, _Z, _F, _I, _I, _P, _P, _Z]),
    "fzh_lorenzo_quant_inv_bin": (_I, [_P, _Z, _Z, _Z, _Z, _P, _P, _P, _Z, _P, _Z]),
    "fzh_lorenzo_quant_inv_strict_f32": (_I, [_P, _P, _Z, _Z, _Z, _Z, _P, _P, _P, _Z, _F, _I, _I, _P, _P, _P, _Z,
                                              _P, _P, _Z]),
    "fzh_lorenzo_quant_inv_apply_f32": (_I, [_P, _P, _Z, _Z, _Z, _Z, _P, _P, _P, _Z, _F, _I, _I, _P, _P, _Z]),
    "fzh_value_base_ws_bytes": (_Z, [_Z]),
    "fzh_value_base_f32": (_I, [_P, _P, _Z, _I, _F, _P, _P, _Z]),
    "fzh_value_base_f32_noalloc": (_I, [_P, _P, _Z, _I, _F, _P]),
    "fzh_quant_rel_params": (None, [_F, _P]),
    "fzh_quant_rel_fwd_f32": (_I, [_P, _P, _Z, _F, _F, _U32, _I, _P, _P, _P, _P, _Z, _P, _Z]),
    "fzh_quant_rel_inv_f32": (_I, [_P, _P, _Z, _F, _I, _P, _P, _P, _Z, _P]),
    "fzh_lorenzo_quant_fwd_f64_ws_bytes": (_Z, [_Z, _Z, _Z, _Z]),
    "fzh_lorenzo_quant_fwd_f64": (_I, [_P, _P, _Z, _Z, _Z, _Z, _D, _P, _U32, _I, _I, _P, _P, _P, _P, _Z, _P, _Z]),
    "fzh_lorenzo_quant_inv_f64_ws_bytes": (_Z, [_Z, _Z, _Z, _Z, _Z]),
    "fzh_lorenzo_quant_inv_f64": (_I, [_P, _P, _Z, _Z, _Z, _Z, _P, _P, _P, _Z, _D, _I, _I, _P, _P, _Z]),
    "fzh_value_base_f64_ws_bytes": (_Z, [_Z]),
    "fzh_value_base_f64": (_I, [_P, _P, _Z, _I, _F, _P, _P, _Z]),
    "fzh_bitshuffle_enc": (_I, [_P, _P, _P, _Z, _U32, _I]),
    "fzh_bitshuffle_dec": (_I, [_P, _P, _P, _Z, _U32, _I]),
    "fzh_rze_enc_bound": (_Z, [_Z]),
    "fzh_rze_enc_ws_bytes": (_Z, [_Z]),
    "fzh_rze_enc": (_I, [_P, _P, _Z, _P, _Z, _P, _P, _Z]),
    "fzh_rze_enc_shuffled": (_I, [_P, _P, _Z, _I, _P, _Z, _P, _P, _Z]),
    "fzh_rze_dec_shuffled": (_I, [_P, _P, _Z, _P, _Z, _I, _P, _P, _Z]),
    "fzh_rze_dec_ws_bytes": (_Z, [_Z]),
    "fzh_rze_dec": (_I, [_P, _P, _Z, _P, _Z, _P, _P, _Z]),
    "fzh_rle_enc_bound": (_Z, [_Z, _I]),
    "fzh_rle_enc_ws_bytes": (_Z, [_Z]),
    "fzh_rle_enc": (_I, [_P, _P, _Z, _I, _P, _Z, _P, _P, _Z]),
    "fzh_rle_dec_ws_bytes": (_Z, [_Z]),
    "fzh_rle_dec": (_I, [_P, _P, _Z, _I, _P, _Z, _P, _P, _Z]),
    "fzh_histogram": (_I, [_P, _P, _Z, _I, _U32, _P]),
    "fzh_huffman_revbook_bytes": (_Z, [_U32, _I]),
    "fzh_huffman_build_book": (_I, [_P, _U32, _I, _P, _P]),
    "fzh_huffman_build_book_device": (_I, [_P, _P, _U32, _I, C.c_uint64, _P, _P, _P]),
    "fzh_huffman_enc_bound": (_Z, [_Z, _U32, _I]),
    "fzh_huffman_enc_ws_bytes": (_Z, [_Z]),
    "fzh_huffman_encode": (_I, [_P, _P, _Z, _I, _U32, _P, _P, _P, _Z, _P, _P, _Z]),
    "fzh_huffman_encode_fine": (_I, [_P, _P, _Z, _I, _U32, _P, _P, _P, _Z, _P, _P, _Z]),
    "fzh_huffman_decode": (_I, [_P, _P, _Z, _I, _P, _Z, _P]),
    "fzh_quant_abs_fwd_f32": (_I, [_P, _P, _Z, _F, _F, _U32, _I, _I, _P, _P, _P, _P, _Z, _P, _Z]),
    "fzh_quant_abs_fwd_inplace_f32": (_I, [_P, _P, _Z, _F, _F, _U32, _P]),
    "fzh_quant_abs_inv_inplace_f32": (_I, [_P, _P, _Z, _F, _U32, _P]),
    "fzh_quant_abs_fwd_ws_bytes": (_Z, [_Z]),
    "fzh_quant_abs_inv_f32": (_I, [_P, _P, _Z, _F, _I, _I, _P, _P, _P, _Z, _P]),
    "fzh_max_u32_ws_bytes": (_Z, []),
    "fzh_max_u32": (_I, [_P, _P, _Z, _I, _P, _P, _Z]),
    "fzh_bitpack_enc": (_I, [_P, _P, _Z, _I, _I, _P]),
    "fzh_bitpack_dec": (_I, [_P, _P, _Z, _I, _I, _P]),
    "fzh_gather": (_I, [_P, _P, _I]),
    "fzh_copy": (_I, [_P, _P, _P, _Z]),
    "fzh_concat_ws_bytes": (_Z, []),
    "fzh_concat_bound": (C.c_uint64, [_P, _I]),
    "fzh_concat": (_I, [_P, _P, _I, _P, C.c_uint64, _P, _P, _Z]),
    "fzh_readback": (_I, [_P, _P, _I, _P]),
    "fzh_concat_readback": (_I, [_P, _P, _I, _P, C.c_uint64, _P, _P, _I, _P]),
    # pipeline C ABI (include/fzp.h)
    "fzp_create": (_I, [_Z, _F, C.POINTER(_P)]),
    "fzp_destroy": (_I, [_P]),
    "fzp_add_lorenzo": (_I, [_P, _F, _I, _U32, _F, _I, _I, _F, C.POINTER(_I)]),
    "fzp_set_lorenzo_strict": (_I, [_P, _I, _I, _F]),
    "fzp_add_lorenzo_f64": (_I, [_P, _F, _I, _U32, _F, _I, _I, _F, C.POINTER(_I)]),
    "fzp_stage_abs_error_bound_f64": (_I, [_P, _I, C.POINTER(_D)]),
    "fzp_add_bitshuffle": (_I, [_P, _U32, _I, C.POINTER(_I)]),
    "fzp_add_rze": (_I, [_P, _U32, _I, C.POINTER(_I)]),
    "fzp_add_rle": (_I, [_P, _I, C.POINTER(_I)]),
    "fzp_add_huffman": (_I, [_P, _I, _U32, C.POINTER(_I)]),
    "fzp_add_quantizer": (_I, [_P, _F, _I, _U32, _F, _I, _I, _F, C.POINTER(_I)]),
    "fzp_set_quantizer_outliers": (_I, [_P, _I, _F, _I]),
    "fzp_add_bitpack": (_I, [_P, _I, _I, C.POINTER(_I)]),
    "fzp_connect": (_I, [_P, _I, _I, C.c_char_p]),
    "fzp_set_dims": (_I, [_P, _Z, _Z, _Z]),
    "fzp_finalize": (_I, [_P]),
    "fzp_enable_graph_mode": (_I, [_P, _I]),
    "fzp_enable_profiling": (_I, [_P, _I]),
    "fzp_set_stage_fusion": (_I, [_P, _I]),
    "fzp_set_pool_managed_decomp_output": (_I, [_P, _I]),
    "fzp_max_compressed_size": (_I, [_P, _Z, C.POINTER(_Z)]),
    "fzp_compress": (_I, [_P, _P, _P, _Z, C.POINTER(_P), C.POINTER(_Z)]),
    "fzp_decompress": (_I, [_P, _P, _P, _Z, C.POINTER(_P), C.POINTER(_Z)]),
    "fzp_compress_into": (_I, [_P, _P, _P, _Z, _P, _Z, C.POINTER(_Z)]),
    "fzp_decompress_into": (_I, [_P, _P, _P, _Z, _P, _Z, C.POINTER(_Z)]),
    "fzp_write_file": (_I, [_P, C.c_char_p]),
    "fzp_decompress_file": (_I, [C.c_char_p, _P, C.POINTER(_P), C.POINTER(_Z)]),
    "fzp_last_perf": (_I, [_P, C.POINTER(_F), C.POINTER(_F), C.POINTER(_Z), C.POINTER(_Z)]),
    "fzp_last_stage_perf": (_I, [_P, _I, C.c_char_p, _Z, C.POINTER(_F)]),
    "fzp_num_outputs": (_I, [_P, C.POINTER(_I)]),
    "fzp_output_info": (_I, [_P, _I, C.c_char_p, _Z, C.POINTER(_Z)]),
    "fzp_stage_abs_error_bound": (_I, [_P, _I, C.POINTER(_F)]),
    "fzb_create": (_I, [_P, C.POINTER(_P)]),
    "fzb_destroy": (_I, [_P]),
    "fzb_set_fields": (_I, [_P, _P, _Z]),
    "fzb_compress": (_I, [_P, _P]),
    "fzb_output": (_I, [_P, C.POINTER(_P), C.POINTER(C.c_uint64)]),
    "fzb_field_result": (_I, [_P, _Z, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "fzb_totals": (_I, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "fzb_set_launch": (_I, [_P, _P]),
    "fzb_get_launch": (_I, [_P, _P]),
    "fzb_autotune": (_I, [_P, _P, _I, _P]),
    "fzb_enable_graph_mode": (_I, [_P, _I]),
    "fzb_enable_profiling": (_I, [_P, _I]),
    "fzb_stage_ms": (_I, [_P, _I, C.c_char_p, _Z, C.POINTER(_F)]),
    "fzs_features_f32": (_I, [_P, _P, _Z, _U32, _U32, _U32, _P, _P]),
    "fzs_mlp_param_count": (_Z, []),
    "fzs_mlp_forward": (_I, [_P, _P, _P, _P, _Z, _P]),
    "fzs_mlp_train_ws_bytes": (_Z, [_Z]),
    "fzs_mlp_train_step": (_I, [_P, _P, _P, _P, _P, _Z, _F, _P, _P, _P, _Z]),
    "fzs_mlp_adam": (_I, [_P, _P, _P, _P, _P, _U32, _F, _F, _F, _F, _F]),
    "fzs_label_ws_bytes": (_Z, [_Z, _U32, _U32, _U32]),
    "fzs_label_sizes": (_I, [_P, _P, _Z, _U32, _U32, _U32, _P, _P, _P, _Z, _I]),
    "fzs_select": (_I, [_P, _P, _Z, _P]),
    "fzs_synth_chunks": (_I, [_P, _I, C.c_uint64, C.c_uint64, _P, _U32, _U32, _U32, _P, _P, _P]),
    "fzs_gather_chunks": (_I, [_P, _P, C.c_uint64, C.c_uint64, C.c_uint64, _U32, _U32, _U32, _P]),
    "fzs_scatter_chunks": (_I, [_P, _P, C.c_uint64, C.c_uint64, C.c_uint64, _U32, _U32, _U32, _P]),
    "fzs_create": (_I, [_U32, _U32, _U32, C.POINTER(_P)]),
    "fzs_destroy": (_I, [_P]),
    "fzs_set_model": (_I, [_P, _P, _P]),
    "fzs_compress_field": (_I, [_P, _P, _P, C.c_uint64, C.c_uint64, C.c_uint64, _F, _I, _I, C.POINTER(_P),
                                C.POINTER(_Z)]),
    "fzs_decompress_field": (_I, [_P, _P, _P, _Z, _P, _Z]),
    "fzs_last_choices": (_I, [_P, C.POINTER(_P), C.POINTER(_Z)]),
    "fzs_last_timing": (_I, [_P, C.POINTER(_F), C.POINTER(_F), C.POINTER(_F)]),
    "fzh_kernel_timing_enable": (_I, [C.c_int]),
    "fzh_kernel_timing_get": (_I, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "fzp_add_difference": (_I, [_P, _I, _I, _Z, C.POINTER(_I)]),
    "fzp_set_huffman_fine": (_I, [_P, _I, _I]),
    "fzp_set_huffman_device_book": (_I, [_P, _I, _I]),
    "fzp_add_zigzag": (_I, [_P, _I, C.POINTER(_I)]),
    "fzp_add_negabinary": (_I, [_P, _I, C.POINTER(_I)]),
    "fzp_add_lorenzo_int": (_I, [_P, _I, C.POINTER(_I)]),
    "fzh_zigzag": (_I, [_P, _P, _P, _Z, _I, _I]),
    "fzh_negabinary": (_I, [_P, _P, _P, _Z, _I, _I]),
    "fzh_diff_fwd": (_I, [_P, _P, _P, _Z, _I, _Z, _I]),
    "fzh_diff_inv_ws_bytes": (_Z, [_Z, _I, _Z]),
    "fzh_diff_inv": (_I, [_P, _P, _P, _Z, _I, _Z, _I, _P, _Z]),
    "fzh_lorenzo_int_fwd": (_I, [_P, _P, _P, _Z, _Z, _Z, _I]),
    "fzh_lorenzo_int_inv_ws_bytes": (_Z, [_Z, _Z, _Z]),
    "fzh_lorenzo_int_inv": (_I, [_P, _P, _P, _Z, _Z, _Z, _I, _P, _Z]),
    "fzh_scan_ws_bytes": (C.c_size_t, [C.c_uint64]),
    "fzh_scan_u32": (_I, [_P, _P, _P, C.c_uint64, _P, C.c_int, _P, C.c_size_t]),
    "fzh_periodic_window_f32": (_I, [_P, _P] + [C.c_uint64] * 6 + [_P] + [C.c_uint64] * 3),
    "fzh_synth_power_law_f32": (_I, [_P, _P, C.c_uint64, C.c_uint64, C.c_uint64, _F, C.c_uint64]),
    "fzh_device_alloc": (_I, [C.POINTER(_P), _Z]),
    "fzh_device_free": (_I, [_P]),
    "fzh_memcpy_htod": (_I, [_P, _P, _Z]),
    "fzh_memcpy_dtoh": (_I, [_P, _P, _Z]),
    "fzh_memset": (_I, [_P, _I, _Z]),
    "fzh_synchronize": (_I, []),
    "fzh_stream_create": (_I, [C.POINTER(_P)]),
    "fzh_stream_destroy": (_I, [_P]),
    "fzh_stream_sync": (_I, [_P]),
    "fzh_event_create": (_I, [C.POINTER(_P)]),
    "fzh_event_destroy": (_I, [_P]),
    "fzh_event_record": (_I, [_P, _P]),
    "fzh_event_elapsed_ms": (_I, [_P, _P, C.POINTER(_F)]),
    "fzh_device_count": (_I, [C.POINTER(_I)]),
}


class FzhError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libfzh.so; raise loudly if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FzhError(f"{LIB_PATH} is missing: build it with `python -m fzgpumodules_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return list(SIGNATURES)


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().fzh_last_error().decode(errors="replace")
        raise FzhError(f"{what or 'fzh call'} failed ({rc}): {msg}")


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc


# ---------------------------------------------------------------------------
# Device memory helper
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A plain hipMalloc'd device buffer (test/bench plumbing)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _P()
        call("fzh_device_alloc", C.byref(p), max(self.nbytes, 16))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            call("fzh_memcpy_htod", b.ptr, a.ctypes.data, a.nbytes)
        return b

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceBuffer":
        b = cls(nbytes)
        call("fzh_memset", b.ptr, 0, max(int(nbytes), 16))
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        if a.nbytes:
            call("fzh_memcpy_htod", self.ptr + offset, a.ctypes.data, a.nbytes)

    def download(self, dtype=np.uint8, count: int | None = None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            assert offset + out.nbytes <= self.nbytes
            call("fzh_memcpy_dtoh", out.ctypes.data, self.ptr + offset, out.nbytes)
        return out

    def free(self) -> None:
        if self.ptr:
            lib().fzh_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def workspace(nbytes: int, stream=None) -> DeviceBuffer:
    """A zero-initialised workspace (fzh_workspace_init)."""
    b = DeviceBuffer(max(int(nbytes), 256))
    call("fzh_workspace_init", stream, b.ptr, b.nbytes)
    return b


def synchronize() -> None:
    call("fzh_synchronize")


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().fzh_device_count(C.byref(n))
    return n.value if rc == 0 else 0


class Event:
    def __init__(self):
        p = _P()
        call("fzh_event_create", C.byref(p))
        self.ptr = p.value

    def record(self, stream=None):
        call("fzh_event_record", self.ptr, stream)

    def elapsed_ms(self, end: "Event") -> float:
        ms = _F(0)
        call("fzh_event_elapsed_ms", self.ptr, end.ptr, C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            lib().fzh_event_destroy(self.ptr)
        except Exception:
            pass
