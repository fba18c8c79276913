"""The drop-in boundary on the CPU: libfzh.so loads, exports every function
include/fzh.h, include/fzp.h, include/fzb.h and include/fzs.h declare, the
ctypes table covers them, and
the host-only entry points (bounds, codebook build, error strings) work
without a GPU.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("fzh.h", "fzp.h", "fzb.h", "fzs.h")]


def declared():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fz[hpbs]_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_the_boundary():
    names = declared()
    for must in ("fzh_lorenzo_quant_fwd_f32", "fzh_lorenzo_quant_inv_f32", "fzh_bitshuffle_enc", "fzh_rze_enc",
                 "fzh_rze_dec", "fzh_rle_enc", "fzh_rle_dec", "fzh_histogram", "fzh_huffman_encode",
                 "fzh_huffman_decode", "fzh_huffman_build_book", "fzh_quant_abs_fwd_f32", "fzh_bitpack_enc",
                 "fzh_gather", "fzh_concat", "fzp_create", "fzp_compress", "fzp_decompress", "fzp_write_file",
                 "fzp_decompress_file", "fzb_create", "fzb_set_fields", "fzb_compress", "fzb_autotune", "fzs_features_f32",
                 "fzs_mlp_forward", "fzs_mlp_train_step", "fzs_mlp_adam", "fzs_label_sizes", "fzs_select"):
        assert must in names, must


def test_library_exports_every_declared_symbol():
    from fzgpumodules_amd import _fzh

    L = C.CDLL(_fzh.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_table_covers_the_headers():
    from fzgpumodules_amd import _fzh

    assert set(declared()) == set(_fzh.declared_symbols())


def test_host_entry_points_without_gpu(orc):
    from fzgpumodules_amd import _fzh

    L = _fzh.lib()
    assert b"gfx950" in L.fzh_version()
    assert L.fzh_rze_enc_bound(0) >= 8
    assert L.fzh_rle_enc_bound(6, 2) == 4 + 12 + 24 + 4
    # host codebook == oracle == the reference's compiled builder (fixtures)
    z = np.load(os.path.join(ROOT, "tests", "golden", "huffman_books.npz"), allow_pickle=False)
    for name in [str(n) for n in z["names"]]:
        f = np.ascontiguousarray(z[f"{name}.freq"])
        eb = int(z[f"{name}.eb"][0])
        bk4 = np.zeros(f.size, np.uint32)
        rev = np.zeros(L.fzh_huffman_revbook_bytes(f.size, eb), np.uint8)
        rc = L.fzh_huffman_build_book(f.ctypes.data, f.size, eb, bk4.ctypes.data, rev.ctypes.data)
        assert rc == 0, L.fzh_last_error()
        np.testing.assert_array_equal(bk4, z[f"{name}.bk4"], err_msg=name)
        np.testing.assert_array_equal(rev, z[f"{name}.rev"], err_msg=name)


def test_invalid_arguments_fail_before_any_launch():
    from fzgpumodules_amd import _fzh

    L = _fzh.lib()
    # ew 3 is rejected by validation (bitshuffle_stage.h:139-150)
    rc = L.fzh_bitshuffle_enc(None, 16, 16, 16384, 16384, 3)
    assert rc == -1
    assert L.fzh_last_error()
    slots = (C.c_byte * 40)()
    assert L.fzh_concat(None, slots, 33, 16, 1 << 20, 16, 16, 4096) == -1
    assert L.fzh_concat_readback(None, slots, 33, 16, 1 << 20, 16, None, 0, None) == -1
