"""The fused concat + readback entry point against fzh_concat followed by
fzh_readback, and against the layout built on the host: [u32 n][u64 size
x n] zero-padded to 16 B, then every slot zero-padded to 16 B
(src/pipeline/compressor.cpp:805-851)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Slot(C.Structure):
    _fields_ = [("src", C.c_void_p), ("size_ptr", C.c_void_p), ("size", C.c_uint64), ("cap", C.c_uint64),
                ("kind", C.c_uint32), ("elem", C.c_uint32)]


class Entry(C.Structure):
    _fields_ = [("src", C.c_void_p), ("bytes", C.c_uint32), ("dst_offset", C.c_uint32)]


def _pad16(b):
    return np.concatenate([b, np.zeros((-b.size) % 16, np.uint8)])


@pytest.mark.parametrize("big", [5000, 300000])
def test_concat_readback_matches_two_calls(big):
    from fzgpumodules_amd._fzh import DeviceBuffer, call, lib
    rng = np.random.default_rng(big)
    n = 5
    hdr = (4 + 8 * n + 15) & ~15
    datas = [rng.integers(0, 256, 1000, dtype=np.uint8), rng.integers(0, 256, 100, dtype=np.uint8),
             rng.integers(0, 256, 64, dtype=np.uint8), np.zeros(16, np.uint8),
             rng.integers(0, 256, big, dtype=np.uint8)]
    sizes = [1000, 37, 9 * 4, 0, big - 3]
    caps = [1000, 100, 64, 16, big]
    bound = hdr + sum((c + 15) & ~15 for c in caps)
    outs = [DeviceBuffer(bound + 64) for _ in range(2)]
    srcs = [DeviceBuffer.from_array(d) for d in datas]
    szp = DeviceBuffer.from_array(np.array([37], np.uint64))
    cntp = DeviceBuffer.from_array(np.array([9, 0], np.uint32))
    totals = [DeviceBuffer(16) for _ in range(2)]
    ws = DeviceBuffer(lib().fzh_concat_ws_bytes())
    rbs = [DeviceBuffer(256) for _ in range(2)]

    def slots(out):
        a = (Slot * n)()
        for i in range(n):
            a[i].src = out.ptr + hdr if i == 0 else srcs[i].ptr
            a[i].cap = caps[i]
            a[i].size = sizes[i] if i in (0, 3, 4) else 0
        a[1].kind, a[1].size_ptr = 1, szp.ptr
        a[2].kind, a[2].size_ptr, a[2].elem = 2, cntp.ptr, 4
        return a

    def entries(out, tot):
        e = (Entry * 3)()
        e[0].src, e[0].bytes, e[0].dst_offset = out.ptr, (4 + 8 * n + 3) & ~3, 0
        e[1].src, e[1].bytes, e[1].dst_offset = tot.ptr, 8, 48
        e[2].src, e[2].bytes, e[2].dst_offset = cntp.ptr, 4, 56
        return e

    for o in outs:  # the in-place slot's bytes
        o.upload(datas[0], offset=hdr)
        o.upload(np.full(bound + 64 - hdr - 1000, 0xEE, np.uint8), offset=hdr + 1000)
    call("fzh_concat", None, C.byref(slots(outs[0])), n, outs[0].ptr, bound, totals[0].ptr, ws.ptr, ws.nbytes)
    call("fzh_readback", None, C.byref(entries(outs[0], totals[0])), 3, rbs[0].ptr)
    call("fzh_concat_readback", None, C.byref(slots(outs[1])), n, outs[1].ptr, bound, totals[1].ptr,
         C.byref(entries(outs[1], totals[1])), 3, rbs[1].ptr)
    head = np.zeros(hdr, np.uint8)
    head[:4] = np.frombuffer(np.uint32(n).tobytes(), np.uint8)
    head[4:4 + 8 * n] = np.array(sizes, np.uint64).view(np.uint8)
    want = np.concatenate([head] + [_pad16(d[:sz]) for d, sz in zip(datas, sizes)])
    for o, t, r in zip(outs, totals, rbs):
        total = int(t.download(np.uint64, 1)[0])
        assert total == want.size
        np.testing.assert_array_equal(o.download(np.uint8, total), want)
        rb = r.download(np.uint8, 60)
        np.testing.assert_array_equal(rb[:4 + 8 * n], want[:4 + 8 * n])
        assert rb[48:56].view(np.uint64)[0] == total
        assert rb[56:60].view(np.uint32)[0] == 9
