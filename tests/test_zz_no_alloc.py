"""fzh_value_base_f32_noalloc (no workspace argument, no allocation; its
partials live in a fixed per-GPU device array) against fzh_value_base_f32 on
the same inputs, through the C ABI.  Two calls in a row share the array: the
first with a large grid over large values, the second with a smaller input
(fewer workgroups) over smaller values, so a partial left over from the
first call would change the second call's result if it were read."""
import numpy as np
import pytest

from fzgpumodules_amd._fzh import DeviceBuffer, call, lib, workspace

pytestmark = pytest.mark.gpu

VB_ZERO_PADDED = 4


def _with_workspace(din, n, mode, eb):
    par = DeviceBuffer(16)
    ws = workspace(lib().fzh_value_base_ws_bytes(n))
    call("fzh_value_base_f32", None, din.ptr, n, mode, np.float32(eb), par.ptr, ws.ptr, ws.nbytes)
    return par.download(np.float32, 4)


def _noalloc(din, n, mode, eb):
    par = DeviceBuffer(16)
    call("fzh_value_base_f32_noalloc", None, din.ptr, n, mode, np.float32(eb), par.ptr)
    return par.download(np.float32, 4)


@pytest.mark.parametrize("mode", [1, 2, 1 | VB_ZERO_PADDED, 2 | VB_ZERO_PADDED])
def test_noalloc_value_base_matches_workspace_entry(gpu, mode):
    rng = np.random.default_rng(20261203 + mode)
    big = (rng.standard_normal(512 * 16384 + 37) * 1000.0 + 250.0).astype(np.float32)  # 512 workgroups
    small = (rng.random(100_003) * 2.0 + 0.5).astype(np.float32)                        # 7 workgroups
    print(f"[no-alloc] mode {mode}: n {big.size} then {small.size}", flush=True)
    for x in (big, small):
        din = DeviceBuffer.from_array(x)
        want = _with_workspace(din, x.size, mode, 1e-3)
        got = _noalloc(din, x.size, mode, 1e-3)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        lo, hi = float(x.min()), float(x.max())
        if mode & VB_ZERO_PADDED:
            lo, hi = min(lo, 0.0), max(hi, 0.0)
        vb = max(abs(lo), abs(hi)) if (mode & 3) == 1 else np.float32(hi) - np.float32(lo)
        assert got[0] == np.float32(vb)


def test_noalloc_value_base_rejects_bad_mode(gpu):
    din = DeviceBuffer.from_array(np.ones(64, np.float32))
    par = DeviceBuffer(16)
    rc = lib().fzh_value_base_f32_noalloc(None, din.ptr, 64, 3, np.float32(1e-3), par.ptr)
    assert rc != 0 and b"mode" in lib().fzh_last_error()
