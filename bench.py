#!/usr/bin/env python3
"""Benchmark: the Lorenzo-2D -> Bitshuffle -> RZE pipeline on one MI355X.

Workload: float32 1800 x 3600 cloud field (Gaussian random field, power
~ k^-3, mapped so ~30% of values are exactly 0, max 1), REL error bound 1e-3,
radius 32768, zigzag codes, outlier capacity 0.10, Bitshuffle (16 KiB blocks,
element width 2), RZE (16 KiB chunks, 4 levels) -- the reference CLI's
`lorenzo->bitshuffle->rze -m rel -e 1e-3` setup.

A step is one Pipeline::compress() of the whole field (input resident in HBM,
output size read back, stream synchronised), its six kernels launched
directly (replaying them as one hipGraph costs about 5 us more per call on
this ROCm stack, so graph mode is left off).  The
JSON line also carries decompress GB/s, the ratio, per-kernel device times
measured with HIP events on the launch stream, the roofline of the dominant
kernel, and the CPU oracle's time for the same chain on the host cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NX, NY = 3600, 1800
USER_EB = 1e-3
CAP = 0.10
# HBM bytes per launch of the chain's main kernels on this workload, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the
# same counters read for a plain copy of known size (FETCH x 1.9996, WRITE x
# 1.0000); profiles/cloud2d_sinusoid1d_turbulence3d_hbm_traffic.md holds the
# counter runs, their commands and the scaling.
MEASURED_TRAFFIC = {
    "minmax_partial_kernel": 25975508,
    "lz2d_fwd_kernel": 39258985,
    "rze_enc_kernel": 16229674,
}
# Reference CLI example, benchmark mean of 10, host time (docs/cli.md:66-70)
REF_COMPRESS_HOST_GBS = 18.85


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--batch-fields", type=int, default=4096,
                    help="fields of the batched workload (0 skips it)")
    ap.add_argument("--selector-train", type=int, default=1000000,
                    help="training chunks of the learned-selector workload, generated on the device (0 skips it)")
    ap.add_argument("--no-cfg1", dest="cfg1", action="store_false", help="skip the 2^24 sinusoid -> RLE workload")
    ap.add_argument("--no-cfg3", dest="cfg3", action="store_false", help="skip the 512^3 turbulence -> Huffman workload")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        log(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
            "(the path is single-GPU; sharding is described in the design notes).")
        sys.exit(2)

    from fzgpumodules_amd import fields
    from fzgpumodules_amd._fzh import DeviceBuffer, Event, call, lib, workspace
    from fzgpumodules_amd.pipeline import Pipeline

    t = time.time()
    x = fields.cloud_2d(NY, NX)
    log(f"[bench] field {NY}x{NX} generated in {time.time() - t:.1f}s sha256={fields.sha256(x)[:16]}")
    nbytes = x.nbytes
    d_in = DeviceBuffer.from_array(x)

    p = Pipeline(nbytes)
    lz = p.add_lorenzo(USER_EB, mode="REL", quant_radius=32768, outlier_capacity=CAP, zigzag=True)
    # strict error bound: reference-identical codes and streams, plus the few
    # values the reference arithmetic would reconstruct outside eb
    p.set_strict(lz, True)
    bs = p.add_bitshuffle(16384, 2)
    rz = p.add_rze(16384, 4)
    p.connect(bs, lz, "codes")
    p.connect(rz, bs)
    p.set_dims(NX, NY, 1)
    p.finalize()
    p.set_pool_managed_decomp_output(True)

    # ---- compress -----------------------------------------------------------
    for _ in range(max(args.warmup, 2)):
        out_ptr, out_n = p.compress(d_in, nbytes)
    call("fzh_synchronize")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out_ptr, out_n = p.compress(d_in, nbytes)
    call("fzh_synchronize")
    t1 = time.perf_counter()
    comp_s = (t1 - t0) / args.steps
    comp_gbs = nbytes / comp_s / 1e9
    ratio = nbytes / out_n
    abs_eb = p.abs_error_bound(lz)
    outs = dict(p.outputs())
    log(f"[bench] compress {comp_s * 1e3:.3f} ms/step  {comp_gbs:.1f} GB/s  ratio {ratio:.2f}  outputs {outs}")

    # copy the compressed blob aside (the pool-owned one is reused by compress)
    blob = DeviceBuffer(out_n)
    tmp = np.empty(out_n, np.uint8)
    call("fzh_memcpy_dtoh", tmp.ctypes.data, out_ptr, out_n)
    blob.upload(tmp)

    # ---- decompress ---------------------------------------------------------
    for _ in range(max(args.warmup, 1)):
        y_ptr, y_n = p.decompress(blob, out_n)
    call("fzh_synchronize")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y_ptr, y_n = p.decompress(blob, out_n)
    call("fzh_synchronize")
    t1 = time.perf_counter()
    dec_s = (t1 - t0) / args.steps
    dec_gbs = nbytes / dec_s / 1e9
    y = np.empty(x.size, np.float32)
    call("fzh_memcpy_dtoh", y.ctypes.data, y_ptr, y.nbytes)
    max_err = float(np.max(np.abs(y.astype(np.float64) - x.ravel().astype(np.float64))))
    log(f"[bench] decompress {dec_s * 1e3:.3f} ms/step  {dec_gbs:.1f} GB/s  max|err| {max_err:.3e} (eb {abs_eb:.3e})")

    # ---- per-stage device time of one eager compress (pipeline profiling) ---
    p.enable_profiling(True)
    for _ in range(3):
        p.compress(d_in, nbytes)
    perf = p.last_perf()
    stages = []
    for i in range(8):
        name = C.create_string_buffer(64)
        ms = C.c_float()
        if lib().fzp_last_stage_perf(p.h, i, name, 64, C.byref(ms)) != 0:
            break
        stages.append((name.value.decode(), ms.value))
    log(f"[bench] eager compress: device {perf['dag_ms'] * 1e3:.1f} us, host {perf['host_ms'] * 1e3:.1f} us; "
        + ", ".join(f"{n} {m * 1e3:.1f} us" for n, m in stages))
    p.enable_profiling(False)

    # ---- per-kernel device time, HIP events on the launch stream ------------
    kern, single = kernel_times(args.kernel_reps, x, d_in, abs_eb)
    # dominant kernel of the compress chain (the main kernel of each stage)
    dom = max(("minmax_partial_kernel", "lz2d_fwd_kernel", "rze_enc_kernel"), key=lambda k: single[k]["ms"])
    d = single[dom]
    roof = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": round(d["gbs"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(d["gbs"] / HBM_PEAK_GBS, 4),
        "avg_launch_us": round(d["ms"] * 1e3, 3),
        "algorithmic_bytes": d["bytes"],
        "traffic": MEASURED_TRAFFIC.get(dom),
    }
    chain = ("value_base", "lorenzo_fwd_2d", "bitshuffle_rze_enc")
    chain_bytes = sum(kern[k]["bytes"] for k in chain)
    chain_ms = sum(kern[k]["ms"] for k in chain)

    # ---- CPU oracle baseline --------------------------------------------------
    cpu = None if args.no_cpu else cpu_baseline(x, args.cpu_runs)

    res = {
        "metric": "compress/decompress GB/s + ratio at rel eb 1e-3; learned-select vs exhaustive",
        "value": round(comp_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(comp_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(comp_gbs / REF_COMPRESS_HOST_GBS, 2),
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "cloud_2d float32 1800x3600 (GRF k^-3, 30% zeros); Lorenzo2D REL 1e-3 r32768 zigzag "
                        "cap0.10 strict-eb -> Bitshuffle(16KiB, ew2) -> RZE(16KiB, 4 levels); one compress() per "
                        "step, direct launches",
            "input_bytes": nbytes,
            "padded_bytes": int(math.ceil(nbytes / 16384) * 16384),
        },
        "eager_compress_device_us": round(perf["dag_ms"] * 1e3, 2),
        "eager_compress_host_us": round(perf["host_ms"] * 1e3, 2),
        "stage_device_us": {n: round(m * 1e3, 2) for n, m in stages},
        "decompress_gbs": round(dec_gbs, 2),
        "decompress_ms_per_step": round(dec_s * 1e3, 4),
        "ratio": round(ratio, 3),
        "compressed_bytes": int(out_n),
        "outputs": outs,
        "abs_eb": abs_eb,
        "max_abs_err": max_err,
        "compress_chain_device_ms": round(chain_ms, 4),
        "compress_chain_roofline_frac": round(chain_bytes / (chain_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": {k: {kk: round(vv, 4) if isinstance(vv, float) else vv for kk, vv in v.items()}
                    for k, v in kern.items()},
        "single_kernels": {k: {kk: round(vv, 5) if isinstance(vv, float) else vv for kk, vv in v.items()}
                           for k, v in single.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    # ---- the GPU blob against the oracle chain at full size ------------------
    # (strict mode adds the patch list; the first four parts are the plain
    # chain's and must equal the oracle's byte for byte)
    if cpu is not None:
        got = split_blob(tmp)
        ref = cpu.pop("_parts")
        res["bit_exact_vs_oracle"] = bool(len(got) == 7 and all(np.array_equal(got[i], ref[i]) for i in range(4)))
        log(f"[bench] compressed parts equal to the oracle chain: {res['bit_exact_vs_oracle']}")
    p.close()
    d_in.free()
    if args.cfg1:
        res["sinusoid_1d"] = chain_workload("1d", args)
    if args.cfg3:
        res["turbulence_3d"] = chain_workload("3d", args)
    if args.batch_fields > 0:
        res["batched"] = batched_workload(args)
    if args.selector_train > 0:
        res["selector"] = selector_workload(args)
    print(json.dumps(res), flush=True)


def kernel_times(reps, x, d_in, abs_eb):
    """Each kernel of the chain on its own, timed with HIP events around
    `reps` back-to-back launches on one stream; algorithmic bytes per launch
    are the compulsory reads + writes of the stage."""
    from fzgpumodules_amd._fzh import DeviceBuffer, Event, call, lib, workspace

    L = lib()
    s = C.c_void_p()
    call("fzh_stream_create", C.byref(s))
    st = s.value
    n = int(math.ceil(x.nbytes / 16384) * 16384) // 4  # padded elements
    nv = x.size
    m = math.ceil(n * float(np.float32(CAP)))
    params = DeviceBuffer(16)
    vb_ws = workspace(L.fzh_value_base_ws_bytes(nv))
    codes = DeviceBuffer(2 * n)
    err = DeviceBuffer(4 * m)
    idx = DeviceBuffer(4 * m)
    cnt = DeviceBuffer(16)
    lz_ws = workspace(L.fzh_lorenzo_quant_fwd_strict_ws_bytes(n, NX, NY, 1))
    mp = max(256, math.ceil(n * 0.01))
    pval = DeviceBuffer(4 * mp)
    pidx = DeviceBuffer(4 * mp)
    pcnt = DeviceBuffer(16)
    shuf = DeviceBuffer(2 * n)
    rze_cap = L.fzh_rze_enc_bound(2 * n)
    rze_out = DeviceBuffer(rze_cap)
    rze_sz = DeviceBuffer(16)
    rze_ws = workspace(L.fzh_rze_enc_ws_bytes(2 * n))
    unshuf = DeviceBuffer(2 * n)
    rzd_ws = workspace(L.fzh_rze_dec_ws_bytes(2 * n))
    status = DeviceBuffer(16)
    call("fzh_memset", status.ptr, 0, 16)
    y = DeviceBuffer(4 * n)
    inv_ws = workspace(L.fzh_lorenzo_quant_inv_ws_bytes(n, NX, NY, 1, m))

    ebx2 = np.float32(2) * np.float32(abs_eb)

    def vb():
        call("fzh_value_base_f32", st, d_in.ptr, nv, 1, np.float32(USER_EB), params.ptr, vb_ws.ptr, vb_ws.nbytes)

    def lzf():
        # the strict variant, as the benchmarked pipeline runs it
        call("fzh_lorenzo_quant_fwd_strict_f32", st, d_in.ptr, n, NX, NY, 1, np.float32(0), params.ptr + 8, 32768,
             1, 2, codes.ptr, err.ptr, idx.ptr, cnt.ptr, m, np.float32(0), np.float32(0), params.ptr, pval.ptr,
             pidx.ptr, pcnt.ptr, mp, lz_ws.ptr, lz_ws.nbytes)

    def bse():
        call("fzh_bitshuffle_enc", st, codes.ptr, shuf.ptr, 2 * n, 16384, 2)

    def rze():
        call("fzh_rze_enc", st, shuf.ptr, 2 * n, rze_out.ptr, rze_cap, rze_sz.ptr, rze_ws.ptr, rze_ws.nbytes)

    def rzd():
        call("fzh_rze_dec", st, rze_out.ptr, int(csz[0]), shuf.ptr, 2 * n, status.ptr, rzd_ws.ptr, rzd_ws.nbytes)

    def bsd():
        call("fzh_bitshuffle_dec", st, shuf.ptr, unshuf.ptr, 2 * n, 16384, 2)

    # the fused pair the pipeline runs for Bitshuffle -> RZE
    def frze():
        call("fzh_rze_enc_shuffled", st, codes.ptr, 2 * n, 2, rze_out.ptr, rze_cap, rze_sz.ptr, rze_ws.ptr,
             rze_ws.nbytes)

    def frzd():
        call("fzh_rze_dec_shuffled", st, rze_out.ptr, int(csz[0]), unshuf.ptr, 2 * n, 2, status.ptr, rzd_ws.ptr,
             rzd_ws.nbytes)

    def patch():
        call("fzh_lorenzo_apply_patches", st, y.ptr, pval.ptr, pidx.ptr, pcnt.ptr, mp, nv)

    # outlier capacity on decompression = the stored count, as a stored
    # stream gives it (set once the forward has run)
    inv_cap = [m]

    def lzi():
        call("fzh_lorenzo_quant_inv_f32", st, unshuf.ptr, n, NX, NY, 1, err.ptr, idx.ptr, cnt.ptr, inv_cap[0], ebx2,
             1, 2, y.ptr, inv_ws.ptr, inv_ws.nbytes)

    # a plain device copy of the input's size: the achievable-bandwidth
    # reference and the calibration point for the FETCH/WRITE counters
    copy_dst = DeviceBuffer(4 * nv)

    def cpy():
        call("fzh_copy", st, copy_dst.ptr, d_in.ptr, 4 * nv)

    def timed(fn):
        for _ in range(3):
            fn()
        e0, e1 = Event(), Event()
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        return e0.elapsed_ms(e1) / reps

    out = {}
    vb()
    lzf()
    bse()
    rze()
    call("fzh_stream_sync", st)
    k = int(cnt.download(np.uint32, 1)[0])
    kp = min(int(pcnt.download(np.uint32, 1)[0]), mp)
    inv_cap[0] = min(k, m)
    csz = rze_sz.download(np.uint64, 1)
    C_rze = int(csz[0])
    # (name, launch, algorithmic bytes per launch, run by the pipeline)
    specs = [
        ("value_base", vb, 4 * nv, True),
        ("lorenzo_fwd_2d", lzf, 4 * NX * NY + 2 * n + 8 * k + 8 * kp + 8, True),
        ("bitshuffle_rze_enc", frze, 2 * n + C_rze, True),
        ("rze_bitshuffle_dec", frzd, C_rze + 2 * n, True),
        ("lorenzo_inv_2d", lzi, 2 * n + 8 * k + 4 + 4 * NX * NY, True),
        ("apply_patches", patch, 8 * kp + 4 * kp, True),
        ("bitshuffle_enc", bse, 4 * n, False),
        ("rze_enc", rze, 2 * n + C_rze, False),
        ("rze_dec", rzd, C_rze + 2 * n, False),
        ("bitshuffle_dec", bsd, 4 * n, False),
        ("copy_reference", cpy, 8 * nv, False),
    ]
    for name, fn, nb, used in specs:
        ms = timed(fn)
        out[name] = {"ms": ms, "bytes": int(nb), "gbs": nb / (ms * 1e-3) / 1e9,
                     "frac": nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "pipeline": used}
        log(f"[bench] kernel {name:16s} {ms * 1e3:8.1f} us  {out[name]['gbs']:7.1f} GB/s  "
            f"{100 * out[name]['frac']:5.1f}% of HBM peak")
    # ---- single kernels: library HIP-event pairs around each launch --------
    # (name, algorithmic bytes per launch or None for helper kernels, whose
    # bytes are not compulsory) per entry point the pipeline runs; helper
    # kernels that a workload's sizes leave out (the offset scans of short
    # streams, the outlier binning of short lists) are reported as not run
    per_entry = [
        (vb, [("minmax_partial_kernel", 4 * nv)]),
        (lzf, [("lz2d_fwd_kernel", 4 * NX * NY + 2 * n + 8 * k + 8 * kp + 8), ("scan_small_kernel", None),
               ("place_kernel", None)]),
        (frze, [("rze_enc_kernel", 2 * n + C_rze), ("scan_small_kernel", None), ("rze_pack_kernel", None)]),
        (frzd, [("rze_offsets_kernel", None), ("rze_dec_kernel", C_rze + 2 * n)]),
        (lzi, [("bin_count_kernel", None), ("scan_small_kernel", None), ("bin_place_kernel", None),
               ("lz2d_inv_kernel", 2 * n + 8 * k + 4 + 4 * NX * NY)]),
        (cpy, [("copy_kernel", 8 * nv)]),
    ]
    single = {}
    tot, cnt_ = C.c_double(), C.c_uint64()
    for fn, names in per_entry:
        call("fzh_kernel_timing_enable", 1)  # clears the totals
        for _ in range(reps):
            fn()
        call("fzh_stream_sync", st)
        for kname, nb in names:
            call("fzh_kernel_timing_get", kname.encode(), C.byref(tot), C.byref(cnt_))
            if cnt_.value == 0:
                if nb is None:
                    log(f"[bench] single kernel {kname:32s} not run by this entry point at these sizes")
                    continue
                raise RuntimeError(f"kernel timing recorded no launch of {kname}")
            ms = tot.value / cnt_.value
            key = kname if kname not in single else f"{kname}@{names[0][0]}"
            single[key] = {"ms": ms, "launches": int(cnt_.value), "bytes": nb}
            if nb:
                single[key]["gbs"] = nb / (ms * 1e-3) / 1e9
                single[key]["frac"] = single[key]["gbs"] / HBM_PEAK_GBS
            log(f"[bench] single kernel {key:32s} {ms * 1e3:8.2f} us"
                + (f"  {single[key]['gbs']:7.1f} GB/s  {100 * single[key]['frac']:5.1f}% of HBM peak" if nb else ""))
    call("fzh_kernel_timing_enable", 0)
    st_ = int(status.download(np.int32, 1)[0])
    if st_ != 0:
        raise RuntimeError("rze_dec reported a malformed stream in the kernel timing loop")
    call("fzh_stream_destroy", st)
    return {kk: vv for kk, vv in out.items() if not kk.startswith("_")}, single


def batched_workload(args):
    """Thousands of uneven fields through one persistent launch per stage:
    sizes log-uniform 64 KiB .. 64 MiB, dimensionality 1D/2D/3D in equal
    thirds with random aspect ratios, each its own seeded Gaussian random
    field with power ~ k^e, e drawn from U[-5,-1] per field (generated on the
    device, fzh_synth_power_law_f32); REL 1e-3, radius 32768, zigzag, cap 0.1,
    strict bound -> Bitshuffle -> RZE; autotuned launch geometry.  Every
    field's bytes are compared with a single-field pipeline run, and the
    smallest fields' with the oracle chain."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import chains
    from fzgpumodules_amd import fields
    from fzgpumodules_amd._fzh import DeviceBuffer, call
    from fzgpumodules_amd.batch import Batch
    from fzgpumodules_amd.pipeline import Pipeline

    t = time.time()
    F = args.batch_fields
    rng = np.random.default_rng(fields.BASE_SEED + 4)
    nbytes = np.exp(rng.uniform(np.log(64 << 10), np.log(64 << 20), F))
    nds = rng.permutation(np.arange(F) % 3 + 1)
    ex = rng.uniform(-5.0, -1.0, F)
    bufs, dims = [], []
    for i in range(F):
        n = max(int(nbytes[i]) // 4, 4096)
        nd = int(nds[i])
        if nd == 1:
            d = (n, 1, 1)
        elif nd == 2:
            a = float(np.exp(rng.uniform(-np.log(4), np.log(4))))
            nx = max(32, int(round(np.sqrt(n * a))))
            d = (nx, max(2, n // nx), 1)
        else:
            a, b = np.exp(rng.uniform(-np.log(2), np.log(2), 2))
            side = n ** (1.0 / 3.0)
            nx, ny = max(8, int(round(side * a))), max(8, int(round(side * b)))
            d = (nx, ny, max(2, n // (nx * ny)))
        buf = DeviceBuffer(4 * d[0] * d[1] * d[2])
        call("fzh_synth_power_law_f32", None, buf.ptr, d[0], d[1], d[2], np.float32(ex[i]),
             (fields.BASE_SEED + 4) * 65536 + i)
        bufs.append(buf)
        dims.append(d)
    call("fzh_synchronize")
    total_in = sum(4 * d[0] * d[1] * d[2] for d in dims)
    log(f"[bench] batched: {F} fields, {total_in / 2**30:.2f} GiB generated in {time.time() - t:.1f}s")
    b = Batch(1e-3, "REL", 32768, 0.1, True, strict=True, patch_capacity=0.01)
    b.set_fields([(buf.ptr, d) for buf, d in zip(bufs, dims)])
    t = time.time()
    tuned = b.autotune(reps=2)
    log(f"[bench] batched: autotune {time.time() - t:.1f}s -> {tuned}")
    b.enable_profiling(True)
    b.compress()
    stages = b.stage_ms()
    b.enable_profiling(False)
    for _ in range(2):  # eager, then capture
        b.compress()
    steps = max(3, min(args.steps, 10))
    t0 = time.perf_counter()
    for _ in range(steps):
        b.compress()
    t1 = time.perf_counter()
    sec = (t1 - t0) / steps
    tin, tout = b.totals()
    # every field against the single-field pipeline, byte for byte
    t = time.time()
    equal = 0
    for i in range(F):
        d = dims[i]
        with Pipeline(4 * d[0] * d[1] * d[2]) as p:
            lz = p.add_lorenzo(1e-3, mode="REL", quant_radius=32768, outlier_capacity=0.1, zigzag=True)
            p.set_strict(lz, True, 0.01)
            bs_ = p.add_bitshuffle(16384, 2)
            rz = p.add_rze(16384, 4)
            p.connect(bs_, lz, "codes")
            p.connect(rz, bs_)
            p.set_dims(*d)
            p.finalize()
            ptr, n = p.compress(bufs[i], 4 * d[0] * d[1] * d[2])
            want = np.empty(n, np.uint8)
            call("fzh_memcpy_dtoh", want.ctypes.data, ptr, n)
        equal += int(np.array_equal(b.field_bytes(i), want))
    log(f"[bench] batched: {equal} of {F} fields equal to single-field pipeline runs ({time.time() - t:.1f}s)")
    # the smallest fields against the oracle chain (parts of the strict result)
    small = sorted(range(F), key=lambda i: dims[i][0] * dims[i][1] * dims[i][2])[:8]
    oracle_equal = 0
    for i in small:
        d = dims[i]
        x = np.empty(d[0] * d[1] * d[2], np.float32)
        call("fzh_memcpy_dtoh", x.ctypes.data, bufs[i].ptr, x.nbytes)
        want, info = chains.lorenzo_bitshuffle_rze(x, d, mode=1, user_eb=1e-3, radius=32768, cap=0.1, zigzag=True)
        got, ref = chains.split(b.field_bytes(i)), chains.split(want)
        ok = len(got) == 7 and all(np.array_equal(got[k], ref[k]) for k in range(4))
        if ok:
            recon = chains.reconstruct(info, d, zigzag=True)
            pidx, pval = chains.strict_patches(x, recon, info["abs_eb"], d)
            ok = np.array_equal(got[5].view(np.uint32), pidx) and np.array_equal(got[4].view(np.float32), pval)
        oracle_equal += int(ok)
    log(f"[bench] batched: {oracle_equal} of {len(small)} smallest fields equal to the oracle chain")
    b.close()
    for buf in bufs:
        buf.free()
    gbs = tin / sec / 1e9
    log(f"[bench] batched: {sec * 1e3:.2f} ms/compress  {gbs:.1f} GB/s  ratio {tin / tout:.2f}  "
        f"stages {[(n, round(m, 3)) for n, m in stages]}")
    stage_ms = {n: round(m, 4) for n, m in stages}
    lz_gbs = (6 * tin / 4) / (stage_ms["lorenzo"] * 1e-3) / 1e9 if stage_ms.get("lorenzo") else None
    return {
        "workload": f"{F} fields, 64 KiB..64 MiB log-uniform, 1D/2D/3D thirds, each a seeded Gaussian random "
                    "field with power ~ k^e, e ~ U[-5,-1] per field; REL 1e-3 strict Lorenzo -> Bitshuffle -> RZE, "
                    "one persistent launch per stage, hipGraph",
        "fields": F,
        "input_bytes": int(tin),
        "compressed_bytes": int(tout),
        "ratio": round(tin / tout, 3),
        "compress_gbs": round(gbs, 2),
        "ms_per_compress": round(sec * 1e3, 3),
        "stage_ms": stage_ms,
        "lorenzo_stage_gbs": round(lz_gbs, 1) if lz_gbs else None,
        "autotuned_launch": tuned,
        "fields_equal_single_field_pipeline": {"checked": F, "equal": equal},
        "smallest_fields_equal_oracle": {"checked": len(small), "equal": oracle_equal},
    }


def selector_workload(args):
    """Learned chain selection on 256 KiB chunks of five regimes (smooth
    k^-4 GRF, white noise, 80% zeros, piecewise constant, Poisson 20), 2D
    256 x 256 and 3D 16 x 64 x 64, each chunk bounded by 1e-3 of its parent
    field's max |x|, all generated on the device from seeds
    (fzs_synth_chunks): exhaustive labels (the labeller: every chain's exact
    compressed size) -> device training -> held-out regret, selection
    throughput, and whole-field compress throughput of the selecting
    compressor (learned; exhaustive, i.e. labeller then compress; every
    chunk with every chain; one fixed chain)."""
    from fzgpumodules_amd import fields
    from fzgpumodules_amd._fzh import DeviceBuffer, call
    from fzgpumodules_amd import selector as S

    t = time.time()
    n_train = max(args.selector_train // 80 * 80, 80)  # whole parent fields of every regime, both dims
    n_test = max(min(n_train // 4, 8192) // 80 * 80, 80)
    batch = 8192
    bases = {nd: S.synth_base(nd, fields.BASE_SEED + 50 + nd) for nd in (2, 3)}
    shapes = {2: (256, 256), 3: (16, 64, 64)}

    def labelled(n_total, seed, keep=False):
        feats, sizes, kept = [], [], []
        for nd in (2, 3):
            n_nd = n_total // 2
            for k, off in enumerate(range(0, n_nd, batch)):
                nb = min(batch, n_nd - off)
                d_c, d_eb = S.synth_chunks(nd, nb, seed * 1000003 + 7919 * nd + k, bases[nd])
                f = S.features(d_c.ptr, nb, shapes[nd], d_eb.ptr)
                feats.append(f.download(np.float32, 32 * nb).reshape(-1, 32))
                f.free()
                sizes.append(S.label_sizes(d_c.ptr, nb, shapes[nd], d_eb.ptr))
                if keep:
                    kept.append((d_c, d_eb, nb, nd))
                else:
                    d_c.free()
                    d_eb.free()
                if k % 16 == 15:
                    log(f"[bench] selector: {nd}D, {off + nb} of {n_nd} chunks generated and labelled")
        return np.concatenate(feats), np.concatenate(sizes), kept

    f_tr, s_tr, _ = labelled(n_train, fields.BASE_SEED + 5)
    gen_s = time.time() - t
    f_te, s_te, dev_te = labelled(n_test, fields.BASE_SEED + 500, keep=True)
    log(f"[bench] selector: {len(f_tr)} training / {len(f_te)} held-out chunks generated on the device and "
        f"labelled in {time.time() - t:.1f}s")
    sel = S.Selector(seed=11)
    t = time.time()
    hist = sel.fit(f_tr, S.regret_costs(s_tr), epochs=20, batch=4096, lr=3e-3, tau=0.02)
    call("fzh_synchronize")
    train_s = time.time() - t
    log(f"[bench] selector: trained in {train_s:.1f}s, loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    d_f = DeviceBuffer.from_array(f_te)
    choice = sel.select(d_f.ptr, len(f_te))
    best = s_te.min(axis=1).astype(np.float64)
    learned = s_te[np.arange(len(choice)), choice].astype(np.float64)
    in_bytes = 262144 * len(f_te)
    single = [float(s_te[:, k].astype(np.float64).sum()) for k in range(4)]
    kfix = int(np.argmin([s_tr[:, k].astype(np.float64).sum() for k in range(4)]))  # best chain on the training set
    st = C.c_void_p()
    call("fzh_stream_create", C.byref(st))

    def timed(fn, reps=3):
        fn()
        call("fzh_synchronize")
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        call("fzh_synchronize")
        return (time.perf_counter() - t0) / reps

    def learned_select():
        for d_c, d_eb, n, nd in dev_te:
            f = S.features(d_c.ptr, n, shapes[nd], d_eb.ptr, stream=st.value)
            sel.select(f.ptr, n, stream=st.value)
            f.free()

    def exhaustive_select():
        for d_c, d_eb, n, nd in dev_te:
            S.label_sizes(d_c.ptr, n, shapes[nd], d_eb.ptr, stream=st.value)

    t_learn = timed(learned_select)
    t_exh = timed(exhaustive_select)
    # whole-field compress: the held-out chunks of each dimensionality as one
    # field (chunks stacked along the slowest axis), the selecting compressor
    # in each mode; the bound is the field's REL 1e-3
    params, norm = sel.get_params(), sel.get_norm()
    comp = {"learned": 0.0, "exhaustive_labeller": 0.0, "fixed": [0.0] * 4}
    comp_bytes = {"learned": 0, "exhaustive_labeller": 0, "fixed": [0] * 4}
    for nd in (2, 3):
        sc = S.SelectingCompressor(shapes[nd])
        sc.set_model(params, norm)
        parts = [(d_c, n) for d_c, _, n, d in dev_te if d == nd]
        for d_c, n in parts:
            shape = (256 * n, 256) if nd == 2 else (16 * n, 64, 64)
            for mode, key in ((S.LEARNED, "learned"), (S.EXHAUSTIVE, "exhaustive_labeller")):
                sc.compress(d_c.ptr, shape, 1e-3, mode)  # warm: chains built, buffers sized
                blob = sc.compress(d_c.ptr, shape, 1e-3, mode)
                comp[key] += sc.timing()["total_ms"]
                comp_bytes[key] += blob.size
            for k in range(4):
                sc.compress(d_c.ptr, shape, 1e-3, S.FIXED, k)
                blob = sc.compress(d_c.ptr, shape, 1e-3, S.FIXED, k)
                comp["fixed"][k] += sc.timing()["total_ms"]
                comp_bytes["fixed"][k] += blob.size
        sc.close()
    call("fzh_stream_destroy", st.value)
    for d_c, d_eb, _, _ in dev_te:
        d_c.free()
        d_eb.free()
    for b in bases.values():
        b.free()
    d_f.free()
    sel.close()
    gbs = lambda ms: round(in_bytes / (ms * 1e-3) / 1e9, 2)
    all_chains_ms = sum(comp["fixed"])
    res = {
        "workload": "256 KiB chunks (2D 256x256, 3D 16x64x64) of 5 regimes generated on the device, eb = 1e-3 "
                    "max|parent|; chains " + ", ".join(S.CHAINS),
        "train_chunks": int(len(f_tr)),
        "heldout_chunks": int(len(f_te)),
        "generate_and_label_seconds": round(gen_s, 2),
        "train_seconds": round(train_s, 2),
        "train_loss_first_last": [round(hist[0], 4), round(hist[-1], 4)],
        "ratio_exhaustive": round(in_bytes / best.sum(), 3),
        "ratio_learned": round(in_bytes / learned.sum(), 3),
        "regret_learned": round(learned.sum() / best.sum() - 1, 5),
        "ratio_single_chain_baseline": round(in_bytes / single[kfix], 3),
        "single_chain_baseline": S.CHAINS[kfix],
        "regret_single_chain": round(single[kfix] / best.sum() - 1, 5),
        "choice_accuracy": round(float(np.mean(choice == np.argmin(s_te, axis=1))), 4),
        "learned_choice_counts": [int(v) for v in np.bincount(choice, minlength=4)],
        "select_gbs_learned": round(in_bytes / t_learn / 1e9, 2),
        "select_gbs_exhaustive": round(in_bytes / t_exh / 1e9, 2),
        "select_speedup": round(t_exh / t_learn, 2),
        "compress_gbs_learned": gbs(comp["learned"]),
        "compress_gbs_exhaustive_labeller": gbs(comp["exhaustive_labeller"]),
        "compress_gbs_exhaustive_all_chains": gbs(all_chains_ms),
        "compress_gbs_single_chain": gbs(comp["fixed"][kfix]),
        "compress_speedup_vs_all_chains": round(all_chains_ms / comp["learned"], 2),
        "compress_ratio_learned": round(in_bytes / comp_bytes["learned"], 3),
        "compress_ratio_single_chain": round(in_bytes / comp_bytes["fixed"][kfix], 3),
    }
    log(f"[bench] selector: {res}")
    return res


def split_blob(blob):
    """Parts of a pipeline result ([u32 n][u64 size x n], 16-byte slots)."""
    n = int(blob[:4].view(np.uint32)[0])
    sizes = blob[4:4 + 8 * n].view(np.uint64).astype(np.int64)
    off = (4 + 8 * n + 15) // 16 * 16
    parts = []
    for sz in sizes:
        parts.append(blob[off:off + sz])
        off += (int(sz) + 15) // 16 * 16
    return parts


def cpu_info():
    """Host CPU: model name, logical CPUs, CPUs this process may run on, and
    the thread share the oracle uses (OMP_NUM_THREADS when set, else the
    affinity set)."""
    model = None
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    if model is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity, "threads": min(share, affinity)}


def cpu_baseline(x, runs):
    """The oracle's value base + Lorenzo2D + Bitshuffle + RZE on the host
    cores (OpenMP over tiles/blocks/chunks), median of `runs` after 1 warm-up,
    on every thread of the process's share and on one thread."""
    import oracle

    info = cpu_info()
    L = oracle.lib()
    n = int(math.ceil(x.nbytes / 16384) * 16384) // 4
    xp = np.zeros(n, np.float32)
    xp[: x.size] = x.ravel()
    m = math.ceil(n * float(np.float32(CAP)))
    codes = np.zeros(n, np.uint16)
    shuf = np.zeros(2 * n, np.uint8)
    out = np.zeros(2 * n + 8 + 4 * (2 * n // 16384 + 1) + 64, np.uint8)
    err = np.zeros(m, np.float32)
    idx = np.zeros(m, np.uint32)
    no = C.c_uint32()

    def run():
        vb = np.float32(L.orc_value_base_f32(xp.ctypes.data, n, 1))
        eb = np.float32(np.float32(USER_EB) * vb)
        r = np.float32(1) / (np.float32(2) * eb)
        return L.orc_chain_l2d_bs_rze(xp.ctypes.data, n, NX, NY, r, 32768, 1, codes.ctypes.data, shuf.ctypes.data,
                                      out.ctypes.data, err.ctypes.data, idx.ctypes.data, m, C.byref(no))

    def timed(threads, k):
        oracle.set_threads(threads)
        sz = run()
        ts = []
        for _ in range(k):
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), sz

    med, sz = timed(info["threads"], runs)
    med1, _ = timed(1, max(3, runs // 2))
    oracle.set_threads(info["threads"])
    gbs, gbs1 = x.nbytes / med / 1e9, x.nbytes / med1 / 1e9
    log(f"[bench] cpu oracle chain: {med * 1e3:.1f} ms  {gbs:.2f} GB/s on {info['threads']} threads; "
        f"{med1 * 1e3:.1f} ms  {gbs1:.3f} GB/s on 1 thread ({info['model']}, nproc {info['nproc']})")
    k = min(int(no.value), m)
    return {
        "value": round(gbs, 3),
        "unit": "GB/s",
        "cores": info["threads"],
        "kind": "port",
        "sample": f"whole 1800x3600 field (value base + Lorenzo2D + Bitshuffle + RZE), median of {runs} runs "
                  f"after 1 warm-up, OpenMP over tiles/blocks/chunks on the process's {info['threads']}-thread "
                  "share of the host",
        "value_1thread": round(gbs1, 4),
        "cpu_model": info["model"],
        "nproc": info["nproc"],
        "affinity_cpus": info["affinity"],
        "_parts": [out[:sz].copy(), err[:k].copy(), idx[:k].copy(), np.array([no.value], np.uint32).view(np.uint8)],
    }


def chain_workload(kind, args):
    """One of the other two single-field workloads, strict bound on:
      1d: sinusoid_1d(2^24), Lorenzo ABS 1e-4 r32768 cap 0.2 -> RLE<u16>
      3d: turbulence_3d(512), Lorenzo REL 1e-4 r512 zigzag cap 0.2 ->
          Huffman<u16> (bklen 1024, coarse partitions of 768 symbols)
    Wall-time compress/decompress GB/s, ratio, max error, the first four
    parts against the oracle chain, per-kernel device times (library
    dispatch timestamps) with each main kernel's algorithmic bytes and HBM
    fraction, and the oracle's time on the host."""
    import oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import chains
    from fzgpumodules_amd import fields
    from fzgpumodules_amd._fzh import DeviceBuffer, call, lib
    from fzgpumodules_amd.pipeline import Pipeline

    t = time.time()
    if kind == "1d":
        x = fields.sinusoid_1d()
        dims = (x.size, 1, 1)
        name = "sinusoid_1d float32 2^24 (8 sinusoids + N(0,1e-3)); Lorenzo1D ABS 1e-4 r32768 cap0.2 strict-eb -> RLE<u16>"
    else:
        x = fields.turbulence_3d(512)
        dims = (512, 512, 512)
        name = ("turbulence_3d float32 512^3 (GRF k^-5/3 + 1% spikes x100); Lorenzo3D REL 1e-4 r512 zigzag cap0.2 "
                "strict-eb -> Huffman<u16> bklen 1024 coarse")
    log(f"[bench] {kind}: field {x.shape} generated in {time.time() - t:.1f}s sha256={fields.sha256(x)[:16]}")
    nbytes = x.nbytes
    nv = x.size
    d_in = DeviceBuffer.from_array(x)
    p = Pipeline(nbytes)
    if kind == "1d":
        lz = p.add_lorenzo(1e-4, mode="ABS", quant_radius=32768, outlier_capacity=0.2)
        co = p.add_rle(2)
    else:
        lz = p.add_lorenzo(1e-4, mode="REL", quant_radius=512, outlier_capacity=0.2, zigzag=True)
        co = p.add_huffman(2, 1024)
        p.set_dims(*dims)
    p.set_strict(lz, True)
    p.connect(co, lz, "codes")
    p.finalize()
    p.set_pool_managed_decomp_output(True)
    steps = max(3, min(args.steps, 20))
    for _ in range(2):
        out_ptr, out_n = p.compress(d_in, nbytes)
    call("fzh_synchronize")
    t0 = time.perf_counter()
    for _ in range(steps):
        out_ptr, out_n = p.compress(d_in, nbytes)
    call("fzh_synchronize")
    comp_s = (time.perf_counter() - t0) / steps
    outs = dict(p.outputs())
    blob_h = np.empty(out_n, np.uint8)
    call("fzh_memcpy_dtoh", blob_h.ctypes.data, out_ptr, out_n)
    blob = DeviceBuffer.from_array(blob_h)
    for _ in range(2):
        y_ptr, y_n = p.decompress(blob, out_n)
    call("fzh_synchronize")
    t0 = time.perf_counter()
    for _ in range(steps):
        y_ptr, y_n = p.decompress(blob, out_n)
    call("fzh_synchronize")
    dec_s = (time.perf_counter() - t0) / steps
    y = np.empty(nv, np.float32)
    call("fzh_memcpy_dtoh", y.ctypes.data, y_ptr, y.nbytes)
    abs_eb = p.abs_error_bound(lz)
    max_err = float(np.max(np.abs(y.astype(np.float64) - x.ravel().astype(np.float64))))
    del y

    # per-kernel device times over `reps` compress + decompress calls
    reps = 10
    call("fzh_kernel_timing_enable", 1)
    for _ in range(reps):
        p.compress(d_in, nbytes)
    for _ in range(reps):
        p.decompress(blob, out_n)
    call("fzh_synchronize")
    n = nv  # no padding: RLE and Huffman need no alignment
    k = outs["outlier_errors"] // 4
    kp = outs["patch_values"] // 4
    Cc = outs["output"]
    nd = "1d" if kind == "1d" else "3d"
    main = {
        "minmax_partial_kernel": 4 * nv,
        f"lz{nd}_fwd_kernel": 4 * nv + 2 * n + 8 * k + 8 * kp + 8,
        f"lz{nd}_inv_kernel": 2 * n + 8 * k + 4 * nv,
    }
    if kind == "1d":
        R = int(blob_h[split_offset(blob_h, 0):][:4].view(np.uint32)[0])
        # the RLE stage's bytes (2n in, the values and lengths out) are
        # charged to the kernel that reads the codes and writes the stream
        main["rle_enc_write_kernel"] = 2 * n + 4 + (2 * R + 3) // 4 * 4 + 4 * R
        helpers = ["place_kernel", "scan_small_kernel", "scan_reduce_kernel", "scan_partials_kernel",
                   "scan_apply_kernel", "rle_enc_count_kernel", "rle_enc_fix_kernel", "rle_dec_prep_kernel",
                   "rle_dec_reduce_kernel",
                   "rle_dec_apply_kernel", "rle_dec_expand_kernel", "concat_readback_kernel",
                   "readback_kernel", "apply_patches_kernel", "bin_count_kernel", "bin_place_kernel"]
    else:
        R = None
        main["hist_kernel"] = 2 * n
        main["huf_enc_kernel"] = 2 * n + Cc
        main["huf_dec_kernel"] = Cc + 2 * n
        helpers = ["minmax_final_kernel", "place_kernel", "scan_small_kernel", "scan_reduce_kernel",
                   "scan_partials_kernel", "scan_apply_kernel", "huf_pack_kernel", "huf_header_kernel",
                   "concat_readback_kernel", "readback_kernel", "apply_patches_kernel", "bin_count_kernel",
                   "bin_place_kernel"]
    kern = {}
    tot, cnt_ = C.c_double(), C.c_uint64()
    for kname in list(main) + helpers:
        call("fzh_kernel_timing_get", kname.encode(), C.byref(tot), C.byref(cnt_))
        if cnt_.value == 0:
            continue
        ms = tot.value / cnt_.value
        e = {"us": round(ms * 1e3, 2), "launches_per_call": round(cnt_.value / reps, 2)}
        if kname in main:
            gbs = main[kname] / (ms * 1e-3) / 1e9
            e.update(bytes=int(main[kname]), gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        kern[kname] = e
        log(f"[bench] {kind}: kernel {kname:24s} {ms * 1e3:9.2f} us x {e['launches_per_call']}"
            + (f"  {e['gbs']:8.1f} GB/s  {100 * e['frac']:5.1f}% of HBM peak" if "gbs" in e else ""))
    call("fzh_kernel_timing_enable", 0)
    p.close()
    d_in.free()
    blob.free()

    # the first four parts against the oracle chain; the oracle's time on the host
    info = cpu_info()
    oracle.set_threads(info["threads"])
    t0 = time.perf_counter()
    if kind == "1d":
        want, oinfo = chains.lorenzo_rle(x)
    else:
        want, oinfo = chains.lorenzo_huffman(x, dims)
    cpu_s = time.perf_counter() - t0
    got, ref = split_blob(blob_h), split_blob(want)
    exact = len(got) == 7 and all(np.array_equal(got[i], ref[i]) for i in range(4))
    res = {
        "workload": name,
        "input_bytes": int(nbytes),
        "compressed_bytes": int(out_n),
        "ratio": round(nbytes / out_n, 3),
        "compress_gbs": round(nbytes / comp_s / 1e9, 2),
        "compress_ms": round(comp_s * 1e3, 4),
        "decompress_gbs": round(nbytes / dec_s / 1e9, 2),
        "decompress_ms": round(dec_s * 1e3, 4),
        "outputs": outs,
        "outliers": int(oinfo["cnt"]),
        "outlier_capacity": chains.max_outliers(n, 0.2),
        "abs_eb": abs_eb,
        "max_abs_err": max_err,
        "within_eb": bool(max_err <= abs_eb),
        "bit_exact_vs_oracle": bool(exact),
        "kernels": kern,
        "cpu_oracle": {"gbs": round(nbytes / cpu_s / 1e9, 3), "threads": info["threads"],
                       "sample": "whole field, one run of the oracle chain (value base, Lorenzo, coder) "
                                 "including its numpy padding copy"},
    }
    if R is not None:
        res["runs"] = R
    log(f"[bench] {kind}: compress {res['compress_gbs']} GB/s, decompress {res['decompress_gbs']} GB/s, ratio "
        f"{res['ratio']}, max|err| {max_err:.3e} (eb {abs_eb:.3e}), outliers {res['outliers']}/"
        f"{res['outlier_capacity']}, equal to the oracle: {exact}, cpu oracle {res['cpu_oracle']['gbs']} GB/s")
    return res


def split_offset(blob, i):
    """Byte offset of part i of a pipeline result."""
    n = int(blob[:4].view(np.uint32)[0])
    sizes = blob[4:4 + 8 * n].view(np.uint64).astype(np.int64)
    off = (4 + 8 * n + 15) // 16 * 16
    for sz in sizes[:i]:
        off += (int(sz) + 15) // 16 * 16
    return off


if __name__ == "__main__":
    main()
