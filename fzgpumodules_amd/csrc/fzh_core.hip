// fzh_core.hip — library plumbing: error strings, workspace init, device
// helpers used by the bindings, and the multi-output gather kernel.
#include "fzh_synth.h"
#include "fzh_internal.h"
#include "fzh_scan.h"
#include "fzh_concat.h"

#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace fzh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

// ---------------------------------------------------------------------------
// Per-kernel timing (opt-in; see launch_k)
// ---------------------------------------------------------------------------
namespace {
struct KtPending {
    std::string name;
    hipEvent_t a, b;
};
std::atomic<bool> g_kt_on{false};
std::mutex g_kt_mu;
std::vector<KtPending> g_kt_pending;
std::vector<hipEvent_t> g_kt_pool;
std::map<std::string, std::pair<double, uint64_t>> g_kt_acc;

hipEvent_t kt_event() {
    if (!g_kt_pool.empty()) {
        hipEvent_t e = g_kt_pool.back();
        g_kt_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Resolves every pending pair (waits for them); caller holds g_kt_mu.
int kt_drain() {
    int rc = FZH_OK;
    for (auto& p : g_kt_pending) {
        float ms = 0.f;
        hipError_t e = hipEventSynchronize(p.b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
        if (e == hipSuccess) {
            auto& acc = g_kt_acc[p.name];
            acc.first += ms;
            acc.second += 1;
        } else if (rc == FZH_OK) {
            rc = fail(FZH_E_RUNTIME, std::string("kernel timing: ") + hipGetErrorString(e));
        }
        g_kt_pool.push_back(p.a);
        g_kt_pool.push_back(p.b);
    }
    g_kt_pending.clear();
    return rc;
}
}  // namespace

KtEvents ktime_events(hipStream_t s, const char* name) {
    KtEvents e;
    if (!g_kt_on.load(std::memory_order_relaxed)) return e;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return e;
    std::lock_guard<std::mutex> g(g_kt_mu);
    hipEvent_t a = kt_event(), b = kt_event();
    if (!a || !b) {
        if (a) g_kt_pool.push_back(a);
        if (b) g_kt_pool.push_back(b);
        return e;
    }
    g_kt_pending.push_back({name, a, b});
    e.a = a;
    e.b = b;
    return e;
}

// Fill kernel: zeroes a byte range with 16-byte stores (no memset node, so it
// replays correctly inside a captured graph).
__global__ void __launch_bounds__(256) fill_zero_kernel(uint8_t* __restrict__ p, size_t n) {
    const size_t n16 = n / 16;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint4* q = reinterpret_cast<uint4*>(p);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        q[i] = make_uint4(0, 0, 0, 0);
    for (size_t i = n16 * 16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

int zero_bytes(hipStream_t s, void* p, size_t n) {
    if (n == 0) return FZH_OK;
    if (reinterpret_cast<uintptr_t>(p) % 16 != 0) {
        FZH_TRY_HIP(hipMemsetAsync(p, 0, n, s));
        return FZH_OK;
    }
    const size_t blocks = std::min<size_t>(cdiv(cdiv(n, 16), 256), 2048);
    fill_zero_kernel<<<(unsigned)std::max<size_t>(blocks, 1), 256, 0, s>>>((uint8_t*)p, n);
    FZH_CHECK_LAUNCH("fill_zero_kernel");
    return FZH_OK;
}

// One workgroup per segment, 16-byte copies when both ends are aligned
// (src/pipeline/concat_kernel.cu:27-61 semantics).
__global__ void __launch_bounds__(256) gather_kernel(const fzh_copy_desc* __restrict__ d, int n) {
    const int seg = blockIdx.x;
    if (seg >= n) return;
    const uint8_t* src = (const uint8_t*)d[seg].src;
    uint8_t* dst = (uint8_t*)d[seg].dst;
    const size_t bytes = d[seg].bytes;
    if (!src || !dst || !bytes) return;
    if (((uintptr_t)src | (uintptr_t)dst) % 16 == 0) {
        const size_t nw = bytes / 16;
        for (size_t i = threadIdx.x; i < nw; i += blockDim.x) gst_u4(dst + 16 * i, gld_u4(src + 16 * i));
        for (size_t i = nw * 16 + threadIdx.x; i < bytes; i += blockDim.x) gp(dst)[i] = gp(src)[i];
    } else {
        for (size_t i = threadIdx.x; i < bytes; i += blockDim.x) gp(dst)[i] = gp(src)[i];
    }
}

// Grid-stride device copy with 16-byte accesses (both ends 16-byte aligned).
__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t nw) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) copy_tail_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                        size_t from, size_t n) {
    const size_t i = from + threadIdx.x;
    if (i < n) dst[i] = src[i];
}


// One wave: sizes, 16-byte-aligned offsets, the header bytes and the total.
// ws: u64 off[n] then u64 size[n].
__global__ void __launch_bounds__(64)
concat_header_kernel(ConcatArgs a, int n, uint8_t* __restrict__ out, uint64_t* __restrict__ total,
                     uint64_t* __restrict__ ws) {
    const int l = threadIdx.x;
    const uint64_t hdr = (4 + 8 * (uint64_t)n + 15) & ~uint64_t(15);
    const uint64_t sz = l < n ? slot_size(a.slot[l]) : 0;
    const uint64_t padded = (sz + 15) & ~uint64_t(15);
    uint64_t incl = padded;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t t = __shfl_up(incl, off, 64);
        if (l >= off) incl += t;
    }
    if (l < n) {
        ws[l] = hdr + incl - padded;
        ws[FZH_CONCAT_MAX_SLOTS + l] = sz;
        uint32_t* w = reinterpret_cast<uint32_t*>(out + 4 + 8 * l);
        w[0] = (uint32_t)sz;
        w[1] = (uint32_t)(sz >> 32);
    }
    if (l == 0) *reinterpret_cast<uint32_t*>(out) = (uint32_t)n;
    // zero the header pad
    for (uint64_t b = 4 + 8 * (uint64_t)n + l; b < hdr; b += 64) out[b] = 0;
    if (l == 63) *total = hdr + incl;
}

// blockIdx.y = slot, blockIdx.x = 16 KiB piece; pieces past the slot's
// padded size exit.  Pad bytes of the last 16-byte word are written as 0.
__global__ void __launch_bounds__(256)
concat_copy_kernel(ConcatArgs a, uint8_t* __restrict__ out, const uint64_t* __restrict__ ws) {
    const int sidx = blockIdx.y;
    const uint64_t sz = ws[FZH_CONCAT_MAX_SLOTS + sidx];
    const uint64_t nw = (sz + 15) / 16;
    const uint64_t w0 = (uint64_t)blockIdx.x * 1024;
    if (w0 >= nw) return;
    const uint8_t* src = reinterpret_cast<const uint8_t*>(a.slot[sidx].src);
    uint4* dst = reinterpret_cast<uint4*>(out + ws[sidx]);
    if (reinterpret_cast<const void*>(dst) == reinterpret_cast<const void*>(src)) {
        // produced in place: only the pad of the last 16-byte word is due
        if (w0 == 0 && threadIdx.x == 0 && (sz & 15)) {
            uint8_t* d8 = out + ws[sidx];
            for (uint64_t b = sz; b < nw * 16; b++) d8[b] = 0;
        }
        return;
    }
    const uint64_t full = sz / 16;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t w = w0 + threadIdx.x + 256 * k;
        if (w < full) {
            dst[w] = reinterpret_cast<const uint4*>(src)[w];
        } else if (w < nw) {
            uint32_t v[4] = {0, 0, 0, 0};
            for (uint64_t b = 16 * w; b < sz; b++) v[(b & 15) >> 2] |= (uint32_t)src[b] << (8 * (b & 3));
            dst[w] = make_uint4(v[0], v[1], v[2], v[3]);
        }
    }
}

// Concat and readback in one launch (body in fzh_concat.h).
__global__ void __launch_bounds__(256)
concat_readback_kernel(ConcatArgs a, int n, uint8_t* __restrict__ out, uint64_t* __restrict__ total,
                       ReadbackArgs r, int nr, uint8_t* __restrict__ mapped) {
    concat_readback_body(a, n, out, total, r, nr, mapped, blockIdx.x, gridDim.x, SizeOverride{});
}

// out[z][y][x] = base[(z0 + z) % bz][(y0 + y) % by][(x0 + x) % bx]
__global__ void __launch_bounds__(256)
periodic_window_kernel(const float* __restrict__ base, uint64_t bx, uint64_t by, uint64_t bz, uint64_t x0,
                       uint64_t y0, uint64_t z0, float* __restrict__ out, uint64_t nx, uint64_t ny, uint64_t total) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (uint64_t)gridDim.x * 256) {
        const uint64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
        out[i] = base[(((z0 + z) % bz) * by + (y0 + y) % by) * bx + (x0 + x) % bx];
    }
}

__global__ void __launch_bounds__(256)
synth_power_law_kernel(float* __restrict__ out, uint64_t nx, uint64_t ny, uint64_t nz, uint64_t total, int d,
                       int octaves, float ea, uint64_t seed) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256) {
        const uint64_t x = e % nx, y = (e / nx) % ny, z = e / (nx * ny);
        out[e] = power_law_value(x, y, z, d, octaves, ea, seed);
    }
}

// The entries' words flattened over the wave's lanes (one load round trip
// for all of them).
__global__ void __launch_bounds__(64) readback_kernel(ReadbackArgs a, int n, uint8_t* __restrict__ dst) {
    uint32_t all = 0;
    for (int i = 0; i < n; i++) all += a.e[i].bytes / 4;
    for (uint32_t g = threadIdx.x; g < all; g += 64) {
        int i = 0;
        uint32_t base = 0;
        while (base + a.e[i].bytes / 4 <= g) base += a.e[i++].bytes / 4;
        const uint32_t w = g - base;
        reinterpret_cast<uint32_t*>(dst + a.e[i].dst_offset)[w] = reinterpret_cast<const uint32_t*>(a.e[i].src)[w];
    }
}

}  // namespace fzh

using namespace fzh;

extern "C" {

const char* fzh_last_error(void) { return g_last_error.c_str(); }
const char* fzh_version(void) { return "fzgpumodules_amd 0.1 (gfx950)"; }

int fzh_workspace_init(void* stream, void* d_ws, size_t ws_bytes) {
    if (ws_bytes && !d_ws) return fail(FZH_E_INVALID, "fzh_workspace_init: null workspace");
    return zero_bytes(S(stream), d_ws, ws_bytes);
}

int fzh_gather(void* stream, const fzh_copy_desc* d_descs, int n_segs) {
    if (n_segs <= 0) return FZH_OK;
    gather_kernel<<<n_segs, 256, 0, S(stream)>>>(d_descs, n_segs);
    FZH_CHECK_LAUNCH("gather_kernel");
    return FZH_OK;
}

int fzh_copy(void* stream, void* d_dst, const void* d_src, size_t bytes) {
    if (bytes == 0) return FZH_OK;
    if (!d_dst || !d_src) return fail(FZH_E_INVALID, "copy: null pointer");
    if (((uintptr_t)d_dst | (uintptr_t)d_src) % 16 != 0) return fail(FZH_E_INVALID, "copy: pointers must be 16-byte aligned");
    const size_t nw = bytes / 16;
    if (nw) {
        const size_t blocks = std::min<size_t>(cdiv(nw, 256 * 4), 8192);
        launch_k("copy_kernel", copy_kernel, dim3((unsigned)blocks), dim3(256), 0, S(stream), (const uint4*)d_src,
                 (uint4*)d_dst, nw);
    }
    if (bytes % 16)
        copy_tail_kernel<<<1, 256, 0, S(stream)>>>((const uint8_t*)d_src, (uint8_t*)d_dst, nw * 16, bytes);
    FZH_CHECK_LAUNCH("copy_kernel");
    return FZH_OK;
}

int fzh_periodic_window_f32(void* stream, const float* d_base, uint64_t bx, uint64_t by, uint64_t bz, uint64_t x0,
                            uint64_t y0, uint64_t z0, float* d_out, uint64_t nx, uint64_t ny, uint64_t nz) {
    if (!d_base || !d_out) return fail(FZH_E_INVALID, "periodic window: null pointer");
    if (!bx || !by || !bz) return fail(FZH_E_INVALID, "periodic window: empty base");
    const uint64_t total = nx * ny * nz;
    if (total == 0) return FZH_OK;
    const unsigned blocks = (unsigned)std::min<uint64_t>(cdiv(total, 256), 65536);
    periodic_window_kernel<<<blocks, 256, 0, S(stream)>>>(d_base, bx, by, bz, x0, y0, z0, d_out, nx, ny, total);
    FZH_CHECK_LAUNCH("periodic_window_kernel");
    return FZH_OK;
}

int fzh_synth_power_law_f32(void* stream, float* d_out, uint64_t nx, uint64_t ny, uint64_t nz, float expo,
                            uint64_t seed) {
    if (!d_out) return fail(FZH_E_INVALID, "synth: null pointer");
    const uint64_t total = nx * ny * nz;
    if (total == 0) return FZH_OK;
    const int d = nz > 1 ? 3 : (ny > 1 ? 2 : 1);
    const int octaves = power_law_octaves(std::max(nx, std::max(ny, nz)));
    const float ea = -0.5f * (expo + (float)d);  // log2 of the amplitude ratio between octaves o+1 and o
    const unsigned blocks = (unsigned)std::min<uint64_t>(cdiv(total, 256), 65536);
    synth_power_law_kernel<<<blocks, 256, 0, S(stream)>>>(d_out, nx, ny, nz, total, d, octaves, ea, seed);
    FZH_CHECK_LAUNCH("synth_power_law_kernel");
    return FZH_OK;
}

int fzh_kernel_timing_enable(int on) {
    std::lock_guard<std::mutex> g(fzh::g_kt_mu);
    int rc = fzh::kt_drain();
    fzh::g_kt_acc.clear();
    fzh::g_kt_on.store(on != 0);
    return rc;
}

int fzh_kernel_timing_get(const char* name, double* total_ms, uint64_t* count) {
    if (!name || !total_ms || !count) return fail(FZH_E_INVALID, "kernel timing: null argument");
    std::lock_guard<std::mutex> g(fzh::g_kt_mu);
    if (int rc = fzh::kt_drain()) return rc;
    auto it = fzh::g_kt_acc.find(name);
    *total_ms = it == fzh::g_kt_acc.end() ? 0.0 : it->second.first;
    *count = it == fzh::g_kt_acc.end() ? 0 : it->second.second;
    return FZH_OK;
}

size_t fzh_scan_ws_bytes(uint64_t n) { return scan_ws_bytes(n); }

int fzh_scan_u32(void* stream, uint32_t* d_in, uint32_t* d_out, uint64_t n, uint32_t* d_total, int flags,
                 void* d_ws, size_t ws_bytes) {
    if ((!d_in || !d_out) && n) return fail(FZH_E_INVALID, "scan: null pointer");
    if (!d_total) return fail(FZH_E_INVALID, "scan: null total");
    if (flags & ~(FZH_SCAN_WRAP | FZH_SCAN_CLEAR)) return fail(FZH_E_INVALID, "scan: unknown flags");
    if ((flags & FZH_SCAN_CLEAR) && d_in == d_out && n) return fail(FZH_E_INVALID, "scan: clear needs d_in != d_out");
    if (n > (uint64_t)SCAN_SEG * SCAN_MAX_SEGS) return fail(FZH_E_INVALID, "scan: more than 2^28 elements");
    if (n > SCAN_SMALL && (!d_ws || ws_bytes < scan_ws_bytes(n))) return fail(FZH_E_CAPACITY, "scan: workspace too small");
    hipStream_t s = S(stream);
    int rc;
    switch (flags) {
        case 0: rc = scan_u32<false, false>(s, d_in, d_out, n, nullptr, d_total, d_ws); break;
        case FZH_SCAN_WRAP: rc = scan_u32<true, false>(s, d_in, d_out, n, nullptr, d_total, d_ws); break;
        case FZH_SCAN_CLEAR: rc = scan_u32<false, true>(s, d_in, d_out, n, nullptr, d_total, d_ws); break;
        default: rc = scan_u32<true, true>(s, d_in, d_out, n, nullptr, d_total, d_ws); break;
    }
    if (rc) return rc;
    FZH_CHECK_LAUNCH("scan");
    return FZH_OK;
}

size_t fzh_concat_ws_bytes(void) { return 2 * FZH_CONCAT_MAX_SLOTS * sizeof(uint64_t); }

uint64_t fzh_concat_bound(const fzh_concat_slot* slots, int n) {
    uint64_t t = (4 + 8 * (uint64_t)n + 15) & ~uint64_t(15);
    for (int i = 0; i < n; i++) t += (slots[i].cap + 15) & ~uint64_t(15);
    return t;
}

int fzh_concat(void* stream, const fzh_concat_slot* slots, int n, void* d_out, uint64_t out_cap,
               uint64_t* d_total, void* d_ws, size_t ws_bytes) {
    if (n < 0 || n > FZH_CONCAT_MAX_SLOTS) return fail(FZH_E_INVALID, "concat: 0..32 slots");
    if (!d_out || !d_total || !d_ws) return fail(FZH_E_INVALID, "concat: null pointer");
    if (ws_bytes < fzh_concat_ws_bytes()) return fail(FZH_E_CAPACITY, "concat: workspace too small");
    if (reinterpret_cast<uintptr_t>(d_out) % 16 != 0 || reinterpret_cast<uintptr_t>(d_ws) % 8 != 0)
        return fail(FZH_E_INVALID, "concat: output must be 16-byte aligned");
    ConcatArgs a;
    std::memset(&a, 0, sizeof(a));
    uint64_t max_words = 0;
    for (int i = 0; i < n; i++) {
        const fzh_concat_slot& s = slots[i];
        if (s.kind > 2) return fail(FZH_E_INVALID, "concat: slot kind must be 0, 1 or 2");
        if (s.kind != 0 && !s.size_ptr) return fail(FZH_E_INVALID, "concat: null size pointer");
        if (s.cap > 0 && (!s.src || reinterpret_cast<uintptr_t>(s.src) % 16 != 0))
            return fail(FZH_E_INVALID, "concat: slot sources must be 16-byte aligned");
        if (s.kind == 0 && s.size > s.cap) return fail(FZH_E_INVALID, "concat: slot size exceeds its capacity");
        a.slot[i] = s;
        max_words = std::max<uint64_t>(max_words, (s.cap + 15) / 16);
    }
    if (fzh_concat_bound(slots, n) > out_cap) return fail(FZH_E_CAPACITY, "concat: output buffer too small");
    concat_header_kernel<<<1, 64, 0, S(stream)>>>(a, n, (uint8_t*)d_out, d_total, (uint64_t*)d_ws);
    if (n > 0 && max_words > 0) {
        const dim3 grid((unsigned)cdiv(max_words, 1024), (unsigned)n);
        concat_copy_kernel<<<grid, 256, 0, S(stream)>>>(a, (uint8_t*)d_out, (const uint64_t*)d_ws);
    }
    FZH_CHECK_LAUNCH("concat");
    return FZH_OK;
}

int fzh_concat_readback(void* stream, const fzh_concat_slot* slots, int n, void* d_out, uint64_t out_cap,
                        uint64_t* d_total, const fzh_readback_entry* entries, int n_entries, void* d_mapped_dst) {
    ConcatArgs a;
    ReadbackArgs r;
    unsigned grid = 1;
    if (int rc = build_concat_args(slots, n, d_out, out_cap, d_total, entries, n_entries, d_mapped_dst, a, r, grid))
        return rc;
    launch_k("concat_readback_kernel", concat_readback_kernel, dim3(grid), dim3(256), 0, S(stream), a, n,
             (uint8_t*)d_out, d_total, r, n_entries, (uint8_t*)d_mapped_dst);
    FZH_CHECK_LAUNCH("concat_readback_kernel");
    return FZH_OK;
}

}  // extern "C"

namespace fzh {

int read_small(hipStream_t s, const void* d_src, size_t bytes, void* h_dst) {
    constexpr size_t kCap = 64 * 1024;
    thread_local void* host = nullptr;  // this thread's staging buffer: allocated once, never released
    if (bytes == 0) return FZH_OK;
    const bool fast = bytes <= kCap && bytes % 4 == 0 && reinterpret_cast<uintptr_t>(d_src) % 4 == 0;
    if (fast && !host &&
        hipHostMalloc(&host, kCap, hipHostMallocMapped | hipHostMallocPortable | hipHostMallocCoherent) != hipSuccess)
        host = nullptr;
    if (!fast || !host) {
        FZH_TRY_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
        FZH_TRY_HIP(hipStreamSynchronize(s));
        return FZH_OK;
    }
    void* dev = nullptr;
    FZH_TRY_HIP(hipHostGetDevicePointer(&dev, host, 0));
    ReadbackArgs a;
    std::memset(&a, 0, sizeof(a));
    a.e[0] = fzh_readback_entry{d_src, (uint32_t)bytes, 0};
    readback_kernel<<<1, 64, 0, s>>>(a, 1, (uint8_t*)dev);
    FZH_CHECK_LAUNCH("readback_kernel");
    FZH_TRY_HIP(hipStreamSynchronize(s));
    std::memcpy(h_dst, host, bytes);
    return FZH_OK;
}

}  // namespace fzh

extern "C" {

int fzh_readback(void* stream, const fzh_readback_entry* entries, int n, void* d_mapped_dst) {
    if (n <= 0) return FZH_OK;
    if (n > FZH_READBACK_MAX) return fail(FZH_E_INVALID, "readback: at most 32 entries");
    if (!d_mapped_dst || reinterpret_cast<uintptr_t>(d_mapped_dst) % 4 != 0)
        return fail(FZH_E_INVALID, "readback: destination must be 4-byte aligned");
    ReadbackArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int i = 0; i < n; i++) {
        const fzh_readback_entry& e = entries[i];
        if (!e.src || e.bytes % 4 != 0 || e.dst_offset % 4 != 0 || reinterpret_cast<uintptr_t>(e.src) % 4 != 0)
            return fail(FZH_E_INVALID, "readback: entries must be non-null 4-byte multiples");
        a.e[i] = e;
    }
    readback_kernel<<<1, 64, 0, S(stream)>>>(a, n, (uint8_t*)d_mapped_dst);
    FZH_CHECK_LAUNCH("readback_kernel");
    return FZH_OK;
}

int fzh_device_alloc(void** d_ptr, size_t bytes) {
    FZH_TRY_HIP(hipMalloc(d_ptr, bytes ? bytes : 16));
    return FZH_OK;
}
int fzh_device_free(void* d_ptr) {
    FZH_TRY_HIP(hipFree(d_ptr));
    return FZH_OK;
}
int fzh_memcpy_htod(void* d, const void* h, size_t n) {
    if (n) FZH_TRY_HIP(hipMemcpy(d, h, n, hipMemcpyHostToDevice));
    return FZH_OK;
}
int fzh_memcpy_dtoh(void* h, const void* d, size_t n) {
    if (n) FZH_TRY_HIP(hipMemcpy(h, d, n, hipMemcpyDeviceToHost));
    return FZH_OK;
}
int fzh_memset(void* d, int v, size_t n) {
    if (n) FZH_TRY_HIP(hipMemset(d, v, n));
    return FZH_OK;
}
int fzh_synchronize(void) {
    FZH_TRY_HIP(hipDeviceSynchronize());
    return FZH_OK;
}
int fzh_stream_create(void** s) {
    hipStream_t st;
    FZH_TRY_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = st;
    return FZH_OK;
}
int fzh_stream_destroy(void* s) {
    FZH_TRY_HIP(hipStreamDestroy(S(s)));
    return FZH_OK;
}
int fzh_stream_sync(void* s) {
    FZH_TRY_HIP(hipStreamSynchronize(S(s)));
    return FZH_OK;
}
int fzh_event_create(void** ev) {
    hipEvent_t e;
    FZH_TRY_HIP(hipEventCreate(&e));
    *ev = e;
    return FZH_OK;
}
int fzh_event_destroy(void* ev) {
    FZH_TRY_HIP(hipEventDestroy((hipEvent_t)ev));
    return FZH_OK;
}
int fzh_event_record(void* ev, void* s) {
    FZH_TRY_HIP(hipEventRecord((hipEvent_t)ev, S(s)));
    return FZH_OK;
}
int fzh_event_elapsed_ms(void* a, void* b, float* ms) {
    FZH_TRY_HIP(hipEventSynchronize((hipEvent_t)b));
    FZH_TRY_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return FZH_OK;
}
int fzh_device_count(int* count) {
    FZH_TRY_HIP(hipGetDeviceCount(count));
    return FZH_OK;
}

}  // extern "C"
