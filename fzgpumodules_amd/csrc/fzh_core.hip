// fzh_core.hip — library plumbing: error strings, workspace init, device
// helpers used by the bindings, and the multi-output gather kernel.
#include "fzh_internal.h"

#include <cstring>

namespace fzh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
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
        for (size_t i = threadIdx.x; i < nw; i += blockDim.x)
            reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
        for (size_t i = nw * 16 + threadIdx.x; i < bytes; i += blockDim.x) dst[i] = src[i];
    } else {
        for (size_t i = threadIdx.x; i < bytes; i += blockDim.x) dst[i] = src[i];
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
