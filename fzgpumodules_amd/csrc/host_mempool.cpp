// host_mempool.cpp — fz::MemoryPool over a hipMemPool, and the FZM CRC32.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "fz/mempool.h"
#include "fz/fzm_format.h"

namespace fz {

namespace {
void hip_or_throw(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

MemoryPool::MemoryPool(const MemoryPoolConfig& cfg) {
    if (cfg.device_id >= 0) device_ = cfg.device_id;
    else hip_or_throw(hipGetDevice(&device_), "hipGetDevice");
    const char* env = std::getenv("FZ_FORCE_MEMPOOL_FALLBACK");
    fallback_ = cfg.force_fallback || (env && env[0] && env[0] != '0');
    if (!fallback_) {
        int supported = 0;
        if (hipDeviceGetAttribute(&supported, hipDeviceAttributeMemoryPoolsSupported, device_) != hipSuccess ||
            !supported) {
            fallback_ = true;
        } else {
            hipMemPoolProps props = {};
            props.allocType = hipMemAllocationTypePinned;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = device_;
            if (hipMemPoolCreate(&pool_, &props) != hipSuccess) {
                (void)hipGetLastError();
                pool_ = nullptr;
                fallback_ = true;
            } else {
                // Keep freed blocks resident: steady-state calls reuse them.
                uint64_t thr = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool_, hipMemPoolAttrReleaseThreshold, &thr);
                if (cfg.input_data_size > 0) {
                    // Warm the pool to about pool_multiplier x the input so the
                    // first compress does not grow it piecemeal.
                    const size_t warm = (size_t)((double)cfg.input_data_size * cfg.pool_multiplier);
                    void* p = nullptr;
                    if (hipMallocFromPoolAsync(&p, warm, pool_, nullptr) == hipSuccess) {
                        (void)hipFreeAsync(p, nullptr);
                        (void)hipStreamSynchronize(nullptr);
                    } else {
                        (void)hipGetLastError();
                    }
                }
            }
        }
    }
}

MemoryPool::~MemoryPool() {
    (void)hipDeviceSynchronize();
    for (auto& kv : live_) (void)hipFree(kv.first);
    for (auto& kv : pdev_) (void)hipFree(kv.first);
    for (auto& kv : ppin_) (void)hipHostFree(kv.first);
    if (pool_) (void)hipMemPoolDestroy(pool_);
}

void* MemoryPool::allocate(size_t bytes, hipStream_t stream, const std::string& tag, bool persistent) {
    if (persistent) return allocatePersistentDevice(bytes, tag);
    bytes = bytes ? bytes : 1;
    void* p = nullptr;
    if (!fallback_) {
        if (hipMallocFromPoolAsync(&p, bytes, pool_, stream) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    } else {
        if (stream) (void)hipStreamSynchronize(stream);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    live_[p] = bytes;
    current_ += bytes;
    if (current_ > peak_) peak_ = current_;
    return p;
}

void MemoryPool::free(void* ptr, hipStream_t stream) {
    if (!ptr) return;
    auto it = live_.find(ptr);
    if (it == live_.end()) {
        if (pdev_.count(ptr)) freePersistentDevice(ptr);
        return;
    }
    current_ -= it->second;
    live_.erase(it);
    if (!fallback_) {
        (void)hipFreeAsync(ptr, stream);
    } else {
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(ptr);
    }
}

void* MemoryPool::allocatePersistentDevice(size_t bytes, const std::string&) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    pdev_[p] = bytes;
    persistent_dev_ += bytes;
    return p;
}

void MemoryPool::freePersistentDevice(void* ptr) {
    auto it = pdev_.find(ptr);
    if (it == pdev_.end()) return;
    persistent_dev_ -= it->second;
    pdev_.erase(it);
    (void)hipFree(ptr);
}

void* MemoryPool::allocatePersistentPinned(size_t bytes, const std::string&) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    ppin_[p] = bytes;
    persistent_pin_ += bytes;
    return p;
}

void MemoryPool::freePersistentPinned(void* ptr) {
    auto it = ppin_.find(ptr);
    if (it == ppin_.end()) return;
    persistent_pin_ -= it->second;
    ppin_.erase(it);
    (void)hipHostFree(ptr);
}

void MemoryPool::setReleaseThreshold(uint64_t bytes) {
    if (pool_) (void)hipMemPoolSetAttribute(pool_, hipMemPoolAttrReleaseThreshold, &bytes);
}

// IEEE 802.3 CRC32, slicing-by-8 tables built on first use
// (the checksum of compressor_io.cpp:25-52).
namespace {
struct Crc32Tables {
    uint32_t t[8][256];
    Crc32Tables() {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1u) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
            t[0][i] = c;
        }
        for (uint32_t i = 0; i < 256; i++)
            for (int s = 1; s < 8; s++) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xFFu];
    }
};
const Crc32Tables& crc_tables() {
    static const Crc32Tables tabs;
    return tabs;
}
}  // namespace

uint32_t crc32Feed(uint32_t crc, const void* data, size_t len) {
    const auto& T = crc_tables().t;
    const uint8_t* p = (const uint8_t*)data;
    while (len >= 8) {
        uint32_t lo, hi;
        std::memcpy(&lo, p, 4);
        std::memcpy(&hi, p + 4, 4);
        lo ^= crc;
        crc = T[7][lo & 0xFF] ^ T[6][(lo >> 8) & 0xFF] ^ T[5][(lo >> 16) & 0xFF] ^ T[4][lo >> 24] ^
              T[3][hi & 0xFF] ^ T[2][(hi >> 8) & 0xFF] ^ T[1][(hi >> 16) & 0xFF] ^ T[0][hi >> 24];
        p += 8;
        len -= 8;
    }
    while (len--) crc = T[0][(crc ^ *p++) & 0xFFu] ^ (crc >> 8);
    return crc;
}

}  // namespace fz
