// fz/mempool.h — stream-ordered device memory pool.
//
// Replaces the reference's MemoryPool (include/mem/mempool.h:82-228,
// src/mem/mempool.cpp) with a hipMemPool: allocations are
// hipMallocFromPoolAsync / hipFreeAsync on the caller's stream, the release
// threshold is pinned to "never trim" so steady-state calls do not re-fault
// pages, and persistent allocations (stage scratch that lives across calls)
// are tracked separately.  Setting FZ_FORCE_MEMPOOL_FALLBACK (or
// MemoryPoolConfig::force_fallback) selects plain hipMalloc/hipFree with
// stream synchronisation instead.  Not thread-safe: one pipeline per thread.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>

namespace fz {

enum class MemoryStrategy { MINIMAL = 0, PREALLOCATE = 1 };

struct MemoryPoolConfig {
    size_t input_data_size = 0;
    float pool_multiplier = 3.0f;
    int device_id = -1;          // -1 = current device
    bool force_fallback = false;
    MemoryPoolConfig() = default;
    MemoryPoolConfig(size_t in, float mult) : input_data_size(in), pool_multiplier(mult) {}
};

class MemoryPool {
  public:
    explicit MemoryPool(const MemoryPoolConfig& cfg = MemoryPoolConfig());
    ~MemoryPool();
    MemoryPool(const MemoryPool&) = delete;
    MemoryPool& operator=(const MemoryPool&) = delete;

    // Stream-ordered allocation; returns nullptr on failure.
    void* allocate(size_t bytes, hipStream_t stream, const std::string& tag = "", bool persistent = false);
    void free(void* ptr, hipStream_t stream);

    // Long-lived stage scratch (device / pinned host).
    void* allocatePersistentDevice(size_t bytes, const std::string& tag = "");
    void freePersistentDevice(void* ptr);
    void* allocatePersistentPinned(size_t bytes, const std::string& tag = "");
    void freePersistentPinned(void* ptr);

    bool isFallbackMode() const { return fallback_; }
    void setReleaseThreshold(uint64_t bytes);
    size_t getCurrentUsage() const { return current_; }
    size_t getPeakUsage() const { return peak_; }
    size_t getPersistentDeviceBytes() const { return persistent_dev_; }
    size_t getPersistentPinnedBytes() const { return persistent_pin_; }
    int deviceId() const { return device_; }

  private:
    hipMemPool_t pool_ = nullptr;
    bool fallback_ = false;
    int device_ = 0;
    size_t current_ = 0, peak_ = 0, persistent_dev_ = 0, persistent_pin_ = 0;
    std::unordered_map<void*, size_t> live_, pdev_, ppin_;
};

}  // namespace fz
