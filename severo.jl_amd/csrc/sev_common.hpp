// Shared device/host helpers for libsevero_hip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

namespace sev {

constexpr int kWave = 64;

// Error carried across the C-ABI as a status code plus a message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
    kOk = 0,
    kErrArg = -1,
    kErrHip = -2,
    kErrConvergence = -3,
    kErrNoMem = -4,
    kErrNoDevice = -5,
    kErrInternal = -6,
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char *what, const char *file, int line) {
    if (e != hipSuccess) {
        char buf[512];
        std::snprintf(buf, sizeof buf, "%s failed at %s:%d: %s", what, file, line,
                      hipGetErrorString(e));
        fail(e == hipErrorOutOfMemory ? kErrNoMem : kErrHip, buf);
    }
}
#define SEV_HIP(x) ::sev::hip_check((x), #x, __FILE__, __LINE__)
#define SEV_LAUNCH_CHECK() ::sev::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)
#define SEV_REQUIRE(cond, msg)                                \
    do {                                                      \
        if (!(cond)) ::sev::fail(::sev::kErrArg, (msg));      \
    } while (0)

// Grow-only device buffer.
template <class T>
struct DevBuf {
    T *ptr = nullptr;
    size_t count = 0;  // capacity in elements
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    T *ensure(size_t n) {
        if (n > count || ptr == nullptr) {
            release();
            size_t bytes = (n > 0 ? n : 1) * sizeof(T);
            SEV_HIP(hipMalloc((void **)&ptr, bytes));
            count = n > 0 ? n : 1;
        }
        return ptr;
    }
    T *get() const { return ptr; }
    void swap(DevBuf &o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(count, o.count);
    }
};

// Grow-only pinned host buffer (hipHostMalloc), for small copies that must
// not go through the pageable staging path.
template <class T>
struct HostPinned {
    T *ptr = nullptr;
    size_t count = 0;
    HostPinned() = default;
    HostPinned(const HostPinned &) = delete;
    HostPinned &operator=(const HostPinned &) = delete;
    ~HostPinned() {
        if (ptr) (void)hipHostFree(ptr);
    }
    T *ensure(size_t n) {
        if (n > count || ptr == nullptr) {
            if (ptr) (void)hipHostFree(ptr);
            ptr = nullptr;
            count = 0;
            SEV_HIP(hipHostMalloc((void **)&ptr, (n > 0 ? n : 1) * sizeof(T), hipHostMallocDefault));
            count = n > 0 ? n : 1;
        }
        return ptr;
    }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline int device_cu_count() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        SEV_HIP(hipGetDevice(&dev));
        SEV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (cus <= 0) cus = 256;
    }
    return cus;
}

inline int device_lds_per_block() {
    static int lds = 0;
    if (lds == 0) {
        int dev = 0;
        SEV_HIP(hipGetDevice(&dev));
        SEV_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (lds <= 0) lds = 65536;
    }
    return lds;
}

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Butterfly sum; the value of lane 0 is returned to every lane so that all
// lanes agree bit for bit (the butterfly pairs differ per lane).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return __shfl(v, 0, kWave);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return __shfl(v, 0, kWave);
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Counter-based uniform in [0,1): identical to the host definition used by
// the synthetic generator.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ double u01(uint64_t seed_mixed, uint32_t stream,
                                               uint32_t index) {
    uint64_t h = mix64(seed_mixed ^ (((uint64_t)stream << 32) | (uint64_t)index));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace sev
