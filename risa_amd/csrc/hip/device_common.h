// Shared pieces of the HIP side of librisa_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../host/common.h"

#define RISA_HIP_CHECK(expr)                                                                     \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            ::risa::fail(RISA_ERR_HIP, "HIP call failed at %s:%d: %s", __FILE__, __LINE__,      \
                         hipGetErrorString(_e));                                                 \
    } while (0)

namespace risa {

// Makes `device` current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int device)
    {
        RISA_HIP_CHECK(hipGetDevice(&prev));
        if (prev != device) RISA_HIP_CHECK(hipSetDevice(device));
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// Owning device buffer.
template <class T>
struct DeviceBuffer {
    T* ptr = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void resize(size_t n)
    {
        release();
        if (n == 0) return;
        void* p = nullptr;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess)
            fail(RISA_ERR_ALLOC, "hipMalloc of %zu bytes failed", n * sizeof(T));
        ptr = (T*)p;
        count = n;
    }
    void upload(const T* host, size_t n, hipStream_t s = nullptr)
    {
        if (n > count) fail(RISA_ERR_ARG, "upload beyond buffer");
        RISA_HIP_CHECK(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
        RISA_HIP_CHECK(hipStreamSynchronize(s));
    }
    void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Waits for all work on `device` (the current device is restored).  Destroy
// calls run this before releasing buffers, graphs or streams a launch on
// any stream may still use: handles do not know their callers' streams.
inline void drain_device(int device)
{
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) return;
    if (prev != device) (void)hipSetDevice(device);
    (void)hipDeviceSynchronize();
    if (prev != device && prev >= 0) (void)hipSetDevice(prev);
}

inline void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(RISA_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// IEEE binary16 conversions on the device.
__device__ __forceinline__ float half_bits_to_float(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }
__device__ __forceinline__ uint16_t float_to_half_bits(float v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }

}  // namespace risa
