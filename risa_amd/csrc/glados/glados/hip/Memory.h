// HIP memory managers for Images: device memory (hipMalloc on the Image's
// device) and pinned host memory (hipHostMalloc, the staging side of H2D/D2H
// copies).  Roles as glados/cuda/DeviceMemoryManager.h and
// HostMemoryManager.h; copies are plain hipMemcpy (stages do their own
// asynchronous copies on their streams).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace glados {
namespace hip {

inline void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Sets the device for the lifetime of the object, restoring the previous one.
class ScopedDevice {
public:
    explicit ScopedDevice(int d)
    {
        check(hipGetDevice(&prev_), "hipGetDevice");
        if (prev_ != d) check(hipSetDevice(d), "hipSetDevice");
    }
    ~ScopedDevice() { (void)hipSetDevice(prev_); }

private:
    int prev_ = 0;
};

template <class T>
class DeviceMemoryManager {
public:
    using value_type = T;
    using size_type = std::size_t;
    struct Deleter {
        void operator()(T* p) const noexcept { (void)hipFree(p); }
    };
    using pointer_type_1D = std::unique_ptr<T[], Deleter>;

    int device() const noexcept { return device_; }
    void setDevice(int d) noexcept { device_ = d; }

    pointer_type_1D make_ptr(size_type n) const
    {
        ScopedDevice sd(device_);
        void* p = nullptr;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw std::bad_alloc();
        return pointer_type_1D(static_cast<T*>(p));
    }
    void copy(pointer_type_1D& dst, const pointer_type_1D& src, size_type n) const
    {
        check(hipMemcpy(dst.get(), src.get(), n * sizeof(T), hipMemcpyDefault), "hipMemcpy");
    }

protected:
    int device_ = 0;
};

// Elements between the starts of two pieces of a pool allocation: n values
// rounded up to 256 bytes, so every piece keeps the alignment copies and
// kernels expect.  Pieces are adjacent (stride n) when n values fill whole
// 256-byte blocks.
template <class T>
inline std::size_t piece_stride(std::size_t n)
{
    return (n * sizeof(T) + 255) / 256 * 256 / sizeof(T);
}

template <class T>
class HostMemoryManager {
public:
    using value_type = T;
    using size_type = std::size_t;
    // A buffer is its own allocation, or (slab set) a piece of a shared one
    // that is freed with its last piece.
    struct Deleter {
        std::shared_ptr<void> slab;
        void operator()(T* p) const noexcept
        {
            if (!slab) (void)hipHostFree(p);
        }
    };
    using pointer_type_1D = std::unique_ptr<T[], Deleter>;

    int device() const noexcept { return device_; }
    void setDevice(int d) noexcept { device_ = d; }

    pointer_type_1D make_ptr(size_type n) const
    {
        void* p = nullptr;
        if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
        return pointer_type_1D(static_cast<T*>(p));
    }
    // `count` buffers of n values carved in address order from one pinned
    // allocation (a pool's buffers): a copy of adjacent buffers can then go
    // as one transfer.
    std::vector<pointer_type_1D> make_ptrs(size_type count, size_type n) const
    {
        const size_type stride = piece_stride<T>(n);
        void* p = nullptr;
        if (hipHostMalloc(&p, count * stride * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
        const std::shared_ptr<void> slab(p, [](void* q) { (void)hipHostFree(q); });
        std::vector<pointer_type_1D> v;
        v.reserve(count);
        for (size_type i = 0; i < count; i++) v.emplace_back(static_cast<T*>(p) + i * stride, Deleter{slab});
        return v;
    }
    void copy(pointer_type_1D& dst, const pointer_type_1D& src, size_type n) const
    {
        check(hipMemcpy(dst.get(), src.get(), n * sizeof(T), hipMemcpyDefault), "hipMemcpy");
    }

protected:
    int device_ = 0;
};

}  // namespace hip
}  // namespace glados
