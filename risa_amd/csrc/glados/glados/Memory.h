// Memory managers: the allocation policy an Image inherits
// (glados/cuda/DeviceMemoryManager.h, HostMemoryManager.h keep the same
// roles).  A manager names value_type, size_type and pointer_type_1D, makes
// and copies buffers, and remembers the device its buffers live on.
//
// This header holds the plain host manager (operator new[]); the HIP
// managers (device memory, pinned host memory) are in glados/hip/Memory.h.
#pragma once

#include <cstddef>
#include <cstring>
#include <memory>

namespace glados {

template <class T>
class PlainHostMemoryManager {
public:
    using value_type = T;
    using size_type = std::size_t;
    struct Deleter {
        void operator()(T* p) const noexcept { delete[] p; }
    };
    using pointer_type_1D = std::unique_ptr<T[], Deleter>;

    int device() const noexcept { return device_; }
    void setDevice(int d) noexcept { device_ = d; }

    pointer_type_1D make_ptr(size_type n) const { return pointer_type_1D(new T[n]); }
    void copy(pointer_type_1D& dst, const pointer_type_1D& src, size_type n) const
    {
        std::memcpy(dst.get(), src.get(), n * sizeof(T));
    }

protected:
    int device_ = 0;
};

}  // namespace glados
