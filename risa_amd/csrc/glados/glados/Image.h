// Image: one frame (fan sinogram, parallel sinogram or slice) flowing
// through the pipeline, and MemoryPool: the per-stage buffer pools it comes
// from and returns to.
//
// Interface kept as glados/Image.h:56-210 (data(), size(), index(), plane(),
// device(), start(), valid(); an invalid Image is the end-of-stream
// sentinel; a valid pooled Image goes back to its pool when destroyed) and
// glados/MemoryPool.h:68-125 (registerStage, requestMemory blocks while the
// stage's pool is empty, returnMemory, freeMemory).  Differences: the pool
// singleton is created thread-safely on first use, Images not taken from a
// pool free their own buffer, and a buffer returned after its stage freed the
// pool is released instead of re-pooled.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <stdexcept>
#include <utility>
#include <vector>

namespace glados {

template <class Manager>
class Image;

template <class Manager>
class MemoryPool {
public:
    using image_type = Image<Manager>;
    using pointer_type = typename Manager::pointer_type_1D;
    using size_type = typename Manager::size_type;

    static MemoryPool* instance()
    {
        static MemoryPool pool;
        return &pool;
    }

    // Allocates `count` buffers of `elements` values on `device` for one
    // stage (in one allocation, in address order, where the manager can);
    // returns the stage's pool id.  The first request gets the lowest
    // address.
    int registerStage(size_type count, size_type elements, int device = 0)
    {
        Slot s;
        s.elements = elements;
        s.device = device;
        Manager mm;
        mm.setDevice(device);
        s.free = make_buffers(mm, count, elements, 0);
        std::reverse(s.free.begin(), s.free.end());   // taken from the back
        std::lock_guard<std::mutex> lock(m_);
        slots_.push_back(std::move(s));
        return (int)slots_.size() - 1;
    }

    // Takes a buffer from stage `id`'s pool, waiting while it is empty.
    image_type requestMemory(int id)
    {
        std::unique_lock<std::mutex> lock(m_);
        Slot& s = slot(id);
        cv_.wait(lock, [&] { return !s.free.empty() || s.released; });
        if (s.released) throw std::runtime_error("MemoryPool: stage pool already freed");
        pointer_type p = std::move(s.free.back());
        s.free.pop_back();
        const size_type n = s.elements;
        const int dev = s.device;
        lock.unlock();
        image_type img(n, 0, 0, std::move(p), dev);
        img.setMemPoolIdx(id);
        return img;
    }

    void returnMemory(image_type&& img)
    {
        const int id = img.memPoolIdx();
        pointer_type p = img.release();
        if (id < 0 || !p) return;
        std::lock_guard<std::mutex> lock(m_);
        Slot& s = slot(id);
        if (!s.released) s.free.push_back(std::move(p));
        cv_.notify_all();
    }

    // Releases the stage's pooled buffers; buffers still in flight are
    // released when they come back.
    void freeMemory(int id)
    {
        std::lock_guard<std::mutex> lock(m_);
        Slot& s = slot(id);
        s.released = true;
        s.free.clear();
        cv_.notify_all();
    }

    size_type available(int id)
    {
        std::lock_guard<std::mutex> lock(m_);
        return slot(id).free.size();
    }

private:
    template <class M>
    static auto make_buffers(const M& mm, size_type count, size_type n, int) -> decltype(mm.make_ptrs(count, n))
    {
        return mm.make_ptrs(count, n);
    }
    template <class M>
    static std::vector<pointer_type> make_buffers(const M& mm, size_type count, size_type n, long)
    {
        std::vector<pointer_type> v;
        for (size_type i = 0; i < count; i++) v.push_back(mm.make_ptr(n));
        return v;
    }
    struct Slot {
        std::vector<pointer_type> free;
        size_type elements = 0;
        int device = 0;
        bool released = false;
    };
    Slot& slot(int id)
    {
        if (id < 0 || id >= (int)slots_.size()) throw std::out_of_range("MemoryPool: unknown stage id");
        return slots_[(size_t)id];
    }
    MemoryPool() = default;
    std::mutex m_;
    std::condition_variable cv_;
    std::vector<Slot> slots_;
};

template <class Manager>
class Image : public Manager {
public:
    using manager_type = Manager;
    using value_type = typename Manager::value_type;
    using pointer_type = typename Manager::pointer_type_1D;
    using size_type = typename Manager::size_type;
    using clock = std::chrono::high_resolution_clock;

    // The invalid Image: the end-of-stream sentinel.
    Image() noexcept = default;

    // A standalone Image owning a new buffer (not pooled) or the given one.
    explicit Image(size_type size, size_type idx = 0, size_type plane = 0, pointer_type data = nullptr, int device = 0)
        : size_(size), index_(idx), plane_(plane), valid_(true)
    {
        Manager::setDevice(device);
        data_ = data ? std::move(data) : Manager::make_ptr(size_);
    }

    // Deep copy into a new, unpooled buffer (glados/Image.h:72-83).
    Image(const Image& o)
        : Manager(o), size_(o.size_), index_(o.index_), plane_(o.plane_), valid_(o.valid_), start_(o.start_)
    {
        if (o.data_) {
            data_ = Manager::make_ptr(size_);
            Manager::copy(data_, o.data_, size_);
        }
    }

    Image(Image&& o) noexcept
        : Manager(std::move(o)), size_(o.size_), index_(o.index_), plane_(o.plane_), pool_(o.pool_),
          data_(std::move(o.data_)), valid_(o.valid_), start_(o.start_)
    {
        o.valid_ = false;
        o.pool_ = -1;
    }

    Image& operator=(Image&& o) noexcept
    {
        if (this != &o) {
            giveBack();
            Manager::operator=(std::move(o));
            size_ = o.size_;
            index_ = o.index_;
            plane_ = o.plane_;
            pool_ = o.pool_;
            data_ = std::move(o.data_);
            valid_ = o.valid_;
            start_ = o.start_;
            o.valid_ = false;
            o.pool_ = -1;
        }
        return *this;
    }

    Image& operator=(const Image&) = delete;

    ~Image() { giveBack(); }

    size_type size() const noexcept { return size_; }
    size_type index() const noexcept { return index_; }
    size_type plane() const noexcept { return plane_; }
    bool valid() const noexcept { return valid_; }
    value_type* data() const noexcept { return data_.get(); }
    const pointer_type& container() const noexcept { return data_; }
    clock::time_point start() const noexcept { return start_; }
    int memPoolIdx() const noexcept { return pool_; }

    void setIdx(size_type i) noexcept { index_ = i; }
    void setPlane(size_type p) noexcept { plane_ = p; }
    void setMemPoolIdx(int i) noexcept { pool_ = i; }
    void setStart(clock::time_point t) noexcept { start_ = t; }
    double duration() const { return std::chrono::duration<double, std::milli>(clock::now() - start_).count(); }
    void invalid() noexcept { valid_ = false; }

    // Hands the buffer out (used by the pool); the Image becomes invalid.
    pointer_type release() noexcept
    {
        valid_ = false;
        return std::move(data_);
    }

private:
    void giveBack() noexcept
    {
        if (pool_ >= 0 && data_) {
            try {
                MemoryPool<Manager>::instance()->returnMemory(std::move(*this));
            } catch (...) {
            }
        }
        pool_ = -1;
    }

    size_type size_ = 0, index_ = 0, plane_ = 0;
    int pool_ = -1;
    pointer_type data_{};
    bool valid_ = false;
    clock::time_point start_{};
};

}  // namespace glados
