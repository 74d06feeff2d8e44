// Bounded blocking FIFO between pipeline threads (glados/Queue.h:56-89 keeps
// the same push/take contract and the default bound of 10 items, :48).
// push() blocks while the queue is full, take() while it is empty; that is
// one half of the pipeline's backpressure (the memory pools are the other).
#pragma once

#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <queue>
#include <utility>

namespace glados {

template <class T>
class Queue {
public:
    explicit Queue(std::size_t limit = 10) : limit_(limit == 0 ? 1 : limit) {}

    Queue(const Queue&) = delete;
    Queue& operator=(const Queue&) = delete;

    void push(T&& item)
    {
        std::unique_lock<std::mutex> lock(m_);
        not_full_.wait(lock, [&] { return q_.size() < limit_; });
        q_.push(std::move(item));
        lock.unlock();
        not_empty_.notify_one();
    }

    T take()
    {
        std::unique_lock<std::mutex> lock(m_);
        not_empty_.wait(lock, [&] { return !q_.empty(); });
        T item = std::move(q_.front());
        q_.pop();
        lock.unlock();
        not_full_.notify_one();
        return item;
    }

    // Takes the front item if there is one, without waiting.
    bool try_take(T& out)
    {
        std::unique_lock<std::mutex> lock(m_);
        if (q_.empty()) return false;
        out = std::move(q_.front());
        q_.pop();
        lock.unlock();
        not_full_.notify_one();
        return true;
    }

    std::size_t size() const
    {
        std::lock_guard<std::mutex> lock(m_);
        return q_.size();
    }

    std::size_t limit() const { return limit_; }

private:
    mutable std::mutex m_;
    std::condition_variable not_full_, not_empty_;
    std::queue<T> q_;
    std::size_t limit_;
};

}  // namespace glados
