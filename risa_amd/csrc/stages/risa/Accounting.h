// Frame accounting of the copy stages: lost-frame counting from index gaps
// (H2D, Copy/H2D.cu:86-110) and the running reconstruction rate (D2H,
// Copy/D2H.cu:82-96).  Host-only; thread-compatible (each stage calls it from
// its process(), which the pipeline drives from one thread).
#pragma once

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <limits>

namespace risa {

// Counts the indices missing between consecutive arrivals: an arrival with
// index i after index l adds i - l - 1 to the lost count (H2D.cu:102-103),
// and the loss is reported against the last index seen (H2D.cu:81,107).
// Unlike the reference, whose last index starts at 0 (so a first index of 0
// counts -1 into its unsigned total), the count starts before index 0: a
// stream 0, 1, 2, ... without gaps loses nothing.  Out-of-order arrivals
// subtract, as the reference's arithmetic does.  The time between two
// arrivals feeds the best/worst inter-arrival figures (H2D.cu:89-100).
class FrameLossCounter {
public:
    void arrived(std::size_t index)
    {
        const auto now = std::chrono::steady_clock::now();
        if (count_ > 0) {
            const double dt = std::chrono::duration<double>(now - last_time_).count();
            if (dt < best_) best_ = dt;
            if (dt > worst_) worst_ = dt;
        }
        last_time_ = now;
        lost_ += (long long)index - last_index_ - 1;
        last_index_ = (long long)index;
        count_++;
    }
    long long lost() const { return lost_; }
    std::size_t arrivals() const { return count_; }
    long long last_index() const { return last_index_; }
    // lost / last index * 100 (H2D.cu:81), 0 before a positive index
    double loss_percent() const { return last_index_ > 0 ? (double)lost_ / (double)last_index_ * 100.0 : 0.0; }
    double best_case_seconds() const { return count_ > 1 ? best_ : 0.0; }
    double worst_case_seconds() const { return worst_; }

private:
    long long lost_ = 0;
    long long last_index_ = -1;
    std::size_t count_ = 0;
    double best_ = std::numeric_limits<double>::max(), worst_ = 0.0;
    std::chrono::steady_clock::time_point last_time_{};
};

// Reconstruction rate.  Every `interval` images the rate over that window is
// folded into a running mean (D2H.cu:85-91: window rate = interval / elapsed,
// mean_n = (mean_{n-1} (n - 1) + rate) / n); the window timer restarts at
// index 0 and after each window.  images_per_second() is the whole-run
// figure (images / time from the first arrival to the last), which is
// defined before the first full window too.
class RateMeter {
public:
    explicit RateMeter(std::size_t interval = 10000) : interval_(interval == 0 ? 1 : interval) {}

    void arrived(std::size_t index)
    {
        const auto now = std::chrono::steady_clock::now();
        if (count_ == 0) first_ = now;
        if (index == 0 || count_ == 0) window_start_ = now;
        if ((count_ % interval_) == interval_ - 1) {
            const double secs = std::chrono::duration<double>(now - window_start_).count();
            last_window_rate_ = secs > 0 ? (double)interval_ / secs : 0.0;
            mean_rate_ = (mean_rate_ * windows_ + last_window_rate_) / (windows_ + 1.0);
            windows_ += 1.0;
            window_start_ = now;
        }
        last_ = now;
        count_++;
    }
    std::size_t images() const { return count_; }
    // the reference's running mean of window rates (0 before a full window)
    double reconstruction_rate() const { return mean_rate_; }
    double last_window_rate() const { return last_window_rate_; }
    std::size_t windows() const { return (std::size_t)windows_; }
    double images_per_second() const
    {
        const double secs = std::chrono::duration<double>(last_ - first_).count();
        return count_ > 1 && secs > 0 ? (double)(count_ - 1) / secs : 0.0;
    }

private:
    std::size_t interval_;
    std::size_t count_ = 0;
    double windows_ = 0.0, mean_rate_ = 0.0, last_window_rate_ = 0.0;
    std::chrono::steady_clock::time_point first_{}, last_{}, window_start_{};
};

}  // namespace risa
