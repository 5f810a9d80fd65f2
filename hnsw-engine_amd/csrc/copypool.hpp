// A few host threads that copy one buffer together with the calling thread (plain C++, no HIP:
// the large-result copies of hostcopy.hpp use it, and a CPU test builds it alone).
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace hm {

// Owned by one index handle, whose calls do not overlap, so copy() is never entered twice at
// once.
class CopyPool {
  public:
    explicit CopyPool(int workers) {
        for (int i = 0; i < workers; ++i) th_.emplace_back([this, i] { run(i + 1); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    CopyPool(const CopyPool &) = delete;
    CopyPool &operator=(const CopyPool &) = delete;

    // dst[0, n) = src[0, n), split in (workers + 1) 64-byte aligned parts; returns when done.
    void copy(void *dst, const void *src, size_t n) {
        const size_t parts = th_.size() + 1;
        {
            std::lock_guard<std::mutex> g(m_);
            d_ = (char *)dst;
            s_ = (const char *)src;
            n_ = n;
            per_ = ((n + parts - 1) / parts + 63) & ~size_t(63);  // covers n, never 0
            pending_ = (int)th_.size();
            ++gen_;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [&] { return pending_ == 0; });
    }

  private:
    void part(size_t p) {
        const size_t a = p * per_;
        if (a >= n_) return;
        memcpy(d_ + a, s_ + a, std::min(n_, a + per_) - a);
    }
    void run(size_t p) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> l(m_);
        for (;;) {
            cv_.wait(l, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            l.unlock();
            part(p);
            l.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }

    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    bool stop_ = false;
    uint64_t gen_ = 0;
    int pending_ = 0;
    char *d_ = nullptr;
    const char *s_ = nullptr;
    size_t n_ = 0, per_ = 0;
};

}  // namespace hm
