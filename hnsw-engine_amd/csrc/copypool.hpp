// A few host threads that work through one job together with the calling thread: a buffer copy
// or any function of (part, parts) (plain C++, no HIP: the large-result copies and the query
// packing of the host entry points use it, and a CPU test builds it alone).
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace hm {

// Owned by one index handle, whose calls do not overlap, so run() and copy() are never entered
// twice at once.
class CopyPool {
  public:
    explicit CopyPool(int workers) {
        for (int i = 0; i < workers; ++i) th_.emplace_back([this, i] { run_worker(i + 1); });
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

    size_t parts() const { return th_.size() + 1; }

    // f(p, parts()) for every p in [0, parts()), part 0 on the calling thread; returns when all
    // are done.
    void run(const std::function<void(size_t, size_t)> &f) {
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f;
            pending_ = (int)th_.size();
            ++gen_;
        }
        cv_.notify_all();
        f(0, parts());
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

    // dst[0, n) = src[0, n), split in parts() 64-byte aligned pieces.
    void copy(void *dst, const void *src, size_t n) {
        const size_t per = ((n + parts() - 1) / parts() + 63) & ~size_t(63);  // covers n, never 0
        char *d = (char *)dst;
        const char *s = (const char *)src;
        run([&](size_t p, size_t) {
            const size_t a = p * per;
            if (a < n) memcpy(d + a, s + a, std::min(n, a + per) - a);
        });
    }

  private:
    void run_worker(size_t p) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> l(m_);
        for (;;) {
            cv_.wait(l, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            const std::function<void(size_t, size_t)> *f = job_;
            l.unlock();
            (*f)(p, parts());
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
    const std::function<void(size_t, size_t)> *job_ = nullptr;
};

}  // namespace hm
