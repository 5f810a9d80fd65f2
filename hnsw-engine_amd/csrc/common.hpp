// Host-side helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <sys/mman.h>
#include <unistd.h>

#include <string>

namespace hm {

void set_error(const std::string &msg);
int fail_hip(hipError_t e, const char *what);

#define HM_TRY(expr, what)                            \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return fail_hip(_e, what); \
    } while (0)

// Grow-only device buffer.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    // Reallocates when `count` exceeds the capacity; a new allocation is zero-filled when
    // `zero` is set (contents are not preserved).
    hipError_t ensure(size_t count, bool zero = false, hipStream_t s = nullptr) {
        if (count <= cap && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = count ? count : 1;
        hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
        if (e != hipSuccess) return e;
        cap = want;
        if (zero) e = hipMemsetAsync(p, 0, want * sizeof(T), s);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Grow-only pinned host buffer (staging for host-to-device copies at full link speed).
template <class T>
struct PinBuf {
    T *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t count) {
        if (count <= cap && p) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        size_t want = count ? count : 1;
        hipError_t e = hipHostMalloc((void **)&p, want * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) return e;
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Maps every page of a host output range ahead of the device-to-host copy that fills it. Called
// while the kernels run: the pages of a fresh allocation otherwise fault in one at a time
// inside the copy, after the kernels have ended (about 0.15 ms per MB of output). Only pages
// wholly inside the range are populated (MADV_POPULATE_WRITE maps them and changes no data). A
// performance hint only: where the call is refused (an older kernel, memory that is not
// ordinary host memory) nothing is done and the copy maps the pages itself.
inline void prefault_output(void *p, size_t bytes) {
    static const uintptr_t page = (uintptr_t)sysconf(_SC_PAGESIZE);
    if (!p || bytes < page) return;
    const uintptr_t a = ((uintptr_t)p + page - 1) & ~(page - 1);
    const uintptr_t e = ((uintptr_t)p + bytes) & ~(page - 1);
    if (e > a) (void)madvise((void *)a, e - a, MADV_POPULATE_WRITE);
}

inline int ceil_log2(int64_t v) {
    int r = 0;
    while ((int64_t(1) << r) < v) ++r;
    return r;
}

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace hm
