// Device-to-host copies of large results into the caller's (pageable) memory: the device writes
// 4 MB chunks into two pinned staging halves while host threads copy the previous chunk out.
// A pageable hipMemcpy stages through the runtime's own buffers and copies out on one thread
// (about 13 GB/s here); this keeps the DMA at pinned speed and spreads the host copy over four
// threads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <memory>

#include "common.hpp"
#include "copypool.hpp"

namespace hm {

// Pinned staging halves, their events and the copy threads of one handle (made on first use).
struct Staging {
    static constexpr size_t kChunk = size_t(4) << 20;
    // smaller copies go straight through (staging 0.8 MB results measured no gain)
    static constexpr size_t kMin = size_t(2) << 20;
    PinBuf<uint8_t> buf;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::unique_ptr<CopyPool> pool;

    CopyPool &team() {
        if (!pool) pool.reset(new CopyPool(3));
        return *pool;
    }
    hipError_t ready() {
        hipError_t e = buf.ensure(2 * kChunk);
        for (int i = 0; i < 2 && e == hipSuccess; ++i)
            if (!ev[i]) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
        if (e == hipSuccess) team();
        return e;
    }
    void release() {
        pool.reset();
        for (hipEvent_t &e : ev) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
        buf.release();
    }

    // dst (host, pageable) = src (device) over `bytes`, ordered after the work queued on s. A
    // copy below kMin is only queued (the caller's next synchronisation completes it); a larger
    // one returns when dst holds the data.
    hipError_t copy_out(void *dst, const void *src, size_t bytes, hipStream_t s) {
        if (bytes < kMin) return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
        hipError_t e = ready();
        if (e != hipSuccess) return e;
        const size_t nch = (bytes + kChunk - 1) / kChunk;
        auto len = [&](size_t i) { return std::min(kChunk, bytes - i * kChunk); };
        for (size_t i = 0; i <= nch; ++i) {
            // chunk i into half i % 2 (that half's previous chunk, i - 2, was copied out in the
            // previous step), then chunk i - 1 out of the other half
            if (i < nch) {
                e = hipMemcpyAsync(buf.p + (i & 1) * kChunk, (const uint8_t *)src + i * kChunk,
                                   len(i), hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipEventRecord(ev[i & 1], s);
                if (e != hipSuccess) {
                    (void)hipStreamSynchronize(s);
                    return e;
                }
            }
            if (i >= 1) {
                const size_t j = i - 1;
                e = hipEventSynchronize(ev[j & 1]);
                if (e != hipSuccess) {
                    (void)hipStreamSynchronize(s);
                    return e;
                }
                pool->copy((uint8_t *)dst + j * kChunk, buf.p + (j & 1) * kChunk, len(j));
            }
        }
        return hipSuccess;
    }
};

}  // namespace hm
