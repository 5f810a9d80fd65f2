// Distance keys: order-preserving integer images of distances, so that (distance, id) tuples
// compare as Python tuples do in the reference (numba_engine.py:116-148) with integer compares.
#pragma once
#include <stdint.h>

#include "search.hpp"
#include "wave.hpp"

namespace hm {

constexpr uint64_t kInfKey = ~0ull;
constexpr uint32_t kInfId = ~0u;

// Order-preserving map of a finite double to uint64 (-0.0 and +0.0 map alike, as Python's tuple
// comparison treats them).
HM_DEV uint64_t key_of_double(double d) {
    if (d == 0.0) d = 0.0;
    uint64_t b = (uint64_t)__double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

HM_DEV double double_of_key(uint64_t k, int metric) {
    uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double d = __longlong_as_double((long long)b);
    // ip returns -acc, and an exactly-zero acc gives -0.0 in the reference.
    if (metric == kIP && d == 0.0) d = -0.0;
    return d;
}

HM_DEV uint64_t key_of_exact(int32_t d, int metric) {
    return metric == kIP ? (uint64_t)((uint32_t)d ^ 0x80000000u) : (uint64_t)(uint32_t)d;
}

HM_DEV double double_of_exact(uint64_t k, int metric) {
    if (metric == kIP) {
        int32_t v = (int32_t)((uint32_t)k ^ 0x80000000u);
        return v == 0 ? -0.0 : (double)v;
    }
    return (double)(uint32_t)k;
}

HM_DEV bool less(uint64_t ak, uint32_t ai, uint64_t bk, uint32_t bi) {
    return ak < bk || (ak == bk && ai < bi);
}

// Sorts (k, id)[0..n) ascending in place; p2 is a power of two >= n and the arrays hold p2
// entries (the tail is padded with +inf). Bitonic network over LDS, one wave.
HM_DEV void bitonic_sort(uint64_t *k, uint32_t *id, int n, int p2) {
    const int lane = lane_id();
    for (int i = n + lane; i < p2; i += 64) {
        k[i] = kInfKey;
        id[i] = kInfId;
    }
    wave_sync();
    for (int size = 2; size <= p2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < (p2 >> 1); t += 64) {
                const int i = 2 * t - (t & (stride - 1));
                const int j = i + stride;
                const bool asc = (i & size) == 0;
                const uint64_t ak = k[i], bk = k[j];
                const uint32_t ai = id[i], bi = id[j];
                if (less(bk, bi, ak, ai) == asc) {
                    k[i] = bk; id[i] = bi;
                    k[j] = ak; id[j] = ai;
                }
            }
            wave_sync();
        }
    }
}

}  // namespace hm
