// Host-side packing of byte masks (nonzero: allowed) into the bitsets the kernels read.
#pragma once
#include <emmintrin.h>
#include <stdint.h>

#include <vector>

namespace hm {

// Rows of n bytes into rows of ceil(n / 32) words, bit i % 32 of word i / 32 for node i: whole
// words from two 16-byte compares and byte masks (a per-byte branch takes 30 ms for 8 masks of
// 1M nodes on the host, this 1.4 ms).
inline std::vector<uint32_t> pack_masks(const uint8_t *pool, int64_t rows, int64_t n) {
    const int64_t words = (n + 31) / 32;
    std::vector<uint32_t> out((size_t)(rows * words), 0u);
    const __m128i zero = _mm_setzero_si128();
    for (int64_t r = 0; r < rows; ++r) {
        const uint8_t *p = pool + r * n;
        uint32_t *o = out.data() + r * words;
        const int64_t full = n / 32;
        for (int64_t w = 0; w < full; ++w) {
            const __m128i lo = _mm_loadu_si128((const __m128i *)(p + 32 * w));
            const __m128i hi = _mm_loadu_si128((const __m128i *)(p + 32 * w + 16));
            const uint32_t zlo = (uint32_t)_mm_movemask_epi8(_mm_cmpeq_epi8(lo, zero));
            const uint32_t zhi = (uint32_t)_mm_movemask_epi8(_mm_cmpeq_epi8(hi, zero));
            o[w] = ~(zlo | (zhi << 16));
        }
        for (int64_t i = full * 32; i < n; ++i)
            if (p[i]) o[i / 32] |= 1u << (i % 32);
    }
    return out;
}

}  // namespace hm
