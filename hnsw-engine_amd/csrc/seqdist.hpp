// The reference's float64 distance (numba_engine.py:57-70) of one row per lane: float64, left
// to right, no fused multiply-add: l2 = sum (x_i - q_i)^2, ip = -sum x_i q_i, cosine =
// 1 - sum x_i q_i. The row is read in 16-byte chunks, eight at a time, and the next eight are
// in flight while the current ones are summed, so a row of up to 128 bytes costs one memory
// latency, and a longer one about one per 128 bytes rather than one per chunk. The query (dim
// float64 or float32 values in global memory, the same for every lane) is read through the
// scalar cache into scalar registers, off the chain of dependent adds (SQ = false: read by
// ordinary loads, for a query in LDS or one per group of lanes); a float32 query value converts
// to float64 exactly, as the reference's does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "search.hpp"
#include "wave.hpp"

namespace hm {

// Adds the elements of one 16-byte chunk (16 bytes of U8 codes or 4 floats), element i0 first,
// the ones at dim and beyond skipped. L2: (x - q)^2 terms, else x q terms (the metric is a
// template argument so that each element computes only its own terms).
// q[i] as float64, by a scalar load (the address is wave-uniform and the data read-only).
// SQ false: the query lies in LDS or differs between lanes (an ordinary load).
template <class QT, bool SQ = true> HM_DEV double qval(const QT *q, int i) {
    if constexpr (SQ) {
        using CP = const __attribute__((address_space(4))) QT *;
        return (double)((CP)q)[i];
    } else {
        return (double)q[i];
    }
}

template <bool U8, bool L2, class QT, bool SQ = true>
HM_DEV void seq_chunk(double &acc, uint4 w, const QT *q, int i0, int dim) {
#pragma clang fp contract(off)
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    constexpr int kPer = U8 ? 16 : 4;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int i = i0 + e;
        if (i >= dim) break;
        const double x = U8 ? (double)((ws[e >> 2] >> (8 * (e & 3))) & 255u)
                            : (double)__uint_as_float(ws[e]);
        const double qi = qval<QT, SQ>(q, i);
        if constexpr (L2) {
            const double d = x - qi;
            acc = acc + d * d;
        } else {
            acc = acc + x * qi;
        }
    }
}

// seq_chunk for a chunk wholly inside the row: no per-element bounds checks, so the chunk's
// query values load together (one wait) instead of one scalar round trip per element.
template <bool U8, bool L2, class QT, bool SQ = true>
HM_DEV void seq_chunk_full(double &acc, uint4 w, const QT *q, int i0) {
#pragma clang fp contract(off)
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    constexpr int kPer = U8 ? 16 : 4;
    double qv[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) qv[e] = qval<QT, SQ>(q, i0 + e);
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const double x = U8 ? (double)((ws[e >> 2] >> (8 * (e & 3))) & 255u)
                            : (double)__uint_as_float(ws[e]);
        if constexpr (L2) {
            const double d = x - qv[e];
            acc = acc + d * d;
        } else {
            acc = acc + x * qv[e];
        }
    }
}

// The sum over one row for one metric family (L2: sum (x - q)^2; else sum x q).
template <bool U8, bool L2, class QT, bool SQ = true>
HM_DEV double seq_sum(const void *row, const QT *q, int dim) {
#pragma clang fp contract(off)
    double acc = 0.0;
    if (!U8 && (dim & 3) != 0) {  // float rows not 16-byte aligned: element by element
        const float *r = (const float *)row;
        for (int i = 0; i < dim; ++i) {
            const double qi = qval<QT, SQ>(q, i);
            if constexpr (L2) {
                const double d = (double)r[i] - qi;
                acc = acc + d * d;
            } else {
                acc = acc + (double)r[i] * qi;
            }
        }
    } else {
        constexpr int kPer = U8 ? 16 : 4;
        // Chunks in flight: a whole 8-bit row up to 128 values (one memory round trip); float
        // rows stream eight chunks ahead.
        constexpr int kB = 8;
        const uint4 *r4 = (const uint4 *)row;
        const int nch = (dim + kPer - 1) / kPer;
        uint4 cur[kB];
#pragma unroll
        for (int t = 0; t < kB; ++t)
            if (t < nch) cur[t] = r4[t];
        for (int b0 = 0; b0 < nch; b0 += kB) {
            uint4 nxt[kB];
#pragma unroll
            for (int t = 0; t < kB; ++t)
                if (b0 + kB + t < nch) nxt[t] = r4[b0 + kB + t];
#pragma unroll
            for (int t = 0; t < kB; ++t) {
                const int i0 = (b0 + t) * kPer;
                if (i0 + kPer <= dim) seq_chunk_full<U8, L2, QT, SQ>(acc, cur[t], q, i0);
                else if (i0 < dim) seq_chunk<U8, L2, QT, SQ>(acc, cur[t], q, i0, dim);
            }
#pragma unroll
            for (int t = 0; t < kB; ++t) cur[t] = nxt[t];
        }
    }
    return acc;
}

// The metric is the same for every lane, so this branch is uniform: each lane runs the loop of
// its metric only.
template <bool U8, class QT, bool SQ = true>
HM_DEV double seq_row(const void *row, const QT *q, int dim, int metric) {
#pragma clang fp contract(off)
    if (metric == kL2) return seq_sum<U8, true, QT, SQ>(row, q, dim);
    const double acc = seq_sum<U8, false, QT, SQ>(row, q, dim);
    return metric == kIP ? -acc : 1.0 - acc;
}

}  // namespace hm
