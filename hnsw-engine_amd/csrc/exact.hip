// Exact k-nearest-neighbour search on the GPU (the reference's oracle, oracle.py:35-126, and the
// masked brute force of benchmarks/run_filtered.py:76-106): the k smallest (distance, id)
// pairs of each query over all (allowed) rows, distances in float64, ties to the lower id.
//
// 8-bit integer data and queries (values 0..255, l2 or ip) use int8 matrix cores. Values are
// shifted to -128..127; v_mfma_i32_32x32x32_i8 accumulates x'.q' exactly in int32 and
// |x-q|^2 = |x'|^2 + |q'|^2 - 2 x'.q' (ip: x.q = x'.q' + 128 (sum x' + sum q') + 128^2 dim) is an
// exact integer, so no error margin and no re-rank are needed: the float64 value of that integer
// is the reference's float64 distance bit for bit. Two consumers:
//   * the range search of exact_tile.hip (dim <= 256, k <= 128, >= 8192 rows): a sampled bound
//     per query, then one scan that records only the rows the bound admits;
//   * the chunked kernels below (every other case, and the queries the range search flags):
//     every work-group owns a block of queries and a chunk of rows and keeps, per query, a small
//     LDS buffer of candidates below a threshold (the k-th smallest (key, id) seen so far); when
//     a buffer nears capacity it is sorted and cut to k. Per-chunk top-k lists are merged per
//     query by a second kernel.
// Anything else runs the chunked kernels with float64 on the vector units, each (query, row) pair
// summed left to right without fusion; cosine divides rows and queries by their float64 norms
// first (oracle.py:228-229).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/hnsw_mi355x.h"
#include "common.hpp"
#include "exact_tile.hpp"
#include "index.hpp"
#include "keys.hpp"
#include "maskbits.hpp"
#include "wave.hpp"

typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(16))) int i32x16;

namespace hm {

struct ExactArgs {
    int64_t n;
    int dim, dimp, metric, k, nq;
    // float64 producer
    const float *x32, *q32;
    const double *xn, *qn;  // cosine: rows / queries divided by their norms
    // int8 producer: rows read as bytes ^ flip (shifted values), queries (nq, dimp) shifted and
    // zero padded; per-row/query sums of the shifted values
    I8Rows rows;
    const int8_t *q8;
    const int32_t *xsq, *qsq, *xsum, *qsum;
    const uint32_t *allowed;  // bitset or null
    int64_t chunk;
    int nchunks;
    int cb;  // per-query buffer capacity (power of two)
    uint64_t *part_key;
    uint32_t *part_id;
};

// ---------------------------------------------------------------------------- top-k buffers

struct TopK {
    uint64_t *key;  // [qb * cb]
    uint32_t *id;
    int *cnt;       // [qb]
    uint64_t *tk;   // [qb] threshold
    uint32_t *ti;
    int cb, k;
};

HM_DEV void topk_insert(const TopK &T, int q, uint64_t key, uint32_t id) {
    if (!less(key, id, T.tk[q], T.ti[q])) return;
    const int pos = atomicAdd(&T.cnt[q], 1);
    T.key[q * T.cb + pos] = key;
    T.id[q * T.cb + pos] = id;
}

// Sorts query q's buffer and keeps its k smallest; the k-th becomes the threshold. One wave.
HM_DEV void topk_compact(const TopK &T, int q) {
    const int m = min(T.cnt[q], T.cb);
    uint64_t *k = T.key + q * T.cb;
    uint32_t *id = T.id + q * T.cb;
    bitonic_sort(k, id, m, T.cb);
    const int keep = min(m, T.k);
    if (lane_id() == 0) {
        T.cnt[q] = keep;
        if (keep == T.k) {
            T.tk[q] = k[T.k - 1];
            T.ti[q] = id[T.k - 1];
        }
    }
    wave_sync();
}

HM_DEV TopK topk_carve(uint8_t *lds, int qb, int cb, int k) {
    TopK T;
    T.key = (uint64_t *)lds;
    T.id = (uint32_t *)(lds + (size_t)qb * cb * 8);
    T.cnt = (int *)(lds + (size_t)qb * cb * 12);
    T.tk = (uint64_t *)(lds + (size_t)qb * cb * 12 + qb * 4 + ((qb * 4) % 8));
    T.ti = (uint32_t *)((uint8_t *)T.tk + qb * 8);
    T.cb = cb;
    T.k = k;
    return T;
}

HM_DEV int topk_bytes(int qb, int cb) { return qb * cb * 12 + qb * 4 + ((qb * 4) % 8) + qb * 12; }

HM_DEV bool row_allowed(const ExactArgs &a, int64_t j) {
    return a.allowed == nullptr || ((a.allowed[j >> 5] >> (j & 31)) & 1u);
}

// Writes each query's k smallest of this chunk to the partial lists.
HM_DEV void topk_flush(const ExactArgs &a, const TopK &T, int q0, int qcount, int wave,
                       int nwaves, int chunk) {
    for (int q = wave; q < qcount; q += nwaves) {
        topk_compact(T, q);
        const int m = T.cnt[q];
        const size_t base = ((size_t)(q0 + q) * a.nchunks + chunk) * a.k;
        for (int t = lane_id(); t < a.k; t += 64) {
            a.part_key[base + t] = t < m ? T.key[q * T.cb + t] : kInfKey;
            a.part_id[base + t] = t < m ? T.id[q * T.cb + t] : kInfId;
        }
    }
}

// ---------------------------------------------------------------------------- int8 producer

// Work-group: 1 or 2 waves (blockDim / 64), 32 queries each (one 32x32 MFMA row block), all
// rows of one chunk in blocks of 32. Operands: lane l holds row/column l % 32 and bytes [16 (l / 32), +16) of each
// 32-byte slice of the dimension, identically for A and B, so the products pair up.
template <int MAXCH>
__global__ __launch_bounds__(128) void exact_i8_kernel(ExactArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    const int lane = lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int qb = 32 * nwaves;
    const int q0 = blockIdx.x * qb;
    const int chunk = blockIdx.y;
    const int qcount = min(qb, a.nq - q0);
    TopK T = topk_carve(lds, qb, a.cb, a.k);
    for (int q = threadIdx.x; q < qb; q += blockDim.x) {
        T.cnt[q] = 0;
        T.tk[q] = kInfKey;
        T.ti[q] = kInfId;
    }
    __syncthreads();
    const int nch = a.dimp >> 5;
    const int half = lane >> 5, col = lane & 31;
    // A operand (queries) stays in registers for the whole chunk.
    const int my_q = q0 + wave * 32 + col;
    i32x4 av[MAXCH];
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
        av[c] = i32x4{0, 0, 0, 0};
        if (c < nch && my_q < a.nq)
            av[c] = *(const i32x4 *)(a.q8 + (size_t)my_q * a.dimp + c * 32 + half * 16);
    }
    // Query terms for the 16 accumulator rows this lane sees: i = (r%4) + 4 half + 8 (r/4).
    int qterm[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int qi = q0 + wave * 32 + (r & 3) + 4 * half + 8 * (r >> 2);
        qterm[r] = qi < a.nq ? (a.metric == kL2 ? a.qsq[qi] : a.qsum[qi]) : 0;
    }
    const int64_t r_begin = (int64_t)chunk * a.chunk;
    const int64_t r_end = min(a.n, r_begin + a.chunk);
    uint64_t tk[16];
    uint32_t ti[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        tk[r] = kInfKey;
        ti[r] = kInfId;
    }
    for (int64_t r0 = r_begin; r0 < r_end; r0 += 32) {
        // Keep every buffer able to absorb one block (at most 32 inserts per query).
        bool need = false;
        for (int q = wave * 32; q < wave * 32 + 32; ++q) need |= T.cnt[q] > T.cb - 32;
        if (need) {
            for (int q = wave * 32; q < wave * 32 + 32; ++q)
                if (T.cnt[q] > T.cb - 32) topk_compact(T, q);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int ql = wave * 32 + (r & 3) + 4 * half + 8 * (r >> 2);
                tk[r] = T.tk[ql];
                ti[r] = T.ti[ql];
            }
        }
        const int64_t j = r0 + col;
        const bool jv = j < r_end;
        i32x4 bv[MAXCH];
#pragma unroll
        for (int c = 0; c < MAXCH; ++c) {
            bv[c] = i32x4{0, 0, 0, 0};
            const int off = c * 32 + half * 16;
            if (c < nch && jv && off < a.rows.ld) {
                const uint4 w = *(const uint4 *)(a.rows.x + (size_t)j * a.rows.ld + off);
                const uint32_t f = a.rows.flip;
                bv[c] = i32x4{(int)(w.x ^ f), (int)(w.y ^ f), (int)(w.z ^ f), (int)(w.w ^ f)};
            }
        }
        i32x16 acc = {0};
#pragma unroll
        for (int c = 0; c < MAXCH; ++c)
            if (c < nch) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[c], bv[c], acc, 0, 0, 0);
        if (jv && row_allowed(a, j)) {
            const int32_t xt = a.metric == kL2 ? a.xsq[j] : a.xsum[j];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int ql = wave * 32 + (r & 3) + 4 * half + 8 * (r >> 2);
                int32_t d;
                if (a.metric == kL2) d = xt + qterm[r] - 2 * acc[r];
                else d = -(acc[r] + 128 * (xt + qterm[r]) + 16384 * a.dim);
                const uint64_t key = key_of_exact(d, a.metric);
                if (q0 + ql < a.nq && less(key, (uint32_t)j, tk[r], ti[r]))
                    topk_insert(T, ql, key, (uint32_t)j);
            }
        }
        wave_sync();
    }
    __syncthreads();
    topk_flush(a, T, q0, qcount, wave, nwaves, chunk);
}

// ---------------------------------------------------------------------------- float64 producer

// Work-group: 256 threads; qb queries (64 or 16) x (4096 / qb) rows per step, 4 x 4 pairs per
// thread, the dimension in slabs of 16 staged in LDS as float64; each pair's sum runs over the
// dimension in order.
__global__ __launch_bounds__(256) void exact_f64_kernel(ExactArgs a, int qb) {
#pragma clang fp contract(off)
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    constexpr int S = 16;
    const int qgroups = qb / 4, rgroups = 256 / qgroups, step = 4 * rgroups;
    double *qs = (double *)lds;  // [qb][S]
    double *xs = qs + qb * S;    // [step][S]
    TopK T = topk_carve(lds + (size_t)(qb + step) * S * 8, qb, a.cb, a.k);
    const int t = threadIdx.x, wave = t >> 6;
    const int q0 = blockIdx.x * qb;
    const int chunk = blockIdx.y;
    const int qcount = min(qb, a.nq - q0);
    for (int q = t; q < qb; q += 256) {
        T.cnt[q] = 0;
        T.tk[q] = kInfKey;
        T.ti[q] = kInfId;
    }
    const int qg = t / rgroups, rg = t % rgroups;
    const int64_t r_begin = (int64_t)chunk * a.chunk;
    const int64_t r_end = min(a.n, r_begin + a.chunk);
    __syncthreads();
    for (int64_t r0 = r_begin; r0 < r_end; r0 += step) {
        // Room for one step (at most `step` inserts per query).
        for (int q = wave; q < qcount; q += 4)
            if (T.cnt[q] > T.cb - step) topk_compact(T, q);
        __syncthreads();
        double acc[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
        for (int s0 = 0; s0 < a.dim; s0 += S) {
            for (int e = t; e < qb * S; e += 256) {
                const int row = e / S, s = e % S, dd = s0 + s;
                const int qi = q0 + row;
                double qv = 0.0;
                if (dd < a.dim && qi < a.nq)
                    qv = a.qn ? a.qn[(size_t)qi * a.dim + dd] : (double)a.q32[(size_t)qi * a.dim + dd];
                qs[row * S + s] = qv;
            }
            for (int e = t; e < step * S; e += 256) {
                const int row = e / S, s = e % S, dd = s0 + s;
                const int64_t xj = r0 + row;
                double xv = 0.0;
                if (dd < a.dim && xj < r_end)
                    xv = a.xn ? a.xn[(size_t)xj * a.dim + dd] : (double)a.x32[(size_t)xj * a.dim + dd];
                xs[row * S + s] = xv;
            }
            __syncthreads();
            const int slab = min(S, a.dim - s0);
            for (int s = 0; s < slab; ++s) {
                double qv[4], xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) qv[u] = qs[(qg * 4 + u) * S + s];
#pragma unroll
                for (int v = 0; v < 4; ++v) xv[v] = xs[(rg * 4 + v) * S + s];
                if (a.metric == kL2) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const double d = xv[v] - qv[u];
                            acc[u][v] = acc[u][v] + d * d;
                        }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][v] = acc[u][v] + xv[v] * qv[u];
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ql = qg * 4 + u;
            if (q0 + ql >= a.nq) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t j = r0 + rg * 4 + v;
                if (j >= r_end || !row_allowed(a, j)) continue;
                const double d = a.metric == kL2 ? acc[u][v] : (a.metric == kIP ? -acc[u][v] : 1.0 - acc[u][v]);
                topk_insert(T, ql, key_of_double(d), (uint32_t)j);
            }
        }
        __syncthreads();
    }
    topk_flush(a, T, q0, qcount, wave, 4, chunk);
}

// ---------------------------------------------------------------------------- merge

__global__ __launch_bounds__(64) void exact_merge_kernel(const uint64_t *pk, const uint32_t *pi,
                                                         int nchunks, int k, int p2, int exact_int,
                                                         int metric, int64_t *ids, double *dists) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    uint64_t *key = (uint64_t *)lds;
    uint32_t *id = (uint32_t *)(lds + (size_t)p2 * 8);
    const int q = blockIdx.x, lane = lane_id();
    const int m = nchunks * k;
    for (int t = lane; t < m; t += 64) {
        key[t] = pk[(size_t)q * m + t];
        id[t] = pi[(size_t)q * m + t];
    }
    wave_sync();
    bitonic_sort(key, id, m, p2);
    for (int t = lane; t < k; t += 64) {
        const bool ok = key[t] != kInfKey;
        ids[(size_t)q * k + t] = ok ? (int64_t)id[t] : -1;
        dists[(size_t)q * k + t] =
            !ok ? __builtin_huge_val() : (exact_int ? double_of_exact(key[t], metric) : double_of_key(key[t], metric));
    }
}

// Row norms for cosine (float64, left to right) and the normalised float64 copy.
__global__ void normalise_kernel(const float *x, int64_t rows, int dim, double *out) {
#pragma clang fp contract(off)
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double s = 0.0;
    for (int i = 0; i < dim; ++i) {
        const double v = (double)x[(size_t)r * dim + i];
        s = s + v * v;
    }
    const double nrm = sqrt(s);
    for (int i = 0; i < dim; ++i) out[(size_t)r * dim + i] = (double)x[(size_t)r * dim + i] / nrm;
}

// 8-bit rows -> shifted int8 (zero padded to dimp) plus sum of squares and plain sum. One wave
// per row (four rows per 256-thread block), consecutive lanes on consecutive values.
__global__ __launch_bounds__(256) void shift_i8_kernel(const float *x, int64_t rows, int dim,
                                                       int dimp, int8_t *out, int32_t *sq,
                                                       int32_t *sum) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int lane = lane_id();
    int32_t s2 = 0, s1 = 0;
    for (int i = lane; i < dimp; i += 64) {
        const int v = i < dim ? (int)x[(size_t)r * dim + i] - 128 : 0;
        out[(size_t)r * dimp + i] = (int8_t)v;
        s2 += v * v;
        s1 += v;
    }
    s2 = wave_sum(s2);
    s1 = wave_sum(s1);
    if (lane == 0) {
        sq[r] = s2;
        sum[r] = s1;
    }
}

// Copies rows of (k int64 ids, k doubles) from the compact list back to their queries.
__global__ void exact_scatter_kernel(const int64_t *sid, const double *sd, const int32_t *qidx,
                                     int64_t m, int k, int64_t *ids, double *dists) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * k) return;
    const int64_t r = e / k;
    const int t = (int)(e % k);
    ids[(int64_t)qidx[r] * k + t] = sid[e];
    dists[(int64_t)qidx[r] * k + t] = sd[e];
}

}  // namespace hm

using namespace hm;

namespace {

bool u8_values(const float *v, size_t count) {
    for (size_t i = 0; i < count; ++i) {
        const float x = v[i];
        if (!(x >= 0.0f && x <= 255.0f) || x != floorf(x)) return false;
    }
    return true;
}

template <class T>
struct Scoped {
    T *p = nullptr;
    hipError_t alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
    }
    ~Scoped() {
        if (p) (void)hipFree(p);
    }
};

constexpr int64_t kTileMinRows = 8192;     // smaller inputs go to the chunked kernels
constexpr int kTileQueriesPerLaunch = 16384;

// Kernel times of the last range search on this thread (HIP events on its stream): bound pass
// plus threshold, range pass, collect; and the number of queries re-run on the chunked kernels.
struct ExactTiming {
    double bound_ms = 0, range_ms = 0, collect_ms = 0;
    int64_t redo = 0;
};
thread_local ExactTiming g_timing;

struct Events {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Events() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// The chunked kernels: per-chunk top-k lists in LDS, merged per query. Works for every k, dim
// and metric; `a` holds the device inputs (int8 or float64 producer) and a.nq queries.
int run_chunked(ExactArgs a, bool int8, const hipDeviceProp_t &prop, int64_t *d_ids,
                double *d_dists, hipStream_t s) {
    const int64_t n = a.n, nq = a.nq;
    const int k = a.k;
    if (nq == 0) return HM_OK;
    // Queries per work-group: as many as the per-query buffers allow in LDS.
    int qb = 0, cb = 0, lds = 0;
    for (int cand : int8 ? std::vector<int>{64, 32} : std::vector<int>{64, 16}) {
        const int step_rows = int8 ? 32 : 4096 / cand;
        const int c = pow2_at_least(k + step_rows);
        const int need = cand * c * 12 + cand * 4 + cand * 12 + 8 +
                         (int8 ? 0 : (cand + step_rows) * 16 * 8);
        if (need <= 160 * 1024) {
            qb = cand;
            cb = c;
            lds = need;
            break;
        }
    }
    if (!qb) return invalid("k too large for the exact search");
    a.cb = cb;
    const int qblocks = (int)((nq + qb - 1) / qb);
    int nchunks = (int)std::max<int64_t>(
        1, std::min<int64_t>((4 * prop.multiProcessorCount + qblocks - 1) / qblocks, (n + 2047) / 2048));
    nchunks = std::max(1, std::min(nchunks, 8192 / k));
    a.chunk = round_up((n + nchunks - 1) / nchunks, 64);
    nchunks = (int)((n + a.chunk - 1) / a.chunk);
    a.nchunks = nchunks;
    const int m = nchunks * k;
    const int p2 = pow2_at_least(m);
    if (p2 * 12 > 160 * 1024) return invalid("k too large for the merge");
    Scoped<uint64_t> pk;
    Scoped<uint32_t> pi;
    HM_TRY(pk.alloc((size_t)nq * m), "alloc");
    HM_TRY(pi.alloc((size_t)nq * m), "alloc");
    a.part_key = pk.p;
    a.part_id = pi.p;
    if (int8) {
        const int nch = a.dimp / 32;
        void (*fn)(ExactArgs) = nch <= 4    ? exact_i8_kernel<4>
                                : nch <= 8  ? exact_i8_kernel<8>
                                : nch <= 16 ? exact_i8_kernel<16>
                                            : exact_i8_kernel<32>;
        if (lds > 64 * 1024)
            HM_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds),
                   "attribute");
        hipLaunchKernelGGL(fn, dim3(qblocks, nchunks), dim3(2 * qb), lds, s, a);
    } else {
        if (lds > 64 * 1024)
            HM_TRY(hipFuncSetAttribute((const void *)exact_f64_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds),
                   "attribute");
        hipLaunchKernelGGL(exact_f64_kernel, dim3(qblocks, nchunks), dim3(256), lds, s, a, qb);
    }
    HM_TRY(hipGetLastError(), "exact launch");
    const int mlds = p2 * 12;
    if (mlds > 64 * 1024)
        HM_TRY(hipFuncSetAttribute((const void *)exact_merge_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, mlds),
               "attribute");
    hipLaunchKernelGGL(exact_merge_kernel, dim3((unsigned)nq), dim3(64), mlds, s, pk.p, pi.p, nchunks,
                       k, p2, int8 ? 1 : 0, a.metric, d_ids, d_dists);
    HM_TRY(hipGetLastError(), "merge launch");
    // The partial lists are freed on return: finish with them first.
    HM_TRY(hipStreamSynchronize(s), "exact search");
    return HM_OK;
}

// Geometry of one range-search launch over m rows and nqb queries.
struct TilePlan {
    int nch = 0, qb = 0, T = 0, groups = 0, nchunks = 0, sample_every = 1;
    int E = 8;      // records per (query, chunk, half)
    int cap = 0;    // candidates per query in the collect kernel
    int64_t chunk = 0;
};

// Tile-shape overrides of the range search (hm_set_exact_tuning; 0: automatic), process-wide.
std::atomic<int> g_tile_qb{0}, g_tile_sample{0};
// Float data on the float64 chunked kernels only (hm_set_exact_tuning with query_blocks -1), for
// tests that compare the two paths.
std::atomic<bool> g_float_chunked{false};

TilePlan plan_tile(int64_t m, int64_t nqb, int nch, int k) {
    TilePlan p;
    p.nch = nch;
    const int qbmax = nch <= 4 ? 4 : 2;
    p.qb = nqb <= 256 ? 1 : (nqb <= 512 || qbmax == 2 ? 2 : 4);
    if (g_tile_qb.load() > 0) p.qb = g_tile_qb.load();
    p.qb = std::min(p.qb, qbmax);
    p.groups = (int)((nqb + kTileWaves * p.qb * 32 - 1) / (kTileWaves * p.qb * 32));
    // About four work-groups per compute unit of MI355X (256 CUs), chunks of >= 4096 rows.
    p.nchunks = (int)std::max<int64_t>(1, std::min<int64_t>(1024 / p.groups, m / 4096));
    p.nchunks = std::min(p.nchunks, kTileMaxChunks);
    p.chunk = round_up((m + p.nchunks - 1) / p.nchunks, kTileRows);
    p.nchunks = (int)((m + p.chunk - 1) / p.chunk);
    const int nsb = (int)(p.chunk / kTileRows);
    // A sixth of each chunk (at least two superblocks): measured on 10,000 x 1M x 128 (bound,
    // range and collect together), every 2nd / 3rd / 4th / 5th / 6th / 8th / 16th superblock
    // 1.74 / 1.58 / 1.42-1.56 / 1.40 / 1.39-1.50 / 1.39-1.52 / 1.62 ms: a sparser sample makes
    // the bound pass cheaper but its looser threshold costs more in the range pass (more 16-row
    // groups reach it) and in the collect.
    p.sample_every = std::max(1, std::min(6, nsb / 2));
    if (g_tile_sample.load() > 0) p.sample_every = std::min(g_tile_sample.load(), nsb);
    p.T = k <= 32 ? 4 : 8;
    // The threshold admits about k * sample_every rows per query (the k-th of a 1/sample_every
    // sample; twice that for an unlucky sample), spread over 2 * nchunks record lists: room for
    // twice the mean plus four standard deviations of a Poisson count.
    const double mu = 2.0 * k * p.sample_every / (2.0 * p.nchunks);
    p.E = (int)std::min(256.0, std::max(8.0, std::ceil(mu + 4.0 * std::sqrt(mu) + 4.0)));
    p.cap = std::min(8192, std::max(1024, pow2_at_least(4 * k * p.sample_every)));
    return p;
}

// Padded width of the float range search for dim (multiples of 8 floats, of 32 above 128), or 0
// when it does not apply.
int ftile_width(int dim) {
    int64_t kp = round_up(dim, 8);
    if (kp > 128) kp = round_up(kp, 32);
    return kp <= 256 ? (int)kp : 0;
}

TilePlan plan_ftile(int64_t m, int64_t nqb, int kp, int k) {
    TilePlan p;
    p.qb = ftile_qb(kp);
    p.groups = (int)((nqb + kTileWaves * p.qb * 32 - 1) / (kTileWaves * p.qb * 32));
    p.nchunks = (int)std::max<int64_t>(1, std::min<int64_t>(1024 / p.groups, m / 4096));
    p.nchunks = std::min(p.nchunks, kTileMaxChunks);
    p.chunk = round_up((m + p.nchunks - 1) / p.nchunks, kTileRows);
    p.nchunks = (int)((m + p.chunk - 1) / p.chunk);
    const int sbr = kp <= 64 ? 128 : (kp <= 128 ? 64 : 32);  // rows per superblock
    const int nsb = (int)(p.chunk / sbr);
    // Every sixth superblock, as for 8-bit rows: on the 400K x 100 angular set (10,000 queries)
    // every 4th / 6th / 8th / 12th 8.23 / 7.94 / 7.94 / 8.10 ms for the three kernels.
    p.sample_every = std::max(1, std::min(6, nsb / 2));
    if (g_tile_sample.load() > 0) p.sample_every = std::min(g_tile_sample.load(), nsb);
    p.T = k <= 32 ? 4 : 8;
    // As for the 8-bit plan; the margin admits a few more rows than the bare sample would.
    const double mu = 2.0 * k * p.sample_every / (2.0 * p.nchunks);
    p.E = (int)std::min(256.0, std::max(8.0, std::ceil(mu + 4.0 * std::sqrt(mu) + 4.0)));
    p.cap = std::min(8192, std::max(1024, pow2_at_least(4 * k * p.sample_every)));
    return p;
}

// Float data: the range search over float32 matrix-core tiles (exact_ftile.hip); the queries it
// flags re-run on the float64 chunked kernels. Sets done = false (and returns HM_OK) when the
// input is outside what the tiles cover (too few rows, too wide, or magnitudes near float32
// overflow), for the caller to run the chunked kernels instead.
int exact_float_tiles(const ExactArgs &a, int kp, const hipDeviceProp_t &prop, int64_t *d_ids,
                      double *d_dists, hipStream_t s, bool &done) {
    done = false;
    const int64_t n = a.n, nq = a.nq;
    const int dim = a.dim, k = a.k, metric = a.metric;
    int64_t m = n;
    Scoped<int32_t> idmap;
    Scoped<int64_t> dcount;
    Scoped<float> xc;
    const float *xs = a.x32;
    int64_t ld = dim;
    int xdim = dim;
    if (a.allowed) {
        HM_TRY(idmap.alloc(n), "alloc");
        HM_TRY(dcount.alloc(1), "alloc");
        HM_TRY(subset_ids_launch(a.allowed, n, idmap.p, dcount.p, s), "subset launch");
        HM_TRY(hipMemcpyAsync(&m, dcount.p, 8, hipMemcpyDeviceToHost, s), "copy count");
        HM_TRY(hipStreamSynchronize(s), "subset count");
        if (m < kTileMinRows || m < k) return HM_OK;
    }
    if (a.allowed || dim % 4 != 0 || (uintptr_t)a.x32 % 16 != 0) {
        HM_TRY(xc.alloc((size_t)m * kp), "alloc");
        HM_TRY(fgather_rows_launch(a.x32, dim, a.allowed ? idmap.p : nullptr, m, kp, xc.p, s),
               "gather launch");
        xs = xc.p;
        ld = kp;
        xdim = kp;
    }
    Scoped<float> cinit, rscale, qp;
    Scoped<double> qnorm;
    Scoped<unsigned long long> maxsq;
    HM_TRY(cinit.alloc(m), "alloc");
    if (metric == HM_COSINE) HM_TRY(rscale.alloc(m), "alloc");
    HM_TRY(maxsq.alloc(1), "alloc");
    HM_TRY(qp.alloc((size_t)nq * kp), "alloc");
    HM_TRY(qnorm.alloc(nq), "alloc");
    HM_TRY(hipMemsetAsync(maxsq.p, 0, 8, s), "clear");
    HM_TRY(fprep_rows_launch(xs, ld, xdim, m, metric, cinit.p, rscale.p, maxsq.p, s), "prep launch");
    HM_TRY(fprep_queries_launch(a.q32, nq, dim, kp, metric, qp.p, qnorm.p, s), "prep launch");
    // Magnitudes: the float32 scores must stay far from overflow.
    unsigned long long msq_bits = 0;
    std::vector<double> hq((size_t)nq);
    HM_TRY(hipMemcpyAsync(&msq_bits, maxsq.p, 8, hipMemcpyDeviceToHost, s), "copy norms");
    HM_TRY(hipMemcpyAsync(hq.data(), qnorm.p, (size_t)nq * 8, hipMemcpyDeviceToHost, s), "copy norms");
    HM_TRY(hipStreamSynchronize(s), "prep");
    double msq;
    memcpy(&msq, &msq_bits, 8);
    const double qmax = *std::max_element(hq.begin(), hq.end());
    if (!(msq < 1e30) || !(qmax * std::sqrt(msq) < 1e30)) return HM_OK;

    Scoped<int32_t> bnd, trig, rec_mask, rec_row, rec_cnt, flags;
    Scoped<float> V;
    HM_TRY(flags.alloc(nq), "alloc");
    size_t cap_bnd = 0, cap_rec = 0, cap_cnt = 0, cap_q = 0;
    Events ev;
    for (hipEvent_t &x : ev.e) HM_TRY(hipEventCreate(&x), "event");
    double times[3] = {0, 0, 0};
    for (int64_t b0 = 0; b0 < nq; b0 += kTileQueriesPerLaunch) {
        const int64_t nqb = std::min<int64_t>(kTileQueriesPerLaunch, nq - b0);
        const TilePlan p = plan_ftile(m, nqb, kp, k);
        const int per_query = p.nchunks * 2 * p.T;
        if (per_query < 2 * k) return HM_OK;
        const size_t slots = (size_t)nqb * p.nchunks * 2;
        if (slots * p.T > cap_bnd) {
            HM_TRY(bnd.alloc(slots * p.T), "alloc");
            cap_bnd = slots * p.T;
        }
        if (slots * p.E > cap_rec) {
            HM_TRY(rec_mask.alloc(slots * p.E), "alloc");
            HM_TRY(rec_row.alloc(slots * p.E), "alloc");
            cap_rec = slots * p.E;
        }
        if (slots > cap_cnt) {
            HM_TRY(rec_cnt.alloc(slots), "alloc");
            cap_cnt = slots;
        }
        if ((size_t)nqb > cap_q) {
            HM_TRY(trig.alloc(nqb), "alloc");
            HM_TRY(V.alloc(nqb), "alloc");
            cap_q = nqb;
        }
        FTileArgs t;
        memset(&t, 0, sizeof(t));
        t.x = xs;
        t.ld = ld;
        t.dim = xdim;
        t.rscale = rscale.p;
        t.cinit = cinit.p;
        t.n = m;
        t.q = qp.p + (size_t)b0 * kp;
        t.nq = (int)nqb;
        t.chunk = p.chunk;
        t.nchunks = p.nchunks;
        t.sample_every = p.sample_every;
        t.bnd = bnd.p;
        t.trig = V.p;
        t.E = p.E;
        t.rec_mask = rec_mask.p;
        t.rec_row = rec_row.p;
        t.rec_cnt = rec_cnt.p;
        const dim3 grid(p.nchunks, p.groups);
        HM_TRY(hipEventRecord(ev.e[0], s), "event");
        HM_TRY(ftile_launch(0, kp, p.T, t, grid, s), "bound launch");
        HM_TRY(threshold_launch(bnd.p, per_query, k, trig.p, (int)nqb, s), "threshold launch");
        HM_TRY(fmargin_launch(trig.p, qnorm.p + b0, maxsq.p, metric, kp, dim, (int)nqb, V.p, s),
               "margin launch");
        HM_TRY(hipEventRecord(ev.e[1], s), "event");
        HM_TRY(ftile_launch(1, kp, 1, t, grid, s), "range launch");
        HM_TRY(hipEventRecord(ev.e[2], s), "event");
        FCollectArgs c;
        memset(&c, 0, sizeof(c));
        c.rec_mask = rec_mask.p;
        c.rec_row = rec_row.p;
        c.rec_cnt = rec_cnt.p;
        c.idmap = a.allowed ? idmap.p : nullptr;
        c.x = a.x32;
        c.q = a.q32 + (size_t)b0 * dim;
        c.n = m;
        c.nchunks = p.nchunks;
        c.E = p.E;
        c.k = k;
        c.metric = metric;
        c.dim = dim;
        c.cap = p.cap;
        c.ids = d_ids + b0 * k;
        c.dists = d_dists + b0 * k;
        c.flag = flags.p + b0;
        HM_TRY(fcollect_launch(c, (int)nqb, s), "collect launch");
        HM_TRY(hipEventRecord(ev.e[3], s), "event");
        HM_TRY(hipEventSynchronize(ev.e[3]), "range search");
        for (int i = 0; i < 3; ++i) {
            float ms = 0.f;
            HM_TRY(hipEventElapsedTime(&ms, ev.e[i], ev.e[i + 1]), "event time");
            times[i] += ms;
        }
    }
    g_timing.bound_ms = times[0];
    g_timing.range_ms = times[1];
    g_timing.collect_ms = times[2];
    std::vector<int32_t> hflag((size_t)nq);
    HM_TRY(hipMemcpyAsync(hflag.data(), flags.p, (size_t)nq * 4, hipMemcpyDeviceToHost, s), "copy flags");
    HM_TRY(hipStreamSynchronize(s), "range search");
    std::vector<int32_t> redo;
    for (int64_t i = 0; i < nq; ++i)
        if (hflag[(size_t)i]) redo.push_back((int32_t)i);
    g_timing.redo = (int64_t)redo.size();
    done = true;
    if (redo.empty()) return HM_OK;

    // Flagged queries: the float64 chunked kernels over all rows (with the bitset when masked).
    const int64_t nr = (int64_t)redo.size();
    Scoped<int32_t> ridx;
    Scoped<float> rq;
    Scoped<double> dxn, dqn, rd;
    Scoped<int64_t> rids;
    HM_TRY(ridx.alloc(nr), "alloc");
    HM_TRY(rq.alloc((size_t)nr * dim), "alloc");
    HM_TRY(rids.alloc((size_t)nr * k), "alloc");
    HM_TRY(rd.alloc((size_t)nr * k), "alloc");
    HM_TRY(hipMemcpyAsync(ridx.p, redo.data(), (size_t)nr * 4, hipMemcpyHostToDevice, s), "copy list");
    HM_TRY(fgather_rows_launch(a.q32, dim, ridx.p, nr, dim, rq.p, s), "gather launch");
    ExactArgs r = a;
    r.q32 = rq.p;
    r.nq = (int)nr;
    if (metric == HM_COSINE) {
        HM_TRY(dxn.alloc((size_t)n * dim), "alloc");
        HM_TRY(dqn.alloc((size_t)nr * dim), "alloc");
        hipLaunchKernelGGL(normalise_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           a.x32, n, dim, dxn.p);
        hipLaunchKernelGGL(normalise_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s,
                           rq.p, nr, dim, dqn.p);
        r.xn = dxn.p;
        r.qn = dqn.p;
    }
    const int st = run_chunked(r, false, prop, rids.p, rd.p, s);
    if (st != HM_OK) return st;
    hipLaunchKernelGGL(exact_scatter_kernel, dim3((unsigned)((nr * k + 255) / 256)), dim3(256), 0, s,
                       rids.p, rd.p, ridx.p, nr, k, d_ids, d_dists);
    HM_TRY(hipGetLastError(), "scatter launch");
    HM_TRY(hipStreamSynchronize(s), "exact search");
    return HM_OK;
}

// Per-device buffers of the 8-bit exact search, kept between calls. A search holds the lock for
// its whole run and waits for its stream before returning, so the next holder finds them idle.
struct ExactScratch {
    std::mutex mu;
    DevBuf<int8_t> x8, q8, cx8;
    DevBuf<int32_t> xsq, xsum, qsq, qsum, cinit, idmap, cxsq, cxsum;
    DevBuf<int32_t> bnd, trig, rec_mask, rec_row, rec_cnt, flags;
    DevBuf<int64_t> count;
};

ExactScratch &exact_scratch(int device) {
    static std::mutex m;
    static std::map<int, std::unique_ptr<ExactScratch>> all;
    std::lock_guard<std::mutex> g(m);
    std::unique_ptr<ExactScratch> &p = all[device];
    if (!p) p.reset(new ExactScratch());
    return *p;
}

// 8-bit rows and their per-row terms. cinit (the range search's accumulator starts over
// n + tile_pad_rows(n) rows) is set when the rows are ready for the tiles as they are: ld equal
// to the tiles' row width and readable up to n + tile_pad_rows(n).
struct I8Data {
    I8Rows rows;
    const int32_t *xsq, *xsum, *cinit;
};

// Exact search over 8-bit rows with float32 queries holding 0..255 integers: the range search
// over int8 matrix-core tiles (exact_tile.hip) when it applies, the chunked kernels otherwise and
// for the queries it flags. The caller holds S's lock.
int exact_int8(const I8Data &d, int64_t n, int dim, int metric, const float *d_q, int64_t nq,
               int k, const uint32_t *d_bits, const hipDeviceProp_t &prop, ExactScratch &S,
               int64_t *d_ids, double *d_dists, hipStream_t s) {
    const bool tile_ok = dim <= 256 && k <= 128 && n >= kTileMinRows;
    const int nch = tile_ok ? pow2_at_least((dim + 31) / 32) : 0;
    ExactArgs a;
    memset(&a, 0, sizeof(a));
    a.n = n;
    a.dim = dim;
    a.metric = metric;
    a.k = k;
    a.nq = (int)nq;
    a.allowed = d_bits;
    a.dimp = tile_ok ? 32 * nch : (int)round_up(dim, 32);
    a.rows = d.rows;
    a.xsq = d.xsq;
    a.xsum = d.xsum;
    HM_TRY(S.q8.ensure((size_t)nq * a.dimp), "alloc");
    HM_TRY(S.qsq.ensure(nq), "alloc");
    HM_TRY(S.qsum.ensure(nq), "alloc");
    hipLaunchKernelGGL(shift_i8_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, d_q, nq,
                       dim, a.dimp, S.q8.p, S.qsq.p, S.qsum.p);
    HM_TRY(hipGetLastError(), "shift launch");
    a.q8 = S.q8.p;
    a.qsq = S.qsq.p;
    a.qsum = S.qsum.p;
    if (!tile_ok) return run_chunked(a, true, prop, d_ids, d_dists, s);

    // The rows the tiles scan: as given when ready for them; otherwise (masked: the allowed
    // rows, compacted in id order) gathered into shifted rows of the tile width, padded.
    int64_t m = n;
    I8Rows trows = d.rows;
    const int32_t *tcinit = d.cinit;
    if (d_bits) {
        HM_TRY(S.idmap.ensure(n), "alloc");
        HM_TRY(S.count.ensure(1), "alloc");
        HM_TRY(subset_ids_launch(d_bits, n, S.idmap.p, S.count.p, s), "subset launch");
        HM_TRY(hipMemcpyAsync(&m, S.count.p, 8, hipMemcpyDeviceToHost, s), "copy count");
        HM_TRY(hipStreamSynchronize(s), "subset count");
        if (m < kTileMinRows || m < k) return run_chunked(a, true, prop, d_ids, d_dists, s);
    }
    if (d_bits || !tcinit || d.rows.ld != a.dimp) {
        HM_TRY(S.cx8.ensure((size_t)(m + tile_pad_rows(m)) * a.dimp), "alloc");
        HM_TRY(S.cxsq.ensure(m), "alloc");
        HM_TRY(S.cxsum.ensure(m), "alloc");
        HM_TRY(gather_rows_launch(d.rows, d.xsq, d.xsum, d_bits ? S.idmap.p : nullptr, m, a.dimp,
                                  S.cx8.p, S.cxsq.p, S.cxsum.p, s),
               "gather launch");
        trows = I8Rows{(const uint8_t *)S.cx8.p, a.dimp, 0u};
        HM_TRY(S.cinit.ensure(m + tile_pad_rows(m)), "alloc");
        HM_TRY(cinit_launch(S.cxsq.p, S.cxsum.p, m, m + tile_pad_rows(m), metric, S.cinit.p, s),
               "cinit launch");
        tcinit = S.cinit.p;
    }
    HM_TRY(S.flags.ensure(nq), "alloc");

    Events ev;
    for (hipEvent_t &x : ev.e) HM_TRY(hipEventCreate(&x), "event");
    double times[3] = {0, 0, 0};
    for (int64_t b0 = 0; b0 < nq; b0 += kTileQueriesPerLaunch) {
        const int64_t nqb = std::min<int64_t>(kTileQueriesPerLaunch, nq - b0);
        const TilePlan p = plan_tile(m, nqb, nch, k);
        const int per_query = p.nchunks * 2 * p.T;
        if (per_query < 2 * k) return run_chunked(a, true, prop, d_ids, d_dists, s);
        const size_t slots = (size_t)nqb * p.nchunks * 2;
        HM_TRY(S.bnd.ensure(slots * p.T), "alloc");
        HM_TRY(S.rec_mask.ensure(slots * p.E), "alloc");
        HM_TRY(S.rec_row.ensure(slots * p.E), "alloc");
        HM_TRY(S.rec_cnt.ensure(slots), "alloc");
        HM_TRY(S.trig.ensure(nqb), "alloc");
        TileArgs t;
        memset(&t, 0, sizeof(t));
        t.rows = trows;
        t.cinit = tcinit;
        t.q8 = S.q8.p + (size_t)b0 * a.dimp;
        t.nq = (int)nqb;
        t.chunk = p.chunk;
        t.nchunks = p.nchunks;
        t.sample_every = p.sample_every;
        t.bnd = S.bnd.p;
        t.trig = S.trig.p;
        t.E = p.E;
        t.rec_mask = S.rec_mask.p;
        t.rec_row = S.rec_row.p;
        t.rec_cnt = S.rec_cnt.p;
        const dim3 grid(p.nchunks, p.groups);
        // The bound pass with eight list entries per query block keeps fewer query blocks per
        // wave at 128 dimensions (register budget); its output does not depend on the grouping.
        const int qb_bound = (p.T == 8 && p.nch >= 4) ? std::min(p.qb, 2) : p.qb;
        const dim3 grid_bound(p.nchunks, (unsigned)((nqb + kTileWaves * qb_bound * 32 - 1) /
                                                    (kTileWaves * qb_bound * 32)));
        HM_TRY(hipEventRecord(ev.e[0], s), "event");
        HM_TRY(tile_launch(0, p.nch, qb_bound, p.T, t, grid_bound, s), "bound launch");
        HM_TRY(threshold_launch(S.bnd.p, per_query, k, S.trig.p, (int)nqb, s), "threshold launch");
        HM_TRY(hipEventRecord(ev.e[1], s), "event");
        HM_TRY(tile_launch(1, p.nch, p.qb, p.T, t, grid, s), "range launch");
        HM_TRY(hipEventRecord(ev.e[2], s), "event");
        CollectArgs c;
        memset(&c, 0, sizeof(c));
        c.rec_mask = S.rec_mask.p;
        c.rec_row = S.rec_row.p;
        c.rec_cnt = S.rec_cnt.p;
        c.rows = trows;
        c.q8 = S.q8.p + (size_t)b0 * a.dimp;
        c.qld = a.dimp;
        c.idmap = d_bits ? S.idmap.p : nullptr;
        c.n = m;
        c.nchunks = p.nchunks;
        c.E = p.E;
        c.k = k;
        c.metric = metric;
        c.dim = dim;
        c.cap = p.cap;
        c.ids = d_ids + b0 * k;
        c.dists = d_dists + b0 * k;
        c.flag = S.flags.p + b0;
        HM_TRY(collect_launch(c, (int)nqb, s), "collect launch");
        HM_TRY(hipEventRecord(ev.e[3], s), "event");
        HM_TRY(hipEventSynchronize(ev.e[3]), "range search");
        for (int i = 0; i < 3; ++i) {
            float ms = 0.f;
            HM_TRY(hipEventElapsedTime(&ms, ev.e[i], ev.e[i + 1]), "event time");
            times[i] += ms;
        }
    }
    g_timing.bound_ms = times[0];
    g_timing.range_ms = times[1];
    g_timing.collect_ms = times[2];
    std::vector<int32_t> hflag((size_t)nq);
    HM_TRY(hipMemcpyAsync(hflag.data(), S.flags.p, (size_t)nq * 4, hipMemcpyDeviceToHost, s),
           "copy flags");
    HM_TRY(hipStreamSynchronize(s), "range search");
    std::vector<int32_t> redo;
    for (int64_t i = 0; i < nq; ++i)
        if (hflag[(size_t)i]) redo.push_back((int32_t)i);
    g_timing.redo = (int64_t)redo.size();
    if (redo.empty()) return HM_OK;

    // Flagged queries: the chunked kernels over the full rows (with the bitset when masked).
    const int64_t nr = (int64_t)redo.size();
    Scoped<int32_t> ridx, rqsq, rqsum;
    Scoped<int8_t> rq8;
    Scoped<int64_t> rids;
    Scoped<double> rd;
    HM_TRY(ridx.alloc(nr), "alloc");
    HM_TRY(rq8.alloc((size_t)nr * a.dimp), "alloc");
    HM_TRY(rqsq.alloc(nr), "alloc");
    HM_TRY(rqsum.alloc(nr), "alloc");
    HM_TRY(rids.alloc((size_t)nr * k), "alloc");
    HM_TRY(rd.alloc((size_t)nr * k), "alloc");
    HM_TRY(hipMemcpyAsync(ridx.p, redo.data(), (size_t)nr * 4, hipMemcpyHostToDevice, s), "copy list");
    HM_TRY(gather_rows_launch(I8Rows{(const uint8_t *)S.q8.p, a.dimp, 0u}, S.qsq.p, S.qsum.p,
                              ridx.p, nr, a.dimp, rq8.p, rqsq.p, rqsum.p, s),
           "gather launch");
    ExactArgs r = a;
    r.q8 = rq8.p;
    r.qsq = rqsq.p;
    r.qsum = rqsum.p;
    r.nq = (int)nr;
    const int st = run_chunked(r, true, prop, rids.p, rd.p, s);
    if (st != HM_OK) return st;
    hipLaunchKernelGGL(exact_scatter_kernel, dim3((unsigned)((nr * k + 255) / 256)), dim3(256), 0, s,
                       rids.p, rd.p, ridx.p, nr, k, d_ids, d_dists);
    HM_TRY(hipGetLastError(), "scatter launch");
    HM_TRY(hipStreamSynchronize(s), "exact search");
    return HM_OK;
}

// The exact search with every buffer on the device. Integer data (int8): rows shifted to int8
// into per-device scratch, then exact_int8; anything else: the float32 range search when it
// applies, the float64 chunked kernels otherwise.
int exact_device(const float *d_data, int64_t n, int dim, int metric, const float *d_q,
                 int64_t nq, int k, const uint32_t *d_bits, bool int8, int64_t *d_ids,
                 double *d_dists, hipStream_t s, int device) {
    g_timing = ExactTiming();
    hipDeviceProp_t prop;
    HM_TRY(hipGetDeviceProperties(&prop, device), "device properties");
    if (!int8) {
        ExactArgs a;
        memset(&a, 0, sizeof(a));
        a.n = n;
        a.dim = dim;
        a.metric = metric;
        a.k = k;
        a.nq = (int)nq;
        a.allowed = d_bits;
        a.x32 = d_data;
        a.q32 = d_q;
        const int kp = ftile_width(dim);
        if (kp > 0 && k <= 128 && n >= kTileMinRows && !g_float_chunked.load()) {
            bool done = false;
            const int st = exact_float_tiles(a, kp, prop, d_ids, d_dists, s, done);
            if (st != HM_OK || done) return st;
        }
        Scoped<double> dxn, dqn;
        if (metric == HM_COSINE) {
            HM_TRY(dxn.alloc((size_t)n * dim), "alloc");
            HM_TRY(dqn.alloc((size_t)nq * dim), "alloc");
            hipLaunchKernelGGL(normalise_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               d_data, n, dim, dxn.p);
            hipLaunchKernelGGL(normalise_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s,
                               d_q, nq, dim, dqn.p);
            a.xn = dxn.p;
            a.qn = dqn.p;
        }
        return run_chunked(a, false, prop, d_ids, d_dists, s);
    }
    ExactScratch &S = exact_scratch(device);
    std::lock_guard<std::mutex> lock(S.mu);
    // Rows shifted to int8 at the width the tiles read, padded (a mask or the chunked kernels
    // take them as they are).
    const bool tile_ok = dim <= 256 && k <= 128 && n >= kTileMinRows;
    const int ld = tile_ok ? 32 * pow2_at_least((dim + 31) / 32) : (int)round_up(dim, 32);
    const int64_t rows = tile_ok ? n + tile_pad_rows(n) : n;
    HM_TRY(S.x8.ensure((size_t)rows * ld), "alloc");
    HM_TRY(S.xsq.ensure(n), "alloc");
    HM_TRY(S.xsum.ensure(n), "alloc");
    hipLaunchKernelGGL(shift_i8_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, d_data, n,
                       dim, ld, S.x8.p, S.xsq.p, S.xsum.p);
    HM_TRY(hipGetLastError(), "shift launch");
    const int32_t *cinit = nullptr;
    if (tile_ok && !d_bits) {
        HM_TRY(S.cinit.ensure(rows), "alloc");
        HM_TRY(cinit_launch(S.xsq.p, S.xsum.p, n, rows, metric, S.cinit.p, s), "cinit launch");
        cinit = S.cinit.p;
    }
    const I8Data d{I8Rows{(const uint8_t *)S.x8.p, ld, 0u}, S.xsq.p, S.xsum.p, cinit};
    return exact_int8(d, n, dim, metric, d_q, nq, k, d_bits, prop, S, d_ids, d_dists, s);
}

// The exact search on an index handle's rows (queries, mask and outputs on the device): the
// handle's raw 8-bit rows and cached row terms when it holds them and the queries are 0..255
// integers, its float32 rows otherwise.
int exact_handle(hm_index *h, const float *d_q, int64_t nq, int k, const uint32_t *d_bits,
                 bool int8, int64_t *d_ids, double *d_dists) {
    if (!int8) {
        return exact_device(h->vec.p, h->n, h->dim, h->metric, d_q, nq, k, d_bits, false, d_ids,
                            d_dists, h->stream, h->device);
    }
    g_timing = ExactTiming();
    hipDeviceProp_t prop;
    HM_TRY(hipGetDeviceProperties(&prop, h->device), "device properties");
    ExactScratch &S = exact_scratch(h->device);
    std::lock_guard<std::mutex> lock(S.mu);
    const I8Data d{I8Rows{h->u8.p, h->dimp, 0x80808080u}, h->ex_xsq.p, h->ex_xsum.p,
                   h->ex_cinit.p};
    return exact_int8(d, h->n, h->dim, h->metric, d_q, nq, k, d_bits, prop, S, d_ids, d_dists,
                      h->stream);
}

int check_exact(int64_t n, int32_t dim, int32_t metric, int32_t k) {
    if (n < 1 || dim < 1) return invalid("data must be non-empty with at least one dimension");
    if (k < 1 || k > n) return invalid("k out of range");
    if (metric < 0 || metric > 2) return invalid("unknown metric");
    if (n >= (int64_t(1) << 31) - 1 - kTilePadRows) return invalid("too many rows");
    return HM_OK;
}

}  // namespace

extern "C" int hm_exact(const float *data, int64_t n, int32_t dim, int32_t metric,
                        const float *queries, int64_t nq, int32_t k, const uint8_t *allowed,
                        int device, int64_t *ids, double *dists) {
    if (!data || !queries || !ids || !dists) return invalid("null argument");
    const int st = check_exact(n, dim, metric, k);
    if (st != HM_OK) return st;
    if (nq == 0) return HM_OK;
    HM_TRY(hipSetDevice(device), "hipSetDevice");
    const bool int8 = metric != HM_COSINE && dim <= 1024 && u8_values(data, (size_t)n * dim) &&
                      u8_values(queries, (size_t)nq * dim);
    Scoped<float> dx, dq;
    Scoped<uint32_t> bits;
    Scoped<int64_t> oid;
    Scoped<double> od;
    hipStream_t s = nullptr;
    HM_TRY(dx.alloc((size_t)n * dim), "alloc");
    HM_TRY(dq.alloc((size_t)nq * dim), "alloc");
    HM_TRY(oid.alloc((size_t)nq * k), "alloc");
    HM_TRY(od.alloc((size_t)nq * k), "alloc");
    HM_TRY(hipMemcpy(dx.p, data, (size_t)n * dim * 4, hipMemcpyHostToDevice), "copy data");
    HM_TRY(hipMemcpy(dq.p, queries, (size_t)nq * dim * 4, hipMemcpyHostToDevice), "copy queries");
    if (allowed) {
        const std::vector<uint32_t> b = pack_masks(allowed, 1, n);
        HM_TRY(bits.alloc(b.size()), "alloc");
        HM_TRY(hipMemcpy(bits.p, b.data(), b.size() * 4, hipMemcpyHostToDevice), "copy mask");
    }
    prefault_output(ids, (size_t)nq * k * 8);
    prefault_output(dists, (size_t)nq * k * 8);
    const int rc = exact_device(dx.p, n, dim, metric, dq.p, nq, k, allowed ? bits.p : nullptr, int8,
                                oid.p, od.p, s, device);
    if (rc != HM_OK) return rc;
    HM_TRY(hipMemcpy(ids, oid.p, (size_t)nq * k * 8, hipMemcpyDeviceToHost), "copy ids");
    HM_TRY(hipMemcpy(dists, od.p, (size_t)nq * k * 8, hipMemcpyDeviceToHost), "copy dists");
    return HM_OK;
}

extern "C" int hm_exact_dev(const float *d_data, int64_t n, int32_t dim, int32_t metric,
                            const float *d_queries, int64_t nq, int32_t k,
                            const uint32_t *d_allowed_bits, int exact_u8, int64_t *d_ids,
                            double *d_dists, void *stream) {
    if (!d_data || !d_queries || !d_ids || !d_dists) return invalid("null argument");
    const int st = check_exact(n, dim, metric, k);
    if (st != HM_OK) return st;
    if (nq == 0) return HM_OK;
    int device = 0;
    HM_TRY(hipGetDevice(&device), "hipGetDevice");
    const bool int8 = exact_u8 != 0 && metric != HM_COSINE && dim <= 1024;
    return exact_device(d_data, n, dim, metric, d_queries, nq, k, d_allowed_bits, int8, d_ids,
                        d_dists, (hipStream_t)stream, device);
}

extern "C" int hm_set_exact_tuning(int32_t query_blocks, int32_t sample_every) {
    if (!(query_blocks >= -1 && query_blocks <= 4 && query_blocks != 3) || sample_every < 0 ||
        sample_every > 64)
        return invalid("tuning values out of range");
    g_float_chunked.store(query_blocks == -1);
    g_tile_qb.store(std::max(0, query_blocks));
    g_tile_sample.store(sample_every);
    return HM_OK;
}

extern "C" int hm_exact_timing(double *out) {
    if (!out) return invalid("null argument");
    out[0] = g_timing.bound_ms;
    out[1] = g_timing.range_ms;
    out[2] = g_timing.collect_ms;
    out[3] = (double)g_timing.redo;
    return HM_OK;
}

extern "C" int hm_index_exact(hm_index *h, const float *queries, int64_t nq, int32_t k,
                              const uint8_t *allowed, int64_t *ids, double *dists) {
    if (!h || !queries || !ids || !dists) return invalid("null argument");
    if (h->n == 0) return invalid("cannot search an empty index");
    const int st = check_exact(h->n, h->dim, h->metric, k);
    if (st != HM_OK) return st;
    if (nq == 0) return HM_OK;
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    const bool int8 = h->exact_u8 && u8_values(queries, (size_t)nq * h->dim);
    Scoped<float> dq;
    Scoped<uint32_t> bits;
    Scoped<int64_t> oid;
    Scoped<double> od;
    HM_TRY(dq.alloc((size_t)nq * h->dim), "alloc");
    HM_TRY(oid.alloc((size_t)nq * k), "alloc");
    HM_TRY(od.alloc((size_t)nq * k), "alloc");
    HM_TRY(hipMemcpyAsync(dq.p, queries, (size_t)nq * h->dim * 4, hipMemcpyHostToDevice, h->stream),
           "copy queries");
    // The packed mask outlives the search: its asynchronous copy may still read it.
    std::vector<uint32_t> b;
    if (allowed) {
        b = pack_masks(allowed, 1, h->n);
        HM_TRY(bits.alloc(b.size()), "alloc");
        HM_TRY(hipMemcpyAsync(bits.p, b.data(), b.size() * 4, hipMemcpyHostToDevice, h->stream),
               "copy mask");
    }
    prefault_output(ids, (size_t)nq * k * 8);  // while the queries upload
    prefault_output(dists, (size_t)nq * k * 8);
    const int rc = exact_handle(h, dq.p, nq, k, allowed ? bits.p : nullptr, int8, oid.p, od.p);
    if (rc != HM_OK) return rc;
    HM_TRY(hipMemcpyAsync(ids, oid.p, (size_t)nq * k * 8, hipMemcpyDeviceToHost, h->stream), "copy ids");
    HM_TRY(hipMemcpyAsync(dists, od.p, (size_t)nq * k * 8, hipMemcpyDeviceToHost, h->stream),
           "copy dists");
    HM_TRY(hipStreamSynchronize(h->stream), "exact search");
    return HM_OK;
}

extern "C" int hm_index_exact_dev(hm_index *h, const float *d_queries, int64_t nq, int32_t k,
                                  const uint32_t *d_allowed_bits, int queries_u8, int64_t *d_ids,
                                  double *d_dists) {
    if (!h || !d_queries || !d_ids || !d_dists) return invalid("null argument");
    if (h->n == 0) return invalid("cannot search an empty index");
    const int st = check_exact(h->n, h->dim, h->metric, k);
    if (st != HM_OK) return st;
    if (nq == 0) return HM_OK;
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    return exact_handle(h, d_queries, nq, k, d_allowed_bits, h->exact_u8 && queries_u8 != 0,
                        d_ids, d_dists);
}
