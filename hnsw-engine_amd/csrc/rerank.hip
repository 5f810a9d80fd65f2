// Exact re-ranking (numba_engine.py:506-530, _rerank_batch): each query's candidate ids (-1
// skipped) are re-scored with the sequential float64 distance to the float32 vectors and the
// best k by (distance, id) are kept. One wave per query; the scored list is sorted in LDS.
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "../../include/hnsw_mi355x.h"
#include "index.hpp"
#include "keys.hpp"
#include "seqdist.hpp"
#include "wave.hpp"

namespace hm {

__global__ __launch_bounds__(64) void rerank_kernel(const float *vec, int dim, int metric,
                                                    const float *queries, const int64_t *cand,
                                                    int fetch, int k, int p2, int64_t *out_ids,
                                                    double *out_d) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    uint64_t *keys = (uint64_t *)lds;
    uint32_t *ids = (uint32_t *)(lds + (size_t)p2 * 8);
    const int lane = lane_id();
    const int qi = blockIdx.x;
    int valid = 0;
    for (int j = lane; j < fetch; j += 64) {
        const int64_t e = cand[(size_t)qi * fetch + j];
        if (e >= 0) valid += 1;
    }
    valid = wave_sum(valid);
    // Compact the valid ids in candidate order, then score them.
    int w = 0;
    for (int b = 0; b < fetch; b += 64) {
        const int j = b + lane;
        const int64_t e = j < fetch ? cand[(size_t)qi * fetch + j] : -1;
        const uint64_t m = ballot(e >= 0);
        if (e >= 0) {
            const int d = w + rank_below(m);
            ids[d] = (uint32_t)e;
            keys[d] = key_of_double(
                seq_row<false>(vec + (size_t)e * dim, queries + (size_t)qi * dim, dim, metric));
        }
        w += popc(m);
    }
    wave_sync();
    bitonic_sort(keys, ids, valid, p2);
    for (int i = lane; i < k; i += 64) {
        if (i < valid) {
            out_ids[(size_t)qi * k + i] = ids[i];
            out_d[(size_t)qi * k + i] = double_of_key(keys[i], metric);
        } else {
            out_ids[(size_t)qi * k + i] = -1;
            out_d[(size_t)qi * k + i] = __builtin_huge_val();
        }
    }
}

// Launches rerank_kernel over device candidates (nq, fetch) into device outputs (nq, k).
hipError_t rerank_device(hm_index *h, const float *d_q, const int64_t *d_cand, int64_t nq,
                         int fetch, int k, int64_t *d_ids, double *d_d, hipStream_t s) {
    int p2 = 1;
    while (p2 < fetch) p2 <<= 1;
    const int lds = p2 * 12;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)rerank_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(rerank_kernel, dim3((unsigned)nq), dim3(64), lds, s, h->vec.p, h->dim,
                       h->metric, d_q, d_cand, fetch, k, p2, d_ids, d_d);
    return hipGetLastError();
}

}  // namespace hm

using namespace hm;

extern "C" int hm_rerank(hm_index *h, const float *queries, int64_t nq, const int64_t *candidates,
                         int32_t fetch, int32_t k, int64_t *ids, double *dists) {
    if (!h || !queries || !candidates || !ids || !dists) return invalid("null argument");
    if (nq < 0 || fetch < 0 || k < 1) return invalid("bad rerank sizes");
    if (nq == 0) return HM_OK;
    for (int64_t i = 0; i < nq * fetch; ++i)
        if (candidates[i] < -1 || candidates[i] >= h->n)
            return invalid("candidate ids outside [-1, n)");
    int p2 = 1;
    while (p2 < fetch) p2 <<= 1;
    if (p2 * 12 > 160 * 1024) return invalid("rerank fetch too large");
    hipStream_t s = h->stream;
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    HM_TRY(h->q32.ensure((size_t)nq * h->dim), "alloc");
    HM_TRY(h->cand.ensure((size_t)nq * std::max(fetch, 1)), "alloc");
    HM_TRY(h->ids.ensure((size_t)nq * k), "alloc");
    HM_TRY(h->dists.ensure((size_t)nq * k), "alloc");
    HM_TRY(hipMemcpyAsync(h->q32.p, queries, (size_t)nq * h->dim * 4, hipMemcpyHostToDevice, s),
           "copy");
    if (fetch)
        HM_TRY(hipMemcpyAsync(h->cand.p, candidates, (size_t)nq * fetch * 8, hipMemcpyHostToDevice,
                              s),
               "copy");
    HM_TRY(rerank_device(h, h->q32.p, h->cand.p, nq, fetch, k, h->ids.p, h->dists.p, s),
           "rerank launch");
    prefault_output(ids, (size_t)nq * k * 8);  // while the re-rank runs
    prefault_output(dists, (size_t)nq * k * 8);
    HM_TRY(hipMemcpyAsync(ids, h->ids.p, (size_t)nq * k * 8, hipMemcpyDeviceToHost, s), "copy");
    HM_TRY(hipMemcpyAsync(dists, h->dists.p, (size_t)nq * k * 8, hipMemcpyDeviceToHost, s), "copy");
    HM_TRY(hipStreamSynchronize(s), "rerank");
    return HM_OK;
}
