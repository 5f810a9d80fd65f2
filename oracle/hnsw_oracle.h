/*
 * CPU restatement of the hnsw-engine search path (test infrastructure; see hnsw_oracle.c).
 */
#ifndef HNSW_ORACLE_H
#define HNSW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { HO_L2 = 0, HO_IP = 1, HO_COSINE = 2 };

/* The reference's native arrays (numba_engine.py:814-825), row strides in elements. */
typedef struct {
    int64_t n;
    int dim;
    int64_t entry_point;
    int max_level;
    const int32_t *links0;      /* (n, stride0) */
    const int32_t *count0;      /* (n,) */
    const int64_t *upper_offset; /* (n,) */
    const int32_t *links_up;    /* (rows, stride_up) */
    const int32_t *count_up;    /* (rows,) */
    int stride0, stride_up;
} ho_graph;

double ho_distance(int metric, const double *q, const void *vectors, int vec_u8, int dim,
                   int64_t j);

/* Algorithm 5 for every query row (numba_engine.py:422-503), with per-query k and ef and an
 * optional per-query mask (mask_index[row] < 0: unfiltered). Queries are float32 or, for the
 * uint8-codes search (numba_engine.py:674-701), float64. Outputs are (nq, k_stride) arrays. */
int ho_search_batch(const ho_graph *g, const void *vectors, int vec_u8, int metric,
                    const void *queries, int queries_f64, int64_t nq, const int64_t *k_per_query,
                    const int64_t *ef_per_query, int64_t k_stride, const uint8_t *mask_pool,
                    const int64_t *mask_index, int64_t *out_ids, double *out_dists,
                    int64_t *out_dcount, int64_t *out_hops, int threads);

int ho_rerank_batch(const float *vectors, int dim, int metric, const float *queries, int64_t nq,
                    const int64_t *candidates, int64_t fetch, int64_t k, int64_t *out_ids,
                    double *out_dists, int threads);

int ho_exact_search(const float *data, int64_t n, int dim, int metric, const float *queries,
                    int64_t nq, int64_t k, const uint8_t *allowed, int64_t *out_ids,
                    double *out_dists, int threads);

#ifdef __cplusplus
}
#endif
#endif
