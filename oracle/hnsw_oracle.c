/*
 * CPU restatement of the hnsw-engine search path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle that the HIP engine is checked against and
 * the timed CPU baseline of bench.py. Nothing in the product library links, loads or calls it.
 *
 * Every function restates the reference's Numba code line by line in semantics (not in form):
 *   ho_distance           numba_engine.py:57-70      fp64, left to right, no FMA contraction
 *   search_layer          numba_engine.py:98-148     Algorithm 2 (plain)
 *   search_layer0_filt    numba_engine.py:151-198    Algorithm 2, filtered layer 0 (rule 9)
 *   ho_search_batch       numba_engine.py:422-503    Algorithm 5 per query, -1/inf padding
 *   ho_rerank_batch       numba_engine.py:506-530    exact re-score of candidate ids
 *   ho_exact_search       oracle.py:35-126           exact top-k, ties to the lower id
 *                          benchmarks/run_filtered.py:76-106 (masked subset variant)
 *
 * Build with -O2 -ffp-contract=off and without -ffast-math so that every distance is the
 * sequential, unfused float64 sum the reference computes.
 *
 * Parity pinning (which check fixes which function) is listed in tests/test_oracle_c.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "hnsw_oracle.h"

/* ------------------------------------------------------------------------------ distance */

/* numba_engine.py:57-70: acc += (f64(x_i) - q_i)^2 for l2; acc += f64(x_i) * q_i otherwise;
 * ip returns -acc, cosine 1 - acc. */
static double dist_f32(int metric, const double *q, const float *x, int dim) {
    double acc = 0.0;
    if (metric == HO_L2) {
        for (int i = 0; i < dim; ++i) {
            double d = (double)x[i] - q[i];
            acc += d * d;
        }
        return acc;
    }
    for (int i = 0; i < dim; ++i) acc += (double)x[i] * q[i];
    return metric == HO_IP ? -acc : 1.0 - acc;
}

/* The same kernel over uint8 codes (numba_engine.py:674-701 runs _search_batch on codes). */
static double dist_u8(int metric, const double *q, const uint8_t *x, int dim) {
    double acc = 0.0;
    if (metric == HO_L2) {
        for (int i = 0; i < dim; ++i) {
            double d = (double)x[i] - q[i];
            acc += d * d;
        }
        return acc;
    }
    for (int i = 0; i < dim; ++i) acc += (double)x[i] * q[i];
    return metric == HO_IP ? -acc : 1.0 - acc;
}

double ho_distance(int metric, const double *q, const void *vectors, int vec_u8, int dim,
                   int64_t j) {
    if (vec_u8) return dist_u8(metric, q, (const uint8_t *)vectors + (size_t)j * dim, dim);
    return dist_f32(metric, q, (const float *)vectors + (size_t)j * dim, dim);
}

/* ------------------------------------------------------------------------------ heaps */

typedef struct {
    double d;
    int64_t id;
} pair_t;

/* Tuple order of Python's (distance, id): distance first, then id. */
static inline int pair_lt(pair_t a, pair_t b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

typedef struct {
    pair_t *v;
    int64_t n, cap;
} heap_t;

static void heap_reserve(heap_t *h, int64_t need) {
    if (need <= h->cap) return;
    int64_t cap = h->cap ? h->cap : 64;
    while (cap < need) cap *= 2;
    h->v = (pair_t *)realloc(h->v, (size_t)cap * sizeof(pair_t));
    h->cap = cap;
}

/* Min-heap on (d, id): the candidate queue. */
static void minheap_push(heap_t *h, pair_t x) {
    heap_reserve(h, h->n + 1);
    int64_t i = h->n++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!pair_lt(x, h->v[p])) break;
        h->v[i] = h->v[p];
        i = p;
    }
    h->v[i] = x;
}

static pair_t minheap_pop(heap_t *h) {
    pair_t top = h->v[0];
    pair_t last = h->v[--h->n];
    int64_t i = 0, n = h->n;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        pair_t best = last;
        if (l < n && pair_lt(h->v[l], best)) { m = l; best = h->v[l]; }
        if (r < n && pair_lt(h->v[r], best)) { m = r; }
        if (m == i) break;
        h->v[i] = h->v[m];
        i = m;
    }
    if (n > 0) h->v[i] = last;
    return top;
}

/* Max-heap on (d, id): the result set; its top is the worst (largest) tuple, which is what the
 * reference's heap of (-d, -id) keeps at index 0. */
static void maxheap_push(heap_t *h, pair_t x) {
    heap_reserve(h, h->n + 1);
    int64_t i = h->n++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!pair_lt(h->v[p], x)) break;
        h->v[i] = h->v[p];
        i = p;
    }
    h->v[i] = x;
}

static void maxheap_pop(heap_t *h) {
    pair_t last = h->v[--h->n];
    int64_t i = 0, n = h->n;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        pair_t best = last;
        if (l < n && pair_lt(best, h->v[l])) { m = l; best = h->v[l]; }
        if (r < n && pair_lt(best, h->v[r])) { m = r; }
        if (m == i) break;
        h->v[i] = h->v[m];
        i = m;
    }
    if (n > 0) h->v[i] = last;
}

static int cmp_pair(const void *a, const void *b) {
    pair_t x = *(const pair_t *)a, y = *(const pair_t *)b;
    if (pair_lt(x, y)) return -1;
    if (pair_lt(y, x)) return 1;
    return 0;
}

/* ------------------------------------------------------------------------------ search */

typedef struct {
    const ho_graph *g;
    const void *vectors;
    int vec_u8;
    int metric;
    uint32_t *visited;
    uint32_t epoch;
    heap_t cand, res;
    pair_t *out;
    int64_t out_cap;
    int64_t dcount, hops;
} ctx_t;

static uint32_t next_epoch(ctx_t *c) {
    /* numba_engine.py:89-95 */
    c->epoch += 1;
    if (c->epoch >= 4294967295u) {
        memset(c->visited, 0, (size_t)c->g->n * sizeof(uint32_t));
        c->epoch = 1;
    }
    return c->epoch;
}

/* numba_engine.py:73-86: the used slots of a node's row on a layer. */
static const int32_t *links_of(const ho_graph *g, int64_t node, int layer, int32_t *count) {
    if (layer == 0) {
        *count = g->count0[node];
        return g->links0 + (size_t)node * g->stride0;
    }
    int64_t row = g->upper_offset[node] + layer - 1;
    *count = g->count_up[row];
    return g->links_up + (size_t)row * g->stride_up;
}

/* Sorted output of the result heap into c->out; returns its length. */
static int64_t drain_sorted(ctx_t *c) {
    int64_t n = c->res.n;
    if (n > c->out_cap) {
        c->out = (pair_t *)realloc(c->out, (size_t)n * sizeof(pair_t));
        c->out_cap = n;
    }
    memcpy(c->out, c->res.v, (size_t)n * sizeof(pair_t));
    qsort(c->out, (size_t)n, sizeof(pair_t), cmp_pair);
    return n;
}

/* numba_engine.py:98-148. `entry` holds `n_entry` pairs; the sorted result is left in c->out. */
static int64_t search_layer(ctx_t *c, const double *q, const pair_t *entry, int64_t n_entry,
                            int64_t ef, int layer) {
    uint32_t epoch = next_epoch(c);
    c->cand.n = 0;
    c->res.n = 0;
    for (int64_t i = 0; i < n_entry; ++i) {
        minheap_push(&c->cand, entry[i]);
        maxheap_push(&c->res, entry[i]);
    }
    while (c->res.n > ef) maxheap_pop(&c->res);
    for (int64_t i = 0; i < n_entry; ++i) c->visited[entry[i].id] = epoch;

    while (c->cand.n > 0) {
        pair_t top = minheap_pop(&c->cand);
        if (top.d > c->res.v[0].d) break;
        c->hops += 1;
        int32_t count;
        const int32_t *nb = links_of(c->g, top.id, layer, &count);
        double worst = c->res.v[0].d;
        for (int32_t i = 0; i < count; ++i) {
            int64_t e = nb[i];
            if (c->visited[e] == epoch) continue;
            c->visited[e] = epoch;
            double d = ho_distance(c->metric, q, c->vectors, c->vec_u8, c->g->dim, e);
            c->dcount += 1;
            if (c->res.n < ef || d < worst) {
                pair_t p = {d, e};
                minheap_push(&c->cand, p);
                maxheap_push(&c->res, p);
                if (c->res.n > ef) maxheap_pop(&c->res);
                worst = c->res.v[0].d;
            }
        }
    }
    return drain_sorted(c);
}

/* numba_engine.py:151-198: only allowed nodes enter the results; every accepted node enters the
 * candidates; the search stops only when the results are full. */
static int64_t search_layer0_filt(ctx_t *c, const double *q, const pair_t *entry,
                                  int64_t n_entry, int64_t ef, const uint8_t *allowed) {
    uint32_t epoch = next_epoch(c);
    c->cand.n = 0;
    c->res.n = 0;
    for (int64_t i = 0; i < n_entry; ++i) {
        minheap_push(&c->cand, entry[i]);
        if (allowed[entry[i].id]) maxheap_push(&c->res, entry[i]);
    }
    while (c->res.n > ef) maxheap_pop(&c->res);
    for (int64_t i = 0; i < n_entry; ++i) c->visited[entry[i].id] = epoch;

    while (c->cand.n > 0) {
        pair_t top = minheap_pop(&c->cand);
        if (c->res.n == ef && top.d > c->res.v[0].d) break;
        c->hops += 1;
        double worst = c->res.n > 0 ? c->res.v[0].d : INFINITY;
        int32_t count = c->g->count0[top.id];
        const int32_t *nb = c->g->links0 + (size_t)top.id * c->g->stride0;
        for (int32_t i = 0; i < count; ++i) {
            int64_t e = nb[i];
            if (c->visited[e] == epoch) continue;
            c->visited[e] = epoch;
            double d = ho_distance(c->metric, q, c->vectors, c->vec_u8, c->g->dim, e);
            c->dcount += 1;
            if (c->res.n < ef || d < worst) {
                pair_t p = {d, e};
                minheap_push(&c->cand, p);
                if (allowed[e]) {
                    maxheap_push(&c->res, p);
                    if (c->res.n > ef) maxheap_pop(&c->res);
                    worst = c->res.v[0].d;
                }
            }
        }
    }
    return drain_sorted(c);
}

/* One row of numba_engine.py:447-502. */
static void search_one(ctx_t *c, const double *q, int64_t k, int64_t ef, const uint8_t *allowed,
                       int64_t *ids, double *dists) {
    const ho_graph *g = c->g;
    for (int64_t i = 0; i < k; ++i) {
        ids[i] = -1;
        dists[i] = INFINITY;
    }
    if (g->n == 0 || g->entry_point < 0) return;
    pair_t cur = {ho_distance(c->metric, q, c->vectors, c->vec_u8, g->dim, g->entry_point),
                  g->entry_point};
    c->dcount += 1;
    for (int layer = g->max_level; layer > 0; --layer) {
        int64_t m = search_layer(c, q, &cur, 1, 1, layer);
        /* ef = 1 keeps exactly one result */
        (void)m;
        cur = c->out[0];
    }
    int64_t efk = ef > k ? ef : k;
    int64_t m = allowed ? search_layer0_filt(c, q, &cur, 1, efk, allowed)
                        : search_layer(c, q, &cur, 1, efk, 0);
    int64_t take = m < k ? m : k;
    for (int64_t i = 0; i < take; ++i) {
        ids[i] = c->out[i].id;
        dists[i] = c->out[i].d;
    }
}

static void ctx_init(ctx_t *c, const ho_graph *g, const void *vectors, int vec_u8, int metric) {
    memset(c, 0, sizeof(*c));
    c->g = g;
    c->vectors = vectors;
    c->vec_u8 = vec_u8;
    c->metric = metric;
    c->visited = (uint32_t *)calloc((size_t)(g->n > 0 ? g->n : 1), sizeof(uint32_t));
}

static void ctx_free(ctx_t *c) {
    free(c->visited);
    free(c->cand.v);
    free(c->res.v);
    free(c->out);
}

int ho_search_batch(const ho_graph *g, const void *vectors, int vec_u8, int metric,
                    const void *queries, int queries_f64, int64_t nq, const int64_t *k_per_query,
                    const int64_t *ef_per_query, int64_t k_stride, const uint8_t *mask_pool,
                    const int64_t *mask_index, int64_t *out_ids, double *out_dists,
                    int64_t *out_dcount, int64_t *out_hops, int threads) {
    int dim = g->dim;
    int status = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        ctx_t c;
        ctx_init(&c, g, vectors, vec_u8, metric);
        double *q = (double *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(double));
        if (!c.visited || !q) status = -1;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t row = 0; row < nq; ++row) {
            if (!c.visited || !q) continue;
            /* numba_engine.py:448 q = queries[row].astype(float64) */
            if (queries_f64) {
                memcpy(q, (const double *)queries + (size_t)row * dim, (size_t)dim * sizeof(double));
            } else {
                const float *qf = (const float *)queries + (size_t)row * dim;
                for (int i = 0; i < dim; ++i) q[i] = (double)qf[i];
            }
            const uint8_t *allowed = NULL;
            if (mask_pool && mask_index && mask_index[row] >= 0)
                allowed = mask_pool + (size_t)mask_index[row] * (size_t)g->n;
            int64_t d0 = c.dcount, h0 = c.hops;
            search_one(&c, q, k_per_query[row], ef_per_query[row], allowed,
                       out_ids + (size_t)row * k_stride, out_dists + (size_t)row * k_stride);
            if (out_dcount) out_dcount[row] = c.dcount - d0;
            if (out_hops) out_hops[row] = c.hops - h0;
        }
        free(q);
        ctx_free(&c);
    }
    return status;
}

/* ------------------------------------------------------------------------------ re-rank */

/* numba_engine.py:506-530: exact fp64 distances to each row's candidate ids (-1 skipped), sorted
 * by (distance, id), best k kept. */
int ho_rerank_batch(const float *vectors, int dim, int metric, const float *queries, int64_t nq,
                    const int64_t *candidates, int64_t fetch, int64_t k, int64_t *out_ids,
                    double *out_dists, int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        pair_t *scored = (pair_t *)malloc((size_t)(fetch > 0 ? fetch : 1) * sizeof(pair_t));
        double *q = (double *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t row = 0; row < nq; ++row) {
            for (int i = 0; i < dim; ++i) q[i] = (double)queries[(size_t)row * dim + i];
            int64_t m = 0;
            for (int64_t j = 0; j < fetch; ++j) {
                int64_t e = candidates[(size_t)row * fetch + j];
                if (e >= 0) {
                    scored[m].d = dist_f32(metric, q, vectors + (size_t)e * dim, dim);
                    scored[m].id = e;
                    ++m;
                }
            }
            qsort(scored, (size_t)m, sizeof(pair_t), cmp_pair);
            for (int64_t i = 0; i < k; ++i) {
                out_ids[(size_t)row * k + i] = i < m ? scored[i].id : -1;
                out_dists[(size_t)row * k + i] = i < m ? scored[i].d : INFINITY;
            }
        }
        free(scored);
        free(q);
    }
    return 0;
}

/* ------------------------------------------------------------------------------ exact kNN */

/* The exact top-k of oracle.py:35-126 and of the masked brute force of
 * benchmarks/run_filtered.py:76-106: every (allowed) row scored exactly in float64, ordered by
 * (distance, id), best k kept, -1/inf padding when fewer than k rows are allowed.
 *
 * Under cosine the rows and the query are divided by their float64 norms first, as
 * oracle.py:228-229 does; the result is then 1 - x.q. A bounded max-heap of k keeps memory at
 * O(k) per thread. */
int ho_exact_search(const float *data, int64_t n, int dim, int metric, const float *queries,
                    int64_t nq, int64_t k, const uint8_t *allowed, int64_t *out_ids,
                    double *out_dists, int threads) {
    double *norm = NULL;
    if (metric == HO_COSINE) {
        norm = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        for (int64_t j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = 0; i < dim; ++i) {
                double v = (double)data[(size_t)j * dim + i];
                s += v * v;
            }
            norm[j] = sqrt(s);
        }
    }
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        heap_t h = {0};
        heap_reserve(&h, k + 1);
        double *q = (double *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(double));
        double *row = (double *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int64_t r = 0; r < nq; ++r) {
            for (int i = 0; i < dim; ++i) q[i] = (double)queries[(size_t)r * dim + i];
            if (metric == HO_COSINE) {
                double s = 0.0;
                for (int i = 0; i < dim; ++i) s += q[i] * q[i];
                double nq_ = sqrt(s);
                for (int i = 0; i < dim; ++i) q[i] = q[i] / nq_;
            }
            h.n = 0;
            for (int64_t j = 0; j < n; ++j) {
                if (allowed && !allowed[j]) continue;
                double d;
                if (metric == HO_COSINE) {
                    for (int i = 0; i < dim; ++i)
                        row[i] = (double)data[(size_t)j * dim + i] / norm[j];
                    double acc = 0.0;
                    for (int i = 0; i < dim; ++i) acc += row[i] * q[i];
                    d = 1.0 - acc;
                } else {
                    d = dist_f32(metric, q, data + (size_t)j * dim, dim);
                }
                pair_t p = {d, j};
                if (h.n < k) {
                    maxheap_push(&h, p);
                } else if (pair_lt(p, h.v[0])) {
                    maxheap_pop(&h);
                    maxheap_push(&h, p);
                }
            }
            int64_t m = h.n;
            qsort(h.v, (size_t)m, sizeof(pair_t), cmp_pair);
            for (int64_t i = 0; i < k; ++i) {
                out_ids[(size_t)r * k + i] = i < m ? h.v[i].id : -1;
                out_dists[(size_t)r * k + i] = i < m ? h.v[i].d : INFINITY;
            }
        }
        free(h.v);
        free(q);
        free(row);
    }
    free(norm);
    return 0;
}
