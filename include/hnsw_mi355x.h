/*
 * C ABI of the MI355X search engine for hnsw-engine graphs (libhnsw_mi355x.so).
 *
 * Every entry point takes plain pointers and sizes and returns 0 on success or a negative
 * status; hm_last_error() then holds the message (the Python layer re-raises it as the
 * reference's exception type). Host-pointer entry points copy to and from HBM themselves;
 * the *_dev variants take device pointers and a HIP stream; the search ones allocate nothing
 * per call beyond the scratch the handle keeps between calls. Calls on one handle share that
 * scratch: each search call's stream waits on the device for the handle's previous call (a
 * HIP event the handle records after each call), so calls on different streams never overlap
 * on one handle.
 *
 * Reference interfaces replaced (paths relative to the hnsw-engine repository):
 *   hm_open            NumbaEngine.attach_native      src/hnsw_engine/numba_engine.py:827-868
 *   hm_check_graph     _validate_native               src/hnsw_engine/numba_engine.py:878-927
 *   hm_search          NumbaEngine.batch_search       src/hnsw_engine/numba_engine.py:652-672
 *                      (-> _search_batch :422-503, _search_layer :98-148,
 *                          _search_layer0_filtered :151-198)
 *   hm_attach_codes    the codes QuantizedIndex keeps src/hnsw_engine/quantization.py:95-102
 *   hm_search_codes    NumbaEngine.batch_search_codes src/hnsw_engine/numba_engine.py:674-701
 *   hm_rerank          NumbaEngine.rerank             src/hnsw_engine/numba_engine.py:703-713
 *                      (-> _rerank_batch :506-530)
 *   hm_exact           exact_search                   src/hnsw_engine/oracle.py:35-126
 *   hm_exact_dev       (masked: exact_subset_search   benchmarks/run_filtered.py:76-106)
 *   hm_build           NumbaEngine.insert             src/hnsw_engine/numba_engine.py:609-643
 *                      (-> _insert_batch :287-419; host C++, not a GPU kernel)
 */
#ifndef HNSW_MI355X_H
#define HNSW_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { HM_L2 = 0, HM_IP = 1, HM_COSINE = 2 };

enum {
    HM_OK = 0,
    HM_EINVAL = -1,  /* bad argument (shape, dtype, range) */
    HM_EDEVICE = -2, /* HIP error */
    HM_ENOMEM = -3
};

/* The reference's native arrays (numba_engine.py:814-825), row-major, as save_native writes
 * them: links0 is (n, 2M+1), links_up is (rows, M+1); the first count slots of a row are used. */
typedef struct {
    int64_t n;
    int32_t dim;
    int32_t M;
    int32_t metric;
    int32_t max_level;
    int64_t entry_point;
    int64_t rows;
    const float *vectors;        /* (n, dim) */
    const int64_t *levels;       /* (n,) */
    const int32_t *links0;       /* (n, 2M+1) */
    const int32_t *count0;       /* (n,) */
    const int64_t *upper_offset; /* (n,) */
    const int32_t *links_up;     /* (rows, M+1) */
    const int32_t *count_up;     /* (rows,) */
} hm_native_graph;

typedef struct hm_index hm_index;

/* Per-batch search options. k and ef are per query (arrays of nq) or, when the array is null,
 * k_all / ef_all for every query; the engine searches layer 0 with max(ef, k) as the reference
 * does (numba_engine.py:473,487). Masks: a pool of n_masks boolean rows of n bytes
 * (mask_pool, n_masks x n); mask_index gives each query its row (-1: unfiltered) or, when
 * null, mask_all applies to every query (-1: unfiltered). Outputs are (nq, kstride) with
 * -1 / +inf padding; kstride >= max k. learned = 1 applies the stop model of
 * hm_set_stop_model to unfiltered batches with ef <= 64 (a stopped beam returns its current
 * top-k; every other batch, or a model that never fires, is the plain fixed-ef search). */
typedef struct {
    int64_t nq;
    const int32_t *k;
    const int32_t *ef;
    int32_t k_all;
    int32_t ef_all;
    int32_t kstride;
    const uint8_t *mask_pool;
    int64_t n_masks;
    const int32_t *mask_index;
    int32_t mask_all;
    int32_t learned;
} hm_search_opts;

/* Uploads a validated graph to HBM (validation is the caller's: the kernels do not
 * bounds-check, as the reference's do not; numba_engine.py:878-927). Rows of links are
 * repacked with -1 beyond their count and repeated links dropped (a repeat is always skipped
 * as visited, so results do not change). When every vector value is an integer in 0..255 an
 * exact uint8 copy is kept as well. */
int hm_open(const hm_native_graph *graph, int device, hm_index **out);

/* The reference's structural checks of native arrays before use (the kernels do not
 * bounds-check): levels, entry point, link counts and ids on layer 0 and the upper layers, the
 * tiling of upper rows by the nodes' offsets, and upper links to nodes present on that layer.
 * HM_EINVAL with the reference's message otherwise. Host only.
 *   replaces _validate_native / _check_links   src/hnsw_engine/numba_engine.py:878-927 */
int hm_check_graph(const hm_native_graph *graph);
void hm_close(hm_index *index);

/* 1 if the index keeps the exact 8-bit copy (integer-valued 0..255 vectors, metric l2/ip). */
int hm_has_exact_u8(const hm_index *index);

/* Algorithm 5 for every query (host buffers). queries: (nq, dim) float32, prepared as the
 * facade prepares them. ids: int64 (nq, kstride); dists: float64 (nq, kstride); dcount and
 * hops: int64 (nq,) per-query distance computations and expansions, counted as the reference
 * counts them (numba_engine.py:130,139,451), or null. */
int hm_search(hm_index *index, const float *queries, const hm_search_opts *opts, int64_t *ids,
              double *dists, int64_t *dcount, int64_t *hops);

/* The same search with every buffer already in HBM (queries float32 (nq, dim); outputs
 * ids int64, dists float64 (nq, kstride), dcount/hops int32 (nq,) or null). Masks are device
 * bitsets: mask_bits holds n_masks rows of ceil(n/32) uint32 words; mask_index (device, nq) or
 * mask_all as in hm_search_opts; k/ef pointers in opts must be device pointers when set.
 * exact_u8 = 1 selects the exact 8-bit kernel: the caller guarantees that every query value
 * is an integer in 0..255 (hm_queries_exact checks a host batch) and the index has the copy.
 * d_fail (device, nq int32, or null) is set to 1 for a query whose candidate overflow exceeded
 * the per-wave spill (its row is then padding; run it again through hm_search, which retries
 * with room for every node). stream: a hipStream_t (null: the default stream). */
int hm_search_dev(hm_index *index, const float *d_queries, const hm_search_opts *opts,
                  const uint32_t *d_mask_bits, int exact_u8, int64_t *d_ids, double *d_dists,
                  int32_t *d_dcount, int32_t *d_hops, int32_t *d_fail, void *stream);

/* hm_search_dev replayed from a HIP graph: the first call with a given set of arguments (and
 * tuning, stop model and scratch buffers) runs the batch directly and captures its launches;
 * later calls with the same arguments launch the captured graph on `stream` (which must not be
 * the null stream), skipping the host-side set-up and per-kernel launch costs. No counterpart in
 * the reference (its engine has no launches); for serving loops that reuse fixed buffers. Any
 * change of argument, tuning or scratch recaptures. Results equal hm_search_dev's. */
int hm_search_graph_dev(hm_index *index, const float *d_queries, const hm_search_opts *opts,
                        const uint32_t *d_mask_bits, int exact_u8, int64_t *d_ids,
                        double *d_dists, int32_t *d_dcount, int32_t *d_hops, int32_t *d_fail,
                        void *stream);

/* hm_search_dev for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. Same
 * arguments and results as hm_search_dev; it calls no hipMalloc / hipFree. Its scratch (work
 * queue, visited bitsets, clearing lists, candidate spill, schedule arrays, and the dcount / hops
 * / fail arrays the caller leaves null) is carved from one static 64 MiB device array per GPU,
 * held from the library's load to the end of the process: the batch runs on as many wave slots
 * as that array holds (about 470 at 1M nodes, so it is slower than hm_search_dev on large
 * graphs), and a batch or graph too large for it is refused with HM_EINVAL. The array is shared
 * by every call on that GPU, from any handle: calls that overlap on one GPU race (calls on one
 * handle are ordered, see below); each call clears what must start at zero. Ordering: every
 * call on a handle waits (on the device) for the work its previous call on that handle
 * enqueued, whatever the streams. */
int hm_search_dev_noalloc(hm_index *index, const float *d_queries, const hm_search_opts *opts,
                          const uint32_t *d_mask_bits, int exact_u8, int64_t *d_ids,
                          double *d_dists, int32_t *d_dcount, int32_t *d_hops, int32_t *d_fail,
                          void *stream);

/* 1 if every value of the host batch is an integer in 0..255 (exact 8-bit kernel eligible). */
int hm_queries_exact(const float *queries, int64_t count);

/* uint8 codes (n, dim) of a ScalarQuantizer, searched by hm_search_codes. */
int hm_attach_codes(hm_index *index, const uint8_t *codes);

/* Algorithm 5 over the codes with float64 code-space queries (nq, dim), unfiltered. */
int hm_search_codes(hm_index *index, const double *queries, const hm_search_opts *opts,
                    int64_t *ids, double *dists, int64_t *dcount, int64_t *hops);

/* QuantizedIndex.batch_search with rerank > 0 in one call: the codes traversal of
 * hm_search_codes (opts: uniform ef, k_all = kstride = fetch, no filter) keeps its fetch
 * candidates in HBM, the exact re-rank of hm_rerank scores them there, and only the best k per
 * query come back: ids int64 / dists float64 (nq, k); dcount / hops per query as in
 * hm_search, or null. code_queries: float64 (nq, dim) code-space queries; queries: the same
 * queries prepared as float32 (nq, dim), for the exact distances.
 *   replaces QuantizedIndex.batch_search   src/hnsw_engine/quantization.py:114-138
 *   (batch_search_codes + rerank, numba_engine.py:674-713, without the host round trip) */
int hm_search_codes_rerank(hm_index *index, const double *code_queries, const float *queries,
                           const hm_search_opts *opts, int32_t k, int64_t *ids, double *dists,
                           int64_t *dcount, int64_t *hops);

/* Exact float64 distances of each row's candidate ids (-1 skipped) to the float32 vectors,
 * best k by (distance, id); -1 / +inf padding. candidates: int64 (nq, fetch). */
int hm_rerank(hm_index *index, const float *queries, int64_t nq, const int64_t *candidates,
              int32_t fetch, int32_t k, int64_t *ids, double *dists);

/* Exact k nearest neighbours by (distance, id) of each query among the rows of data (float32
 * (n, dim)), optionally restricted to allowed rows (bool bytes (n,), or null). The float64
 * distances are the reference's; the shortlist runs on the matrix cores. */
int hm_exact(const float *data, int64_t n, int32_t dim, int32_t metric, const float *queries,
             int64_t nq, int32_t k, const uint8_t *allowed, int device, int64_t *ids,
             double *dists);

/* hm_exact with every buffer already in HBM, on the current device: d_data float32 (n, dim),
 * d_queries float32 (nq, dim), d_allowed_bits ceil(n/32) uint32 words (bit i of word i/32 set:
 * row i allowed) or null, outputs int64 / float64 (nq, k). exact_u8 = 1 selects the int8
 * matrix-core path: the caller guarantees that every data and query value is an integer in
 * 0..255 (metric l2 or ip); 0 selects the float path (float32 matrix-core shortlist, float64
 * ranking). The int8 path converts the rows per call into scratch that it keeps per device
 * between calls (calls on one device take turns); the call synchronises `stream` (a
 * hipStream_t, null: the default stream) before it returns. */
int hm_exact_dev(const float *d_data, int64_t n, int32_t dim, int32_t metric,
                 const float *d_queries, int64_t nq, int32_t k, const uint32_t *d_allowed_bits,
                 int exact_u8, int64_t *d_ids, double *d_dists, void *stream);

/* The exact search over the vectors an index already keeps in HBM (the masked brute force of a
 * filter planner; benchmarks/run_filtered.py:76-106 is the reference's exact_subset_search):
 * queries float32 (nq, dim) host, allowed bool bytes (n,) host or null, outputs host int64 /
 * float64 (nq, k). Same results as hm_exact on the index's vectors. */
int hm_index_exact(hm_index *index, const float *queries, int64_t nq, int32_t k,
                   const uint8_t *allowed, int64_t *ids, double *dists);

/* hm_index_exact with queries, mask and outputs in HBM (the layouts of hm_exact_dev), on the
 * index's stream, synchronised before return (the reference's oracle.exact_search run over an
 * index's vectors, oracle.py:35-126). An index of 0..255 integer rows keeps them as bytes with
 * their row terms, so the int8 path reads them in place: queries_u8 = 1 declares every query
 * value an integer in 0..255. */
int hm_index_exact_dev(hm_index *index, const float *d_queries, int64_t nq, int32_t k,
                       const uint32_t *d_allowed_bits, int queries_u8, int64_t *d_ids,
                       double *d_dists);

/* Tile shape of the range searches (no counterpart in the reference; 0 restores the automatic
 * choice), process-wide: query_blocks (1, 2 or 4) blocks of 32 queries per wave of the int8
 * search, and sample_every, the bound pass's sampling of one superblock in that many (both the
 * int8 and the float32 search). Neither changes results: a looser threshold only admits more rows
 * to the exact ranking. query_blocks = -1 sends float data to the float64 chunked kernels instead
 * of the float32 range search (same results; for comparing the two). */
int hm_set_exact_tuning(int32_t query_blocks, int32_t sample_every);

/* Instrumentation (no counterpart in the reference): kernel times in ms of the last exact
 * search on the calling thread that took a range-search path (int8 or float32), out[0] bound
 * pass plus threshold, out[1] range pass, out[2] collect, and out[3] the number of queries it
 * re-ran on the chunked kernels (all 0 when the last search took another path). */
int hm_exact_timing(double *out);

/* Sequential insertion (host C++, the reference's _insert_batch): inserts vectors[start..n)
 * with their levels into the native arrays, which hold capacity for all n nodes and
 * sum(levels) upper rows; state = {entry_point, max_level} in and out. selection: 0 simple,
 * 1 heuristic. upper_offset must already hold each node's first upper row (-1 for level 0). */
int hm_build(int64_t start, int64_t n, int32_t dim, int32_t metric, int32_t M,
             int32_t ef_construction, int32_t selection, int32_t extend_candidates,
             int32_t keep_pruned, const float *vectors, const int64_t *levels, int32_t *links0,
             int32_t *count0, const int64_t *upper_offset, int32_t *links_up, int32_t *count_up,
             int64_t *state, int64_t *dcount);

/* Learned stop (no counterpart in the reference; an extension that only ever stops a beam
 * earlier than the plain fixed-ef search). A model is hm_model_size() floats: feature means
 * (16), inverse standard deviations (16), W1 (16 x 64, row i = feature i), b1 (64), W2 (64), b2;
 * weights are used rounded to bfloat16, accumulation is float32. */
int hm_model_size(void);

/* Installs (params) or removes (null) the stop model: a beam stops after an expansion when the
 * model's probability that the current top-k is final exceeds threshold, once at least
 * min_steps layer-0 expansions ran. */
int hm_set_stop_model(hm_index *index, const float *params, float threshold, int32_t min_steps);

/* Caps the stopped beam at max_steps layer-0 expansions (0: no cap), a second part of the stop
 * rule (no counterpart in the reference): the launch's time follows its longest queries, and a
 * cap chosen on training traces bounds them. Applies only while a stop model is installed and
 * to learned = 1 batches. */
int hm_set_stop_cap(hm_index *index, int32_t max_steps);

/* Plain register-beam search (unfiltered, k <= ef <= 64, device buffers) that records, for
 * each query and each of its first `steps` layer-0 expansions, the 16 stop features, the number
 * of its current top-k whose distance is <= d_thr[query] (the exact k-th distance plus the
 * recall tolerance) and |R|: d_trace is (nq, steps, 18) float32. d_len: expansions per query;
 * d_final: hits of the final result. */
int hm_trace_dev(hm_index *index, const float *d_queries, int64_t nq, int32_t k, int32_t ef,
                 int exact_u8, const double *d_thr, float *d_trace, int32_t steps,
                 int32_t *d_len, int32_t *d_final, void *stream);

/* hm_trace_dev with host buffers (the library allocates and frees the device copies). */
int hm_trace(hm_index *index, const float *queries, int64_t nq, int32_t k, int32_t ef,
             const double *thr, float *trace, int32_t steps, int32_t *len, int32_t *final_hits);

/* Training kernels (host buffers; x: (n, 16) raw features, y: (n,) 0/1 labels):
 * hm_mlp_grad -- mean binary cross-entropy gradient (positives weighted by pos_weight) of a
 * model, and the loss; hm_mlp_forward -- logits; hm_mlp_train -- Adam over shuffled mini-batches
 * (deterministic for a seed), params in/out, epoch_loss (epochs floats) or null. */
int hm_mlp_grad(const float *x, const float *y, int64_t n, const float *params, float pos_weight,
                float *grad, float *loss);
int hm_mlp_forward(const float *x, int64_t n, const float *params, float *logits);
int hm_mlp_train(const float *x, const float *y, int64_t n, float *params, int32_t epochs,
                 int32_t batch, float lr, float pos_weight, uint64_t seed, float *epoch_loss);

/* Tuning knobs (0 restores the automatic choice): log2 of the visited-hash slots per wave of the
 * register-beam kernel and of the general kernel (larger hashes spill to the global bitset
 * later but leave fewer waves per compute unit), a cap on resident waves per compute unit, a
 * cap on the probe distance of the 16-bit visited slots (beyond it the visited set moves to the
 * global bitset), one_launch = 1 to run a register-beam batch with
 * more queries than wave slots in one launch instead of all descents first (16 lanes per query)
 * and layer 0 longest-first, and descent_cap (1..192): the visited entries one layer of that
 * batched descent may hold before the query's descent moves into the layer-0 launch. None of
 * them changes results. */
int hm_set_tuning(hm_index *index, int32_t fast_hash_log2, int32_t hash_log2,
                  int32_t waves_per_cu, int32_t probe_limit, int32_t one_launch,
                  int32_t descent_cap);

/* Event counters since the last reset (out: 4 int32): out[0] layer searches whose visited set
 * outgrew the LDS hash, out[1] queries whose evicted ties outgrew the fixed tie list of the
 * register beam, out[2] layer searches of the general register beam whose candidates spilled
 * to global memory, out[3] queries hm_search ran again (after a tie overflow, or candidates
 * beyond the spill) on the general register beam with room for every node. Synchronises the
 * device. */
int hm_read_stats(hm_index *index, int32_t *out, int reset);

const char *hm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
