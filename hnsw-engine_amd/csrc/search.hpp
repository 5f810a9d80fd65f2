// Shared definitions between the search kernels and the host launch code.
#pragma once
#include <stdint.h>

namespace hm {

enum Metric : int { kL2 = 0, kIP = 1, kCosine = 2 };

// Distance kernels the search can run with (chosen per batch on the host):
//   kExactU8  vectors and queries are 8-bit integers 0..255; distances are exact integers
//             computed with packed byte dot products (any order is exact, so the result equals
//             the reference's float64 sum bit for bit);
//   kSeqF32   float32 vectors, float64 query, float64 sequential unfused sum (the reference's
//             numba_engine.py:57-70 order);
//   kSeqU8    uint8 codes, float64 code-space query, the same sequential sum
//             (numba_engine.py:674-701).
enum DistMode : int { kExactU8 = 0, kSeqF32 = 1, kSeqU8 = 2 };

struct SearchArgs {
    // graph, repacked: rows of w0 / w1 int32 with -1 beyond the used count
    const int32_t *links0;
    const int32_t *links_up;
    const int32_t *upper_off;  // (n,) row of layer 1 in links_up, -1 for level-0 nodes
    int64_t n;
    int w0, w1;
    int entry, max_level;
    // vectors
    const float *vec_f32;   // (n, dim)
    const uint8_t *vec_u8;  // (n, dimp): exact 8-bit copy (kExactU8) or codes (kSeqU8)
    int dim, dimp;
    int metric;
    // queries: float32 (kExactU8, kSeqF32) or float64 (kSeqU8), (nq, dim)
    const float *q_f32;
    const double *q_f64;
    const uint8_t *q_u8;  // exact mode: query bytes (dim per query) in place of q_f32, or null
    int nq;
    const int32_t *ef_q;  // per query, already max(ef, k); null: ef_all
    const int32_t *k_q;   // per query; null: k_all
    int ef_all, k_all;
    int kstride;
    // filters: a pool of bitsets (mask_words uint32 each); per-query index or one for all
    const uint32_t *mask_pool;
    int64_t mask_words;
    const int32_t *mask_idx;  // per query (-1: unfiltered); null: mask_all
    int mask_all;             // -1: unfiltered
    // outputs
    int64_t *out_ids;
    double *out_d;
    int32_t *out_dc;
    int32_t *out_hops;
    int32_t *out_fail;  // 1: candidate spill overflowed, rerun with a larger spill
    // Two-phase schedule of the register beam: descend_kernel (descend_seq_kernel for the
    // float64 distances) runs every query's upper-layer
    // descent and stores its layer-0 entry (key, id) and counters (pre_hops = -1: not done,
    // the layer-0 launch descends itself); phase 2 starts layer 0 from them, taking queries in
    // `order` (longest expected first). phase 0: one launch does both.
    int phase;
    const int32_t *order;  // queue position -> query index (null: identity)
    // Per position: log-buckets by which the order moves the query earlier (null: none); a
    // filtered query's work grows as its mask's selectivity falls.
    const uint8_t *order_boost;
    uint64_t *pre_key;
    uint32_t *pre_id;
    int32_t *pre_dc, *pre_hops;
    int descent_cap;  // descend_kernel: visited entries per layer before a query falls back
    // persistent work queue and per-wave scratch (slot = global wave index)
    int32_t *queue;
    uint32_t *bitsets;
    int64_t bitset_words;
    uint32_t *vlist;  // per wave slot: vlist_cap ids set in its bitset (clearing list)
    int id_bits;  // node ids are < 2^id_bits
    int probe_limit;  // > 0: cap on the probe distance of 16-bit visited slots
    int vlist_cap;
    uint64_t *spill_key;
    uint32_t *spill_id;
    int64_t spill_cap;
    int r_cap, c_cap, h_log2;
    int lds_per_wave;
    int slots;
    int fast;  // 1: unfiltered, every ef <= 64 (<= 128 when nr = 2) -> the register beam
    int nr;    // register beam: R entries per lane (2: the constant shape with ef <= 128)
    int beam;  // 1: the general register beam (filters, codes, any shape; ef <= 64 nr)
    int nc;    // general register beam: held candidates per lane
    // queue position (after `order`) -> query index; null: the identity. A batch split into
    // kernel classes launches each class over its own subset.
    const int32_t *subset;
    // learned stopping (register-beam kernel only): model parameters (layout below) and the
    // logit above which a beam stops; at least stop_tmin layer-0 expansions run first, at most
    // stop_tmax (0: no cap)
    const float *stop_w;
    float stop_logit;
    int stop_tmin, stop_tmax;
    // traces of the plain register-beam search: per query up to trace_steps rows of
    // kTraceCols floats, the number of expansions and the final tie-aware hit count
    float *trace;
    int trace_steps;
    const double *trace_thr;  // per query: exact k-th distance plus the recall tolerance
    int32_t *trace_len;
    int32_t *trace_final;
    // event counters (atomic adds): [0] layer searches whose visited set moved from the LDS
    // hash to the global bitset, [1] queries whose evicted ties outgrew the register beam,
    // [2] layer searches of the general register beam whose candidates spilled to global memory
    int32_t *stats;
};

// descend_kernel (phase 1 of the two-phase schedule): queries per block of 256 threads, visited
// slots per query, and the largest fill of those slots it allows (3/4, so every probe ends).
constexpr int kDescentGroups = 16;
constexpr int kDescentSlots = 256;
constexpr int kDescentMaxCap = 192;

// Learned-stop model parameters (float32, weights already rounded to bfloat16):
// [0,16) feature means, [16,32) inverse standard deviations, [32, 32+16*64) W1 (row i: feature
// i's weights to the 64 hidden units), then b1 (64), W2 (64), b2 (1).
constexpr int kFeatures = 16;
constexpr int kHidden = 64;
constexpr int kModelW1 = 32;
constexpr int kModelB1 = kModelW1 + kFeatures * kHidden;
constexpr int kModelW2 = kModelB1 + kHidden;
constexpr int kModelB2 = kModelW2 + kHidden;
constexpr int kModelSize = kModelB2 + 1;
constexpr int kModelWords = kFeatures / 2 + 1;        // packed bfloat16 words per hidden unit
constexpr int kTraceCols = kFeatures + 2;  // features, hits, |R|

// Bytes of the visited hash with 1 << h_log2 slots: 16-bit quotient slots when the ids have at
// most 10 bits more than the slot index, else 32-bit slots.
__host__ __device__ inline int visited_hash_bytes(int h_log2, int id_bits) {
    return (id_bits - h_log2 <= 10 ? 2 : 4) << h_log2;
}

// Bytes of LDS one wave of the register-beam kernel needs: query, (64 nr)-entry R scatter buffer
// with its flags (shared with the new-neighbour staging; each with a spare entry that takes the
// stores of entries not kept, so those stores need no lane mask), 64-entry tie list and the
// visited hash (layer 0 with 1 << h_log2 slots, upper layers with at most 512). 128 dimensions,
// 1M ids and 4,096 slots: 10,144 bytes (16 waves per CU).
inline int lds_bytes_fast(int mode, int dim, int dimp, int h_log2, int id_bits, int nr = 1) {
    auto up = [](int v) { return (v + 15) & ~15; };
    int q = mode == kExactU8 ? up(dimp) : up(dim * 8);
    const int upper = h_log2 < 9 ? h_log2 : 9;
    const int h0 = visited_hash_bytes(h_log2, id_bits), hu = visited_hash_bytes(upper, id_bits);
    const int hash = h0 > hu ? h0 : hu;
    return q + nr * 64 * 8 + 16 + 64 * 4 + 64 * 8 + 64 * 4 + nr * 64 * 4 + 16 + hash;
}

// Bytes of LDS one wave of the general register beam needs: query, new ids and row sums, R
// (64 nr entries of key, id and expanded flag), C (64 nc entries of key and id), the radix
// histogram of the candidate refill, the first spilled entry, and the visited hash.
inline int lds_bytes_beam(int mode, int dim, int dimp, int h_log2, int id_bits, int nr, int nc) {
    auto up = [](int v) { return (v + 15) & ~15; };
    const int q = mode == kExactU8 ? up(dimp) : up(dim * 8);
    const int upper = h_log2 < 9 ? h_log2 : 9;
    const int h0 = visited_hash_bytes(h_log2, id_bits), hu = visited_hash_bytes(upper, id_bits);
    return q + 64 * 4 + 64 * 4 + 64 * nr * 16 + 64 * nc * 12 + 256 * 4 + 16 + (h0 > hu ? h0 : hu);
}

// Bytes of LDS one wave needs for the given capacities.
inline int lds_bytes_per_wave(int mode, int dim, int dimp, int r_cap, int c_cap, int h_log2,
                              int id_bits) {
    auto up = [](int v) { return (v + 15) & ~15; };
    int q = mode == kExactU8 ? up(dimp) : up(dim * 8);
    int r = 2 * (up(r_cap * 8) + up(r_cap * 4));
    int c = 2 * (up(c_cap * 8) + up(c_cap * 4));
    int stage = 2 * (64 * 8 + 64 * 4) + 64 * 8 + 64 * 4 + 64 * 4 + 64 * 4;
    const int upper = h_log2 < 9 ? h_log2 : 9;
    const int h0 = visited_hash_bytes(h_log2, id_bits), hu = visited_hash_bytes(upper, id_bits);
    int hash = h0 > hu ? h0 : hu;
    int misc = 256 * 4 + 64;
    return q + r + c + stage + hash + misc;
}

}  // namespace hm
