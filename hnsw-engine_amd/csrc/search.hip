// Batched HNSW search on CDNA4: one wavefront owns one query (Algorithm 5 of the reference,
// numba_engine.py:422-503), fed by a persistent work queue.
//
// Exactness. The reference pops candidates in (distance, id) order, scans a node's links in
// stored order and accepts neighbour j iff |R| < ef or d_j < worst(R) (numba_engine.py:126-145),
// updating R after every accept. For one expansion that serial rule has a closed parallel form:
//
//     accept_j  <=>  #{r in R : d_r <= d_j} + #{i < j new : d_i <= d_j} < ef
//
// (an earlier rejected i with d_i <= d_j forces j's rejection anyway, so counting all earlier
// new neighbours is equivalent). R after the expansion is the ef smallest (d, id) of R and the
// accepted set, which is what evicting the largest tuple after each push yields. The filtered
// variant (numba_engine.py:151-198) counts only allowed i and merges only allowed accepts into R;
// its break needs |R| == ef. Links rows wider than 64 are processed in order, 64 at a time, with
// R and C updated between chunks, which is again the serial order.
//
// State of a query (per wave, in LDS): R (sorted (key, id), <= ef), C (sorted, the candidates
// not yet expanded, capacity c_cap, overflow spilled to global memory with a radix-select
// refill), a visited hash (overflow: a per-wave global bitset), staging for one expansion.
// Candidates with d > worst(R) once R is full can never be expanded (the first such pop ends the
// search), so they are dropped at merge time; d == worst is kept.
#include <utility>

#include "search.hpp"
#include "wave.hpp"
#include "keys.hpp"
#include "seqdist.hpp"

namespace hm {

// ---------------------------------------------------------------------------- LDS layout

struct Lds {
    uint8_t *qb;    // exact: dimp query bytes; sequential: dim doubles
    uint64_t *rk[2];
    uint32_t *ri[2];
    uint64_t *ck[2];
    uint32_t *ci[2];
    uint64_t *sak, *srk;  // accepted (all / allowed) of one expansion, sorted
    uint32_t *sai, *sri;
    uint64_t *nbk;  // new neighbours of one expansion, in link order
    uint32_t *nbi;
    int32_t *nbacc;
    uint32_t *nbal;
    uint32_t *hash;
    uint32_t *hist;  // 256 bins for the refill's radix select
    uint64_t *misc;  // [0..1]: first spilled entry
    // general register beam: R (64 nr entries: key, id, expanded flag) and C (64 nc entries)
    // as the merges leave them, sorted
    uint64_t *bufk, *ck_b;
    uint32_t *bufi, *buff, *ci_b;
};

HM_DEV Lds carve(uint8_t *base, const SearchArgs &a, int mode) {
    auto up = [](int v) { return (v + 15) & ~15; };
    Lds L;
    uint8_t *p = base;
    L.qb = p;
    p += mode == kExactU8 ? up(a.dimp) : up(a.dim * 8);
    for (int b = 0; b < 2; ++b) {
        L.rk[b] = (uint64_t *)p;
        p += up(a.r_cap * 8);
        L.ri[b] = (uint32_t *)p;
        p += up(a.r_cap * 4);
    }
    for (int b = 0; b < 2; ++b) {
        L.ck[b] = (uint64_t *)p;
        p += up(a.c_cap * 8);
        L.ci[b] = (uint32_t *)p;
        p += up(a.c_cap * 4);
    }
    L.sak = (uint64_t *)p; p += 64 * 8;
    L.sai = (uint32_t *)p; p += 64 * 4;
    L.srk = (uint64_t *)p; p += 64 * 8;
    L.sri = (uint32_t *)p; p += 64 * 4;
    L.nbk = (uint64_t *)p; p += 64 * 8;
    L.nbi = (uint32_t *)p; p += 64 * 4;
    L.nbacc = (int32_t *)p; p += 64 * 4;
    L.nbal = (uint32_t *)p; p += 64 * 4;
    L.hash = (uint32_t *)p;
    {
        const int upper = a.h_log2 < 9 ? a.h_log2 : 9;
        const int h0 = visited_hash_bytes(a.h_log2, a.id_bits);
        const int hu = visited_hash_bytes(upper, a.id_bits);
        p += h0 > hu ? h0 : hu;
    }
    L.hist = (uint32_t *)p; p += 256 * 4;
    L.misc = (uint64_t *)p;
    return L;
}

// Compact layout of the register-beam kernel (fields it does not use stay null).
HM_DEV Lds carve_fast(uint8_t *base, const SearchArgs &a, int mode) {
    auto up = [](int v) { return (v + 15) & ~15; };
    Lds L = {};
    uint8_t *p = base;
    L.qb = p;
    p += mode == kExactU8 ? up(a.dimp) : up(a.dim * 8);
    L.rk[0] = (uint64_t *)p; p += 64 * 8 * a.nr + 16;
    L.ck[0] = (uint64_t *)p; p += 64 * 8;
    L.ri[0] = (uint32_t *)p; p += 64 * 4;
    L.ci[0] = (uint32_t *)p; p += 64 * 4;
    L.nbal = (uint32_t *)p; p += 64 * 4 * a.nr + 16;
    // One expansion reads the new ids (nbi) and the partial sums (nbacc) before its merge
    // stages R in rk[0] / ri[0] / nbal, so they share that space.
    L.nbi = L.ri[0];
    L.nbacc = (int32_t *)L.nbal;
    L.hash = (uint32_t *)p;
    return L;
}

// carve_fast for the constant shape (128 query bytes, NR R entries per lane): every size is known
// at compile time, so the LDS addresses become instruction offsets instead of registers.
template <int NR> HM_DEV Lds carve_fast_m16(uint8_t *base) {
    Lds L = {};
    uint8_t *p = base;
    L.qb = p; p += 128;
    L.rk[0] = (uint64_t *)p; p += 64 * 8 * NR + 16;
    L.ck[0] = (uint64_t *)p; p += 64 * 8;
    L.ri[0] = (uint32_t *)p; p += 64 * 4;
    L.ci[0] = (uint32_t *)p; p += 64 * 4;
    L.nbal = (uint32_t *)p; p += 64 * 4 * NR + 16;
    L.nbi = L.ri[0];
    L.nbacc = (int32_t *)L.nbal;
    L.hash = (uint32_t *)p;
    return L;
}

// Layout of the general register beam (layer0_beam).
HM_DEV Lds carve_beam(uint8_t *base, const SearchArgs &a, int mode) {
    auto up = [](int v) { return (v + 15) & ~15; };
    Lds L = {};
    uint8_t *p = base;
    L.qb = p;
    p += mode == kExactU8 ? up(a.dimp) : up(a.dim * 8);
    L.nbi = (uint32_t *)p; p += 64 * 4;
    L.nbacc = (int32_t *)p; p += 64 * 4;
    L.bufk = (uint64_t *)p; p += 64 * a.nr * 8;
    L.bufi = (uint32_t *)p; p += 64 * a.nr * 4;
    L.buff = (uint32_t *)p; p += 64 * a.nr * 4;
    L.ck_b = (uint64_t *)p; p += 64 * a.nc * 8;
    L.ci_b = (uint32_t *)p; p += 64 * a.nc * 4;
    L.hist = (uint32_t *)p; p += 256 * 4;
    L.misc = (uint64_t *)p; p += 16;
    L.hash = (uint32_t *)p;
    return L;
}

// Compile-time shape of the fast kernels: kShapeAny reads the row widths from the arguments;
// kShapeM16D128 is the common build (M = 16: 32 links per layer-0 row, 16 per upper row; 128
// one-byte dimensions), where every loop bound and stride is a constant.
constexpr int kShapeAny = 0;
constexpr int kShapeM16D128 = 1;
constexpr int kShapeM16D128Wide = 2;   // the same shape with two R entries per lane (ef <= 128)
constexpr int kShapeM16D128Wide4 = 3;  // four R entries per lane (ef <= 256)
constexpr int kShapeM16D128Wide8 = 4;  // eight R entries per lane (ef <= 512)

template <int SH> constexpr int nr_of() {
    return SH == kShapeM16D128Wide ? 2
           : SH == kShapeM16D128Wide4 ? 4
           : SH == kShapeM16D128Wide8 ? 8 : 1;
}

// v[s] for a wave-uniform slot index s of a register array (a select chain, no scratch).
template <int NR, typename T> HM_DEV T slot_of(const T (&v)[NR], int s) {
    T r = v[0];
#pragma unroll
    for (int i = 1; i < NR; ++i)
        if (s == i) r = v[i];
    return r;
}
template <int SH> HM_DEV int w0_of(const SearchArgs &a) { return SH != kShapeAny ? 32 : a.w0; }
template <int SH> HM_DEV int w1_of(const SearchArgs &a) { return SH != kShapeAny ? 16 : a.w1; }

// Quotient visited slots (16 bits) are used when the ids have at most this many bits more than
// the table index.
constexpr int kQuotientMaxRemainder = 10;

// The constant-shape kernels also assume quotient slots at layer 0.
inline bool shape_m16d128(const SearchArgs &a, int mode) {
    return mode == kExactU8 && a.w0 == 32 && a.w1 == 16 && a.dimp == 128 &&
           a.id_bits - a.h_log2 <= kQuotientMaxRemainder;
}

// Number of entries of sorted (k, i)[0..n) strictly below (bk, bi).
HM_DEV int lower_count(const uint64_t *k, const uint32_t *id, int n, uint64_t bk, uint32_t bi) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (less(k[mid], id[mid], bk, bi)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Number of entries of sorted k[0..n) with key <= bk (distance-only comparison).
HM_DEV int upper_count_key(const uint64_t *k, int n, uint64_t bk) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (k[mid] <= bk) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------- distances

HM_DEV int32_t exact_partial(uint4 x, uint4 q, int metric) {
    uint32_t xq = 0;
    xq = __builtin_amdgcn_udot4(x.x, q.x, xq, false);
    xq = __builtin_amdgcn_udot4(x.y, q.y, xq, false);
    xq = __builtin_amdgcn_udot4(x.z, q.z, xq, false);
    xq = __builtin_amdgcn_udot4(x.w, q.w, xq, false);
    if (metric != kL2) return (int32_t)xq;
    uint32_t xx = 0;
    xx = __builtin_amdgcn_udot4(x.x, x.x, xx, false);
    xx = __builtin_amdgcn_udot4(x.y, x.y, xx, false);
    xx = __builtin_amdgcn_udot4(x.z, x.z, xx, false);
    xx = __builtin_amdgcn_udot4(x.w, x.w, xx, false);
    return (int32_t)xx - 2 * (int32_t)xq;
}

// Partial sums of dist_exact for a power-of-two chunk count 2^LG per row (LG < 0: 2^lg at
// run time): 64 >> lg rows per pass, each lane one 16-byte chunk, kBatch passes of loads in
// flight; the group's first lane stores the row's sum into L.nbacc.
template <int LG>
HM_DEV void exact_rows(const SearchArgs &a, const Lds &L, int nnew, int lg_rt) {
    const int lane = lane_id();
    const int lg = LG >= 0 ? LG : lg_rt;
    const int cpv = 1 << lg;
    const int c = lane & (cpv - 1);
    const uint4 q4 = ((const uint4 *)L.qb)[c];
    const int vpi = 64 >> lg;
    constexpr int kBatch = 4;
    for (int v0 = 0; v0 < nnew; v0 += vpi * kBatch) {
        // All ids first, then the row loads: every load is in flight before the first is used.
        // Lanes past the last row load that row again (their sums are not stored), so the loads
        // take no per-lane branch.
        uint32_t ids[kBatch];
#pragma unroll
        for (int it = 0; it < kBatch; ++it)
            ids[it] = L.nbi[min(v0 + it * vpi + (lane >> lg), nnew - 1)];
        uint4 xs[kBatch];
#pragma unroll
        for (int it = 0; it < kBatch; ++it) {
            if (v0 + it * vpi < nnew)
                xs[it] = *(const uint4 *)(a.vec_u8 + ((size_t)ids[it] << (lg + 4)) + (size_t)c * 16);
        }
#pragma unroll
        for (int it = 0; it < kBatch; ++it) {
            if (v0 + it * vpi >= nnew) break;
            int v = v0 + it * vpi + (lane >> lg);
            int32_t part = v < nnew ? exact_partial(xs[it], q4, a.metric) : 0;
            part = group_sum(part, cpv);
            if (c == 0 && v < nnew) L.nbacc[v] = part;
        }
    }
}

// Partial sums of dist_exact for rows of any chunk count (784 dimensions: 49 chunks): 16 lanes
// per row, 4 rows per pass, lane j summing chunks j, j + 16, j + 32, ...; the loads of two passes
// (8 rows) are all in flight before the first is used. The group's first lane stores the row's
// sum into L.nbacc.
HM_DEV void exact_rows_strided(const SearchArgs &a, const Lds &L, int nnew) {
    const int lane = lane_id();
    const int j = lane & 15;
    const int cpv = a.dimp >> 4;
    const uint4 *qv = (const uint4 *)L.qb;
    constexpr int kIn = 4, kPass = 2;
    for (int v0 = 0; v0 < nnew; v0 += 4 * kPass) {
        int32_t part[kPass] = {};
        for (int c0 = 0; c0 < cpv; c0 += 16 * kIn) {
            uint4 xs[kPass][kIn];
#pragma unroll
            for (int ps = 0; ps < kPass; ++ps) {
                if (v0 + 4 * ps >= nnew) break;  // wave-uniform
                const int v = min(v0 + 4 * ps + (lane >> 4), nnew - 1);
                const uint4 *row = (const uint4 *)(a.vec_u8 + (size_t)L.nbi[v] * a.dimp);
#pragma unroll
                for (int t = 0; t < kIn; ++t) {
                    const int c = c0 + j + 16 * t;
                    if (c < cpv) xs[ps][t] = row[c];
                }
            }
#pragma unroll
            for (int ps = 0; ps < kPass; ++ps) {
                if (v0 + 4 * ps >= nnew) break;
#pragma unroll
                for (int t = 0; t < kIn; ++t) {
                    const int c = c0 + j + 16 * t;
                    if (c < cpv) part[ps] += exact_partial(xs[ps][t], qv[c], a.metric);
                }
            }
        }
#pragma unroll
        for (int ps = 0; ps < kPass; ++ps) {
            const int v = v0 + 4 * ps + (lane >> 4);
            const int32_t sum = group_sum(part[ps], 16);
            if (j == 0 && v < nnew) L.nbacc[v] = sum;
        }
    }
}

// Exact integer distances of the nnew ids in L.nbi to the 8-bit query. Lanes 0..nnew-1 get
// their key. Every partial sum is an exact integer, so any summation order equals the
// reference's float64 sum.
// Constant shape (128-byte rows): 8 lanes per row, 8 rows per pass, at most 4 passes; each
// pass's row sums reach lanes 0..nnew-1 by ds_bpermute (lane i takes row i's sum from lane
// 8 (i % 8) of pass i / 8), with no round trip through LDS memory.
// NP passes (the pass count is wave-uniform: one straight-line form per count), so that the
// passes' sums interleave and their ds_bpermute results are awaited once.
template <int NP>
HM_DEV int32_t rows128_sums(const SearchArgs &a, const Lds &L, int nnew) {
    const int lane = lane_id();
    const int c = lane & 7;
    const uint4 q4 = ((const uint4 *)L.qb)[c];
    uint32_t ids[NP];
#pragma unroll
    for (int it = 0; it < NP; ++it) ids[it] = L.nbi[min(it * 8 + (lane >> 3), nnew - 1)];
    uint4 xs[NP];
#pragma unroll
    for (int it = 0; it < NP; ++it)
        xs[it] = *(const uint4 *)(a.vec_u8 + ((size_t)ids[it] << 7) + c * 16);
    int32_t row[NP];
#pragma unroll
    for (int it = 0; it < NP; ++it) row[it] = exact_partial(xs[it], q4, a.metric);
#pragma unroll
    for (int it = 0; it < NP; ++it) row[it] = group_sum(row[it], 8);
    const int src = c << 5;  // byte address of lane 8 (lane % 8)
    int32_t got[NP];
#pragma unroll
    for (int it = 0; it < NP; ++it) got[it] = __builtin_amdgcn_ds_bpermute(src, row[it]);
    int32_t s = got[0];
#pragma unroll
    for (int it = 1; it < NP; ++it) s = (lane >> 3) == it ? got[it] : s;
    return s;
}

HM_DEV uint64_t dist_exact_rows128(const SearchArgs &a, const Lds &L, int nnew, int32_t qq) {
    const int lane = lane_id();
    const int np = (nnew + 7) >> 3;  // 1..4
    const int32_t s = np == 1   ? rows128_sums<1>(a, L, nnew)
                      : np == 2 ? rows128_sums<2>(a, L, nnew)
                      : np == 3 ? rows128_sums<3>(a, L, nnew)
                                : rows128_sums<4>(a, L, nnew);
    uint64_t key = a.metric == kL2 ? key_of_exact(s + qq, kL2) : key_of_exact(-s, a.metric);
    __asm__ volatile("" : "+v"(key));  // computed by every lane: the select below stays a select
    return lane < nnew ? key : kInfKey;
}

template <int SH>
HM_DEV uint64_t dist_exact(const SearchArgs &a, const Lds &L, int nnew, int32_t qq) {
    const int lane = lane_id();
    const int cpv = a.dimp >> 4;
    const uint4 *qv = (const uint4 *)L.qb;
    if constexpr (SH != kShapeAny) {
        return dist_exact_rows128(a, L, nnew, qq);
    } else if ((cpv & (cpv - 1)) == 0 && cpv <= 64) {
        exact_rows<-1>(a, L, nnew, __builtin_ctz(cpv));
    } else {
        exact_rows_strided(a, L, nnew);
    }
    wave_sync();
    uint64_t key = kInfKey;
    if (lane < nnew) {
        int32_t s = L.nbacc[lane];
        key = a.metric == kL2 ? key_of_exact(s + qq, kL2) : key_of_exact(-s, a.metric);
    }
    return key;
}

// numba_engine.py:57-70 in its exact operation order: one lane per neighbour, float64,
// left to right, no fused multiply-add (seq_row).
template <bool U8>
HM_DEV uint64_t dist_seq(const SearchArgs &a, const Lds &L, int nnew, int qi) {
    const int lane = lane_id();
    if (lane >= nnew) return kInfKey;
    const uint32_t id = L.nbi[lane];
    if constexpr (U8)
        return key_of_double(seq_row<true>(a.vec_u8 + (size_t)id * a.dimp,
                                           a.q_f64 + (size_t)qi * a.dim, a.dim, a.metric));
    else
        return key_of_double(seq_row<false>(a.vec_f32 + (size_t)id * a.dim,
                                            a.q_f32 + (size_t)qi * a.dim, a.dim, a.metric));
}

// Distances of the nnew new ids to query qi (lanes 0..nnew-1 get their key).
template <int MODE, int SH = kShapeAny>
HM_DEV uint64_t distances(const SearchArgs &a, const Lds &L, int nnew, int32_t qq, int qi) {
    if constexpr (MODE == kExactU8) return dist_exact<SH>(a, L, nnew, qq);
    else return dist_seq<MODE == kSeqU8>(a, L, nnew, qi);
}

template <int MODE>
HM_DEV double decode(uint64_t k, int metric) {
    if constexpr (MODE == kExactU8) return double_of_exact(k, metric);
    else return double_of_key(k, metric);
}

// ---------------------------------------------------------------------------- visited set

// Visited set of one layer search: an open-addressing hash in LDS (linear probing) until it is
// half full, then a per-wave global bitset. Whole-id mode: one 32-bit slot per id (id + 1).
// Quotient mode (when the ids have at most 10 bits more than the table index): 16-bit slots,
// two per word. The id is first mapped by a bijection of [0, 2^B) (multiplication
// by an odd constant modulo 2^B, B = max(id bits, hlog)); the top hlog bits of the image are the
// home slot, the low r = B - hlog bits (the remainder) are stored with the probe distance from
// home: slot = (remainder << (16 - r)) | (distance + 1). A slot therefore determines its id
// exactly, and two entries are equal iff their ids are. A distance beyond the field moves the
// set to the bitset like a full table does.
struct Visited {
    int32_t *stats;
    uint32_t *bits;
    uint32_t *list;  // ids set in the bitset during this layer search (for clearing)
    int list_cap, list_n;  // list_n > list_cap: too many, clear the whole bitset
    int hlog;    // active hash size is 1 << hlog slots
    int count;   // ids inserted in this layer
    int id_bits;  // ids are < 2^id_bits
    int rbits;    // quotient mode: remainder bits
    int probe_limit;  // > 0: quotient mode moves to the bitset beyond this probe distance
    bool use_bits, dirty;
    bool quotient;  // the current layer search uses 16-bit slots
};

constexpr uint32_t kHashMul = 2654435761u;
constexpr uint32_t inverse_odd(uint32_t a) {  // Newton iteration: 3, 6, 12, 24, 48 bits
    uint32_t x = a;
    for (int i = 0; i < 5; ++i) x *= 2u - a * x;
    return x;
}
constexpr uint32_t kHashMulInv = inverse_odd(kHashMul);
static_assert(kHashMul * kHashMulInv == 1u, "kHashMul must be invertible modulo 2^32");


// Zeroes the bitset words this layer search set: through the id list when it held them all,
// otherwise the whole bitset.
HM_DEV void visited_clear(const SearchArgs &a, Visited &V) {
    const int lane = lane_id();
    if (V.list_n <= V.list_cap) {
        for (int i = lane; i < V.list_n; i += 64) V.bits[V.list[i] >> 5] = 0u;
    } else {
        for (int64_t w = lane; w < a.bitset_words; w += 64) V.bits[w] = 0u;
    }
    __threadfence();
    V.dirty = false;
    V.list_n = 0;
}

// Appends the ids of the lanes with `isnew` to the clearing list (bitset mode only).
HM_DEV void visited_note(Visited &V, bool isnew, uint32_t id) {
    if (!V.use_bits) return;
    const uint64_t m = ballot(isnew);
    if (V.list_n + popc(m) <= V.list_cap) {
        if (isnew) V.list[V.list_n + rank_below(m)] = id;
        V.list_n += popc(m);
    } else {
        V.list_n = V.list_cap + 1;
    }
}

HM_DEV void visited_reset(const SearchArgs &a, const Lds &L, Visited &V, int hlog) {
    const int lane = lane_id();
    if (V.dirty) visited_clear(a, V);
    V.use_bits = false;
    V.hlog = hlog;
    V.count = 0;
    V.rbits = V.id_bits > hlog ? V.id_bits - hlog : 0;
    V.quotient = V.rbits <= kQuotientMaxRemainder;
    const int words = V.quotient ? (1 << hlog) / 2 : (1 << hlog);  // a multiple of 4 (hlog >= 6)
    for (int i = 4 * lane; i < words; i += 256) *(uint4 *)&L.hash[i] = make_uint4(0, 0, 0, 0);
    wave_sync();
}

// Moves every hashed id into the bitset (called by the whole wave).
HM_DEV void visited_to_bits(const Lds &L, Visited &V) {
    const int lane = lane_id();
    V.use_bits = true;
    V.dirty = true;
    V.list_n = 0;
    const uint32_t hmask = (1u << V.hlog) - 1u;
    const int r = V.rbits, db = 16 - r;
    for (int i0 = 0; i0 < (1 << V.hlog); i0 += 64) {
        const uint32_t slot = (uint32_t)(i0 + lane);
        uint32_t v;  // id + 1, 0 for an empty slot
        if (V.quotient) {
            const uint32_t t = (L.hash[slot >> 1] >> ((slot & 1u) << 4)) & 0xFFFFu;
            const uint32_t home = (slot - ((t & ((1u << db) - 1u)) - 1u)) & hmask;
            const uint32_t h = (home << r) | (t >> db);
            const uint32_t bmask = (uint32_t)((1ull << (V.hlog + r)) - 1ull);
            v = t ? ((h * kHashMulInv) & bmask) + 1u : 0u;
        } else {
            v = L.hash[slot];
        }
        if (v) atomicOr(&V.bits[(v - 1) >> 5], 1u << ((v - 1) & 31));
        visited_note(V, v != 0u, v - 1u);
    }
    __threadfence();
    if (lane == 0 && V.stats) atomicAdd(&V.stats[0], 1);
}

// Moves the set to the bitset once the hash would be more than half full.
HM_DEV void visited_make_room(const Lds &L, Visited &V, int incoming) {
    if (V.use_bits || V.count + incoming <= (1 << V.hlog) / 2) return;
    visited_to_bits(L, V);
}

// True if `id` was not yet visited in this layer; marks it.
HM_DEV bool visit_bits(const Visited &V, uint32_t id) {
    const uint32_t bit = 1u << (id & 31);
    return (atomicOr(&V.bits[id >> 5], bit) & bit) == 0u;
}

HM_DEV bool visit(const Lds &L, const Visited &V, uint32_t id) {
    if (V.use_bits) return visit_bits(V, id);
    const uint32_t v = id + 1u, mask = (1u << V.hlog) - 1u;
    uint32_t pos = (id * 2654435761u) >> (32 - V.hlog);
    for (;;) {
        uint32_t old = atomicCAS(&L.hash[pos], 0u, v);
        if (old == 0u) return true;
        if (old == v) return false;
        pos = (pos + 1u) & mask;
    }
}

// Quotient-mode visit of the active lanes' ids (distinct within one call); called by every lane.
// A lane walks from its home slot comparing each slot with its own entry at that distance; it
// claims the first empty slot by a CAS on the word holding it (the word is guessed empty first:
// a failed CAS returns its contents, and a stale guess can only show fewer claimed slots than
// there are, since a claimed slot never changes). A lane that runs out of distance bits moves
// the whole set to the bitset (this call's insertions included) and retries there. Returns
// true for a newly inserted id.
HM_DEV bool visit_quotient(const Lds &L, Visited &V, uint32_t id, bool active) {
    const int r = V.rbits, db = 16 - r;
    const uint32_t hmask = (1u << V.hlog) - 1u;
    const uint32_t bmask = (uint32_t)((1ull << (V.hlog + r)) - 1ull);
    const uint32_t h = (id * kHashMul) & bmask;
    const uint32_t home = h >> r, rem = h & ((1u << r) - 1u);
    uint32_t maxd = (1u << db) - 2u;
    if (V.probe_limit > 0 && (uint32_t)V.probe_limit < maxd) maxd = (uint32_t)V.probe_limit;
    bool isnew = false, full = false;
    if (active) {
        // The probe position advances incrementally: word index, the slot's half of it (shift 0
        // or 16) and this lane's content at the current distance (the distance field is the low
        // bits, so the next distance's content is one more).
        const uint32_t wmask = hmask >> 1;
        const uint32_t last = (rem << db) | (maxd + 1u);  // content at the largest distance
        uint32_t word = home >> 1, sh = (home & 1u) << 4, mine = (rem << db) | 1u;
        uint32_t w = 0u;  // contents (or guess) of the word holding the slot
        bool go;
        do {  // branch-free step, one LDS operation: claim a slot that looks empty, else read
            const uint32_t cur = (w >> sh) & 0xFFFFu;
            const uint32_t old = atomicCAS(&L.hash[word], w, cur == 0u ? w | (mine << sh) : w);
            const uint32_t now = (old >> sh) & 0xFFFFu;
            const bool claimed = old == w && cur == 0u;
            const bool other = !claimed && now != mine && now != 0u;
            const bool step = other && mine != last;
            isnew = claimed;
            full = other && mine == last;
            go = step || (!claimed && now == 0u);  // next slot, or this one again with its word
            const bool next_word = step && sh != 0u;  // the slot after an odd one starts a word
            mine += step ? 1u : 0u;
            sh = step ? sh ^ 16u : sh;
            word = (word + (next_word ? 1u : 0u)) & wmask;
            w = next_word ? 0u : old;
        } while (go);
    }
    if (ballot(full)) {
        visited_to_bits(L, V);
        if (full) isnew = visit(L, V, id);
    }
    return isnew;
}

// True for the active lanes whose id was not yet visited in this layer; marks them. Called by
// every lane of the wave.
HM_DEV bool visit_lanes(const Lds &L, Visited &V, uint32_t id, bool active) {
    if (!V.quotient || V.use_bits) return active && visit(L, V, id);
    return visit_quotient(L, V, id, active);
}

// visit_lanes for a layer search known to use quotient slots (the constant-shape kernels at
// layer 0): two forms instead of three keep the probe loop's control flow small.
HM_DEV bool visit_lanes_quotient(const Lds &L, Visited &V, uint32_t id, bool active) {
    if (V.use_bits) return active && visit_bits(V, id);
    return visit_quotient(L, V, id, active);
}

// ---------------------------------------------------------------------------- query state

struct Query {
    int qi;  // index of the query in the batch
    int ef, k;
    const uint32_t *mask;  // null: every node allowed
    int32_t qq;
    int dcount, hops;
};

HM_DEV bool allowed(const Query &Q, uint32_t id) {
    return Q.mask == nullptr || ((Q.mask[id >> 5] >> (id & 31)) & 1u);
}

// Greedy ef = 1 search on an upper layer (numba_engine.py:98-148 with ef = 1): the result is
// the single best node; a neighbour is accepted iff strictly closer than the current best, so
// the candidate after an expansion is the first occurrence of the row's minimum when that beats
// the current best, and the search ends after the first expansion that does not improve.
// uoff is cur_id's first upper row: each expansion reads it for every new neighbour alongside
// their vectors, so the next node's row needs no dependent load.
template <int MODE, int SH>
HM_DEV void greedy_layer(const SearchArgs &a, const Lds &L, Visited &V, Query &Q, int layer,
                         uint64_t &cur_k, uint32_t &cur_id, int32_t &uoff) {
    const int lane = lane_id();
    visited_reset(a, L, V, a.h_log2 < 9 ? a.h_log2 : 9);
    visit_lanes(L, V, cur_id, lane == 0);
    V.count = 1;
    wave_sync();
    for (;;) {
        Q.hops += 1;
        const int64_t row = (int64_t)uoff + layer - 1;
        bool improved = false;
        const int w1 = w1_of<SH>(a);
        for (int base = 0; base < w1; base += 64) {
            const int j = base + lane;
            const int32_t nb = j < w1 ? a.links_up[row * w1 + j] : -1;
            visited_make_room(L, V, 64);
            const bool isnew = visit_lanes(L, V, (uint32_t)nb, nb >= 0);
            visited_note(V, isnew, (uint32_t)nb);
            const uint64_t m = ballot(isnew);
            const int nnew = popc(m);
            if (nnew == 0) continue;
            if (isnew) L.nbi[rank_below(m)] = (uint32_t)nb;
            wave_sync();
            V.count += nnew;
            const int32_t nuo = lane < nnew ? a.upper_off[L.nbi[lane]] : 0;
            uint64_t key = distances<MODE, SH>(a, L, nnew, Q.qq, Q.qi);
            Q.dcount += nnew;
            if constexpr (MODE == kExactU8) {
                // Exact keys fit in 32 bits: the row minimum by DPP, its first lane by ballot.
                const uint32_t k32 = lane < nnew ? (uint32_t)key : ~0u;
                const uint32_t m = wave_min_u32(k32);
                if ((uint64_t)m < cur_k) {
                    const int idx = __ffsll((long long)ballot(lane < nnew && k32 == m)) - 1;
                    cur_k = m;
                    cur_id = L.nbi[idx];
                    uoff = __builtin_amdgcn_readlane(nuo, idx);
                    improved = true;
                }
            } else {
                uint32_t idx = lane < nnew ? (uint32_t)lane : kInfId;
                wave_min_pair(key, idx);
                if (key < cur_k) {
                    cur_k = key;
                    cur_id = L.nbi[idx];
                    uoff = __builtin_amdgcn_readlane(nuo, (int)idx);
                    improved = true;
                }
            }
            wave_sync();
        }
        if (!improved) break;
    }
}

// Selects a prefix of the spilled candidates (by (key, id)) of size 1..c_cap: returns the
// boundary (bk, bi) such that the entries strictly below it form that prefix. Radix select over
// the 96-bit composite (key, id), 8 bits at a time, most significant first. Entries with a key
// above wk take no part; when at most cap entries are left the boundary is (kInfKey, kInfId),
// which selects them all.
HM_DEV void spill_boundary(const Lds &L, const uint64_t *sk, const uint32_t *si, int w, int cap,
                           uint64_t wk, uint64_t &bk, uint32_t &bi) {
    const int lane = lane_id();
    uint64_t pk = 0;
    uint32_t pi = 0;
    for (int t = 0; t < 12; ++t) {
        for (int i = lane; i < 256; i += 64) L.hist[i] = 0u;
        wave_sync();
        for (int b = 0; b < w; b += 64) {
            const int j = b + lane;
            if (j >= w) continue;
            const uint64_t k = sk[j];
            const uint32_t id = si[j];
            if (k > wk) continue;
            bool match;
            if (t == 0) match = true;
            else if (t < 8) match = (k >> (64 - 8 * t)) == (pk >> (64 - 8 * t));
            else if (t == 8) match = k == pk;
            else match = k == pk && (id >> (32 - 8 * (t - 8))) == (pi >> (32 - 8 * (t - 8)));
            if (!match) continue;
            const uint32_t d = t < 8 ? (uint32_t)(k >> (56 - 8 * t)) & 255u
                                     : (id >> (24 - 8 * (t - 8))) & 255u;
            atomicAdd(&L.hist[d], 1u);
        }
        wave_sync();
        // Inclusive scan of the 256 bins, 4 per lane.
        const uint32_t h0 = L.hist[4 * lane], h1 = L.hist[4 * lane + 1];
        const uint32_t h2 = L.hist[4 * lane + 2], h3 = L.hist[4 * lane + 3];
        int s = (int)(h0 + h1 + h2 + h3);
        int incl = s;
        for (int o = 1; o < 64; o <<= 1) {
            int v = shfl(incl, lane >= o ? lane - o : lane);
            if (lane >= o) incl += v;
        }
        const int excl = incl - s;
        if (t == 0 && shfl(incl, 63) <= cap) {  // every entry left fits
            bk = kInfKey;
            bi = kInfId;
            return;
        }
        // First bin whose inclusive count exceeds the capacity.
        int dstar = 256, before = 0;
        if (incl > cap && excl <= cap) {
            int c0 = excl;
            const uint32_t hs[4] = {h0, h1, h2, h3};
            for (int q = 0; q < 4; ++q) {
                if (c0 + (int)hs[q] > cap) {
                    dstar = 4 * lane + q;
                    before = c0;
                    break;
                }
                c0 += (int)hs[q];
            }
        }
        const uint64_t found = ballot(dstar < 256);
        const int src = __ffsll((long long)found) - 1;
        dstar = shfl(dstar, src);
        before = shfl(before, src);
        if (t < 8) pk |= (uint64_t)dstar << (56 - 8 * t);
        else pi |= (uint32_t)dstar << (24 - 8 * (t - 8));
        if (before >= 1) break;  // entries below the boundary: 1..cap of them
        wave_sync();
    }
    bk = pk;
    bi = pi;
}

// ---------------------------------------------------------------------------- layer 0

template <int MODE, bool FILT>
HM_DEV bool layer0(const SearchArgs &a, const Lds &L, Visited &V, Query &Q, int slot,
                   uint64_t entry_k, uint32_t entry_id, int &rb, int &rn) {
    const int lane = lane_id();
    const int ef = Q.ef;
    uint64_t *const sk = a.spill_key + (size_t)slot * a.spill_cap;
    uint32_t *const si = a.spill_id + (size_t)slot * a.spill_cap;

    visited_reset(a, L, V, a.h_log2);
    visit_lanes(L, V, entry_id, lane == 0);
    V.count = 1;

    rb = 0;
    rn = (!FILT || allowed(Q, entry_id)) ? 1 : 0;
    int cb = 0, cbeg = 0, cend = 1;
    if (lane == 0) {
        L.rk[0][0] = entry_k;
        L.ri[0][0] = entry_id;
        L.ck[0][0] = entry_k;
        L.ci[0][0] = entry_id;
    }
    wave_sync();
    uint64_t wk = entry_k;  // worst key of R (meaningful when rn > 0)
    int spn = 0;
    uint64_t smk = kInfKey;
    uint32_t smi = kInfId;

    for (;;) {
        if (cbeg == cend) {
            if (spn == 0) break;
            // Refill C from the spill: prune what can never be expanded, then move the smallest
            // prefix (at most c_cap) into LDS and sort it.
            __threadfence();
            const bool full = rn == ef;
            int w = 0;
            for (int b = 0; b < spn; b += 64) {
                const int j = b + lane;
                const bool v = j < spn;
                const uint64_t k = v ? sk[j] : kInfKey;
                const uint32_t id = v ? si[j] : kInfId;
                const bool kp = v && (!full || k <= wk);
                const uint64_t m = ballot(kp);
                if (kp) {
                    const int d = w + rank_below(m);
                    sk[d] = k;
                    si[d] = id;
                }
                w += popc(m);
            }
            __threadfence();
            if (w == 0) break;
            uint64_t bk = kInfKey;
            uint32_t bi = kInfId;
            if (w > a.c_cap) spill_boundary(L, sk, si, w, a.c_cap, kInfKey, bk, bi);
            int nc = 0, nw = 0;
            uint64_t mk = kInfKey;
            uint32_t mi = kInfId;
            for (int b = 0; b < w; b += 64) {
                const int j = b + lane;
                const bool v = j < w;
                const uint64_t k = v ? sk[j] : kInfKey;
                const uint32_t id = v ? si[j] : kInfId;
                const bool sel = v && less(k, id, bk, bi);
                const bool rest = v && !sel;
                const uint64_t m1 = ballot(sel), m2 = ballot(rest);
                if (sel) {
                    const int d = nc + rank_below(m1);
                    L.ck[0][d] = k;
                    L.ci[0][d] = id;
                }
                if (rest) {
                    const int d = nw + rank_below(m2);
                    sk[d] = k;
                    si[d] = id;
                    if (less(k, id, mk, mi)) {
                        mk = k;
                        mi = id;
                    }
                }
                nc += popc(m1);
                nw += popc(m2);
            }
            __threadfence();
            // A refill always moves at least one candidate; anything else would repeat forever.
            if (nc == 0) return false;
            wave_min_pair(mk, mi);
            smk = mk;
            smi = mi;
            spn = nw;
            int p2 = 1;
            while (p2 < nc) p2 <<= 1;
            bitonic_sort(L.ck[0], L.ci[0], nc, p2);
            cb = 0;
            cbeg = 0;
            cend = nc;
        }
        const uint64_t ck = L.ck[cb][cbeg];
        const uint32_t cid = L.ci[cb][cbeg];
        if (FILT) {
            if (rn == ef && ck > wk) break;
        } else if (ck > wk) {
            break;
        }
        cbeg += 1;
        Q.hops += 1;

        for (int base = 0; base < a.w0; base += 64) {
            const int j = base + lane;
            const int32_t nb = j < a.w0 ? a.links0[(size_t)cid * a.w0 + j] : -1;
            visited_make_room(L, V, 64);
            const bool isnew = visit_lanes(L, V, (uint32_t)nb, nb >= 0);
            visited_note(V, isnew, (uint32_t)nb);
            const uint64_t nm = ballot(isnew);
            const int nnew = popc(nm);
            if (nnew == 0) continue;
            if (isnew) L.nbi[rank_below(nm)] = (uint32_t)nb;
            wave_sync();
            V.count += nnew;
            const uint64_t ki = distances<MODE>(a, L, nnew, Q.qq, Q.qi);
            Q.dcount += nnew;
            const uint32_t ii = lane < nnew ? L.nbi[lane] : kInfId;
            const bool al = FILT && lane < nnew ? allowed(Q, ii) : true;
            L.nbk[lane] = ki;
            L.nbal[lane] = al ? 1u : 0u;
            wave_sync();

            // Parallel form of the serial acceptance rule.
            bool acc = false;
            if (lane < nnew) {
                int cnt = upper_count_key(L.rk[rb], rn, ki);
                for (int q = 0; q < lane; ++q) {
                    if (L.nbk[q] <= ki && (!FILT || L.nbal[q])) cnt += 1;
                }
                acc = cnt < ef;
            }
            const uint64_t am = ballot(acc);
            if (am == 0) continue;
            const uint64_t rm = FILT ? ballot(acc && al) : am;
            const int na = popc(am), nr = popc(rm);
            int ra = 0, rr = 0;
            if (acc) {
                for (int q = 0; q < nnew; ++q) {
                    if (!((am >> q) & 1ull)) continue;
                    if (less(L.nbk[q], L.nbi[q], ki, ii)) {
                        ra += 1;
                        if ((rm >> q) & 1ull) rr += 1;
                    }
                }
                L.sak[ra] = ki;
                L.sai[ra] = ii;
                if ((rm >> lane) & 1ull) {
                    L.srk[rr] = ki;
                    L.sri[rr] = ii;
                }
            }
            wave_sync();

            if (nr > 0) {  // R <- ef smallest of R and the accepted allowed entries
                const int nrb = rb ^ 1;
                const int nrn = min(ef, rn + nr);
                if ((rm >> lane) & 1ull) {
                    const int p = lower_count(L.rk[rb], L.ri[rb], rn, ki, ii) + rr;
                    if (p < ef) {
                        L.rk[nrb][p] = ki;
                        L.ri[nrb][p] = ii;
                    }
                }
                for (int q = lane; q < rn; q += 64) {
                    const uint64_t k = L.rk[rb][q];
                    const uint32_t id = L.ri[rb][q];
                    const int p = q + lower_count(L.srk, L.sri, nr, k, id);
                    if (p < ef) {
                        L.rk[nrb][p] = k;
                        L.ri[nrb][p] = id;
                    }
                }
                wave_sync();
                rb = nrb;
                rn = nrn;
                wk = L.rk[rb][rn - 1];
            }

            {  // C <- live candidates and all accepted entries, pruned, overflow spilled
                const int live = cend - cbeg;
                const uint64_t *ck0 = L.ck[cb] + cbeg;
                const uint32_t *ci0 = L.ci[cb] + cbeg;
                const bool full = rn == ef;
                const int total = live + na;
                const int cut_p =
                    full ? upper_count_key(ck0, live, wk) + upper_count_key(L.sak, na, wk) : total;
                const int cut_s = spn > 0 ? lower_count(ck0, ci0, live, smk, smi) +
                                                lower_count(L.sak, L.sai, na, smk, smi)
                                          : total;
                const int keep = min(a.c_cap, min(cut_p, cut_s));
                const int nsp = cut_p - keep;
                if (nsp > 0 && (int64_t)spn + nsp > a.spill_cap) return false;
                const int ncb = cb ^ 1;
                if (acc) {
                    const int p = lower_count(ck0, ci0, live, ki, ii) + ra;
                    if (p < keep) {
                        L.ck[ncb][p] = ki;
                        L.ci[ncb][p] = ii;
                    } else if (p < cut_p) {
                        sk[spn + p - keep] = ki;
                        si[spn + p - keep] = ii;
                        if (p == keep) {
                            L.misc[0] = ki;
                            L.misc[1] = ii;
                        }
                    }
                }
                for (int q = lane; q < live; q += 64) {
                    const uint64_t k = ck0[q];
                    const uint32_t id = ci0[q];
                    const int p = q + lower_count(L.sak, L.sai, na, k, id);
                    if (p < keep) {
                        L.ck[ncb][p] = k;
                        L.ci[ncb][p] = id;
                    } else if (p < cut_p) {
                        sk[spn + p - keep] = k;
                        si[spn + p - keep] = id;
                        if (p == keep) {
                            L.misc[0] = k;
                            L.misc[1] = id;
                        }
                    }
                }
                wave_sync();
                if (nsp > 0) {
                    const uint64_t fk = L.misc[0];
                    const uint32_t fi = (uint32_t)L.misc[1];
                    if (less(fk, fi, smk, smi)) {
                        smk = fk;
                        smi = fi;
                    }
                    spn += nsp;
                }
                cb = ncb;
                cbeg = 0;
                cend = keep;
            }
        }
    }
    return true;
}


// ---------------------------------------------------------------------------- layer 0, register beam

// Unfiltered layer 0 with ef <= 64, state in registers. Lane i holds R[i] (sorted (key, id));
// a wave-uniform 64-bit mask marks the entries already expanded. In the unfiltered search every
// candidate that can still be expanded is either an unexpanded entry of R or an entry evicted
// from a full R whose key equals the current worst key (anything evicted with a larger key can
// never be popped before the search ends): the first are popped in R order, then the evicted
// ties in id order (they were evicted as the largest tuples, so their ids exceed every R entry
// with that key). When the worst key drops, the ties die. The search ends when no candidate is
// left, which is when the reference's first pop beyond the worst would break.
HM_DEV uint64_t readlane_u64(uint64_t v, int j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

HM_DEV uint32_t readlane_u32(uint32_t v, int j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, j);
}

// Per-expansion features of the beam state, for the learned stop and its traces (float math).
HM_DEV float slog(float x) { return x < 0.f ? -__log2f(1.f - x) : __log2f(1.f + x); }

template <int MODE>
HM_DEV float fdist(uint64_t k, int metric) {
    if constexpr (MODE == kExactU8) {
        if (metric == kIP) return (float)(int32_t)((uint32_t)k ^ 0x80000000u);
        return (float)(uint32_t)k;
    } else {
        return (float)double_of_key(k, metric);
    }
}

struct StopState {
    int t = 0, dc0 = 0, since = 0, last_new = 0, last_acc = 0;
    uint64_t last_kth = kInfKey;
};

// The key of the candidate the search would expand next (the first unexpanded entry of R, else
// the head of the tie list); the worst key of R when none is left.

// The model as the search kernel holds it: per lane (hidden unit) kModelWords words in
// registers for the whole launch: words 0..7 the W1 column as bfloat16 pairs, word 8 b1 (low
// half) and W2 (high half). Registers rather than LDS keep the stop kernel at the plain
// kernel's LDS footprint, so as many waves fit per compute unit.
struct StopModel {
    uint32_t w[kModelWords];
    float b2;
};

HM_DEV uint32_t bf16_bits(float x) {  // the value is already a bfloat16: keep its high half
    return __float_as_uint(x) >> 16;
}

HM_DEV void load_model(const SearchArgs &a, StopModel &M) {
    const int lane = lane_id();
#pragma unroll
    for (int i = 0; i < kFeatures; i += 2)
        M.w[i / 2] = bf16_bits(a.stop_w[kModelW1 + i * kHidden + lane]) |
                     (bf16_bits(a.stop_w[kModelW1 + (i + 1) * kHidden + lane]) << 16);
    M.w[8] = bf16_bits(a.stop_w[kModelB1 + lane]) | (bf16_bits(a.stop_w[kModelW2 + lane]) << 16);
    M.b2 = a.stop_w[kModelB2];
}

// 16 -> 64 (ReLU) -> 1 with float32 accumulation on normalised features (mean and inverse
// standard deviation from the model, uniform across lanes); lane j owns hidden unit j.
HM_DEV float model_logit(const SearchArgs &a, const StopModel &M, const float *f) {
    const uint32_t *w = M.w;
    float z = __uint_as_float(w[8] << 16);
#pragma unroll
    for (int i = 0; i < kFeatures; i += 2) {
        const float x0 = (f[i] - a.stop_w[i]) * a.stop_w[kFeatures + i];
        const float x1 = (f[i + 1] - a.stop_w[i + 1]) * a.stop_w[kFeatures + i + 1];
        z += __uint_as_float(w[i / 2] << 16) * x0;
        z += __uint_as_float(w[i / 2] & 0xffff0000u) * x1;
    }
    const float y = __uint_as_float(w[8] & 0xffff0000u) * fmaxf(z, 0.f);
    // Sum over the 64 hidden units: DPP sums within each row of 16 lanes, then the four row sums.
    const float r = __int_as_float(group_sum_f32(y));
    const float s01 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r), 0)) +
                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r), 16));
    const float s23 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r), 32)) +
                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r), 48));
    return (s01 + s23) + M.b2;
}

template <int MODE>
HM_DEV void features(const SearchArgs &a, const Query &Q, const StopState &S, uint64_t rk,
                     int rn, uint64_t live, uint64_t next_key, float *f) {
    const int kk = max(1, min(Q.k, rn));
    const float dbest = fdist<MODE>(readlane_u64(rk, 0), a.metric);
    const float dk = fdist<MODE>(readlane_u64(rk, kk - 1), a.metric);
    const float dw = fdist<MODE>(readlane_u64(rk, rn - 1), a.metric);
    const float dn = next_key == kInfKey ? dw : fdist<MODE>(next_key, a.metric);
    const float inv_span = __frcp_rn(fabsf(dk - dbest) + 1.f);
    const uint64_t topk = kk >= 64 ? ~0ull : ((1ull << kk) - 1ull);
    // Every feature is wave-uniform: computed by the vector unit, kept in scalar registers.
    auto u = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
    f[0] = (float)S.t;
    f[1] = u(__log2f(1.f + (float)S.dc0));
    f[2] = (float)Q.ef;
    f[3] = (float)Q.k;
    f[4] = u((float)rn * __frcp_rn((float)Q.ef));
    f[5] = u((float)popc(live) * __frcp_rn((float)Q.ef));
    f[6] = u(slog(dbest));
    f[7] = u(slog(dk));
    f[8] = u(slog(dn));
    f[9] = u((dn - dk) * inv_span);
    f[10] = u((dk - dbest) * __frcp_rn(fabsf(dbest) + 1.f));
    f[11] = u((dw - dk) * inv_span);
    f[12] = (float)S.since;
    f[13] = (float)S.last_acc;
    f[14] = (float)S.last_new;
    f[15] = (float)popc(live & topk);  // top-k entries not yet expanded
}

// AUX: 0 plain, 1 trace (plain search that records features and hits), 2 learned stop.
template <int MODE, int AUX, int SH>
HM_DEV bool layer0_fast(const SearchArgs &a, const Lds &L, Visited &V, Query &Q,
                        uint64_t entry_k, uint32_t entry_id, uint64_t &rk, uint32_t &ri,
                        int &rn, int qi, const StopModel &M) {
    const int lane = lane_id();
    StopState S;
    const int ef = Q.ef;
    constexpr int kTies = 64;
    uint64_t *const tk = L.ck[0];  // evicted ties, sorted by id, [th, tn)
    uint32_t *const ti = L.ci[0];
    visited_reset(a, L, V, a.h_log2);
    visit_lanes(L, V, entry_id, lane == 0);
    V.count = 1;
    wave_sync();
    rk = lane == 0 ? entry_k : kInfKey;
    ri = lane == 0 ? entry_id : kInfId;
    rn = 1;
    uint64_t expd = 0;
    int th = 0, tn = 0;
    for (;;) {
        const uint64_t valid = rn >= 64 ? ~0ull : ((1ull << rn) - 1ull);
        const uint64_t live = valid & ~expd;
        uint32_t cid;
        uint64_t ck;
        if (live) {
            const int p = __ffsll((long long)live) - 1;
            cid = readlane_u32(ri, p);
            ck = readlane_u64(rk, p);
            expd |= 1ull << p;
        } else if (th < tn) {
            cid = ti[th];
            ck = tk[th];
            th += 1;
        } else {
            break;
        }
        Q.hops += 1;
        if constexpr (AUX != 0) {
            S.last_new = 0;
            S.last_acc = 0;
        }
        const int w0 = w0_of<SH>(a);
        for (int base = 0; base < w0; base += 64) {
            const int j = base + lane;
            const int32_t nb = j < w0 ? a.links0[(size_t)cid * w0 + j] : -1;
            visited_make_room(L, V, min(64, w0 - base));
            const bool isnew = visit_lanes(L, V, (uint32_t)nb, nb >= 0);
            visited_note(V, isnew, (uint32_t)nb);
            const uint64_t nm = ballot(isnew);
            const int nnew = popc(nm);
            if (nnew == 0) continue;
            if (isnew) L.nbi[rank_below(nm)] = (uint32_t)nb;
            wave_sync();
            V.count += nnew;
            const uint64_t ki = distances<MODE, SH>(a, L, nnew, Q.qq, Q.qi);
            const uint32_t ii = lane < nnew ? L.nbi[lane] : kInfId;
            Q.dcount += nnew;
            if constexpr (AUX != 0) {
                S.dc0 += nnew;
                S.last_new += nnew;
            }
            // accept_i <=> #{r in R: d_r <= d_i} + #{j < i new: d_j <= d_i} < ef. Once R is
            // full, a neighbour at or above its worst key has ef entries of R at or below it
            // (rejected) and is above every neighbour below that key (counts for none of
            // them), so only the lanes strictly below the worst key take part.
            const uint64_t all_new = nnew >= 64 ? ~0ull : ((1ull << nnew) - 1ull);
            const uint64_t worst = readlane_u64(rk, ef - 1);
            const uint64_t part = rn == ef ? ballot(lane < nnew && ki < worst) : all_new;
            int c_r = 0, c_n = 0;
            for (uint64_t m = part; m; m &= m - 1ull) {
                const int q = __ffsll((long long)m) - 1;
                const uint64_t kq = readlane_u64(ki, q);
                const int cnt = popc(ballot(lane < rn && rk <= kq));
                if (lane == q) c_r = cnt;
                if (q < lane && kq <= ki) c_n += 1;
            }
            const bool acc = ((part >> lane) & 1ull) && c_r + c_n < ef;
            const uint64_t am = ballot(acc);
            if (am == 0) continue;
            const int na = popc(am);
            if constexpr (AUX != 0) S.last_acc += na;
            // Merge positions: R entry -> lane + #{accepted below it}; accepted entry ->
            // #{R below it} + #{accepted below it}.
            int pos_r = lane, r_below = 0, a_below = 0;
            for (uint64_t m = am; m; m &= m - 1ull) {
                const int q = __ffsll((long long)m) - 1;
                const uint64_t kq = readlane_u64(ki, q);
                const uint32_t iq = readlane_u32(ii, q);
                if (less(kq, iq, rk, ri)) pos_r += 1;
                const int below = popc(ballot(lane < rn && less(rk, ri, kq, iq)));
                if (lane == q) r_below = below;
                if (acc && less(kq, iq, ki, ii)) a_below += 1;
            }
            const int pos_a = r_below + a_below;
            const int total = rn + na;
            const int nrn = min(ef, total);
            const bool rflag = (expd >> lane) & 1ull;
            if (lane < rn && pos_r < ef) {
                L.rk[0][pos_r] = rk;
                L.ri[0][pos_r] = ri;
                L.nbal[pos_r] = rflag ? 1u : 0u;
            }
            if (acc && pos_a < ef) {
                L.rk[0][pos_a] = ki;
                L.ri[0][pos_a] = ii;
                L.nbal[pos_a] = 0u;
            }
            wave_sync();
            const uint64_t nrk = lane < nrn ? L.rk[0][lane] : kInfKey;
            const uint32_t nri = lane < nrn ? L.ri[0][lane] : kInfId;
            const bool nflag = lane < nrn && L.nbal[lane] != 0u;
            if (total > ef) {
                // Evictions: keep the unexpanded ones whose key equals the new worst.
                const uint64_t nwk = readlane_u64(nrk, ef - 1);
                const bool was_full = rn == ef;
                const uint64_t owk = was_full ? readlane_u64(rk, ef - 1) : kInfKey;
                if (!was_full || nwk < owk) {
                    th = 0;
                    tn = 0;
                }
                // (an R entry and a new entry of the same lane can both be such)
                const bool er = lane < rn && pos_r >= ef && !rflag && rk == nwk;
                const bool ea = acc && pos_a >= ef && ki == nwk;
                const uint64_t emr = ballot(er), ema = ballot(ea);
                if (emr | ema) {
                    const int ne = popc(emr) + popc(ema);
                    // compact the live ties to the front, then append the new ones and sort
                    const int live_t = tn - th;
                    if (live_t + ne > kTies) {
                        if (lane == 0 && a.stats) atomicAdd(&a.stats[1], 1);
                        return false;
                    }
                    uint64_t k0 = 0;
                    uint32_t i0 = 0;
                    if (lane < live_t) {  // live_t <= kTies = 64: one entry per lane
                        k0 = tk[th + lane];
                        i0 = ti[th + lane];
                    }
                    wave_sync();
                    if (lane < live_t) {
                        tk[lane] = k0;
                        ti[lane] = i0;
                    }
                    wave_sync();
                    if (er) {
                        tk[live_t + rank_below(emr)] = rk;
                        ti[live_t + rank_below(emr)] = ri;
                    }
                    if (ea) {
                        tk[live_t + popc(emr) + rank_below(ema)] = ki;
                        ti[live_t + popc(emr) + rank_below(ema)] = ii;
                    }
                    th = 0;
                    tn = live_t + ne;
                    int p2 = 1;
                    while (p2 < tn) p2 <<= 1;
                    bitonic_sort(tk, ti, tn, p2);
                }
            }
            wave_sync();
            rk = nrk;
            ri = nri;
            expd = ballot(nflag);
            rn = nrn;
        }
        if constexpr (AUX != 0) {
            S.t += 1;
            const int kk = max(1, min(Q.k, rn));
            const uint64_t kth = readlane_u64(rk, kk - 1);
            S.since = kth == S.last_kth ? S.since + 1 : 0;
            S.last_kth = kth;
            // The learned stop evaluates (and so computes features) every eighth expansion once
            // stop_tmin expansions ran (every fourth / second: about 3 / 8 % fewer queries/s at the
            // chosen setting); traces record every expansion.
            if (AUX == 2 && a.stop_tmax > 0 && S.t >= a.stop_tmax) break;
            if (AUX == 2 && (S.t < a.stop_tmin || (S.t & 7) != 0)) continue;
            const uint64_t valid_now = rn >= 64 ? ~0ull : ((1ull << rn) - 1ull);
            const uint64_t live_now = valid_now & ~expd;
            uint64_t next_key = kInfKey;
            if (live_now) next_key = readlane_u64(rk, __ffsll((long long)live_now) - 1);
            else if (th < tn) next_key = tk[th];
            float f[kFeatures];
            features<MODE>(a, Q, S, rk, rn, live_now, next_key, f);
            if constexpr (AUX == 1) {
                const double thr = a.trace_thr[qi];
                const int hits = popc(ballot(lane < kk && decode<MODE>(rk, a.metric) <= thr));
                (void)ck;
                if (S.t <= a.trace_steps) {
                    float *row = a.trace + ((size_t)qi * a.trace_steps + (S.t - 1)) * kTraceCols;
                    if (lane < kFeatures) {
                        float v = 0.f;
#pragma unroll
                        for (int i = 0; i < kFeatures; ++i)
                            if (lane == i) v = f[i];
                        row[lane] = v;
                    }
                    if (lane == kFeatures) row[lane] = (float)hits;
                    if (lane == kFeatures + 1) row[lane] = (float)rn;
                }
                if (lane == 0) {
                    a.trace_len[qi] = S.t;
                    a.trace_final[qi] = hits;
                }
            } else if (model_logit(a, M, f) > a.stop_logit) {
                break;
            }
        }
    }
    return true;
}

// layer0_fast for the constant shapes (exact 8-bit distances, whose keys fit in 32 bits): an
// entry of R is one 64-bit word, key << 32 | id, so that (key, id) order is the word's order
// and a key comparison reads the high half. Same algorithm, fewer registers and compares. R
// entry e lives in lane e % 64 of slot e / 64: NR = 1 holds ef <= 64, NR = 2 ef <= 128.
HM_DEV uint32_t key_hi(uint64_t e) { return (uint32_t)(e >> 32); }

template <int AUX, int NR>
HM_DEV bool layer0_packed(const SearchArgs &a, const Lds &L, Visited &V, Query &Q,
                          uint64_t entry_k, uint32_t entry_id, uint64_t *rk_out,
                          uint32_t *ri_out, int &rn, int qi, const StopModel &M) {
    constexpr int MODE = kExactU8;
    constexpr uint64_t kInfEntry = ~0ull;
    const int lane = lane_id();
    StopState S;
    const int ef = Q.ef;
    constexpr int kTies = 64;
    uint64_t *const tk = L.ck[0];  // evicted ties (key, id), sorted, [th, tn)
    uint32_t *const ti = L.ci[0];
    visited_reset(a, L, V, a.h_log2);
    visit_lanes_quotient(L, V, entry_id, lane == 0);
    V.count = 1;
    wave_sync();
    uint64_t re[NR], expd[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) {
        re[s] = kInfEntry;
        expd[s] = 0;
    }
    if (lane == 0) re[0] = (entry_k << 32) | entry_id;
    rn = 1;
    int th = 0, tn = 0;
    // entry lane + 64 s of R exists
    auto in_r = [&](int s, int n) { return lane + 64 * s < n; };
    for (;;) {
        uint32_t cid = 0;
        bool have = false;
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            const int cnt = min(max(rn - 64 * s, 0), 64);
            const uint64_t valid = cnt >= 64 ? ~0ull : ((1ull << cnt) - 1ull);
            const uint64_t live = valid & ~expd[s];
            if (!have && live) {
                const int p = __ffsll((long long)live) - 1;
                cid = readlane_u32((uint32_t)re[s], p);
                expd[s] |= 1ull << p;
                have = true;
            }
        }
        if (!have) {
            if (th >= tn) break;
            cid = ti[th];
            th += 1;
        }
        Q.hops += 1;
        if constexpr (AUX != 0) {
            S.last_new = 0;
            S.last_acc = 0;
        }
        // both half-waves read the row (the same lines), so the load needs no lane mask
        const int32_t raw = a.links0[(size_t)cid * 32 + (lane & 31)];
        const int32_t nb = lane < 32 ? raw : -1;
        visited_make_room(L, V, 32);
        const bool isnew = visit_lanes_quotient(L, V, (uint32_t)nb, nb >= 0);
        visited_note(V, isnew, (uint32_t)nb);
        const uint64_t nm = ballot(isnew);
        const int nnew = popc(nm);
        if (nnew == 0) goto expanded;
        {
            int rkn = rank_below(nm);
            __asm__ volatile("" : "+v"(rkn));  // every lane computes: the store takes no lane mask
            L.nbi[isnew ? rkn : 32 + (lane & 31)] = (uint32_t)nb;
            wave_sync();
            V.count += nnew;
            const uint64_t ki = distances<MODE, kShapeM16D128>(a, L, nnew, Q.qq, Q.qi);
            uint32_t nbl = L.nbi[lane];  // every lane reads (in bounds), so no lane mask
            __asm__ volatile("" : "+v"(nbl));
            const uint64_t ne = lane < nnew ? (ki << 32) | nbl : kInfEntry;
            const uint32_t nk = key_hi(ne);
            Q.dcount += nnew;
            if constexpr (AUX != 0) {
                S.dc0 += nnew;
                S.last_new += nnew;
            }
            // Acceptance as in layer0_fast: only neighbours below the worst key of a full R.
            uint32_t rkey[NR];
#pragma unroll
            for (int s = 0; s < NR; ++s) rkey[s] = key_hi(re[s]);
            const uint64_t all_new = (1ull << nnew) - 1ull;  // nnew <= 32
            const uint32_t worst = readlane_u32(slot_of<NR>(rkey, (ef - 1) >> 6), (ef - 1) & 63);
            const uint64_t part = rn == ef ? lanes_lt_u32(nk, worst) & all_new : all_new;
            uint64_t rvalid[NR];  // lanes of slot s that hold an R entry
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                const int cnt = min(max(rn - 64 * s, 0), 64);
                rvalid[s] = cnt >= 64 ? ~0ull : ((1ull << cnt) - 1ull);
            }
            // One pass over the candidates in link order decides each one (uniformly: #{R keys <=
            // its key} + #{earlier candidates with keys <= its key} < ef) and, when it is
            // accepted, adds it to the merge positions: every R entry above it moves up one, it
            // counts the R entries below it, and every lane counts the accepted entries below its
            // own entry (only the accepted lanes use that count).
            int pos_r[NR];
#pragma unroll
            for (int s = 0; s < NR; ++s) pos_r[s] = lane + 64 * s;
            int r_below = 0, a_below = 0;
            uint64_t am = 0;
            for (uint64_t m = part; m; m &= m - 1ull) {
                const int q = __ffsll((long long)m) - 1;
                const uint64_t eq = readlane_u64(ne, q);
                const uint32_t kq = (uint32_t)(eq >> 32);
                // part & ~m: the candidates before q (m still holds q and those after it)
                int c = popc(lanes_le_u32(nk, kq) & part & ~m);
#pragma unroll
                for (int s = 0; s < NR; ++s) c += popc(lanes_le_u32(rkey[s], kq) & rvalid[s]);
                if (c >= ef) continue;
                am |= 1ull << q;
                int below = 0;
#pragma unroll
                for (int s = 0; s < NR; ++s) {
                    if (eq < re[s]) pos_r[s] += 1;
                    below += popc(lanes_lt_u64(re[s], eq) & rvalid[s]);
                }
                if (lane == q) r_below = below;
                if (eq < ne) a_below += 1;
            }
            if (am == 0) goto expanded;
            const bool acc = (am >> lane) & 1ull;
            const int na = popc(am);
            if constexpr (AUX != 0) S.last_acc += na;
            const int pos_a = r_below + a_below;
            const int total = rn + na;
            const int nrn = min(ef, total);
            bool rflag[NR];
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                rflag[s] = (expd[s] >> lane) & 1ull;
                const int pr = in_r(s, rn) && pos_r[s] < ef ? pos_r[s] : 64 * NR;  // spare slot
                L.rk[0][pr] = re[s];
                L.nbal[pr] = rflag[s] ? 1u : 0u;
            }
            {
                const int pa = acc && pos_a < ef ? pos_a : 64 * NR;
                L.rk[0][pa] = ne;
                L.nbal[pa] = 0u;
            }
            wave_sync();
            uint64_t nre[NR];
            bool nflag[NR];
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                const uint64_t e = L.rk[0][lane + 64 * s];  // every lane reads (in bounds)
                const uint32_t f = L.nbal[lane + 64 * s];
                nre[s] = in_r(s, nrn) ? e : kInfEntry;
                nflag[s] = in_r(s, nrn) && f != 0u;
            }
            if (total > ef) {
                // Evictions: keep the unexpanded ones whose key equals the new worst (an R entry
                // and a new one of the same lane can both be such).
                const uint32_t nwk =
                    key_hi(readlane_u64(slot_of<NR>(nre, (ef - 1) >> 6), (ef - 1) & 63));
                const bool was_full = rn == ef;
                if (!was_full || nwk < worst) {
                    th = 0;
                    tn = 0;
                }
                uint64_t em[NR + 1];
#pragma unroll
                for (int s = 0; s < NR; ++s)
                    em[s] = ballot(in_r(s, rn) && pos_r[s] >= ef && !rflag[s] && rkey[s] == nwk);
                em[NR] = ballot(acc && pos_a >= ef && nk == nwk);
                int nev = 0;
#pragma unroll
                for (int s = 0; s <= NR; ++s) nev += popc(em[s]);
                if (nev) {
                    const int live_t = tn - th;
                    if (live_t + nev > kTies) {
                        if (lane == 0 && a.stats) atomicAdd(&a.stats[1], 1);
                        return false;
                    }
                    uint64_t k0 = 0;
                    uint32_t i0 = 0;
                    if (lane < live_t) {
                        k0 = tk[th + lane];
                        i0 = ti[th + lane];
                    }
                    wave_sync();
                    if (lane < live_t) {
                        tk[lane] = k0;
                        ti[lane] = i0;
                    }
                    wave_sync();
                    int d0 = live_t;
#pragma unroll
                    for (int s = 0; s <= NR; ++s) {
                        if ((em[s] >> lane) & 1ull) {
                            const uint64_t e = s < NR ? re[s < NR ? s : 0] : ne;
                            tk[d0 + rank_below(em[s])] = key_hi(e);
                            ti[d0 + rank_below(em[s])] = (uint32_t)e;
                        }
                        d0 += popc(em[s]);
                    }
                    th = 0;
                    tn = live_t + nev;
                    int p2 = 1;
                    while (p2 < tn) p2 <<= 1;
                    bitonic_sort(tk, ti, tn, p2);
                }
            }
            wave_sync();
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                re[s] = nre[s];
                expd[s] = ballot(nflag[s]);
            }
            rn = nrn;
        }
    expanded:
        if constexpr (AUX != 0) {
            static_assert(NR == 1, "the stop model and traces read a one-slot beam");
            S.t += 1;
            const uint64_t rk = lane < rn ? (uint64_t)key_hi(re[0]) : kInfKey;
            const int kk = max(1, min(Q.k, rn));
            const uint64_t kth = readlane_u64(rk, kk - 1);
            S.since = kth == S.last_kth ? S.since + 1 : 0;
            S.last_kth = kth;
            if (AUX == 2 && a.stop_tmax > 0 && S.t >= a.stop_tmax) break;
            if (AUX == 2 && (S.t < a.stop_tmin || (S.t & 7) != 0)) continue;
            const uint64_t valid_now = rn >= 64 ? ~0ull : ((1ull << rn) - 1ull);
            const uint64_t live_now = valid_now & ~expd[0];
            uint64_t next_key = kInfKey;
            if (live_now) next_key = readlane_u64(rk, __ffsll((long long)live_now) - 1);
            else if (th < tn) next_key = tk[th];
            float f[kFeatures];
            features<MODE>(a, Q, S, rk, rn, live_now, next_key, f);
            if constexpr (AUX == 1) {
                const double thr = a.trace_thr[qi];
                const int hits = popc(ballot(lane < kk && decode<MODE>(rk, a.metric) <= thr));
                if (S.t <= a.trace_steps) {
                    float *row = a.trace + ((size_t)qi * a.trace_steps + (S.t - 1)) * kTraceCols;
                    if (lane < kFeatures) {
                        float v = 0.f;
#pragma unroll
                        for (int i = 0; i < kFeatures; ++i)
                            if (lane == i) v = f[i];
                        row[lane] = v;
                    }
                    if (lane == kFeatures) row[lane] = (float)hits;
                    if (lane == kFeatures + 1) row[lane] = (float)rn;
                }
                if (lane == 0) {
                    a.trace_len[qi] = S.t;
                    a.trace_final[qi] = hits;
                }
            } else if (model_logit(a, M, f) > a.stop_logit) {
                break;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NR; ++s) {
        rk_out[s] = in_r(s, rn) ? (uint64_t)key_hi(re[s]) : kInfKey;
        ri_out[s] = in_r(s, rn) ? (uint32_t)re[s] : kInfId;
    }
    return true;
}

// ---------------------------------------------------------------------------- layer 0, general register beam

// Entries of the general register beam. Exact 8-bit keys are below 2^32, so (key, id) packs into
// one word whose order is the pair's order; a float64 key keeps its 64-bit image and travels
// with the id.
template <bool PACK> struct Ent;

template <> struct Ent<true> {
    uint64_t w;
    static HM_DEV Ent inf() { return Ent{~0ull}; }
    static HM_DEV Ent make(uint64_t k, uint32_t i) { return Ent{(k << 32) | i}; }
    HM_DEV uint64_t key() const { return w >> 32; }
    HM_DEV uint32_t id() const { return (uint32_t)w; }
    HM_DEV bool lt(const Ent &o) const { return w < o.w; }
    HM_DEV Ent at(int l) const { return Ent{readlane_u64(w, l)}; }  // lane l's entry, uniform
    HM_DEV void put(uint64_t *K, uint32_t *, int p) const { K[p] = w; }
    static HM_DEV Ent get(const uint64_t *K, const uint32_t *, int p) { return Ent{K[p]}; }
};

template <> struct Ent<false> {
    uint64_t k;
    uint32_t i;
    static HM_DEV Ent inf() { return Ent{kInfKey, kInfId}; }
    static HM_DEV Ent make(uint64_t k, uint32_t i) { return Ent{k, i}; }
    HM_DEV uint64_t key() const { return k; }
    HM_DEV uint32_t id() const { return i; }
    HM_DEV bool lt(const Ent &o) const { return k < o.k || (k == o.k && i < o.i); }
    HM_DEV Ent at(int l) const { return Ent{readlane_u64(k, l), readlane_u32(i, l)}; }
    HM_DEV void put(uint64_t *K, uint32_t *I, int p) const {
        K[p] = k;
        I[p] = i;
    }
    static HM_DEV Ent get(const uint64_t *K, const uint32_t *I, int p) { return Ent{K[p], I[p]}; }
};

// f(std::integral_constant<int, i>) for i = 0..N-1: slot indices known at compile time, so
// register arrays indexed by them stay in registers.
template <class F, int... I>
HM_DEV void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> HM_DEV void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int N, class E> HM_DEV E ent_slot(const E (&v)[N], int s) {
    E r = v[0];
#pragma unroll
    for (int i = 1; i < N; ++i)
        if (s == i) r = v[i];
    return r;
}

// Lanes of slot s that hold one of the first n entries of a 64-per-slot register array.
HM_DEV uint64_t slot_mask(int n, int s) {
    const int c = min(max(n - 64 * s, 0), 64);
    return c >= 64 ? ~0ull : ((1ull << c) - 1ull);
}

// Layer 0 with R and the other candidates in registers: any metric and distance kernel,
// filtered or not, ef <= 64 NR (numba_engine.py:98-148 and, with FILT, :151-198).
//
// R, the results (allowed nodes only), is sorted (key, id) with entry e in lane e % 64 of slot
// e / 64; a per-slot mask marks the entries already expanded. Every other node that the
// reference would still hold as a candidate is in C: the accepted nodes that are not allowed
// and the unexpanded entries evicted from R. C is sorted too, NC slots in registers, of which
// the first ch are expanded (candidates leave C in order); what does not fit goes to a per-wave
// spill in global memory, every spilled entry larger than every entry of C, and comes back by a
// radix select when C runs dry. Each step expands the smaller of R's first unexpanded entry and
// C's first unexpanded one, and stops when none is left or, with R full, when that candidate
// lies beyond R's worst key: the reference's pop order and break rule. Once R is full a C entry
// beyond the worst key can never be expanded, so merges drop such entries.
//
// One expansion, in link order (rows wider than 64 links 64 at a time): acceptance of new
// neighbour j iff #{r in R: key_r <= key_j} + #{earlier new i (allowed, when filtered): key_i <=
// key_j} < ef (once R is full only keys below its worst take part); R becomes the ef smallest of
// R and the accepted allowed entries; C gains the accepted entries that are not allowed and the
// unexpanded entries pushed out of R.
template <int MODE, bool FILT, int NR, int NC, int SH>
HM_DEV bool layer0_beam(const SearchArgs &a, const Lds &L, Visited &V, Query &Q, int slot,
                        uint64_t entry_k, uint32_t entry_id, uint64_t *rk_out, uint32_t *ri_out,
                        int &rn) {
    using E = Ent<MODE == kExactU8>;
    constexpr int kC = 64 * NC;
    const int lane = lane_id();
    const int ef = Q.ef;
    uint64_t *const sk = a.spill_key + (size_t)slot * a.spill_cap;
    uint32_t *const si = a.spill_id + (size_t)slot * a.spill_cap;
    visited_reset(a, L, V, a.h_log2);
    visit_lanes(L, V, entry_id, lane == 0);
    V.count = 1;
    E r[NR], c[NC];
    uint64_t rexp[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) {
        r[s] = E::inf();
        rexp[s] = 0;
    }
#pragma unroll
    for (int s = 0; s < NC; ++s) c[s] = E::inf();
    int cn = 0, ch = 0;  // C holds entries [0, cn) of which [0, ch) were expanded
    rn = 0;
    {
        const E e0 = E::make(entry_k, entry_id);
        if (!FILT || allowed(Q, entry_id)) {
            if (lane == 0) {
                r[0] = e0;
                e0.put(L.bufk, L.bufi, 0);
            }
            rn = 1;
        } else {
            if (lane == 0) {
                c[0] = e0;
                e0.put(L.ck_b, L.ci_b, 0);
            }
            cn = 1;
        }
        wave_sync();
    }
    int spn = 0;          // spilled candidates
    E smin = E::inf();    // the smallest spilled one
    bool spilled = false;
    // R and C are also kept, sorted, in LDS: entries picked by a wave-uniform index come from
    // there (picking them from the register arrays would index those at run time)
    auto worst_key = [&]() -> uint64_t {  // R's worst key when R is full, else +inf
        return rn == ef ? E::get(L.bufk, L.bufi, ef - 1).key() : kInfKey;
    };
    for (;;) {
        if (ch == cn && spn > 0) {
            // C ran dry: bring back the smallest kC spilled entries. With R full, entries beyond
            // its worst key can never be expanded: the select ignores them and the pass that
            // moves the selected entries to C drops them while it compacts the others in place
            // (they move down, never past one not yet read: 4 x 64 entries are read before any
            // of them is written).
            __threadfence();
            const uint64_t wk = worst_key();
            uint64_t bk = kInfKey;
            uint32_t bi = kInfId;
            spill_boundary(L, sk, si, spn, kC, wk, bk, bi);
            int nc = 0, nw = 0;
            uint64_t mk = kInfKey;
            uint32_t mi = kInfId;
            for (int b = 0; b < spn; b += 256) {
                uint64_t k4[4];
                uint32_t i4[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int j = b + 64 * t + lane;
                    k4[t] = j < spn ? sk[j] : kInfKey;
                    i4[t] = j < spn ? si[j] : kInfId;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bool v = b + 64 * t + lane < spn && k4[t] <= wk;
                    const uint64_t k = k4[t];
                    const uint32_t id = i4[t];
                    const bool sel = v && less(k, id, bk, bi);
                    const bool rest = v && !sel;
                    const uint64_t m1 = ballot(sel), m2 = ballot(rest);
                    if (sel) {
                        L.ck_b[nc + rank_below(m1)] = k;
                        L.ci_b[nc + rank_below(m1)] = id;
                    }
                    if (rest) {
                        sk[nw + rank_below(m2)] = k;
                        si[nw + rank_below(m2)] = id;
                        if (less(k, id, mk, mi)) {
                            mk = k;
                            mi = id;
                        }
                    }
                    nc += popc(m1);
                    nw += popc(m2);
                }
            }
            __threadfence();
            spn = nw;
            smin = E::inf();
            cn = ch = 0;
            if (nc > 0) {  // none when every spilled entry lay beyond the worst key
                wave_min_pair(mk, mi);
                smin = nw > 0 ? E::make(mk, mi) : E::inf();
                int p2 = 1;
                while (p2 < nc) p2 <<= 1;
                bitonic_sort(L.ck_b, L.ci_b, nc, p2);
#pragma unroll
                for (int s = 0; s < NC; ++s) {
                    const int e = lane + 64 * s;
                    c[s] = e < nc ? E::make(L.ck_b[e], L.ci_b[e]) : E::inf();
                    if (e < nc) c[s].put(L.ck_b, L.ci_b, e);  // packed form for the head reads
                }
                cn = nc;
                wave_sync();
            }
        }
        // The next candidate: R's first unexpanded entry or C's, whichever is smaller.
        // (the slot is picked by selects, not by indexing, so that R stays in registers)
        int rs = -1;
        uint64_t rlive = 0;
#pragma unroll
        for (int s = NR - 1; s >= 0; --s) {
            const uint64_t live = slot_mask(rn, s) & ~rexp[s];
            rs = live ? s : rs;
            rlive = live ? live : rlive;
        }
        const int rp = rs >= 0 ? __ffsll((long long)rlive) - 1 : 0;
        const E er = rs >= 0 ? E::get(L.bufk, L.bufi, 64 * rs + rp) : E::inf();
        const E ec = ch < cn ? E::get(L.ck_b, L.ci_b, ch) : E::inf();
        if (rs < 0 && ch >= cn) break;
        E cur;
        if (rs >= 0 && (ch >= cn || er.lt(ec))) {
            cur = er;
#pragma unroll
            for (int s = 0; s < NR; ++s)
                if (s == rs) rexp[s] |= 1ull << rp;
        } else {
            cur = ec;
            ch += 1;
            if (cur.key() > worst_key()) break;  // the reference's break (R entries never are)
        }
        const uint32_t cid = cur.id();
        Q.hops += 1;
        const int w0 = w0_of<SH>(a);
        for (int base = 0; base < w0; base += 64) {
            const int j = base + lane;
            const int32_t nb = j < w0 ? a.links0[(size_t)cid * w0 + j] : -1;
            visited_make_room(L, V, min(64, w0 - base));
            const bool isnew = visit_lanes(L, V, (uint32_t)nb, nb >= 0);
            visited_note(V, isnew, (uint32_t)nb);
            const uint64_t nm = ballot(isnew);
            const int nnew = popc(nm);
            if (nnew == 0) continue;
            if (isnew) L.nbi[rank_below(nm)] = (uint32_t)nb;
            wave_sync();
            V.count += nnew;
            const uint64_t ki = distances<MODE, SH>(a, L, nnew, Q.qq, Q.qi);
            const uint32_t ii = lane < nnew ? L.nbi[lane] : kInfId;
            Q.dcount += nnew;
            const E ne = lane < nnew ? E::make(ki, ii) : E::inf();
            const uint64_t all_new = nnew >= 64 ? ~0ull : ((1ull << nnew) - 1ull);
            const uint64_t alm = FILT ? ballot(lane < nnew && allowed(Q, ii)) : all_new;
            // Acceptance (see above); a neighbour at or above the worst key of a full R is
            // rejected and counts for no neighbour below that key.
            const uint64_t wk = worst_key();
            const uint64_t part = rn == ef ? ballot(lane < nnew && ki < wk) : all_new;
            uint64_t rvalid[NR];
#pragma unroll
            for (int s = 0; s < NR; ++s) rvalid[s] = slot_mask(rn, s);
            int pos_r[NR];
#pragma unroll
            for (int s = 0; s < NR; ++s) pos_r[s] = lane + 64 * s;
            int r_below = 0, a_below = 0;
            uint64_t am = 0, aam = 0;
            for (uint64_t m = part; m; m &= m - 1ull) {
                const int q = __ffsll((long long)m) - 1;
                const E eq = ne.at(q);
                const uint64_t kq = eq.key();
                int cnt = popc(ballot(ki <= kq) & part & alm & ((1ull << q) - 1ull));
#pragma unroll
                for (int s = 0; s < NR; ++s) cnt += popc(ballot(r[s].key() <= kq) & rvalid[s]);
                if (cnt >= ef) continue;
                am |= 1ull << q;
                if (FILT && !((alm >> q) & 1ull)) continue;
                aam |= 1ull << q;
                int below = 0;
#pragma unroll
                for (int s = 0; s < NR; ++s) {
                    if (eq.lt(r[s])) pos_r[s] += 1;
                    below += popc(ballot(r[s].lt(eq)) & rvalid[s]);
                }
                if (lane == q) r_below = below;
                if (eq.lt(ne)) a_below += 1;
            }
            if (am == 0) continue;
            const bool acc = (am >> lane) & 1ull;
            const bool acc_al = (aam >> lane) & 1ull;
            const int naa = popc(aam);
            const int pos_a = r_below + a_below;
            // R <- the ef smallest of R and the accepted allowed entries.
            const int nrn = min(ef, rn + naa);
            E nr[NR];
            uint64_t nexp[NR];
            bool ev_r[NR];
            if (naa > 0) {
#pragma unroll
                for (int s = 0; s < NR; ++s) {
                    if (((rvalid[s] >> lane) & 1ull) && pos_r[s] < ef) {
                        r[s].put(L.bufk, L.bufi, pos_r[s]);
                        L.buff[pos_r[s]] = (uint32_t)((rexp[s] >> lane) & 1ull);
                    }
                }
                if (acc_al && pos_a < ef) {
                    ne.put(L.bufk, L.bufi, pos_a);
                    L.buff[pos_a] = 0u;
                }
                wave_sync();
#pragma unroll
                for (int s = 0; s < NR; ++s) {
                    const int e = lane + 64 * s;
                    const bool in = e < nrn;
                    nr[s] = in ? E::get(L.bufk, L.bufi, e) : E::inf();
                    nexp[s] = ballot(in && L.buff[e] != 0u);
                    ev_r[s] = ((rvalid[s] >> lane) & 1ull) && pos_r[s] >= ef &&
                              !((rexp[s] >> lane) & 1ull);
                }
                wave_sync();
            } else {
#pragma unroll
                for (int s = 0; s < NR; ++s) {
                    nr[s] = r[s];
                    nexp[s] = rexp[s];
                    ev_r[s] = false;
                }
            }
            // C <- C and the new candidates: accepted entries that are not allowed, and the
            // unexpanded entries pushed out of R (the new ones among them included).
            const bool s_new = (acc && !acc_al) || (acc_al && pos_a >= ef);
            uint64_t sm_r[NR];
            int ns = popc(ballot(s_new)), nev = 0;
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                sm_r[s] = ballot(ev_r[s]);
                nev += popc(sm_r[s]);
            }
            ns += nev;
            if (ns > 0) {
                const uint64_t nwk = nrn == ef ? E::get(L.bufk, L.bufi, ef - 1).key() : kInfKey;
                uint64_t clive[NC];
#pragma unroll
                for (int s = 0; s < NC; ++s) clive[s] = slot_mask(cn, s) & ~slot_mask(ch, s);
                int pc[NC];
#pragma unroll
                for (int s = 0; s < NC; ++s) pc[s] = lane + 64 * s - ch;
                int hb_new = 0, xb_new = 0;
                int hb_r[NR], xb_r[NR];
#pragma unroll
                for (int s = 0; s < NR; ++s) hb_r[s] = xb_r[s] = 0;
                // one step per new candidate x: it moves the C entries above it up by one, counts
                // the live C entries below it, and is counted by the new candidates above it
                auto step = [&](const E &x, int q, int holder) {
                    int hb = 0;
#pragma unroll
                    for (int s = 0; s < NC; ++s) {
                        if (x.lt(c[s])) pc[s] += 1;
                        hb += popc(ballot(c[s].lt(x)) & clive[s]);
                    }
                    if (holder < 0) {
                        if (lane == q) hb_new = hb;
                    }
#pragma unroll
                    for (int s = 0; s < NR; ++s)
                        if (holder == s && lane == q) hb_r[s] = hb;
                    if (x.lt(ne)) xb_new += 1;
                    if (nev > 0) {  // entries pushed out of R are new candidates too
#pragma unroll
                        for (int s = 0; s < NR; ++s)
                            if (x.lt(r[s])) xb_r[s] += 1;
                    }
                };
                for (uint64_t m = ballot(s_new); m; m &= m - 1ull) {
                    const int q = __ffsll((long long)m) - 1;
                    step(ne.at(q), q, -1);
                }
                if (nev > 0)
                    static_for<NR>([&](auto si) {
                        constexpr int s = decltype(si)::value;
                        for (uint64_t m = sm_r[s]; m; m &= m - 1ull) {
                            const int q = __ffsll((long long)m) - 1;
                            step(r[s].at(q), q, s);
                        }
                    });
                const int total = cn - ch + ns;
                // sorted, so what lies beyond the new worst key (R full) is a suffix, and so is
                // what is not below the smallest spilled entry
                int cut_p = total, cut_s = total;
                if (nrn == ef || spn > 0) {
                    int np = 0, nsp = 0;
#pragma unroll
                    for (int s = 0; s < NC; ++s) {
                        np += popc(ballot(c[s].key() <= nwk) & clive[s]);
                        nsp += popc(ballot(c[s].lt(smin)) & clive[s]);
                    }
                    np += popc(ballot(s_new && ne.key() <= nwk));
                    nsp += popc(ballot(s_new && ne.lt(smin)));
                    if (nev > 0) {
#pragma unroll
                        for (int s = 0; s < NR; ++s) {
                            np += popc(ballot(r[s].key() <= nwk) & sm_r[s]);
                            nsp += popc(ballot(r[s].lt(smin)) & sm_r[s]);
                        }
                    }
                    if (nrn == ef) cut_p = np;
                    if (spn > 0) cut_s = nsp;
                }
                const int keep = min(kC, min(cut_p, cut_s));
                const int nspill = cut_p - keep;
                if (nspill > 0 && (int64_t)spn + nspill > a.spill_cap) return false;
                if (nspill > 0 && !spilled) {
                    spilled = true;
                    if (lane == 0 && a.stats) atomicAdd(&a.stats[2], 1);
                }
                // place every kept entry in the buffer and every spilled one in the spill; the
                // entry at position keep is the smallest spilled this time
                auto place = [&](const E &x, int p) {
                    if (p < keep) {
                        x.put(L.ck_b, L.ci_b, p);
                    } else if (p < cut_p) {
                        sk[spn + p - keep] = x.key();
                        si[spn + p - keep] = x.id();
                        if (p == keep) {
                            L.misc[0] = x.key();
                            L.misc[1] = x.id();
                        }
                    }
                };
#pragma unroll
                for (int s = 0; s < NC; ++s)
                    if ((clive[s] >> lane) & 1ull) place(c[s], pc[s]);
                if (s_new) place(ne, hb_new + xb_new);
                if (nev > 0) {
#pragma unroll
                    for (int s = 0; s < NR; ++s)
                        if ((sm_r[s] >> lane) & 1ull) place(r[s], hb_r[s] + xb_r[s]);
                }
                wave_sync();
#pragma unroll
                for (int s = 0; s < NC; ++s) {
                    const int e = lane + 64 * s;
                    c[s] = e < keep ? E::get(L.ck_b, L.ci_b, e) : E::inf();
                }
                if (nspill > 0) {
                    const E first = E::make(L.misc[0], (uint32_t)L.misc[1]);
                    if (first.lt(smin)) smin = first;
                    spn += nspill;
                }
                cn = keep;
                ch = 0;
                wave_sync();
            }
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                r[s] = nr[s];
                rexp[s] = nexp[s];
            }
            rn = nrn;
        }
    }
#pragma unroll
    for (int s = 0; s < NR; ++s) {
        const bool in = lane + 64 * s < rn;
        rk_out[s] = in ? r[s].key() : kInfKey;
        ri_out[s] = in ? r[s].id() : kInfId;
    }
    return true;
}

// ---------------------------------------------------------------------------- one query

// One query: qi indexes the batch's per-query arrays, p the schedule arrays (pre_*) of this
// launch. BNR > 0: layer 0 on the general register beam with BNR entries per lane.
template <int MODE, bool FILT, bool FAST, int AUX, int SH, int BNR = 0>
HM_DEV void run_query(const SearchArgs &a, const Lds &L, Visited &V, int qi, int p, int slot,
                      const StopModel &M) {
    const int lane = lane_id();
    Query Q;
    Q.qi = qi;
    Q.ef = a.ef_q ? a.ef_q[qi] : a.ef_all;
    Q.k = a.k_q ? a.k_q[qi] : a.k_all;
    Q.mask = nullptr;
    if (FILT) {
        const int mi = a.mask_idx ? a.mask_idx[qi] : a.mask_all;
        if (mi >= 0) Q.mask = a.mask_pool + (size_t)mi * a.mask_words;
    }
    Q.dcount = 0;
    Q.hops = 0;
    Q.qq = 0;

    // Query into LDS.
    if (MODE == kExactU8) {
        const float *q = a.q_f32 + (size_t)qi * a.dim;
        const uint8_t *q8 = a.q_u8 + (size_t)qi * a.dim;
        int32_t s = 0;
        for (int i = lane; i < a.dimp; i += 64) {
            const int v = i < a.dim ? (a.q_u8 ? (int)q8[i] : (int)q[i]) : 0;
            L.qb[i] = (uint8_t)v;
            s += v * v;
        }
        Q.qq = wave_sum(s);
    } else {
        double *qd = (double *)L.qb;
        for (int i = lane; i < a.dim; i += 64)
            qd[i] = MODE == kSeqU8 ? a.q_f64[(size_t)qi * a.dim + i]
                                   : (double)a.q_f32[(size_t)qi * a.dim + i];
    }
    uint64_t cur_k;
    uint32_t cur_id;
    if (a.phase == 2 && a.pre_hops[p] >= 0) {  // descend_kernel ran the descent
        cur_k = a.pre_key[p];
        cur_id = a.pre_id[p];
        Q.dcount = a.pre_dc[p];
        Q.hops = a.pre_hops[p];
        wave_sync();
    } else {
        if (lane == 0) L.nbi[0] = (uint32_t)a.entry;
        wave_sync();
        // Entry point distance (numba_engine.py:449-451), then greedy descent.
        cur_k = shfl_u64(distances<MODE, SH>(a, L, 1, Q.qq, Q.qi), 0);
        cur_id = (uint32_t)a.entry;
        Q.dcount = 1;
        wave_sync();
        int32_t uoff = a.max_level > 0 ? a.upper_off[cur_id] : 0;
        for (int layer = a.max_level; layer > 0; --layer)
            greedy_layer<MODE, SH>(a, L, V, Q, layer, cur_k, cur_id, uoff);
    }

    int64_t *oi = a.out_ids + (size_t)qi * a.kstride;
    double *od = a.out_d + (size_t)qi * a.kstride;
    int rn = 0;
    bool ok;
    if constexpr (BNR > 0) {
        uint64_t rk[BNR];
        uint32_t ri[BNR];
        ok = layer0_beam<MODE, FILT, BNR, FILT ? 4 : 1, SH>(a, L, V, Q, slot, cur_k, cur_id, rk,
                                                           ri, rn);
        if (!ok) rn = 0;
        const int take = min(Q.k, rn);
#pragma unroll
        for (int s = 0; s < BNR; ++s) {  // lane i of slot s holds R[i + 64 s]
            const int i = lane + 64 * s;
            if (i < a.kstride) {
                const bool own = i < take;
                oi[i] = own ? (int64_t)ri[s] : -1;
                od[i] = own ? decode<MODE>(rk[s], a.metric) : __builtin_huge_val();
            }
        }
        for (int i = 64 * BNR + lane; i < a.kstride; i += 64) {  // k <= ef <= 64 BNR
            oi[i] = -1;
            od[i] = __builtin_huge_val();
        }
    } else if constexpr (FAST) {
        constexpr int NR = nr_of<SH>();
        uint64_t rk[NR];
        uint32_t ri[NR];
        if constexpr (SH != kShapeAny)
            ok = layer0_packed<AUX, NR>(a, L, V, Q, cur_k, cur_id, rk, ri, rn, qi, M);
        else
            ok = layer0_fast<MODE, AUX, SH>(a, L, V, Q, cur_k, cur_id, rk[0], ri[0], rn, qi, M);
        if (!ok) rn = 0;
        const int take = min(Q.k, rn);
#pragma unroll
        for (int s = 0; s < NR; ++s) {  // lane i of slot s holds R[i + 64 s]
            const int i = lane + 64 * s;
            if (i < a.kstride) {
                const bool own = i < take;
                oi[i] = own ? (int64_t)ri[s] : -1;
                od[i] = own ? decode<MODE>(rk[s], a.metric) : __builtin_huge_val();
            }
        }
        for (int i = 64 * NR + lane; i < a.kstride; i += 64) {  // k <= ef <= 64 NR
            oi[i] = -1;
            od[i] = __builtin_huge_val();
        }
    } else {
        int rb = 0;
        ok = layer0<MODE, FILT>(a, L, V, Q, slot, cur_k, cur_id, rb, rn);
        if (!ok) rn = 0;
        const int take = min(Q.k, rn);
        for (int i = lane; i < a.kstride; i += 64) {
            if (i < take) {
                oi[i] = (int64_t)L.ri[rb][i];
                od[i] = decode<MODE>(L.rk[rb][i], a.metric);
            } else {
                oi[i] = -1;
                od[i] = __builtin_huge_val();
            }
        }
    }
    if (lane == 0) {
        a.out_dc[qi] = Q.dcount;
        a.out_hops[qi] = Q.hops;
        a.out_fail[qi] = ok ? 0 : 1;
    }
    wave_sync();
}

template <int MODE, bool FILT, bool FAST, int AUX, int SH, int BNR = 0>
HM_DEV void search_loop(const SearchArgs &a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds_raw[];
    const int slot = blockIdx.x;
    Lds L;
    if constexpr (BNR > 0) L = carve_beam(lds_raw, a, MODE);
    else if constexpr (FAST && SH == kShapeM16D128 && MODE == kExactU8)
        L = carve_fast_m16<1>(lds_raw);  // (the wider beams measured no gain from it)
    else if constexpr (FAST) L = carve_fast(lds_raw, a, MODE);
    else L = carve(lds_raw, a, MODE);
    Visited V;
    V.bits = a.bitsets + (size_t)slot * a.bitset_words;
    V.list = a.vlist + (size_t)slot * a.vlist_cap;
    V.quotient = false;
    V.id_bits = a.id_bits;
    V.rbits = 0;
    V.probe_limit = a.probe_limit;
    V.list_cap = a.vlist_cap;
    V.list_n = 0;
    V.stats = a.stats;
    V.dirty = false;
    V.use_bits = false;
    V.hlog = a.h_log2;
    V.count = 0;
    StopModel M = {};
    if constexpr (AUX == 2) load_model(a, M);
    for (;;) {
        int pos = 0;
        if (lane_id() == 0) pos = atomicAdd(a.queue, 1);
        pos = uniform(shfl(pos, 0));
        if (pos >= a.nq) break;
        const int p = a.order ? uniform(a.order[pos]) : pos;
        const int qi = a.subset ? uniform(a.subset[p]) : p;
        run_query<MODE, FILT, FAST, AUX, SH, BNR>(a, L, V, qi, p, slot, M);
    }
    // Leave the bitset clean for the next launch.
    if (V.dirty) visited_clear(a, V);
}

// General kernel (any ef, filters): LDS-bound, so no occupancy target for the compiler.
template <int MODE, bool FILT>
__global__ __launch_bounds__(64) void search_kernel(SearchArgs a) {
    search_loop<MODE, FILT, false, 0, kShapeAny>(a);
}

// Register-beam kernel: four waves per SIMD (<= 128 VGPRs, no spills at one to four entries per
// lane), which is what the LDS allows anyway with 4096 visited slots (constant shape: 10 KB per
// wave with one entry per lane, 16 waves per CU).
template <int MODE, int AUX, int SH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void search_fast_kernel(
    SearchArgs a) {
    search_loop<MODE, false, true, AUX, SH>(a);
}

// Register-beam kernel with the learned stop (its model in LDS): three waves per SIMD
// (<= 168 VGPRs), the most its LDS allows with 4096 visited slots.
template <int MODE, int SH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SH == kShapeM16D128 ? 4 : 3))) void search_stop_kernel(
    SearchArgs a) {
    search_loop<MODE, false, true, 2, SH>(a);
}

// General register beam (filters, codes, any shape): LDS allows 7-11 waves per CU, so no
// occupancy target for the compiler beyond what the registers of the widest beam need.
template <int MODE, bool FILT, int NR, int SH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void search_beam_kernel(
    SearchArgs a) {
    search_loop<MODE, FILT, false, 0, SH, NR>(a);
}

__global__ void reset_queue_kernel(int32_t *q) { q[0] = 0; }

// Orders queries by their layer-0 entry key, largest first (a counting sort on 256 buckets of
// the key's log2 with 3 fraction bits; order within a bucket is arbitrary): the farther a query
// enters layer 0 from its nearest neighbours, the more expansions it tends to take, and taking
// long queries first shortens a launch's tail. With `relative` (inner product, whose keys are
// 2^31 minus the dot product and so share one log bucket) the key's excess over the batch's
// smallest key is bucketed instead; relative = 2 (float64 keys, key_of_double): that excess is
// a 64-bit number, bucketed by its log2 with 2 fraction bits. One workgroup of 1024 threads; it
// also resets the work queue of the layer-0 launch that follows.
__global__ __launch_bounds__(1024) void order_by_entry_kernel(const uint64_t *key, int nq,
                                                              int relative,
                                                              const uint8_t *boost,
                                                              int32_t *order, int32_t *queue) {
    constexpr int kKeep = 16384;  // batches up to this size keep each query's bucket in LDS
    __shared__ int32_t count[256];
    __shared__ int32_t start[256];
    __shared__ uint8_t kept[kKeep];
    __shared__ unsigned long long kmin;
    const bool keep = nq <= kKeep;
    const int t = threadIdx.x;
    if (t == 0) {
        queue[0] = 0;
        kmin = 0ull;
    }
    if (t < 256) count[t] = 0;
    __syncthreads();
    if (relative) {
        if (t == 0) kmin = ~0ull;
        __syncthreads();
        unsigned long long m = ~0ull;
        for (int i = t; i < nq; i += 1024) m = min(m, (unsigned long long)key[i]);
        atomicMin(&kmin, m);
        __syncthreads();
    }
    const uint64_t base = kmin;
    const bool wide = relative == 2;
    auto bucket = [base, wide](uint64_t k0) {
        const uint64_t k = k0 - base;
        if (wide) {
            if (k == 0) return 255;
            const int e = 63 - __clzll((long long)k);
            const int frac = e >= 2 ? (int)((k >> (e - 2)) & 3u) : (int)((k << (2 - e)) & 3u);
            return 255 - min(255, e * 4 + frac);
        }
        const uint32_t v = (uint32_t)min<uint64_t>(k, 0xFFFFFFFEull) + 1u;  // >= 1
        const int e = 31 - __clz(v);
        const int frac = e >= 3 ? (int)((v >> (e - 3)) & 7u) : (int)((v << (3 - e)) & 7u);
        return 255 - min(255, e * 8 + frac);  // descending key -> ascending bucket
    };
    for (int i = t; i < nq; i += 1024) {
        const int b = boost ? max(0, bucket(key[i]) - (int)boost[i]) : bucket(key[i]);
        if (keep) kept[i] = (uint8_t)b;
        atomicAdd(&count[b], 1);
    }
    __syncthreads();
    if (t < 64) {  // exclusive prefix of the 256 counts: 4 per lane, then a scan over the wave
        const int c0 = count[4 * t], c1 = count[4 * t + 1], c2 = count[4 * t + 2];
        const int own = c0 + c1 + c2 + count[4 * t + 3];
        int incl = own;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = shfl(incl, t >= o ? t - o : t);
            if (t >= o) incl += v;
        }
        const int excl = incl - own;
        start[4 * t] = excl;
        start[4 * t + 1] = excl + c0;
        start[4 * t + 2] = excl + c0 + c1;
        start[4 * t + 3] = excl + c0 + c1 + c2;
    }
    __syncthreads();
    for (int i = t; i < nq; i += 1024) {
        const int b = keep ? (int)kept[i]
                           : (boost ? max(0, bucket(key[i]) - (int)boost[i]) : bucket(key[i]));
        order[atomicAdd(&start[b], 1)] = i;
    }
}

// Upper-layer descent of a whole batch at once (phase 1 of the two-phase schedule) for up to
// 128 one-byte dimensions (rows of 128 bytes) and 16 links per upper row: 16 lanes per query,
// lane j owning link j of the current row, so one greedy hop is one links load, one visited
// probe and one 128-byte vector row per lane and a 16-lane minimum. Same rule and counters as
// greedy_layer (numba_engine.py:98-148 with ef = 1): a hop moves to the first new neighbour
// holding the row's minimum when it is strictly closer, and the layer ends after the first hop
// that does not improve. The visited set of each query's current layer is a 256-slot
// open-addressing table of whole ids (id + 1) in LDS; a layer that would hold more than `cap`
// entries stops the query's descent and flags it (pre_hops = -1, pre_key = all ones, so it
// sorts first), and the layer-0 launch then runs that query's whole descent itself.

__global__ __launch_bounds__(256) void descend_kernel(SearchArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t table[kDescentGroups][kDescentSlots];
    __shared__ __attribute__((aligned(16))) uint8_t qbytes[kDescentGroups][128];
    const int lane = lane_id();
    const int j = lane & 15;
    const int g = threadIdx.x >> 4;
    const int first = lane & 48;  // this query's first lane in the wave
    const int pi = blockIdx.x * kDescentGroups + g;  // position in the launch's schedule arrays
    const bool act = pi < a.nq;
    const int pc = act ? pi : a.nq - 1;  // groups past the batch compute on a valid query
    const int qc = a.subset ? a.subset[pc] : pc;
    uint32_t *const tab = table[g];

    // Query bytes (lane j: bytes 8j..8j+7) and |q|^2, every lane of the block active.
    int32_t qq = 0;
    uint32_t qw[2] = {0u, 0u};
    {
        const float *q = a.q_f32 + (size_t)qc * a.dim;
        const uint8_t *q8 = a.q_u8 + (size_t)qc * a.dim;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int i = 8 * j + t;
            const int v = i < a.dim ? (a.q_u8 ? (int)q8[i] : (int)q[i]) : 0;
            qw[t >> 2] |= (uint32_t)v << (8 * (t & 3));
            qq += v * v;
        }
        *(uint2 *)&qbytes[g][8 * j] = make_uint2(qw[0], qw[1]);
        qq = group_sum(qq, 16);
    }
    auto key_of = [&](int32_t s) -> uint32_t {
        return (uint32_t)(a.metric == kL2 ? key_of_exact(s + qq, kL2) : key_of_exact(-s, a.metric));
    };
    // Entry distance (counted, numba_engine.py:449-451): 8 bytes per lane, summed over the 16.
    const uint32_t entry = (uint32_t)a.entry;
    uint32_t cur_k;
    {
        const uint2 x = *(const uint2 *)(a.vec_u8 + (size_t)entry * 128 + 8 * j);
        uint32_t xq = __builtin_amdgcn_udot4(x.x, qw[0], 0u, false);
        xq = __builtin_amdgcn_udot4(x.y, qw[1], xq, false);
        uint32_t xx = __builtin_amdgcn_udot4(x.x, x.x, 0u, false);
        xx = __builtin_amdgcn_udot4(x.y, x.y, xx, false);
        const int32_t part = a.metric == kL2 ? (int32_t)xx - 2 * (int32_t)xq : (int32_t)xq;
        cur_k = key_of(group_sum(part, 16));
    }
    wave_sync();
    uint32_t cur_id = entry;
    int32_t uoff = a.max_level > 0 ? a.upper_off[entry] : 0;
    int dcount = 1, hops = 0;
    bool flagged = false;
    if (act) {
        for (int layer = a.max_level; layer > 0 && !flagged; --layer) {
#pragma unroll
            for (int s = 0; s < kDescentSlots / 64; ++s)
                *(uint4 *)&tab[64 * s + 4 * j] = make_uint4(0u, 0u, 0u, 0u);
            wave_sync();
            if (j == 0) tab[(cur_id * kHashMul) >> 24] = cur_id + 1u;
            wave_sync();
            int count = 1;
            for (;;) {
                if (count + 16 > a.descent_cap) {
                    flagged = true;
                    break;
                }
                hops += 1;
                const int32_t nb = a.links_up[((int64_t)uoff + layer - 1) * 16 + j];
                bool isnew = false;
                if (nb >= 0) {  // the ids of a row are distinct
                    const uint32_t v = (uint32_t)nb + 1u;
                    uint32_t pos = ((uint32_t)nb * kHashMul) >> 24;
                    for (;;) {
                        const uint32_t old = atomicCAS(&tab[pos], 0u, v);
                        if (old == 0u) {
                            isnew = true;
                            break;
                        }
                        if (old == v) break;
                        pos = (pos + 1u) & (kDescentSlots - 1);
                    }
                }
                const uint32_t gm = (uint32_t)(ballot(isnew) >> first) & 0xFFFFu;
                count += __popc(gm);
                dcount += __popc(gm);
                uint32_t k32 = ~0u;
                int32_t nuo = 0;
                if (isnew) {
                    const uint4 *row = (const uint4 *)(a.vec_u8 + (size_t)nb * 128);
                    const uint4 *qv = (const uint4 *)qbytes[g];
                    uint32_t xq = 0u, xx = 0u;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const uint4 x = row[c];
                        const uint4 y = qv[c];
                        xq = __builtin_amdgcn_udot4(x.x, y.x, xq, false);
                        xq = __builtin_amdgcn_udot4(x.y, y.y, xq, false);
                        xq = __builtin_amdgcn_udot4(x.z, y.z, xq, false);
                        xq = __builtin_amdgcn_udot4(x.w, y.w, xq, false);
                        xx = __builtin_amdgcn_udot4(x.x, x.x, xx, false);
                        xx = __builtin_amdgcn_udot4(x.y, x.y, xx, false);
                        xx = __builtin_amdgcn_udot4(x.z, x.z, xx, false);
                        xx = __builtin_amdgcn_udot4(x.w, x.w, xx, false);
                    }
                    nuo = a.upper_off[nb];
                    k32 = key_of(a.metric == kL2 ? (int32_t)xx - 2 * (int32_t)xq : (int32_t)xq);
                }
                // Minimum over the query's 16 lanes (one DPP row).
                uint32_t m = k32;
                m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0xB1, 0xF, 0xF, true));
                m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0x4E, 0xF, 0xF, true));
                m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0x141, 0xF, 0xF, true));
                m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0x140, 0xF, 0xF, true));
                const uint32_t hit = (uint32_t)(ballot(isnew && k32 == m) >> first) & 0xFFFFu;
                const int src = first + (hit ? __ffs(hit) - 1 : 0);
                const uint32_t nid = (uint32_t)shfl(nb, src);
                const int32_t noff = shfl(nuo, src);
                if (!(m < cur_k)) break;
                cur_k = m;
                cur_id = nid;
                uoff = noff;
            }
            wave_sync();
        }
    }
    if (act && j == 0) {
        a.pre_key[pi] = flagged ? kInfKey : (uint64_t)cur_k;
        a.pre_id[pi] = cur_id;
        a.pre_dc[pi] = dcount;
        a.pre_hops[pi] = flagged ? -1 : hops;
    }
}

// descend_kernel for the float64 sequential distances (codes with float64 queries: U8; float32
// rows with float32 queries: !U8): the same schedule, rule and counters, 16 lanes per query, lane
// j of a query computing link j's distance with seq_row in the reference's order, from the
// query staged in LDS as float64 (a float32 value converts exactly); keys are key_of_double.
// Up to 128 dimensions.
template <bool U8>
__global__ __launch_bounds__(256) void descend_seq_kernel(SearchArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t table[kDescentGroups][kDescentSlots];
    __shared__ __attribute__((aligned(16))) double qlds[kDescentGroups][128];
    const int lane = lane_id();
    const int j = lane & 15;
    const int g = threadIdx.x >> 4;
    const int first = lane & 48;  // this query's first lane in the wave
    const int pi = blockIdx.x * kDescentGroups + g;  // position in the launch's schedule arrays
    const bool act = pi < a.nq;
    const int pc = act ? pi : a.nq - 1;  // groups past the batch compute on a valid query
    const int qc = a.subset ? a.subset[pc] : pc;
    uint32_t *const tab = table[g];
    const double *const qd = qlds[g];
    for (int i = j; i < a.dim; i += 16)
        qlds[g][i] = U8 ? a.q_f64[(size_t)qc * a.dim + i] : (double)a.q_f32[(size_t)qc * a.dim + i];
    wave_sync();
    auto dist = [&](uint32_t id) __attribute__((always_inline)) -> uint64_t {
        const double d = U8 ? seq_row<true, double, false>(a.vec_u8 + (size_t)id * a.dimp, qd,
                                                            a.dim, a.metric)
                            : seq_row<false, double, false>(a.vec_f32 + (size_t)id * a.dim, qd,
                                                             a.dim, a.metric);
        return key_of_double(d);
    };
    // Entry distance (counted, numba_engine.py:449-451): the group's first lane, then shared.
    const uint32_t entry = (uint32_t)a.entry;
    uint64_t cur_k = j == 0 ? dist(entry) : 0ull;
    cur_k = shfl_u64(cur_k, first);
    uint32_t cur_id = entry;
    int32_t uoff = a.max_level > 0 ? a.upper_off[entry] : 0;
    int dcount = 1, hops = 0;
    bool flagged = false;
    if (act) {
        for (int layer = a.max_level; layer > 0 && !flagged; --layer) {
#pragma unroll
            for (int s = 0; s < kDescentSlots / 64; ++s)
                *(uint4 *)&tab[64 * s + 4 * j] = make_uint4(0u, 0u, 0u, 0u);
            wave_sync();
            if (j == 0) tab[(cur_id * kHashMul) >> 24] = cur_id + 1u;
            wave_sync();
            int count = 1;
            for (;;) {
                if (count + 16 > a.descent_cap) {
                    flagged = true;
                    break;
                }
                hops += 1;
                const int32_t nb = a.links_up[((int64_t)uoff + layer - 1) * 16 + j];
                bool isnew = false;
                if (nb >= 0) {  // the ids of a row are distinct
                    const uint32_t v = (uint32_t)nb + 1u;
                    uint32_t pos = ((uint32_t)nb * kHashMul) >> 24;
                    for (;;) {
                        const uint32_t old = atomicCAS(&tab[pos], 0u, v);
                        if (old == 0u) {
                            isnew = true;
                            break;
                        }
                        if (old == v) break;
                        pos = (pos + 1u) & (kDescentSlots - 1);
                    }
                }
                const uint32_t gm = (uint32_t)(ballot(isnew) >> first) & 0xFFFFu;
                count += __popc(gm);
                dcount += __popc(gm);
                uint64_t k64 = kInfKey;
                int32_t nuo = 0;
                if (isnew) {
                    nuo = a.upper_off[nb];
                    k64 = dist((uint32_t)nb);
                }
                // Minimum over the query's 16 lanes.
                uint64_t m = k64;
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) {
                    const uint64_t other = shfl_u64(m, first | ((lane ^ o) & 15));
                    m = other < m ? other : m;
                }
                const uint32_t hit = (uint32_t)(ballot(isnew && k64 == m) >> first) & 0xFFFFu;
                const int src = first + (hit ? __ffs(hit) - 1 : 0);
                const uint32_t nid = (uint32_t)shfl(nb, src);
                const int32_t noff = shfl(nuo, src);
                if (!(m < cur_k)) break;
                cur_k = m;
                cur_id = nid;
                uoff = noff;
            }
            wave_sync();
        }
    }
    if (act && j == 0) {
        a.pre_key[pi] = flagged ? kInfKey : cur_k;
        a.pre_id[pi] = cur_id;
        a.pre_dc[pi] = dcount;
        a.pre_hops[pi] = flagged ? -1 : hops;
    }
}

template __global__ void descend_seq_kernel<true>(SearchArgs);
template __global__ void descend_seq_kernel<false>(SearchArgs);

template __global__ void search_kernel<kExactU8, false>(SearchArgs);
template __global__ void search_kernel<kExactU8, true>(SearchArgs);
template __global__ void search_kernel<kSeqF32, false>(SearchArgs);
template __global__ void search_kernel<kSeqF32, true>(SearchArgs);
template __global__ void search_kernel<kSeqU8, false>(SearchArgs);
template __global__ void search_fast_kernel<kExactU8, 0, kShapeAny>(SearchArgs);
template __global__ void search_fast_kernel<kExactU8, 0, kShapeM16D128>(SearchArgs);
template __global__ void search_fast_kernel<kExactU8, 0, kShapeM16D128Wide>(SearchArgs);
template __global__ void search_fast_kernel<kExactU8, 0, kShapeM16D128Wide4>(SearchArgs);
template __global__ void search_fast_kernel<kExactU8, 0, kShapeM16D128Wide8>(SearchArgs);
template __global__ void search_fast_kernel<kSeqF32, 0, kShapeAny>(SearchArgs);
template __global__ void search_fast_kernel<kSeqU8, 0, kShapeAny>(SearchArgs);
template __global__ void search_fast_kernel<kExactU8, 1, kShapeAny>(SearchArgs);
template __global__ void search_fast_kernel<kSeqF32, 1, kShapeAny>(SearchArgs);
#define HM_BEAM(MODE, FILT, SH)                                          \
    template __global__ void search_beam_kernel<MODE, FILT, 1, SH>(SearchArgs); \
    template __global__ void search_beam_kernel<MODE, FILT, 2, SH>(SearchArgs); \
    template __global__ void search_beam_kernel<MODE, FILT, 4, SH>(SearchArgs); \
    template __global__ void search_beam_kernel<MODE, FILT, 8, SH>(SearchArgs);
HM_BEAM(kExactU8, true, kShapeAny)
HM_BEAM(kExactU8, true, kShapeM16D128)
HM_BEAM(kSeqF32, true, kShapeAny)
HM_BEAM(kExactU8, false, kShapeAny)
HM_BEAM(kSeqF32, false, kShapeAny)
HM_BEAM(kSeqU8, false, kShapeAny)
#undef HM_BEAM
template __global__ void search_stop_kernel<kExactU8, kShapeAny>(SearchArgs);
template __global__ void search_stop_kernel<kExactU8, kShapeM16D128>(SearchArgs);
template __global__ void search_stop_kernel<kSeqF32, kShapeAny>(SearchArgs);

}  // namespace hm

// ---------------------------------------------------------------------------- host launch

namespace hm {

using KernelFn = void (*)(SearchArgs);

static int aux_of(const SearchArgs &a) { return a.trace ? 1 : (a.stop_w ? 2 : 0); }

static KernelFn pick(int mode, bool filt, const SearchArgs &a) {
    const bool fast = a.fast != 0;
    const int aux = aux_of(a);
    const bool m16 = shape_m16d128(a, mode);
    if (aux == 1)
        return mode == kExactU8 ? search_fast_kernel<kExactU8, 1, kShapeAny>
                                : search_fast_kernel<kSeqF32, 1, kShapeAny>;
    if (aux == 2) {
        if (mode != kExactU8) return search_stop_kernel<kSeqF32, kShapeAny>;
        return m16 ? search_stop_kernel<kExactU8, kShapeM16D128>
                   : search_stop_kernel<kExactU8, kShapeAny>;
    }
    if (a.beam) {
#define HM_NR(MODE, FILT, SH)                                                     \
    (a.nr == 1   ? search_beam_kernel<MODE, FILT, 1, SH>                          \
     : a.nr == 2 ? search_beam_kernel<MODE, FILT, 2, SH>                          \
     : a.nr == 4 ? search_beam_kernel<MODE, FILT, 4, SH>                          \
                 : search_beam_kernel<MODE, FILT, 8, SH>)
        if (filt) {
            if (mode == kExactU8)
                return shape_m16d128(a, mode) ? HM_NR(kExactU8, true, kShapeM16D128)
                                              : HM_NR(kExactU8, true, kShapeAny);
            return HM_NR(kSeqF32, true, kShapeAny);
        }
        if (mode == kExactU8) return HM_NR(kExactU8, false, kShapeAny);
        if (mode == kSeqF32) return HM_NR(kSeqF32, false, kShapeAny);
        return HM_NR(kSeqU8, false, kShapeAny);
#undef HM_NR
    }
    if (fast && !filt) {
        if (a.nr == 2) return search_fast_kernel<kExactU8, 0, kShapeM16D128Wide>;
        if (a.nr == 4) return search_fast_kernel<kExactU8, 0, kShapeM16D128Wide4>;
        if (a.nr == 8) return search_fast_kernel<kExactU8, 0, kShapeM16D128Wide8>;
        if (mode == kExactU8)
            return m16 ? search_fast_kernel<kExactU8, 0, kShapeM16D128>
                       : search_fast_kernel<kExactU8, 0, kShapeAny>;
        if (mode == kSeqF32) return search_fast_kernel<kSeqF32, 0, kShapeAny>;
        return search_fast_kernel<kSeqU8, 0, kShapeAny>;
    }
    if (mode == kExactU8) return filt ? search_kernel<kExactU8, true> : search_kernel<kExactU8, false>;
    if (mode == kSeqF32) return filt ? search_kernel<kSeqF32, true> : search_kernel<kSeqF32, false>;
    return search_kernel<kSeqU8, false>;
}

// Blocks (one wave each) that fit on the device at once for this LDS size.
int search_slots(int mode, bool filt, const SearchArgs &a, int lds, int device_cus) {
    KernelFn fn = pick(mode, filt, a);
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, 64, lds) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu * device_cus;
}

hipError_t search_launch(int mode, bool filt, const SearchArgs &a, int blocks, hipStream_t s) {
    KernelFn fn = pick(mode, filt, a);
    if (a.lds_per_wave > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)fn,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           a.lds_per_wave);
        if (e != hipSuccess) return e;
    }
    if (a.phase == 0) {
        hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, a.queue);
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(64), a.lds_per_wave, s, a);
        return hipGetLastError();
    }
    // Two phases: every descent (16 lanes per query), the order, then layer 0 longest first.
    if (a.nq < 1 || a.w1 != 16 || a.dim > 128 || (mode == kExactU8 && a.dimp != 128) ||
        a.descent_cap < 1 || a.descent_cap > kDescentMaxCap)
        return hipErrorInvalidValue;
    const dim3 dgrid((a.nq + kDescentGroups - 1) / kDescentGroups);
    if (mode == kExactU8) hipLaunchKernelGGL(descend_kernel, dgrid, dim3(256), 0, s, a);
    else if (mode == kSeqU8) hipLaunchKernelGGL(descend_seq_kernel<true>, dgrid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(descend_seq_kernel<false>, dgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(order_by_entry_kernel, dim3(1), dim3(1024), 0, s, a.pre_key, a.nq,
                       mode != kExactU8 ? 2 : a.metric == kIP ? 1 : 0, a.order_boost,
                       const_cast<int32_t *>(a.order), a.queue);
    SearchArgs l = a;
    l.phase = 2;
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(64), a.lds_per_wave, s, l);
    return hipGetLastError();
}

}  // namespace hm
