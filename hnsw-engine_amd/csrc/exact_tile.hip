// Exact k-NN on 8-bit integer data with int8 matrix cores, as a bounded range search
// (the reference's oracle, oracle.py:35-126: shortlist, then exact ranking by (distance, id)).
//
// The reference shortlists with a float32 GEMM and a rounding margin, then ranks the shortlist
// in float64. Here every distance is an exact integer (values 0..255 shifted to -128..127,
// v_mfma_i32_32x32x32_i8 accumulating in int32), so the shortlist needs no margin; what it
// needs is a threshold per query that is known to admit the true k nearest before the scan
// starts, so the scan keeps no per-query state beyond one compare. Three kernels:
//
//   bound     a sample of the rows (every sample_every-th 128-row superblock of each chunk):
//             per lane, the T largest "block scores" (the best score among the 16 rows a lane
//             holds for one 32-row MFMA block);
//   threshold per query, V = the k-th largest of those block scores. They belong to disjoint
//             row groups, so at least k rows score >= V, and every true k-nearest row does;
//   range     all rows: a lane whose block score reaches V writes the block's first row and a
//             mask of its 16 rows that reach V to a per-(query, chunk, half) record list
//             (E records; more sets an overflow count);
//   collect   per query: the masked rows, exact distances recomputed from the rows, sort by
//             (d, id), top k.
//             A query whose records overflowed or whose candidates exceed the collect buffer is
//             flagged and re-run by the general kernel.
//
// Scores. The accumulator starts at a per-row value c_r, so a row's score is x'.q' + c_r:
//   l2: c_r = -ceil(|x'|^2 / 2); 2 score + (|x'|^2 mod 2) = 2 x'.q' - |x'|^2, and
//       d = |x'|^2 + |q'|^2 - 2 x'.q' falls as the score rises;
//   ip: c_r = 128 sum(x'); x.q = score + 128 sum(q') + 128^2 dim, d = -x.q.
// For l2 a row with d <= d(V) := |q'|^2 - 2V has 2 score + 1 >= 2V, so score >= V: the range
// pass misses no row the threshold admits.
//
// Rows are read as bytes ^ flip (I8Rows), so the search runs on raw 0..255 rows held by an
// index handle (flip 0x80 per byte) as well as on rows shifted per call (flip 0).
//
// Tiling: a work-group is 8 waves (two per SIMD, so one wave's matrix-core work overlaps the
// other's compares) over QB x 32 queries each (B operands in registers for the whole launch) and
// one chunk of rows, staged 128 rows at a time through double-buffered LDS (XOR-swizzled 16-byte
// slots, so the 32 lanes reading 32 rows hit distinct banks); one barrier per 128 rows. Per
// 32-row block and wave: QB x NCH MFMAs (the next block's operands load meanwhile), then a max
// over each lane's 16 scores and one compare per query block.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "exact_tile.hpp"
#include "keys.hpp"
#include "wave.hpp"

typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(16))) int i32x16;

namespace hm {

HM_DEV int tile_swz(int r) { return (r & 7) ^ ((r >> 3) & 1); }

HM_DEV int32_t max16(const i32x16 &v) {
    const int32_t a = max(max(v[0], v[1]), v[2]);
    const int32_t b = max(max(v[3], v[4]), v[5]);
    const int32_t c = max(max(v[6], v[7]), v[8]);
    const int32_t d = max(max(v[9], v[10]), v[11]);
    const int32_t e = max(max(v[12], v[13]), v[14]);
    return max(max(max(a, b), max(c, d)), max(e, v[15]));
}

// MODE 0: bound pass (sampled superblocks, top-T block scores per lane); MODE 1: range pass.
template <int MODE, int NCH, int QB, int T>
__global__ __launch_bounds__(kTileWaves * 64) void exact_tile_kernel(TileArgs a) {
    constexpr int WPB = kTileWaves;
    constexpr int kTileThreads = WPB * 64;
    constexpr int ROWB = NCH * 32;
    constexpr int SLOTS = 2 * NCH;
    // Superblock: 32 KB of rows (128 or 256 rows), so that one superblock of prefetch covers
    // the memory latency; kTileRows (256) is a multiple of every size.
    constexpr int SBR = NCH >= 8 ? 128 : 256;
    constexpr int SUBS = SBR / 32;
    static_assert(kTileRows % SBR == 0, "chunks are multiples of every superblock size");
    constexpr int ELEMS = SBR * SLOTS;
    constexpr int LPT = (ELEMS + kTileThreads - 1) / kTileThreads;
    __shared__ __attribute__((aligned(16))) uint8_t rows_lds[2][SBR * ROWB];
    __shared__ __attribute__((aligned(16))) int32_t ci_lds[2][SBR];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, col = lane & 31;
    const int chunk = blockIdx.x;
    const int64_t r0 = (int64_t)chunk * a.chunk;
    const int nsb_all = (int)(a.chunk / SBR);
    const int nsb = MODE == 0 ? (nsb_all + a.sample_every - 1) / a.sample_every : nsb_all;

    int qi[QB];
    i32x4 bq[QB][NCH];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        qi[qb] = blockIdx.y * (WPB * QB * 32) + (wave * QB + qb) * 32 + col;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            bq[qb][c] = i32x4{0, 0, 0, 0};
            if (qi[qb] < a.nq)
                bq[qb][c] = *(const i32x4 *)(a.q8 + (size_t)qi[qb] * ROWB + c * 32 + half * 16);
        }
    }
    int32_t V[QB];
    int cnt[QB];
    int32_t top[QB][T];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        cnt[qb] = 0;
        V[qb] = INT_MAX;
        if (MODE == 1 && qi[qb] < a.nq) V[qb] = a.trig[qi[qb]];
#pragma unroll
        for (int t = 0; t < T; ++t) top[qb][t] = INT_MIN;
    }

    // Prefetch registers of the next superblock (plain variables: a private array captured by a
    // lambda stays in scratch memory, and the store to it would wait for the load).
    static_assert(LPT <= 4, "at most four 16-byte loads per thread and superblock");
    uint4 pre0 = uint4{0, 0, 0, 0}, pre1 = pre0, pre2 = pre0, pre3 = pre0;
    int4 pci = int4{0, 0, 0, 0};
    // Rows are ROWB bytes apart and readable up to the end of every chunk (rows at and beyond n
    // are padding whose accumulator start keeps them below any threshold). The shift (^ flip)
    // waits for the stash, so the loads stay in flight through this superblock's matrix work.
    const uint32_t flip = a.rows.flip;
#define HM_TILE_LOAD(S)                                                                       \
    do {                                                                                      \
        const int sb_ = MODE == 0 ? (S) * a.sample_every : (S);                               \
        const int64_t base_ = r0 + (int64_t)sb_ * SBR;                                        \
        const uint8_t *src_ = a.rows.x + (size_t)base_ * ROWB;                                \
        if (tid < ELEMS)                                                                      \
            pre0 = *(const uint4 *)(src_ + (size_t)(tid / SLOTS) * ROWB + (tid % SLOTS) * 16); \
        if (LPT > 1 && tid + kTileThreads < ELEMS)                                            \
            pre1 = *(const uint4 *)(src_ + (size_t)((tid + kTileThreads) / SLOTS) * ROWB +     \
                                    ((tid + kTileThreads) % SLOTS) * 16);                     \
        if (LPT > 2 && tid + 2 * kTileThreads < ELEMS)                                        \
            pre2 = *(const uint4 *)(src_ + (size_t)((tid + 2 * kTileThreads) / SLOTS) * ROWB + \
                                    ((tid + 2 * kTileThreads) % SLOTS) * 16);                 \
        if (LPT > 3 && tid + 3 * kTileThreads < ELEMS)                                        \
            pre3 = *(const uint4 *)(src_ + (size_t)((tid + 3 * kTileThreads) / SLOTS) * ROWB + \
                                    ((tid + 3 * kTileThreads) % SLOTS) * 16);                 \
        if (tid < SBR / 4) pci = *(const int4 *)(a.cinit + base_ + tid * 4);                  \
    } while (0)
    auto stash_one = [&](int buf, int e, uint4 v) {
        v = uint4{v.x ^ flip, v.y ^ flip, v.z ^ flip, v.w ^ flip};
        const int r = e / SLOTS, j = e % SLOTS;
        const int p = (j ^ tile_swz(r)) & (SLOTS - 1);
        *(uint4 *)(rows_lds[buf] + r * ROWB + p * 16) = v;
    };
    auto stash = [&](int buf) {
        if (tid < ELEMS) stash_one(buf, tid, pre0);
        if (LPT > 1 && tid + kTileThreads < ELEMS) stash_one(buf, tid + kTileThreads, pre1);
        if (LPT > 2 && tid + 2 * kTileThreads < ELEMS) stash_one(buf, tid + 2 * kTileThreads, pre2);
        if (LPT > 3 && tid + 3 * kTileThreads < ELEMS) stash_one(buf, tid + 3 * kTileThreads, pre3);
        if (tid < SBR / 4) *(int4 *)(ci_lds[buf] + tid * 4) = pci;
    };
    // Operands of one 32-row block: this lane's row (A) and the accumulator start of the 16
    // rows whose scores it will hold.
    auto operands = [&](int buf, int h2, i32x4 *af, i32x16 &ci) {
        const int r = h2 * 32 + col;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int p = ((2 * c + half) ^ tile_swz(r)) & (SLOTS - 1);
            af[c] = *(const i32x4 *)(rows_lds[buf] + r * ROWB + p * 16);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int4 v = *(const int4 *)(ci_lds[buf] + h2 * 32 + 8 * g + 4 * half);
            ci[4 * g + 0] = v.x;
            ci[4 * g + 1] = v.y;
            ci[4 * g + 2] = v.z;
            ci[4 * g + 3] = v.w;
        }
    };

    // The query operands and thresholds must have landed before the loop: otherwise the wait the
    // compiler places before the first matrix op of every iteration also waits for the next
    // superblock's prefetch (the memory counter is in order), exposing its latency.
    __builtin_amdgcn_s_waitcnt(0);
    if (nsb > 0) HM_TILE_LOAD(0);
    for (int s = 0; s < nsb; ++s) {
        const int buf = s & 1;
        stash(buf);
        __syncthreads();
        if (s + 1 < nsb) HM_TILE_LOAD(s + 1);
        const int sb = MODE == 0 ? s * a.sample_every : s;
        // The next block's operands load while this block's matrix ops run.
        i32x4 af[NCH];
        i32x16 ci;
        operands(buf, 0, af, ci);
#pragma unroll
        for (int h2 = 0; h2 < SUBS; ++h2) {
            i32x16 acc[QB];
#pragma unroll
            for (int c = 0; c < NCH; ++c)
#pragma unroll
                for (int qb = 0; qb < QB; ++qb)
                    acc[qb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[c], bq[qb][c],
                                                                    c == 0 ? ci : acc[qb], 0, 0, 0);
            if (h2 + 1 < SUBS) operands(buf, h2 + 1, af, ci);
            if constexpr (MODE == 0) {
#pragma unroll
                for (int qb = 0; qb < QB; ++qb) {
                    int32_t v = max16(acc[qb]);
#pragma unroll
                    for (int t = 0; t < T; ++t) {
                        const int32_t hi = max(top[qb][t], v);
                        v = min(top[qb][t], v);
                        top[qb][t] = hi;
                    }
                }
            } else {
                int32_t m[QB];
                bool any = false;
#pragma unroll
                for (int qb = 0; qb < QB; ++qb) {
                    m[qb] = max16(acc[qb]);
                    any |= m[qb] >= V[qb];
                }
                if (ballot(any)) {
                    const int32_t row0 = (int32_t)(r0 + (int64_t)sb * SBR + h2 * 32);
#pragma unroll
                    for (int qb = 0; qb < QB; ++qb) {
                        if (m[qb] >= V[qb]) {
                            const size_t slot = ((size_t)qi[qb] * a.nchunks + chunk) * 2 + half;
                            if (cnt[qb] < a.E) {
                                const size_t rec = slot * a.E + cnt[qb];
                                uint32_t mask = 0;
#pragma unroll
                                for (int i = 0; i < 16; ++i)
                                    mask |= (acc[qb][i] >= V[qb] ? 1u : 0u) << i;
                                a.rec_mask[rec] = (int32_t)mask;
                                a.rec_row[rec] = row0;
                            }
                            cnt[qb] += 1;
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        if (qi[qb] >= a.nq) continue;
        const size_t slot = ((size_t)qi[qb] * a.nchunks + chunk) * 2 + half;
        if constexpr (MODE == 0) {
#pragma unroll
            for (int t = 0; t < T; ++t) a.bnd[slot * T + t] = top[qb][t];
        } else {
            a.rec_cnt[slot] = cnt[qb];
        }
    }
#undef HM_TILE_LOAD
}

// V for each query: the k-th largest of its per_query sampled block scores (one wave per
// query; bisection on the value with a wave-wide count). Up to 1024 scores stay in registers
// (16 per lane), so a bisection step is 16 compares and a wave sum with no memory access; more
// go through LDS.
__global__ __launch_bounds__(64) void exact_threshold_kernel(const int32_t *bnd, int per_query,
                                                             int k, int32_t *trig) {
    extern __shared__ __attribute__((aligned(16))) int32_t vals[];
    const int q = blockIdx.x, lane = lane_id();
    if (per_query <= 1024) {
        int32_t v[16];
        int32_t lo = INT_MAX, hi = INT_MIN;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int i = lane + 64 * j;
            v[j] = i < per_query ? bnd[(size_t)q * per_query + i] : INT_MIN;
            if (i < per_query) {
                lo = min(lo, v[j]);
                hi = max(hi, v[j]);
            }
        }
        for (int m = 32; m >= 1; m >>= 1) {
            lo = min(lo, shfl_xor(lo, m));
            hi = max(hi, shfl_xor(hi, m));
        }
        // Largest x with #{v >= x} >= k; #{v >= lo} = per_query >= k. Padding (INT_MIN) is
        // counted only at x = INT_MIN, where the answer is lo = INT_MIN anyway.
        int64_t L = lo, H = hi;
        while (L < H) {
            const int64_t mid = L + (H - L + 1) / 2;
            int c = 0;
#pragma unroll
            for (int j = 0; j < 16; ++j) c += (int64_t)v[j] >= mid ? 1 : 0;
            c = wave_sum(c);
            if (c >= k) L = mid;
            else H = mid - 1;
        }
        if (lane == 0) trig[q] = (int32_t)L;
        return;
    }
    int32_t lo = INT_MAX, hi = INT_MIN;
    for (int i = lane; i < per_query; i += 64) {
        const int32_t v = bnd[(size_t)q * per_query + i];
        vals[i] = v;
        lo = min(lo, v);
        hi = max(hi, v);
    }
    for (int m = 32; m >= 1; m >>= 1) {
        lo = min(lo, shfl_xor(lo, m));
        hi = max(hi, shfl_xor(hi, m));
    }
    wave_sync();
    // Largest x with #{v >= x} >= k; #{v >= lo} = per_query >= k.
    int64_t L = lo, H = hi;
    while (L < H) {
        const int64_t mid = L + (H - L + 1) / 2;
        int c = 0;
        for (int i = lane; i < per_query; i += 64) c += vals[i] >= mid ? 1 : 0;
        c = wave_sum(c);
        if (c >= k) L = mid;
        else H = mid - 1;
    }
    if (lane == 0) trig[q] = (int32_t)L;
}

// The exact distance of a row (bytes ^ flip: shifted values) to a shifted query, over the first
// dim values: l2 sum (x' - q')^2, ip -(sum (x' + 128)(q' + 128)); integers, exact in int32.
HM_DEV int32_t exact_int_dist(const uint8_t *row, uint32_t flip, const int8_t *q, int dim,
                              int metric) {
    int32_t acc = 0;
    for (int j0 = 0; j0 < dim; j0 += 16) {
        const uint4 w = *(const uint4 *)(row + j0);
        const uint4 u = *(const uint4 *)(q + j0);
        const uint32_t xs[4] = {w.x ^ flip, w.y ^ flip, w.z ^ flip, w.w ^ flip};
        const uint32_t qs[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            if (j0 + b >= dim) break;
            const int xv = (int)(int8_t)(xs[b >> 2] >> (8 * (b & 3)));
            const int qv = (int)(int8_t)(qs[b >> 2] >> (8 * (b & 3)));
            if (metric == kL2) acc += (xv - qv) * (xv - qv);
            else acc += (xv + 128) * (qv + 128);
        }
    }
    return metric == kL2 ? acc : -acc;
}

// Per query: every recorded row scoring >= V (the record masks), its exact distance recomputed
// from the row, sorted by (distance, id), the k smallest written out. Sets flag[q] (and writes
// nothing else) when records overflowed or the candidates exceed the buffer. Lane l owns slots
// l, l + 64, l + 128, l + 192 of each group of 256: their counts load together, then record e
// of all four in one round of loads.
__global__ __launch_bounds__(64) void exact_collect_kernel(CollectArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    uint64_t *ck = (uint64_t *)lds;
    uint32_t *cid = (uint32_t *)(lds + (size_t)a.cap * 8);
    int *ncand = (int *)(lds + (size_t)a.cap * 12);
    const int q = blockIdx.x, lane = lane_id();
    const int slots = a.nchunks * 2;
    const int32_t *cnt_q = a.rec_cnt + (size_t)q * slots;
    if (lane == 0) *ncand = 0;
    wave_sync();
    bool over = false;
    for (int base = 0; base < slots; base += 256) {
        int c[4];
        int emax = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = base + j * 64 + lane;
            c[j] = s < slots ? cnt_q[s] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (c[j] > a.E) {
                over = true;
                c[j] = a.E;
            }
            emax = max(emax, c[j]);
        }
        for (int e = 0; e < emax; ++e) {
            int32_t row0[4], mask[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                mask[j] = 0;
                if (e >= c[j]) continue;
                const size_t rec = ((size_t)q * slots + base + j * 64 + lane) * a.E + e;
                row0[j] = a.rec_row[rec];
                mask[j] = a.rec_mask[rec];
            }
            for (int j = 0; j < 4; ++j) {
                const int half = (base + j * 64 + lane) & 1;
                uint32_t m = (uint32_t)mask[j];
                while (m) {
                    const int b = __ffs(m) - 1;
                    m &= m - 1u;
                    const int64_t row = (int64_t)row0[j] + 8 * (b >> 2) + 4 * half + (b & 3);
                    if (row >= a.n) continue;
                    const int pos = atomicAdd(ncand, 1);
                    if (pos < a.cap) cid[pos] = (uint32_t)row;
                }
            }
        }
    }
    wave_sync();
    const int nc = *ncand;
    over = ballot(over) != 0 || nc > a.cap;
    if (over) {
        if (lane == 0) a.flag[q] = 1;
        return;
    }
    const int8_t *qrow = a.q8 + (size_t)q * a.qld;
    for (int t = lane; t < nc; t += 64) {
        const uint32_t row = cid[t];
        const int32_t d = exact_int_dist(a.rows.x + (size_t)row * a.rows.ld, a.rows.flip, qrow,
                                         a.dim, a.metric);
        ck[t] = key_of_exact(d, a.metric);
        cid[t] = a.idmap ? (uint32_t)a.idmap[row] : row;
    }
    wave_sync();
    int p2 = 1;
    while (p2 < nc) p2 <<= 1;
    bitonic_sort(ck, cid, nc, p2);
    for (int t = lane; t < a.k; t += 64) {
        const bool ok = t < nc;
        a.ids[(size_t)q * a.k + t] = ok ? (int64_t)cid[t] : -1;
        a.dists[(size_t)q * a.k + t] = ok ? double_of_exact(ck[t], a.metric) : __builtin_huge_val();
    }
    if (lane == 0) a.flag[q] = 0;
}

// Per-row accumulator start (see the header comment); rows in [n, rows) are padding that can
// never reach a threshold.
__global__ void exact_cinit_kernel(const int32_t *xsq, const int32_t *xsum, int64_t n,
                                   int64_t rows, int metric, int32_t *cinit) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    if (r >= n) cinit[r] = kTilePadInit;
    else cinit[r] = metric == kL2 ? -((xsq[r] + 1) >> 1) : 128 * xsum[r];
}

// Rows of the allowed subset, in id order: counts per 32-row word, an exclusive scan (one
// work-group) and the scatter of the ids.
__global__ __launch_bounds__(1024) void exact_subset_ids_kernel(const uint32_t *bits,
                                                                int64_t words, int64_t n,
                                                                int32_t *idmap, int64_t *count) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (words + 1023) / 1024;
    const int64_t w0 = t * per, w1 = min(words, w0 + per);
    // Bits at or beyond n (in the last word) are ignored.
    auto word = [&](int64_t w) {
        const uint32_t b = bits[w];
        const int64_t valid = n - w * 32;
        return valid >= 32 ? b : (b & ((1u << valid) - 1u));
    };
    int64_t c = 0;
    for (int64_t w = w0; w < w1; ++w) c += __popc(word(w));
    part[t] = c;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t pos = part[t] - c;
    for (int64_t w = w0; w < w1; ++w) {
        uint32_t b = word(w);
        while (b) {
            const int i = __ffs(b) - 1;
            b &= b - 1u;
            idmap[pos++] = (int32_t)(w * 32 + i);
        }
    }
    if (t == 1023) *count = part[1023];
}

// Gathers the subset's rows into shifted rows of dimp bytes, and their per-row terms.
__global__ void exact_gather_rows_kernel(I8Rows src, const int32_t *xsq, const int32_t *xsum,
                                         const int32_t *idmap, int64_t m, int dimp, int8_t *ox8,
                                         int32_t *oxsq, int32_t *oxsum) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int per = dimp / 16;
    if (e >= m * per) return;
    const int64_t r = e / per;
    const int j = (int)(e % per);
    const int64_t from = idmap ? (int64_t)idmap[r] : r;
    uint4 v = uint4{0, 0, 0, 0};
    if (j * 16 < src.ld) {
        v = *(const uint4 *)(src.x + from * src.ld + j * 16);
        v = uint4{v.x ^ src.flip, v.y ^ src.flip, v.z ^ src.flip, v.w ^ src.flip};
    }
    *(uint4 *)(ox8 + r * dimp + j * 16) = v;
    if (j == 0) {
        oxsq[r] = xsq[from];
        oxsum[r] = xsum[from];
    }
}

// Per-row terms over the first dim values of 8-bit rows: one wave per row, four rows per block.
__global__ __launch_bounds__(256) void exact_row_terms_kernel(I8Rows src, int64_t n, int dim,
                                                              int32_t *xsq, int32_t *xsum) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const int lane = lane_id();
    int32_t s2 = 0, s1 = 0;
    for (int i = lane; i < dim; i += 64) {
        const uint8_t b = src.x[r * src.ld + i] ^ (uint8_t)(src.flip & 0xff);
        const int v = (int)(int8_t)b;
        s2 += v * v;
        s1 += v;
    }
    s2 = wave_sum(s2);
    s1 = wave_sum(s1);
    if (lane == 0) {
        xsq[r] = s2;
        xsum[r] = s1;
    }
}

// ---------------------------------------------------------------------------- host launch

template <int MODE, int NCH, int QB, int T>
static hipError_t launch_tile(const TileArgs &a, dim3 grid, hipStream_t s) {
    hipLaunchKernelGGL((exact_tile_kernel<MODE, NCH, QB, T>), grid, dim3(kTileWaves * 64), 0, s, a);
    return hipGetLastError();
}

template <int NCH, int QB>
static hipError_t launch_tile_mode(int mode, int T, const TileArgs &a, dim3 grid, hipStream_t s) {
    if (mode == 1) return launch_tile<1, NCH, QB, 1>(a, grid, s);
    return T <= 4 ? launch_tile<0, NCH, QB, 4>(a, grid, s) : launch_tile<0, NCH, QB, 8>(a, grid, s);
}

hipError_t tile_launch(int mode, int nch, int qb, int T, const TileArgs &a, dim3 grid,
                       hipStream_t s) {
    if (qb == 4) {
        switch (nch) {
            case 1: return launch_tile_mode<1, 4>(mode, T, a, grid, s);
            case 2: return launch_tile_mode<2, 4>(mode, T, a, grid, s);
            case 4: return launch_tile_mode<4, 4>(mode, T, a, grid, s);
        }
    } else if (qb == 2) {
        switch (nch) {
            case 1: return launch_tile_mode<1, 2>(mode, T, a, grid, s);
            case 2: return launch_tile_mode<2, 2>(mode, T, a, grid, s);
            case 4: return launch_tile_mode<4, 2>(mode, T, a, grid, s);
            case 8: return launch_tile_mode<8, 2>(mode, T, a, grid, s);
        }
    } else if (qb == 1) {
        switch (nch) {
            case 1: return launch_tile_mode<1, 1>(mode, T, a, grid, s);
            case 2: return launch_tile_mode<2, 1>(mode, T, a, grid, s);
            case 4: return launch_tile_mode<4, 1>(mode, T, a, grid, s);
            case 8: return launch_tile_mode<8, 1>(mode, T, a, grid, s);
        }
    }
    return hipErrorInvalidValue;
}

hipError_t threshold_launch(const int32_t *bnd, int per_query, int k, int32_t *trig, int nq,
                            hipStream_t s) {
    hipLaunchKernelGGL(exact_threshold_kernel, dim3(nq), dim3(64),
                       per_query <= 1024 ? 0 : per_query * 4, s, bnd, per_query, k, trig);
    return hipGetLastError();
}

hipError_t collect_launch(const CollectArgs &a, int nq, hipStream_t s) {
    const int lds = a.cap * 12 + 16;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)exact_collect_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(exact_collect_kernel, dim3(nq), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t cinit_launch(const int32_t *xsq, const int32_t *xsum, int64_t n, int64_t rows,
                        int metric, int32_t *cinit, hipStream_t s) {
    hipLaunchKernelGGL(exact_cinit_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s,
                       xsq, xsum, n, rows, metric, cinit);
    return hipGetLastError();
}

hipError_t subset_ids_launch(const uint32_t *bits, int64_t n, int32_t *idmap, int64_t *count,
                             hipStream_t s) {
    hipLaunchKernelGGL(exact_subset_ids_kernel, dim3(1), dim3(1024), 0, s, bits, (n + 31) / 32, n,
                       idmap, count);
    return hipGetLastError();
}

hipError_t gather_rows_launch(I8Rows src, const int32_t *xsq, const int32_t *xsum,
                              const int32_t *idmap, int64_t m, int dimp, int8_t *ox8,
                              int32_t *oxsq, int32_t *oxsum, hipStream_t s) {
    const int64_t total = m * (dimp / 16);
    hipLaunchKernelGGL(exact_gather_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256),
                       0, s, src, xsq, xsum, idmap, m, dimp, ox8, oxsq, oxsum);
    return hipGetLastError();
}

hipError_t row_terms_launch(I8Rows src, int64_t n, int dim, int32_t *xsq, int32_t *xsum,
                            hipStream_t s) {
    hipLaunchKernelGGL(exact_row_terms_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, src,
                       n, dim, xsq, xsum);
    return hipGetLastError();
}

}  // namespace hm
