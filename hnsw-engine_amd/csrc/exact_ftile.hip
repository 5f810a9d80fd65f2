// Exact k-NN on float32 data with float32 matrix cores: the reference's oracle (oracle.py:35-126)
// shortlists with a float32 product and a rounding margin (oracle.py:171-181), then ranks the
// shortlist by float64 distance with ties to the lower id. Here the shortlist is a bounded range
// search with the same four kernels as the 8-bit path (exact_tile.hip):
//
//   bound     per lane, the T largest block scores (best of the 16 rows a lane holds for one
//             32-row block) over a sample of each chunk's superblocks;
//   threshold per query, Vh = the k-th largest sampled block score (exact_threshold_kernel on
//             order keys of the floats), then V = Vh - 2e - eps (exact_fmargin_kernel);
//   range     all rows: a lane whose block score reaches V records which of its 16 rows do;
//   collect   per query: the recorded rows scoring >= V, float64 distances computed exactly as
//             the float64 kernel computes them, sorted by (distance, id), the k smallest.
//
// Scores. s = c_r + x.q, c_r = -|x|^2 / 2 for l2 (0 for ip and cosine; cosine rows and queries
// scaled by the float32 reciprocal of their norms), so d = |q|^2 - 2 s (l2), -s (ip), 1 - s
// (cosine) in exact arithmetic. Let e bound |s_computed - s_exact| and eps bound the float64
// distance's own error. At least k sampled rows have s_computed >= Vh, so the k-th smallest
// float64 distance is at most that of exact score Vh - e plus eps, and every row among the k
// nearest has s_computed >= Vh - 2e - eps (in score units). The bound e: the matrix core sums
// K = padded-width products and c_r in float32, error <= gamma_{K+1} (|c_r| + sum |x_i q_i|)
// <= gamma_{K+1} (M^2 / 2 + |q| M) with M the largest row norm (Cauchy-Schwarz), plus the
// rounding of c_r (and, for cosine, of the scaled values: 2u relative each); e is twice that
// (an allowance for any accumulation order or rounding mode of the hardware sum).
//
// Tiling. A work-group is 8 waves, 2 x 32 queries each (1 x 32 above 128 padded dimensions),
// whose B operands (lane l: query l % 32, dimensions [kp/2 * (l / 32), +kp/2)) stay in
// registers; rows are staged superblock by superblock through double-buffered LDS (row stride
// kp + 4 floats: an odd number of 16-byte slots, so the 32 rows a half-wave reads fall in
// distinct bank groups). Per 32-row block and wave: kp / 2 matrix ops per query block.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "exact_tile.hpp"
#include "keys.hpp"
#include "seqdist.hpp"
#include "wave.hpp"

typedef __attribute__((ext_vector_type(16))) float f32x16;

namespace hm {

// Order key of a float: signed integer compare of keys = float compare (no NaNs here).
HM_DEV int32_t okey(float f) {
    const int32_t b = __float_as_int(f);
    return b >= 0 ? b : (b ^ 0x7fffffff);
}

HM_DEV float of_okey(int32_t k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

HM_DEV double wave_sum_d(double v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

HM_DEV float fmax16(const f32x16 &v) {
    const float a = fmaxf(fmaxf(v[0], v[1]), v[2]);
    const float b = fmaxf(fmaxf(v[3], v[4]), v[5]);
    const float c = fmaxf(fmaxf(v[6], v[7]), v[8]);
    const float d = fmaxf(fmaxf(v[9], v[10]), v[11]);
    const float e = fmaxf(fmaxf(v[12], v[13]), v[14]);
    return fmaxf(fmaxf(fmaxf(a, b), fmaxf(c, d)), fmaxf(e, v[15]));
}

template <int KH>
struct FGeom {
    static constexpr int KP = 8 * KH;                              // padded width (floats)
    static constexpr int QB = KP <= 128 ? 2 : 1;                   // query blocks per wave
    static constexpr int SLOTS = KP / 4;                           // 16-byte slots per row
    static constexpr int STRIDE = KP + 4;                          // LDS row stride (floats)
    static constexpr int SBR = KP <= 64 ? 128 : (KP <= 128 ? 64 : 32);  // rows per superblock
    static constexpr int LDS = 2 * SBR * STRIDE * 4 + 2 * SBR * 4;
};

// MODE 0: bound pass; MODE 1: range pass. KH: 16-byte slots per lane half (kp = 8 KH).
template <int MODE, int KH, int T>
__global__ __launch_bounds__(kTileWaves * 64) void exact_ftile_kernel(FTileArgs a) {
    using G = FGeom<KH>;
    constexpr int KP = G::KP, QB = G::QB, SLOTS = G::SLOTS, STRIDE = G::STRIDE, SBR = G::SBR;
    constexpr int THREADS = kTileWaves * 64;
    constexpr int SUBS = SBR / 32;
    constexpr int ELEMS = SBR * SLOTS;
    constexpr int LPT = (ELEMS + THREADS - 1) / THREADS;
    static_assert(LPT <= 4, "at most four 16-byte loads per thread and superblock");
    static_assert(kTileRows % SBR == 0, "chunks are multiples of every superblock size");
    extern __shared__ __attribute__((aligned(16))) float flds[];
    auto rows_lds = [&](int buf) { return flds + buf * (SBR * STRIDE); };
    auto ci_lds = [&](int buf) { return flds + 2 * SBR * STRIDE + buf * SBR; };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, col = lane & 31;
    const int chunk = blockIdx.x;
    const int64_t r0 = (int64_t)chunk * a.chunk;
    const int nsb_all = (int)(a.chunk / SBR);
    const int nsb = MODE == 0 ? (nsb_all + a.sample_every - 1) / a.sample_every : nsb_all;

    int qi[QB];
    float4 bq[QB][KH];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        qi[qb] = blockIdx.y * (kTileWaves * QB * 32) + (wave * QB + qb) * 32 + col;
#pragma unroll
        for (int c = 0; c < KH; ++c) {
            bq[qb][c] = float4{0.f, 0.f, 0.f, 0.f};
            if (qi[qb] < a.nq)
                bq[qb][c] = *(const float4 *)(a.q + (size_t)qi[qb] * KP + half * (KP / 2) + 4 * c);
        }
    }
    float V[QB];
    int cnt[QB];
    float top[QB][T];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        cnt[qb] = 0;
        V[qb] = INFINITY;
        if (MODE == 1 && qi[qb] < a.nq) V[qb] = a.trig[qi[qb]];
#pragma unroll
        for (int t = 0; t < T; ++t) top[qb][t] = -INFINITY;
    }

    // Prefetch registers of the next superblock (plain variables, see exact_tile.hip).
    float4 pre0 = float4{0.f, 0.f, 0.f, 0.f}, pre1 = pre0, pre2 = pre0, pre3 = pre0;
    float rs0 = 1.f, rs1 = 1.f, rs2 = 1.f, rs3 = 1.f;
    float pci = -INFINITY;
    const bool scaled = a.rscale != nullptr;
    auto fetch = [&](int64_t base, int e, float4 &v, float &rs) {
        const int64_t gr = base + e / SLOTS;
        const int sl = e % SLOTS;
        v = float4{0.f, 0.f, 0.f, 0.f};
        if (gr < a.n && 4 * sl < a.dim) v = *(const float4 *)(a.x + gr * a.ld + 4 * sl);
        if (scaled && gr < a.n) rs = a.rscale[gr];
    };
#define HM_FTILE_LOAD(S)                                                               \
    do {                                                                               \
        const int sb_ = MODE == 0 ? (S) * a.sample_every : (S);                        \
        const int64_t base_ = r0 + (int64_t)sb_ * SBR;                                 \
        if (tid < ELEMS) fetch(base_, tid, pre0, rs0);                                 \
        if (LPT > 1 && tid + THREADS < ELEMS) fetch(base_, tid + THREADS, pre1, rs1);  \
        if (LPT > 2 && tid + 2 * THREADS < ELEMS)                                      \
            fetch(base_, tid + 2 * THREADS, pre2, rs2);                                \
        if (LPT > 3 && tid + 3 * THREADS < ELEMS)                                      \
            fetch(base_, tid + 3 * THREADS, pre3, rs3);                                \
        if (tid < SBR) {                                                               \
            pci = -INFINITY;                                                           \
            if (base_ + tid < a.n) pci = a.cinit[base_ + tid];                         \
        }                                                                              \
    } while (0)
    auto stash_one = [&](int buf, int e, float4 v, float rs) {
        const int r = e / SLOTS, sl = e % SLOTS;
        if (scaled) v = float4{v.x * rs, v.y * rs, v.z * rs, v.w * rs};
        *(float4 *)(rows_lds(buf) + r * STRIDE + 4 * sl) = v;
    };
    auto stash = [&](int buf) {
        if (tid < ELEMS) stash_one(buf, tid, pre0, rs0);
        if (LPT > 1 && tid + THREADS < ELEMS) stash_one(buf, tid + THREADS, pre1, rs1);
        if (LPT > 2 && tid + 2 * THREADS < ELEMS) stash_one(buf, tid + 2 * THREADS, pre2, rs2);
        if (LPT > 3 && tid + 3 * THREADS < ELEMS) stash_one(buf, tid + 3 * THREADS, pre3, rs3);
        if (tid < SBR) ci_lds(buf)[tid] = pci;
    };

    __builtin_amdgcn_s_waitcnt(0);
    if (nsb > 0) HM_FTILE_LOAD(0);
    for (int s = 0; s < nsb; ++s) {
        const int buf = s & 1;
        stash(buf);
        __syncthreads();
        if (s + 1 < nsb) HM_FTILE_LOAD(s + 1);
        const int sb = MODE == 0 ? s * a.sample_every : s;
#pragma unroll
        for (int h2 = 0; h2 < SUBS; ++h2) {
            const float *rowp = rows_lds(buf) + (h2 * 32 + col) * STRIDE + half * (KP / 2);
            f32x16 acc[QB];
            {
                f32x16 ci;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const float4 v = *(const float4 *)(ci_lds(buf) + h2 * 32 + 8 * g + 4 * half);
                    ci[4 * g + 0] = v.x;
                    ci[4 * g + 1] = v.y;
                    ci[4 * g + 2] = v.z;
                    ci[4 * g + 3] = v.w;
                }
#pragma unroll
                for (int qb = 0; qb < QB; ++qb) acc[qb] = ci;
            }
#pragma unroll
            for (int c = 0; c < KH; ++c) {
                const float4 af = *(const float4 *)(rowp + 4 * c);
#pragma unroll
                for (int qb = 0; qb < QB; ++qb) {
                    acc[qb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af.x, bq[qb][c].x, acc[qb], 0, 0, 0);
                    acc[qb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af.y, bq[qb][c].y, acc[qb], 0, 0, 0);
                    acc[qb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af.z, bq[qb][c].z, acc[qb], 0, 0, 0);
                    acc[qb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af.w, bq[qb][c].w, acc[qb], 0, 0, 0);
                }
            }
            if constexpr (MODE == 0) {
#pragma unroll
                for (int qb = 0; qb < QB; ++qb) {
                    float v = fmax16(acc[qb]);
#pragma unroll
                    for (int t = 0; t < T; ++t) {
                        const float hi = fmaxf(top[qb][t], v);
                        v = fminf(top[qb][t], v);
                        top[qb][t] = hi;
                    }
                }
            } else {
                float m[QB];
                bool any = false;
#pragma unroll
                for (int qb = 0; qb < QB; ++qb) {
                    m[qb] = fmax16(acc[qb]);
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
        // The next superblock's stash overwrites the other buffer, which every wave finished
        // reading before the barrier at the top of this iteration.
    }
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        if (qi[qb] >= a.nq) continue;
        const size_t slot = ((size_t)qi[qb] * a.nchunks + chunk) * 2 + half;
        if constexpr (MODE == 0) {
#pragma unroll
            for (int t = 0; t < T; ++t) a.bnd[slot * T + t] = okey(top[qb][t]);
        } else {
            a.rec_cnt[slot] = cnt[qb];
        }
    }
#undef HM_FTILE_LOAD
}

// Per row: the accumulator start (l2: -|x|^2 / 2), the cosine scale 1 / |x| (float32), and the
// largest |x|^2 (float64 bits; nonnegative doubles order as their bit patterns). One wave per
// row, grid-stride, one atomic per wave.
__global__ __launch_bounds__(256) void fprep_rows_kernel(const float *x, int64_t ld, int dim,
                                                         int64_t n, int metric, float *cinit,
                                                         float *rscale,
                                                         unsigned long long *max_sq) {
    const int lane = lane_id();
    const int64_t waves = (int64_t)gridDim.x * 4;
    double best = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < n; r += waves) {
        double s = 0.0;
        for (int i = lane; i < dim; i += 64) {
            const double v = (double)x[r * ld + i];
            s += v * v;
        }
        s = wave_sum_d(s);
        best = fmax(best, s);
        if (lane == 0) {
            cinit[r] = metric == kL2 ? (float)(-0.5 * s) : 0.f;
            if (rscale) rscale[r] = (float)(1.0 / sqrt(s));
        }
    }
    if (lane == 0) atomicMax(max_sq, (unsigned long long)__double_as_longlong(best));
}

// Queries padded to kp floats (cosine: times the float32 reciprocal of the norm) and their
// float64 norms. One wave per query.
__global__ __launch_bounds__(256) void fprep_queries_kernel(const float *q, int64_t nq, int dim,
                                                            int kp, int metric, float *qp,
                                                            double *qnorm) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= nq) return;
    const int lane = lane_id();
    double s = 0.0;
    for (int i = lane; i < dim; i += 64) {
        const double v = (double)q[r * dim + i];
        s += v * v;
    }
    s = wave_sum_d(s);
    const float f = metric == kCosine ? (float)(1.0 / sqrt(s)) : 1.f;
    for (int i = lane; i < kp; i += 64) qp[r * kp + i] = i < dim ? q[r * dim + i] * f : 0.f;
    if (lane == 0) qnorm[r] = sqrt(s);
}

// V = Vh - 2e - eps per query (see the file comment), rounded down to float32.
__global__ void exact_fmargin_kernel(const int32_t *trig_key, const double *qnorm,
                                     const unsigned long long *max_sq, int metric, int kp, int dim,
                                     int nq, float *V) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double vh = (double)of_okey(trig_key[q]);
    const double u = 0x1p-24;
    const double M = sqrt(__longlong_as_double((long long)*max_sq)) * (1.0 + 0x1p-40);
    const double Q = qnorm[q] * (1.0 + 0x1p-40);
    double e, eps;
    if (metric == kL2) {
        e = 2.0 * u * (kp + 4) * (0.5 * M * M + Q * M);
        eps = 0x1p-48 * (dim + 4) * (Q + M) * (Q + M);
    } else if (metric == kIP) {
        e = 2.0 * u * (kp + 4) * Q * M;
        eps = 0x1p-48 * (dim + 4) * Q * M;
    } else {
        e = 2.0 * u * (kp + 10);
        eps = 0x1p-46 * (dim + 8);
    }
    e += 0x1p-100;
    V[q] = __double2float_rd(vh - 2.0 * e - eps);
}

// The float64 distance of row `row` to query q, exactly as exact_f64_kernel computes it: left
// to right, unfused; cosine divides by the float64 norms first (normalise_kernel's steps).
HM_DEV double fexact_dist(const float *x, const float *q, int dim, int metric) {
#pragma clang fp contract(off)
    if (metric != kCosine) return seq_row<false, float>(x, q, dim, metric);
    double sx = 0.0, sq = 0.0;
    for (int i = 0; i < dim; ++i) {
        const double v = (double)x[i];
        sx = sx + v * v;
        const double w = qval(q, i);
        sq = sq + w * w;
    }
    const double nx = sqrt(sx), nq = sqrt(sq);
    double acc = 0.0;
    for (int i = 0; i < dim; ++i) acc = acc + ((double)x[i] / nx) * (qval(q, i) / nq);
    return 1.0 - acc;
}

// Per query: every recorded row scoring >= V, its float64 distance, sorted by (distance, id),
// the k smallest written. Flags the query (and writes nothing else) when records overflowed or
// the candidates exceed the buffer.
__global__ __launch_bounds__(64) void exact_fcollect_kernel(FCollectArgs a) {
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
    const float *qrow = a.q + (size_t)q * a.dim;
    for (int t = lane; t < nc; t += 64) {
        const uint32_t row = cid[t];
        const uint32_t id = a.idmap ? (uint32_t)a.idmap[row] : row;
        const double d = fexact_dist(a.x + (size_t)id * a.dim, qrow, a.dim, a.metric);
        ck[t] = key_of_double(d);
        cid[t] = id;
    }
    wave_sync();
    int p2 = 1;
    while (p2 < nc) p2 <<= 1;
    bitonic_sort(ck, cid, nc, p2);
    for (int t = lane; t < a.k; t += 64) {
        const bool ok = t < nc;
        a.ids[(size_t)q * a.k + t] = ok ? (int64_t)cid[t] : -1;
        a.dists[(size_t)q * a.k + t] = ok ? double_of_key(ck[t], a.metric) : __builtin_huge_val();
    }
    if (lane == 0) a.flag[q] = 0;
}

// Rows of a subset (or all rows when idmap is null), padded to kp floats with zeros.
__global__ void fgather_rows_kernel(const float *x, int dim, const int32_t *idmap, int64_t m,
                                    int kp, float *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * kp) return;
    const int64_t r = e / kp;
    const int c = (int)(e % kp);
    const int64_t src = idmap ? (int64_t)idmap[r] : r;
    out[e] = c < dim ? x[src * dim + c] : 0.f;
}

// ---------------------------------------------------------------------------- host launch

int ftile_qb(int kp) { return kp <= 128 ? 2 : 1; }

template <int KH>
static hipError_t launch_ftile_kh(int mode, int T, const FTileArgs &a, dim3 grid, hipStream_t s) {
    const int lds = FGeom<KH>::LDS;
    const void *fn = mode == 1   ? (const void *)exact_ftile_kernel<1, KH, 1>
                     : T <= 4    ? (const void *)exact_ftile_kernel<0, KH, 4>
                                 : (const void *)exact_ftile_kernel<0, KH, 8>;
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    if (mode == 1)
        hipLaunchKernelGGL((exact_ftile_kernel<1, KH, 1>), grid, dim3(kTileWaves * 64), lds, s, a);
    else if (T <= 4)
        hipLaunchKernelGGL((exact_ftile_kernel<0, KH, 4>), grid, dim3(kTileWaves * 64), lds, s, a);
    else
        hipLaunchKernelGGL((exact_ftile_kernel<0, KH, 8>), grid, dim3(kTileWaves * 64), lds, s, a);
    return hipGetLastError();
}

int ftile_lds_bytes(int kp) {
    const int sbr = kp <= 64 ? 128 : (kp <= 128 ? 64 : 32);
    return 2 * sbr * (kp + 4) * 4 + 2 * sbr * 4;
}

hipError_t ftile_launch(int mode, int kp, int T, const FTileArgs &a, dim3 grid, hipStream_t s) {
    switch (kp / 8) {
#define HM_KH(K) \
    case K: return launch_ftile_kh<K>(mode, T, a, grid, s);
        HM_KH(1) HM_KH(2) HM_KH(3) HM_KH(4) HM_KH(5) HM_KH(6) HM_KH(7) HM_KH(8)
        HM_KH(9) HM_KH(10) HM_KH(11) HM_KH(12) HM_KH(13) HM_KH(14) HM_KH(15) HM_KH(16)
        HM_KH(20) HM_KH(24) HM_KH(28) HM_KH(32)
#undef HM_KH
    }
    return hipErrorInvalidValue;
}

hipError_t fprep_rows_launch(const float *x, int64_t ld, int dim, int64_t n, int metric,
                             float *cinit, float *rscale, unsigned long long *max_sq,
                             hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<int64_t>(2048, (n + 3) / 4);
    hipLaunchKernelGGL(fprep_rows_kernel, dim3(blocks), dim3(256), 0, s, x, ld, dim, n, metric,
                       cinit, rscale, max_sq);
    return hipGetLastError();
}

hipError_t fprep_queries_launch(const float *q, int64_t nq, int dim, int kp, int metric,
                                float *qp, double *qnorm, hipStream_t s) {
    hipLaunchKernelGGL(fprep_queries_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, q,
                       nq, dim, kp, metric, qp, qnorm);
    return hipGetLastError();
}

hipError_t fmargin_launch(const int32_t *trig_key, const double *qnorm,
                          const unsigned long long *max_sq, int metric, int kp, int dim, int nq,
                          float *V, hipStream_t s) {
    hipLaunchKernelGGL(exact_fmargin_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s,
                       trig_key, qnorm, max_sq, metric, kp, dim, nq, V);
    return hipGetLastError();
}

hipError_t fcollect_launch(const FCollectArgs &a, int nq, hipStream_t s) {
    const int lds = a.cap * 12 + 16;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)exact_fcollect_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(exact_fcollect_kernel, dim3(nq), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t fgather_rows_launch(const float *x, int dim, const int32_t *idmap, int64_t m, int kp,
                               float *out, hipStream_t s) {
    const int64_t total = m * kp;
    hipLaunchKernelGGL(fgather_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       s, x, dim, idmap, m, kp, out);
    return hipGetLastError();
}

}  // namespace hm
