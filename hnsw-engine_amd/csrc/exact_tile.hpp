// Shared definitions of the int8 matrix-core range search (exact_tile.hip) and its host driver
// (exact.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hm {

constexpr int kTileRows = 256;     // chunks are multiples of this many rows (the largest superblock)
constexpr int kTileWaves = 8;      // waves per work-group: two per SIMD
constexpr int32_t kTilePadInit = -(1 << 30);
// A tile plan over m rows has at most min(kTileMaxChunks, max(1, m / 4096)) chunks, each rounded
// up to a multiple of kTileRows: the range search reads rows (and accumulator starts) up to
// m + tile_pad_rows(m).
constexpr int kTileMaxChunks = 1024;
constexpr int64_t kTilePadRows = (int64_t)kTileMaxChunks * kTileRows;
inline int64_t tile_pad_rows(int64_t m) {
    const int64_t chunks = m / 4096 + 1;
    return (chunks < kTileMaxChunks ? chunks : kTileMaxChunks) * kTileRows;
}

// Rows of 8-bit values as the matrix cores read them: byte ^ flip is the shifted value v - 128
// as int8 (flip 0x80 per byte over raw 0..255 bytes, 0 over bytes already shifted). Bytes at
// and beyond the row width (and rows at and beyond n) read as zero.
struct I8Rows {
    const uint8_t *x;
    int64_t ld;      // row stride in bytes, a multiple of 16
    uint32_t flip;   // 0x80808080 or 0
};

struct TileArgs {
    I8Rows rows;           // ld == 32 * NCH; readable to the end of the last chunk
    const int32_t *cinit;  // accumulator start per row (kTilePadInit beyond the valid rows)
    const int8_t *q8;      // (nq, 32 * NCH) shifted queries, zero padded
    int nq;
    int64_t chunk;  // rows per chunk, a multiple of kTileRows
    int nchunks;
    int sample_every;  // bound pass: superblocks 0, s, 2s, ... of each chunk
    int32_t *bnd;      // bound pass out: (nq, nchunks, 2, T)
    const int32_t *trig;  // range pass: (nq) V
    int E;                // records per (query, chunk, half)
    int32_t *rec_mask;    // (nq, nchunks, 2, E) which of the lane's 16 rows reach V (bit 4g+t:
                          // row 8g + 4 (lane / 32) + t of the block)
    int32_t *rec_row;     // (nq, nchunks, 2, E) first row of the 32-row block
    int32_t *rec_cnt;     // (nq, nchunks, 2) records written (> E: overflow)
};

struct CollectArgs {
    const int32_t *rec_mask, *rec_row, *rec_cnt;
    I8Rows rows;           // the rows scanned (exact distances are recomputed from them)
    const int8_t *q8;      // shifted queries of this launch, row stride qld bytes
    int qld;
    const int32_t *idmap;  // row -> id, or null (identity)
    int64_t n;             // valid rows
    int nchunks, E, k, metric, dim, cap;
    int64_t *ids;
    double *dists;
    int32_t *flag;  // 1: re-run this query on the general kernel
};

// Host launchers (exact_tile.hip); each returns the launch status. The tile grid is
// (nchunks, query groups of 8 * qb * 32).
hipError_t tile_launch(int mode, int nch, int qb, int T, const TileArgs &a, dim3 grid,
                       hipStream_t s);
hipError_t threshold_launch(const int32_t *bnd, int per_query, int k, int32_t *trig, int nq,
                            hipStream_t s);
hipError_t collect_launch(const CollectArgs &a, int nq, hipStream_t s);
// Accumulator starts of rows [0, n) and kTilePadInit for [n, rows).
hipError_t cinit_launch(const int32_t *xsq, const int32_t *xsum, int64_t n, int64_t rows,
                        int metric, int32_t *cinit, hipStream_t s);
hipError_t subset_ids_launch(const uint32_t *bits, int64_t n, int32_t *idmap, int64_t *count,
                             hipStream_t s);
// Gathers the subset's rows (idmap null: rows 0..m-1) into shifted int8 rows of dimp bytes, and
// their per-row terms.
hipError_t gather_rows_launch(I8Rows src, const int32_t *xsq, const int32_t *xsum,
                              const int32_t *idmap, int64_t m, int dimp, int8_t *ox8,
                              int32_t *oxsq, int32_t *oxsum, hipStream_t s);
// Per-row terms of 8-bit rows over their first dim values: sum (v - 128)^2 and sum (v - 128).
hipError_t row_terms_launch(I8Rows src, int64_t n, int dim, int32_t *xsq, int32_t *xsum,
                            hipStream_t s);

// ---------------------------------------------------------------- float data (exact_ftile.hip)

// Range search over float32 rows with float32 matrix cores (v_mfma_f32_32x32x2_f32). A row's
// score is s = c_r + x.q (x, q scaled to unit norm for cosine): c_r = -|x|^2 / 2 for l2, 0
// otherwise, so the score rises as the distance falls. Scores are approximate; the threshold
// carries a margin that covers their rounding error (exact_fmargin_kernel).
struct FTileArgs {
    const float *x;        // rows, row stride ld floats (ld % 4 == 0), dim % 4 == 0 values each
    int64_t ld;
    int dim;
    const float *rscale;   // per row factor applied while staging (cosine), or null
    const float *cinit;    // (n) accumulator start per row
    int64_t n;             // valid rows; rows beyond score -inf
    const float *q;        // (nq, 8 * KH) prepared queries, zero padded
    int nq;
    int64_t chunk;
    int nchunks;
    int sample_every;
    int32_t *bnd;          // bound pass out: (nq, nchunks, 2, T) order keys of block scores
    const float *trig;     // range pass: (nq) threshold V
    int E;
    int32_t *rec_mask;     // (nq, nchunks, 2, E) which of the lane's 16 rows reach V (bit 4g+t:
                           // row 8g + 4 (lane / 32) + t of the block)
    int32_t *rec_row;      // (nq, nchunks, 2, E) first row of the 32-row block
    int32_t *rec_cnt;      // (nq, nchunks, 2) records written (> E: overflow)
};

struct FCollectArgs {
    const int32_t *rec_mask, *rec_row, *rec_cnt;
    const int32_t *idmap;  // compact row -> id, or null
    const float *x;        // original rows (n_all, dim), for the float64 distances
    const float *q;        // original queries of this launch (nq, dim)
    int64_t n;             // rows scanned
    int nchunks, E, k, metric, dim, cap;
    int64_t *ids;
    double *dists;
    int32_t *flag;
};

// Kernel geometry: queries per work-group and the LDS bytes of one launch (padded width kp).
int ftile_qb(int kp);
int ftile_lds_bytes(int kp);
hipError_t ftile_launch(int mode, int kp, int T, const FTileArgs &a, dim3 grid, hipStream_t s);
hipError_t fprep_rows_launch(const float *x, int64_t ld, int dim, int64_t n, int metric,
                             float *cinit, float *rscale, unsigned long long *max_sq,
                             hipStream_t s);
hipError_t fprep_queries_launch(const float *q, int64_t nq, int dim, int kp, int metric,
                                float *qp, double *qnorm, hipStream_t s);
hipError_t fmargin_launch(const int32_t *trig_key, const double *qnorm,
                          const unsigned long long *max_sq, int metric, int kp, int dim, int nq,
                          float *V, hipStream_t s);
hipError_t fcollect_launch(const FCollectArgs &a, int nq, hipStream_t s);
hipError_t fgather_rows_launch(const float *x, int dim, const int32_t *idmap, int64_t m, int kp,
                               float *out, hipStream_t s);

}  // namespace hm
