// C ABI: index upload (the reference's attach_native) and the batch search entry points.
#include <emmintrin.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hnsw_mi355x.h"
#include "common.hpp"
#include "exact_tile.hpp"
#include "maskbits.hpp"
#include "index.hpp"
#include "search.hpp"

namespace hm {

int search_slots(int mode, bool filt, const SearchArgs &a, int lds, int device_cus);
hipError_t rerank_device(hm_index *h, const float *d_q, const int64_t *d_cand, int64_t nq,
                         int fetch, int k, int64_t *d_ids, double *d_d, hipStream_t s);
hipError_t search_launch(int mode, bool filt, const SearchArgs &a, int blocks, hipStream_t s);

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

int fail_hip(hipError_t e, const char *what) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? HM_ENOMEM : HM_EDEVICE;
}

int invalid(const std::string &msg) {
    g_error = msg;
    return HM_EINVAL;
}

}  // namespace hm

using namespace hm;

namespace {

// One batch already in HBM.
struct Batch {
    const float *q32 = nullptr;
    const double *q64 = nullptr;
    const uint8_t *q8 = nullptr;  // exact mode: query bytes in place of q32
    const int32_t *k = nullptr, *ef = nullptr, *midx = nullptr;
    const uint32_t *mask_bits = nullptr;
    int64_t mask_words = 0;
    int k_all = 0, ef_all = 0, mask_all = -1, kstride = 0;
    int max_ef = 0;
    int64_t nq = 0;
    int64_t *ids = nullptr;
    double *dists = nullptr;
    int32_t *dc = nullptr, *hops = nullptr, *fail = nullptr;
    bool learned = false;  // use the handle's stop model (register-beam path only)
    bool u8rows = false;   // kSeqU8 over the index's exact 8-bit rows (not codes)
    const int32_t *subset = nullptr;  // device: queue position -> query index (a kernel class)
    int64_t q_rows = 0;               // with a subset: the query rows it indexes
    const uint8_t *boost = nullptr;   // device, per position: order boost (filtered classes)
    // traces (register-beam path only)
    float *trace = nullptr;
    int trace_steps = 0;
    const double *trace_thr = nullptr;
    int32_t *trace_len = nullptr, *trace_final = nullptr;
};

constexpr int64_t kDefaultSpill = 16384;

// Scratch of hm_search_dev_noalloc, for callers that cannot allocate: 64 MiB per GPU, held from
// the library's load until the process ends. Host code never passes this name to a kernel (on
// the host it is only a stand-in for the symbol): hm_open looks up each device's address of it
// once, and the entry point carves its scratch from that address.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
__device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up once per device.
hipError_t static_scratch_of(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
// Ids per wave slot remembered for clearing its visited bitset (a layer search that sets more
// clears the whole bitset).
constexpr int kVisitedList = 4096;

// Layer-0 kernels a launch can take (decided per launch; none changes results).
enum Kind { kGeneral = 0, kFast = 1, kBeam = 2 };

// Capacities of the on-chip state for a launch whose largest max(ef, k) is max_ef; nr: R
// entries per lane of the register beams.
void plan(SearchArgs &a, int mode, bool filt, const hm_index *h, int max_ef, int kind, int nr) {
    max_ef = std::max(1, max_ef);
    a.fast = kind == kFast ? 1 : 0;
    a.beam = kind == kBeam ? 1 : 0;
    a.nr = nr;
    if (kind != kGeneral) {
        // 4096 slots (8 KB of 16-bit slots for up to 4M ids). Measured on 10,000 queries over
        // 1M x 128 with the longest-first schedule, against 2048 slots (which fit more waves per
        // CU but send 2-10 % of the layer searches to the global bitset): ef 16 / 32 / 48 / 64 /
        // 96 / 128 -9 / -9 / -8 / -17 / -3 / -5 %, and -4 % on 2,500 queries in one launch;
        // 8192 slots are 30-40 % slower (too few waves). The four- and eight-slot beams
        // (ef 129..512, 1,250-1,400 visits) need 4096 slots all the more.
        a.h_log2 = h->tune_fast_hlog ? h->tune_fast_hlog : (max_ef <= 8 ? 10 : 12);
        a.r_cap = 64 * nr;
        a.c_cap = 64;
        if (kind == kFast) {
            a.lds_per_wave = lds_bytes_fast(mode, h->dim, h->dimp, a.h_log2, a.id_bits, nr);
        } else {
            a.nc = filt ? 4 : 1;
            a.lds_per_wave = lds_bytes_beam(mode, h->dim, h->dimp, a.h_log2, a.id_bits, nr, a.nc);
        }
        return;
    }
    a.r_cap = (int)round_up(max_ef, 8);
    a.c_cap = 1 << std::min(12, std::max(6, ceil_log2(2 * (int64_t)max_ef)));
    a.h_log2 = h->tune_hlog ? h->tune_hlog
                            : std::min(12, std::max(10, ceil_log2((int64_t)max_ef * 64)));
    a.lds_per_wave =
        lds_bytes_per_wave(mode, h->dim, h->dimp, a.r_cap, a.c_cap, a.h_log2, a.id_bits);
}

// R entries per lane a register beam needs for max(ef, k) up to max_ef (0: more than 512).
int beam_width(int max_ef) {
    return max_ef <= 64 ? 1 : max_ef <= 128 ? 2 : max_ef <= 256 ? 4 : max_ef <= 512 ? 8 : 0;
}

// Where a launch's scratch comes from: the handle's grow-only buffers (allocated on demand), or
// a fixed region carved in order (no allocation; the launch shrinks its wave slots to fit).
struct Carver {
    uint8_t *base = nullptr;
    size_t cap = 0, used = 0;
    template <class T> T *take(size_t count) {
        const size_t at = (used + 255) & ~size_t(255);
        if (at + count * sizeof(T) > cap) return nullptr;
        used = at + count * sizeof(T);
        return (T *)(base + at);
    }
};

// Zeroes `words` uint32 words (a kernel, not a memset node, so that captured work stays
// correct on replay).
__global__ void clear_words_kernel(uint32_t *p, int64_t words) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// The two-phase schedule's descent kernels: 16 upper links, up to 128 dimensions (one-byte
// rows of exactly 128 bytes for the integer-key mode).
// Widens n float32 values to float64 (exact).
__global__ void widen_kernel(const float *x, int64_t n, double *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)x[i];
}

bool two_phase_shape(const hm_index *h, int mode) {
    if (h->w1 != 16 || h->dim > 128) return false;
    return mode == kExactU8 ? h->dimp == 128 : (mode == kSeqU8 || mode == kSeqF32);
}

// Launches one batch. fixed == null: scratch from the handle's buffers; otherwise every scratch
// array (and the counters / fail flags the batch leaves null) is carved from *fixed, the
// bitsets and event counters cleared by a kernel first, and the wave slots reduced to what fits.
// tier: 2 any kernel; 1 no kFast (a rerun after the register beam's tie list overflowed);
// 0 the general kernel only.
int launch(hm_index *h, int mode, bool filt, const Batch &b0, int64_t spill_cap, int max_slots,
           hipStream_t s, int tier = 2, Carver *fixed = nullptr) {
    Batch b = b0;
    // Float32 queries on an index of 8-bit integers: the float64 sum reads the rows' exact
    // one-byte copy (128 B per 128-dimension row instead of 512 B) with the queries widened to
    // float64. Every term has the same value as over the float32 rows, so every sum, key and
    // result is the same bit for bit. Not for filtered batches (their kernels exist for the
    // float32 rows only), traces, the stop model or the allocation-free scratch. A class subset
    // widens every query row it indexes.
    const int64_t q_rows = b.subset ? b.q_rows : b.nq;
    if (mode == kSeqF32 && h->u8.p && b.q32 && !filt && !fixed && !b.trace && !b.learned &&
        q_rows > 0) {
        const int64_t vals = q_rows * (int64_t)h->dim;
        HM_TRY(h->qwide.ensure((size_t)vals), "query allocation");
        const int grid = (int)std::min<int64_t>(2048, (vals + 255) / 256);
        hipLaunchKernelGGL(widen_kernel, dim3(grid), dim3(256), 0, s, b.q32, vals, h->qwide.p);
        b.q64 = h->qwide.p;
        b.u8rows = true;
        mode = kSeqU8;
    }
    SearchArgs a;
    memset(&a, 0, sizeof(a));
    a.links0 = h->links0.p;
    a.links_up = h->links_up.p;
    a.upper_off = h->upper.p;
    a.n = h->n;
    a.id_bits = std::max(1, ceil_log2(h->n));
    a.probe_limit = h->tune_probe_limit;
    a.w0 = h->w0;
    a.w1 = h->w1;
    a.entry = (int)h->entry;
    a.max_level = h->max_level;
    a.vec_f32 = h->vec.p;
    a.vec_u8 = mode == kExactU8 || b.u8rows ? h->u8.p : h->codes.p;
    a.dim = h->dim;
    a.dimp = h->dimp;
    a.metric = h->metric;
    a.q_f32 = b.q32;
    a.q_u8 = b.q8;
    a.q_f64 = b.q64;
    a.nq = (int)b.nq;
    a.ef_q = b.ef;
    a.k_q = b.k;
    a.ef_all = b.ef_all;
    a.k_all = b.k_all;
    a.kstride = b.kstride;
    a.mask_pool = b.mask_bits;
    a.mask_words = b.mask_words;
    a.mask_idx = b.midx;
    a.mask_all = b.mask_all;
    // Which layer-0 kernel: unfiltered max(ef, k) <= 64 the register beam (the packed form on
    // the constant shape); unfiltered 65..512 on the constant shape the packed register beam with
    // 2, 4 or 8 entries per lane; every other batch up to 512 the general register beam
    // (filters, codes, other shapes); beyond 512, or a rerun after a spill overflow, the general
    // kernel.
    const int nr_ef = beam_width(b.max_ef);
    const bool allow_fast = tier >= 2;
    const bool fast = allow_fast && !filt && b.max_ef <= 64 && (mode != kSeqU8 || b.u8rows);
    const int wide_hlog = h->tune_fast_hlog ? h->tune_fast_hlog : 12;
    const bool wide = allow_fast && !filt && !fast && !b.trace && mode == kExactU8 && nr_ef > 0 &&
                      h->w0 == 32 && h->w1 == 16 && h->dimp == 128 &&
                      a.id_bits - wide_hlog <= 10;
    const bool fast_any = allow_fast && !filt && b.max_ef <= 64;
    const int kind = fast || wide || fast_any ? kFast
                     : tier >= 1 && !b.trace && nr_ef > 0 ? kBeam
                                                          : kGeneral;
    plan(a, mode, filt, h, b.max_ef, kind, kind == kFast ? (wide ? nr_ef : 1) : nr_ef);
    if (fast && b.trace) {
        a.trace = b.trace;
        a.trace_steps = b.trace_steps;
        a.trace_thr = b.trace_thr;
        a.trace_len = b.trace_len;
        a.trace_final = b.trace_final;
    } else if (b.trace) {
        return invalid("traces need an unfiltered batch with ef <= 64");
    } else if (fast && b.learned && h->stop_on) {
        a.stop_w = h->stop_w.p;
        a.stop_logit = h->stop_logit;
        a.stop_tmin = h->stop_tmin;
        a.stop_tmax = h->stop_tmax;
    }
    if (a.lds_per_wave > 160 * 1024) return invalid("ef too large for the on-chip search state");
    int slots = search_slots(mode, filt, a, a.lds_per_wave, h->cus);
    if (h->tune_waves_per_cu > 0) slots = std::min(slots, h->tune_waves_per_cu * h->cus);
    slots = (int)std::min<int64_t>(slots, std::max<int64_t>(1, b.nq));
    if (max_slots > 0) slots = std::min(slots, max_slots);
    a.bitset_words = (h->n + 31) / 32;
    a.vlist_cap = kVisitedList;
    // The register beams of kind kFast never spill candidates.
    if (a.fast) spill_cap = 0;
    a.subset = b.subset;
    a.order_boost = b.boost;
    a.spill_cap = spill_cap;
    if (fixed) {
        Carver &c = *fixed;
        int32_t *queue = c.take<int32_t>(1);
        int32_t *stats = c.take<int32_t>(4);
        if (!b.dc) b.dc = c.take<int32_t>(b.nq);
        if (!b.hops) b.hops = c.take<int32_t>(b.nq);
        if (!b.fail) b.fail = c.take<int32_t>(b.nq);
        // The schedule arrays are reserved whenever the batch might take the two-phase schedule.
        uint64_t *pre_key = nullptr;
        uint32_t *pre_id = nullptr;
        int32_t *pre_dc = nullptr, *pre_hops = nullptr, *order = nullptr;
        if ((a.fast || a.beam) && two_phase_shape(h, mode) && !a.trace && h->tune_schedule == 0) {
            pre_key = c.take<uint64_t>(b.nq);
            pre_id = c.take<uint32_t>(b.nq);
            pre_dc = c.take<int32_t>(b.nq);
            pre_hops = c.take<int32_t>(b.nq);
            order = c.take<int32_t>(b.nq);
            if (!order) return invalid("batch too large for the allocation-free scratch");
        }
        if (!queue || !stats || !b.dc || !b.hops || !b.fail)
            return invalid("batch too large for the allocation-free scratch");
        const size_t per_slot = (size_t)a.bitset_words * 4 + (size_t)kVisitedList * 4 +
                                (size_t)spill_cap * 12 + 3 * 256;
        const size_t room = c.cap > c.used + 256 ? c.cap - c.used - 256 : 0;
        const int fit = (int)std::min<size_t>(room / per_slot, (size_t)INT32_MAX);
        if (fit < 1) return invalid("graph too large for the allocation-free scratch");
        slots = std::min(slots, fit);
        a.bitsets = c.take<uint32_t>((size_t)slots * a.bitset_words);
        a.vlist = c.take<uint32_t>((size_t)slots * kVisitedList);
        a.spill_key = c.take<uint64_t>(std::max<size_t>(1, (size_t)slots * spill_cap));
        a.spill_id = c.take<uint32_t>(std::max<size_t>(1, (size_t)slots * spill_cap));
        if (!a.bitsets || !a.vlist || !a.spill_key || !a.spill_id)
            return invalid("graph too large for the allocation-free scratch");
        a.queue = queue;
        a.stats = stats;
        a.pre_key = pre_key;
        a.pre_id = pre_id;
        a.pre_dc = pre_dc;
        a.pre_hops = pre_hops;
        a.order = order;
        hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, s, (uint32_t *)stats,
                           (int64_t)4);
        const int64_t words = (int64_t)slots * a.bitset_words;
        const int grid = (int)std::min<int64_t>(4096, (words + 1023) / 1024);
        hipLaunchKernelGGL(clear_words_kernel, dim3(std::max(1, grid)), dim3(256), 0, s,
                           a.bitsets, words);
    } else {
        HM_TRY(h->stats.ensure(4, true, s), "stats allocation");
        a.stats = h->stats.p;
        HM_TRY(h->bits.ensure((size_t)slots * a.bitset_words, true, s), "bitset allocation");
        a.bitsets = h->bits.p;
        HM_TRY(h->vlist.ensure((size_t)slots * kVisitedList), "visited list allocation");
        a.vlist = h->vlist.p;
        HM_TRY(h->spill_key.ensure((size_t)slots * spill_cap), "spill allocation");
        HM_TRY(h->spill_id.ensure((size_t)slots * spill_cap), "spill allocation");
        a.spill_key = h->spill_key.p;
        a.spill_id = h->spill_id.p;
        HM_TRY(h->queue.ensure(1), "queue allocation");
        a.queue = h->queue.p;
    }
    a.out_ids = b.ids;
    a.out_d = b.dists;
    a.out_dc = b.dc;
    a.out_hops = b.hops;
    a.out_fail = b.fail;
    a.slots = slots;
    // More queries than wave slots on the register beam (16 upper links, up to 128
    // dimensions): every query's upper-layer descent first, 16 lanes per query (descend_kernel;
    // descend_seq_kernel for the float64 distances), then layer 0 with the queries whose
    // layer-0 entry lies farthest from them first, which shortens the launch's tail. Measured on
    // 10,000 queries over 1M x 128 against one launch: integer queries ef 32 / 48 / 64 / 96
    // -12 / -17 / -16 / -14 %; fractional queries ef 16 / 64 / 128 -22 / -22 / -17 %; all
    // outputs identical.
    if ((a.fast || a.beam) && two_phase_shape(h, mode) && !a.trace && b.nq > slots &&
        h->tune_schedule == 0) {
        a.descent_cap = h->tune_descent_cap > 0 ? h->tune_descent_cap : kDescentMaxCap;
        if (!fixed) {
            HM_TRY(h->pre_key.ensure(b.nq), "schedule allocation");
            HM_TRY(h->pre_id.ensure(b.nq), "schedule allocation");
            HM_TRY(h->pre_dc.ensure(b.nq), "schedule allocation");
            HM_TRY(h->pre_hops.ensure(b.nq), "schedule allocation");
            HM_TRY(h->order.ensure(b.nq), "schedule allocation");
            a.pre_key = h->pre_key.p;
            a.pre_id = h->pre_id.p;
            a.pre_dc = h->pre_dc.p;
            a.pre_hops = h->pre_hops.p;
            a.order = h->order.p;
        }
        a.phase = 2;  // search_launch runs phase 1, the order, then phase 2
    } else {
        a.order = nullptr;  // carved schedule arrays stay unused
    }
    HM_TRY(search_launch(mode, filt, a, slots, s), "search launch");
    return HM_OK;
}

// Calls on one handle share its scratch, so each call's stream first waits for the work the
// handle's previous call enqueued (on whatever stream), and records the handle's event after its
// own. Inside a stream capture neither is done: the captured launches are ordered by the stream
// the graph is later launched on, and hm_search_graph_dev orders the launch itself.
hipError_t order_before(hm_index *h, hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
        return hipSuccess;
    return h->last_done ? hipStreamWaitEvent(s, h->last_done, 0) : hipSuccess;
}

hipError_t order_after(hm_index *h, hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
        return hipSuccess;
    return h->last_done ? hipEventRecord(h->last_done, s) : hipSuccess;
}

// Packs boolean rows (n bytes each) into uint32 bitsets.

int check_opts(const hm_index *h, const hm_search_opts *o) {
    if (!o || o->nq < 0) return invalid("bad search options");
    if (o->kstride < 1) return invalid("kstride must be >= 1");
    if (!o->k && (o->k_all < 1 || o->k_all > o->kstride)) return invalid("k must be in [1, kstride]");
    if (!o->ef && o->ef_all < 1) return invalid("ef must be >= 1");
    if (h->n == 0 || h->entry < 0) return invalid("cannot search an empty index");
    return HM_OK;
}

// Host-buffer search shared by hm_search, hm_search_codes and hm_search_codes_rerank. With
// rerank_q (prepared float32 queries, host) the batch's kstride candidates stay in HBM and the
// exact re-rank scores them there; ids / dists then receive its best rerank_k per query.
int search_host(hm_index *h, int mode, const float *q32, const double *q64,
                const hm_search_opts *o, int64_t *ids, double *dists, int64_t *dcount,
                int64_t *hops, const float *rerank_q = nullptr, int rerank_k = 0,
                const uint8_t *q8 = nullptr) {
    int st = check_opts(h, o);
    if (st) return st;
    const int64_t nq = o->nq;
    if (nq == 0) return HM_OK;
    hipStream_t s = h->stream;
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    HM_TRY(order_before(h, s), "stream order");
    Batch b;
    b.nq = nq;
    b.kstride = o->kstride;
    b.k_all = o->k_all;
    b.mask_all = o->mask_all;
    // Per-query k and max(ef, k) (numba_engine.py:473,487).
    std::vector<int32_t> kv, efv;
    int max_ef = 0;
    if (o->k || o->ef) {
        kv.resize(nq);
        efv.resize(nq);
        for (int64_t i = 0; i < nq; ++i) {
            int k = o->k ? o->k[i] : o->k_all;
            int ef = o->ef ? o->ef[i] : o->ef_all;
            if (k < 1 || k > o->kstride || ef < 1) return invalid("k or ef out of range");
            kv[i] = k;
            efv[i] = std::max(ef, k);
            max_ef = std::max(max_ef, efv[i]);
        }
        HM_TRY(h->kq.ensure(nq), "alloc");
        HM_TRY(h->efq.ensure(nq), "alloc");
        HM_TRY(hipMemcpyAsync(h->kq.p, kv.data(), nq * 4, hipMemcpyHostToDevice, s), "copy k");
        HM_TRY(hipMemcpyAsync(h->efq.p, efv.data(), nq * 4, hipMemcpyHostToDevice, s), "copy ef");
        b.k = h->kq.p;
        b.ef = h->efq.p;
    } else {
        b.ef_all = std::max(o->ef_all, o->k_all);
        max_ef = b.ef_all;
    }
    b.max_ef = max_ef;
    b.learned = o->learned != 0;
    bool filt = false;
    std::vector<uint32_t> bits;
    if (o->mask_pool && o->n_masks > 0 && (o->mask_index || o->mask_all >= 0)) {
        if (mode == kSeqU8) return invalid("the codes search takes no filter");
        filt = true;
        bits = pack_masks(o->mask_pool, o->n_masks, h->n);
        b.mask_words = (h->n + 31) / 32;
        HM_TRY(h->maskbits.ensure(bits.size()), "alloc");
        HM_TRY(hipMemcpyAsync(h->maskbits.p, bits.data(), bits.size() * 4, hipMemcpyHostToDevice, s),
               "copy masks");
        b.mask_bits = h->maskbits.p;
        if (o->mask_index) {
            for (int64_t i = 0; i < nq; ++i)
                if (o->mask_index[i] >= o->n_masks) return invalid("mask index out of range");
            HM_TRY(h->midx.ensure(nq), "alloc");
            HM_TRY(hipMemcpyAsync(h->midx.p, o->mask_index, nq * 4, hipMemcpyHostToDevice, s),
                   "copy mask index");
            b.midx = h->midx.p;
        } else if (o->mask_all >= o->n_masks) {
            return invalid("mask index out of range");
        }
    }
    const size_t qn = (size_t)nq * h->dim;
    if (q8) {  // exact mode: one byte per value (a quarter of the float32 upload)
        HM_TRY(h->q8.ensure(qn), "alloc");
        HM_TRY(hipMemcpyAsync(h->q8.p, q8, qn, hipMemcpyHostToDevice, s), "copy queries");
        b.q8 = h->q8.p;
    } else if (q32) {
        HM_TRY(h->q32.ensure(qn), "alloc");
        HM_TRY(hipMemcpyAsync(h->q32.p, q32, qn * 4, hipMemcpyHostToDevice, s), "copy queries");
        b.q32 = h->q32.p;
    } else {
        HM_TRY(h->q64.ensure(qn), "alloc");
        HM_TRY(hipMemcpyAsync(h->q64.p, q64, qn * 8, hipMemcpyHostToDevice, s), "copy queries");
        b.q64 = h->q64.p;
    }
    const size_t on = (size_t)nq * o->kstride;
    HM_TRY(h->ids.ensure(on), "alloc");
    HM_TRY(h->dists.ensure(on), "alloc");
    HM_TRY(h->dc.ensure(nq), "alloc");
    HM_TRY(h->hops.ensure(nq), "alloc");
    HM_TRY(h->fail.ensure(nq), "alloc");
    b.ids = h->ids.p;
    b.dists = h->dists.p;
    b.dc = h->dc.p;
    b.hops = h->hops.p;
    b.fail = h->fail.p;
    // Queries that need different kernels (per-query ef; filtered and unfiltered ones) run as
    // one launch per kernel class, each over its own subset of the batch: class = (filtered,
    // register-beam width 1 / 2 / 4 / 8, or the general kernel beyond ef 512).
    auto class_of = [&](int64_t i) {
        const bool fq = filt && (o->mask_index ? o->mask_index[i] >= 0 : o->mask_all >= 0);
        const int w = beam_width(efv.empty() ? b.ef_all : efv[i]);
        return (fq ? 5 : 0) + (w == 0 ? 4 : w == 1 ? 0 : w == 2 ? 1 : w == 4 ? 2 : 3);
    };
    // Filtered queries are scheduled by their mask's selectivity s: a filtered beam expands
    // about 1/s times the nodes to meet ef allowed ones (measured on the 1M index at ef 512: 10,200
    // expansions at 5 %, 1,040 at 50 %), so the most selective masks go first, which shortens
    // the launch's tail (a list-scheduling model of that class over 2,048 wave slots: 1.5x).
    // boost = 8 log2(1 / s), in the order kernel's log buckets (eight per doubling).
    std::vector<uint8_t> mask_boost;
    if (filt) {
        const int64_t words = (h->n + 31) / 32;
        mask_boost.resize((size_t)o->n_masks);
        for (int64_t m = 0; m < o->n_masks; ++m) {
            int64_t c = 0;
            for (int64_t w = 0; w < words; ++w) c += __builtin_popcount(bits[(size_t)(m * words + w)]);
            const double sel = c > 0 ? (double)c / (double)h->n : 1.0 / (double)h->n;
            mask_boost[(size_t)m] = (uint8_t)std::min(255.0, std::round(8.0 * std::log2(1.0 / sel)));
        }
    }
    auto boost_of = [&](int64_t i) -> uint8_t {
        if (!filt) return 0;
        const int m = o->mask_index ? o->mask_index[i] : o->mask_all;
        return m >= 0 ? mask_boost[(size_t)m] : 0;
    };
    // Runs the queries `which` (all of the batch when null), class by class; spill: candidate
    // spill entries per wave slot, max_slots: a cap on wave slots (0: none).
    // Host arrays the class launches copy from: they live until the call's final
    // synchronisation (a re-run reuses them only after the first run's results were fetched).
    std::vector<int32_t> order;
    std::vector<uint8_t> boost;
    auto run = [&](const std::vector<int32_t> *which, int64_t spill, int max_slots,
                   int tier) -> int {
        const int64_t count = which ? (int64_t)which->size() : nq;
        // Each query's class once, then a counting sort by class (stable: batch order within
        // a class).
        std::vector<uint8_t> cls((size_t)count);
        std::vector<int64_t> start(11, 0);
        for (int64_t t = 0; t < count; ++t) {
            const int c = class_of(which ? (*which)[t] : t);
            cls[(size_t)t] = (uint8_t)c;
            start[c + 1] += 1;
        }
        int present = 0;
        for (int c = 0; c < 10; ++c) {
            present += start[c + 1] > 0 ? 1 : 0;
            start[c + 1] += start[c];
        }
        // One mask for the whole batch: every query has the same selectivity, nothing to order.
        const bool boosted = filt && o->mask_index;
        if (!which && present == 1 && !boosted)
            return launch(h, mode, filt, b, spill, max_slots, s, tier);
        order.assign((size_t)count, 0);
        {
            std::vector<int64_t> next(start.begin(), start.end() - 1);
            for (int64_t t = 0; t < count; ++t)
                order[(size_t)next[cls[(size_t)t]]++] = (int32_t)(which ? (*which)[t] : t);
        }
        for (int c = 5; c < 10; ++c)  // filtered: most selective masks first (queue order)
            std::stable_sort(order.begin() + start[c], order.begin() + start[c + 1],
                             [&](int32_t x, int32_t y) { return boost_of(x) > boost_of(y); });
        HM_TRY(h->subset.ensure(count), "alloc");
        HM_TRY(hipMemcpyAsync(h->subset.p, order.data(), count * 4, hipMemcpyHostToDevice, s),
               "copy subsets");
        if (boosted) {
            boost.assign((size_t)count, 0);
            for (int64_t t = 0; t < count; ++t) boost[(size_t)t] = boost_of(order[(size_t)t]);
            HM_TRY(h->boost.ensure(count), "alloc");
            HM_TRY(hipMemcpyAsync(h->boost.p, boost.data(), count, hipMemcpyHostToDevice, s),
                   "copy order boosts");
        }
        for (int c = 0; c < 10; ++c) {
            const int64_t cnt = start[c + 1] - start[c];
            if (cnt == 0) continue;
            Batch bc = b;
            bc.nq = cnt;
            bc.subset = h->subset.p + start[c];
            bc.q_rows = count;
            bc.boost = boosted && c >= 5 ? h->boost.p + start[c] : nullptr;
            int mx = 0;
            for (int64_t t = start[c]; t < start[c + 1]; ++t)
                mx = std::max(mx, efv.empty() ? b.ef_all : efv[order[t]]);
            bc.max_ef = mx;
            const int st2 = launch(h, mode, filt && c >= 5, bc, spill, max_slots, s, tier);
            if (st2) {
                (void)hipStreamSynchronize(s);  // the copies from `order` and `boost` are done
                return st2;
            }
        }
        return HM_OK;
    };
    st = run(nullptr, std::min<int64_t>(h->n, kDefaultSpill), 0, 2);
    if (st) return st;
    {
        const size_t out = (size_t)nq * (rerank_q ? rerank_k : o->kstride) * 8;
        prefault_output(ids, out);
        prefault_output(dists, out);
        if (dcount) prefault_output(dcount, (size_t)nq * 8);
        if (hops) prefault_output(hops, (size_t)nq * 8);
    }
    std::vector<int32_t> failv(nq), dcv(nq), hv(nq);
    // The failure flags and everything the caller gets back are queued behind the search and
    // fetched with one synchronisation; a batch with failed queries (rare) runs them again and
    // fetches once more.
    auto fetch = [&]() -> int {
        HM_TRY(hipMemcpyAsync(failv.data(), h->fail.p, nq * 4, hipMemcpyDeviceToHost, s), "copy");
        if (rerank_q) {
            const size_t rn = (size_t)nq * rerank_k;
            HM_TRY(h->q32.ensure((size_t)nq * h->dim), "alloc");
            HM_TRY(h->rids.ensure(rn), "alloc");
            HM_TRY(h->rdists.ensure(rn), "alloc");
            HM_TRY(hipMemcpyAsync(h->q32.p, rerank_q, (size_t)nq * h->dim * 4,
                                  hipMemcpyHostToDevice, s),
                   "copy queries");
            HM_TRY(rerank_device(h, h->q32.p, h->ids.p, nq, o->kstride, rerank_k, h->rids.p,
                                 h->rdists.p, s),
                   "rerank launch");
            HM_TRY(hipMemcpyAsync(ids, h->rids.p, rn * 8, hipMemcpyDeviceToHost, s), "copy ids");
            HM_TRY(hipMemcpyAsync(dists, h->rdists.p, rn * 8, hipMemcpyDeviceToHost, s), "copy");
        } else {  // large results through pinned staging and four host threads
            HM_TRY(h->staging.copy_out(ids, h->ids.p, on * 8, s), "copy ids");
            HM_TRY(h->staging.copy_out(dists, h->dists.p, on * 8, s), "copy dists");
        }
        HM_TRY(hipMemcpyAsync(dcv.data(), h->dc.p, nq * 4, hipMemcpyDeviceToHost, s), "copy");
        HM_TRY(hipMemcpyAsync(hv.data(), h->hops.p, nq * 4, hipMemcpyDeviceToHost, s), "copy");
        HM_TRY(hipStreamSynchronize(s), "search");
        return HM_OK;
    };
    st = fetch();
    if (st) return st;
    // Queries that failed (candidates beyond the spill, or evicted ties beyond the tie list of
    // the kFast register beams) run again, as one batch per class, on the general register beam
    // or the general kernel with room for every node (no query can hold more candidates than
    // there are nodes), on as many wave slots as a spill budget of 4 GiB allows.
    std::vector<int32_t> redo;
    for (int64_t i = 0; i < nq; ++i)
        if (failv[i]) redo.push_back((int32_t)i);
    if (!redo.empty()) {
        h->reruns += (int64_t)redo.size();
        const int64_t budget = int64_t(4) << 30;
        const int slots = (int)std::max<int64_t>(1, std::min<int64_t>(1 << 20, budget / (h->n * 12)));
        st = run(&redo, h->n, slots, 1);
        if (st) return st;
        st = fetch();
        if (st) return st;
        for (int64_t i = 0; i < nq; ++i)
            if (failv[i]) return invalid("candidate spill overflow");
    }
    HM_TRY(order_after(h, s), "stream order");
    for (int64_t i = 0; i < nq; ++i) {
        if (dcount) dcount[i] = dcv[i];
        if (hops) hops[i] = hv[i];
    }
    return HM_OK;
}

// Converts float32 values to bytes when every one is an integer in 0..255 (NaN and -0.0 as
// all_u8_integers treats them), four at a time by 128-bit compares and packs; false otherwise.
bool pack_u8_integers(const float *v, size_t count, uint8_t *out) {
    const __m128 lo = _mm_setzero_ps(), hi = _mm_set1_ps(255.0f);
    __m128 bad = _mm_setzero_ps();
    size_t i = 0;
    for (; i + 16 <= count; i += 16) {
        __m128i w[4];
        for (int t = 0; t < 4; ++t) {
            const __m128 x = _mm_loadu_ps(v + i + 4 * t);
            const __m128 c = _mm_min_ps(_mm_max_ps(x, lo), hi);  // NaN -> 0
            w[t] = _mm_cvttps_epi32(c);
            bad = _mm_or_ps(bad, _mm_cmpneq_ps(_mm_cvtepi32_ps(w[t]), x));
        }
        const __m128i p = _mm_packus_epi16(_mm_packs_epi32(w[0], w[1]), _mm_packs_epi32(w[2], w[3]));
        _mm_storeu_si128((__m128i *)(out + i), p);
    }
    if (_mm_movemask_ps(bad)) return false;
    for (; i < count; ++i) {
        const float x = v[i];
        const float c = std::min(std::max(x, 0.0f), 255.0f);
        if (c != x || (float)(int)c != c) return false;
        out[i] = (uint8_t)(int)c;
    }
    return true;
}

// Branch-free over blocks so that the compiler vectorises it (NaN clamps to 0 and differs).
bool all_u8_integers(const float *v, size_t count) {
    constexpr size_t kBlock = 4096;
    for (size_t i = 0; i < count; i += kBlock) {
        const size_t e = std::min(count, i + kBlock);
        int bad = 0;
        for (size_t j = i; j < e; ++j) {
            const float x = v[j];
            const float c = std::min(std::max(x, 0.0f), 255.0f);
            bad |= (int)(c != x) | (int)((float)(int)c != c);
        }
        if (bad) return false;
    }
    return true;
}

}  // namespace

extern "C" {

int hm_open(const hm_native_graph *g, int device, hm_index **out) {
    if (!g || !out) return invalid("null argument");
    *out = nullptr;
    if (g->n < 0 || g->dim < 1 || g->M < 2 || g->metric < 0 || g->metric > 2)
        return invalid("bad graph header");
    if (g->n >= (int64_t(1) << 31) - 1) return invalid("more than 2^31-2 nodes");
    HM_TRY(hipSetDevice(device), "hipSetDevice");
    hm_index *h = new hm_index();
    h->device = device;
    h->n = g->n;
    h->rows = g->rows;
    h->entry = g->entry_point;
    h->dim = g->dim;
    h->M = g->M;
    h->metric = g->metric;
    h->max_level = g->max_level;
    const int m0 = 2 * g->M;
    h->w0 = (int)round_up(m0, 16);
    h->w1 = (int)round_up(g->M, 16);
    h->dimp = (int)round_up(g->dim, 16);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->cus = prop.multiProcessorCount;
    auto fail = [&](int st) {
        hm_close(h);
        return st;
    };
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(fail_hip(e, "stream"));
    e = hipEventCreateWithFlags(&h->last_done, hipEventDisableTiming);
    if (e != hipSuccess) return fail(fail_hip(e, "event"));
    // Only hm_search_dev_noalloc needs the static scratch; the index opens without it.
    e = static_scratch_of(device, &h->static_scratch);
    if (e != hipSuccess) h->static_scratch_error = hipGetErrorString(e);
    const int64_t n = g->n;
    // Layer-0 rows: the used slots in stored order, repeats dropped, -1 beyond.
    {
        std::vector<int32_t> rows((size_t)n * h->w0, -1);
        for (int64_t i = 0; i < n; ++i) {
            const int32_t *src = g->links0 + (size_t)i * (m0 + 1);
            int32_t *dst = rows.data() + (size_t)i * h->w0;
            int c = std::min(g->count0[i], h->w0), w = 0;
            for (int j = 0; j < c; ++j) {
                bool rep = false;
                for (int t = 0; t < w; ++t) rep |= dst[t] == src[j];
                if (!rep) dst[w++] = src[j];
            }
        }
        if ((e = h->links0.ensure(rows.size())) != hipSuccess) return fail(fail_hip(e, "alloc"));
        e = hipMemcpy(h->links0.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(fail_hip(e, "upload links0"));
    }
    {
        std::vector<int32_t> rows((size_t)g->rows * h->w1, -1);
        for (int64_t r = 0; r < g->rows; ++r) {
            const int32_t *src = g->links_up + (size_t)r * (g->M + 1);
            int32_t *dst = rows.data() + (size_t)r * h->w1;
            int c = std::min(g->count_up[r], h->w1), w = 0;
            for (int j = 0; j < c; ++j) {
                bool rep = false;
                for (int t = 0; t < w; ++t) rep |= dst[t] == src[j];
                if (!rep) dst[w++] = src[j];
            }
        }
        std::vector<int32_t> up((size_t)n);
        for (int64_t i = 0; i < n; ++i) up[i] = (int32_t)g->upper_offset[i];
        if ((e = h->links_up.ensure(rows.size())) != hipSuccess) return fail(fail_hip(e, "alloc"));
        if (!rows.empty()) {
            e = hipMemcpy(h->links_up.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(fail_hip(e, "upload links_up"));
        }
        if ((e = h->upper.ensure(up.size())) != hipSuccess) return fail(fail_hip(e, "alloc"));
        if (n) {
            e = hipMemcpy(h->upper.p, up.data(), up.size() * 4, hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(fail_hip(e, "upload upper_offset"));
        }
    }
    const size_t vn = (size_t)n * g->dim;
    if ((e = h->vec.ensure(vn)) != hipSuccess) return fail(fail_hip(e, "alloc"));
    if (vn) {
        e = hipMemcpy(h->vec.p, g->vectors, vn * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(fail_hip(e, "upload vectors"));
    }
    // Exact 8-bit copy when every value is an integer in 0..255 (zero padded to 16 bytes).
    if (g->metric != HM_COSINE && n > 0 && all_u8_integers(g->vectors, vn)) {
        std::vector<uint8_t> u((size_t)n * h->dimp, 0);
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < g->dim; ++j)
                u[(size_t)i * h->dimp + j] = (uint8_t)g->vectors[(size_t)i * g->dim + j];
        // Padding rows after the data let the exact search's tiles read the rows in place.
        const int64_t rows_padded = n + tile_pad_rows(n);
        const size_t padded = (size_t)rows_padded * h->dimp;
        if ((e = h->u8.ensure(padded, true, h->stream)) != hipSuccess)
            return fail(fail_hip(e, "alloc"));
        e = hipMemcpyAsync(h->u8.p, u.data(), u.size(), hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return fail(fail_hip(e, "upload u8"));
        h->exact_u8 = true;
        // Row terms and accumulator starts of the exact search, computed once.
        if ((e = h->ex_xsq.ensure(n)) != hipSuccess || (e = h->ex_xsum.ensure(n)) != hipSuccess ||
            (e = h->ex_cinit.ensure(rows_padded)) != hipSuccess)
            return fail(fail_hip(e, "alloc"));
        const I8Rows rows{h->u8.p, h->dimp, 0x80808080u};
        if ((e = row_terms_launch(rows, n, h->dim, h->ex_xsq.p, h->ex_xsum.p, h->stream)) != hipSuccess ||
            (e = cinit_launch(h->ex_xsq.p, h->ex_xsum.p, n, rows_padded, h->metric,
                              h->ex_cinit.p, h->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(h->stream)) != hipSuccess)
            return fail(fail_hip(e, "exact row terms"));
    }
    *out = h;
    return HM_OK;
}

void hm_close(hm_index *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->release_all();
    if (h->last_done) (void)hipEventDestroy(h->last_done);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int hm_has_exact_u8(const hm_index *h) { return h && h->exact_u8 ? 1 : 0; }

int hm_queries_exact(const float *q, int64_t count) {
    return q && count >= 0 && all_u8_integers(q, (size_t)count) ? 1 : 0;
}

int hm_search(hm_index *h, const float *queries, const hm_search_opts *o, int64_t *ids,
              double *dists, int64_t *dcount, int64_t *hops) {
    if (!h || !queries || !o || !ids || !dists) return invalid("null argument");
    if (h->exact_u8 && o->nq > 0) {
        // Integer queries 0..255 travel as bytes: checked and packed in one pass into pinned
        // memory, a quarter of the float32 bytes over the link. (Every host-buffer call returns
        // synchronised, so no copy from the staging buffer is still in flight.)
        const size_t count = (size_t)o->nq * h->dim;
        HM_TRY(hipSetDevice(h->device), "hipSetDevice");
        HM_TRY(h->pin_q8.ensure(count), "pinned allocation");
        bool bytes = true;
        if (count >= (size_t(1) << 18)) {  // the handle's copy threads pack a quarter each
            CopyPool &team = h->staging.team();
            std::vector<uint8_t> good(team.parts(), 1);
            team.run([&](size_t p, size_t parts) {
                const size_t per = ((count + parts - 1) / parts + 15) & ~size_t(15);
                const size_t a = std::min(count, p * per), e = std::min(count, a + per);
                good[p] = a == e || pack_u8_integers(queries + a, e - a, h->pin_q8.p + a);
            });
            for (size_t p = 0; p < team.parts(); ++p) bytes = bytes && good[p];
        } else {
            bytes = pack_u8_integers(queries, count, h->pin_q8.p);
        }
        if (bytes)
            return search_host(h, kExactU8, nullptr, nullptr, o, ids, dists, dcount, hops,
                               nullptr, 0, h->pin_q8.p);
    }
    return search_host(h, kSeqF32, queries, nullptr, o, ids, dists, dcount, hops);
}

namespace {

// The batch of hm_search_dev / hm_search_graph_dev (arguments already checked).
int device_batch(hm_index *h, const float *d_queries, const hm_search_opts *o,
                 const uint32_t *d_mask_bits, int64_t *d_ids, double *d_dists, int32_t *d_dcount,
                 int32_t *d_hops, int32_t *d_fail, Batch &b, bool &filt, bool alloc = true) {
    b = Batch();
    b.nq = o->nq;
    b.q32 = d_queries;
    b.kstride = o->kstride;
    b.k_all = o->k_all;
    b.ef_all = std::max(o->ef_all, o->k_all);
    b.max_ef = b.ef_all;
    b.mask_bits = d_mask_bits;
    b.mask_words = (h->n + 31) / 32;
    b.midx = o->mask_index;
    b.mask_all = o->mask_all;
    filt = d_mask_bits != nullptr && (o->mask_index || o->mask_all >= 0);
    b.ids = d_ids;
    b.dists = d_dists;
    b.dc = d_dcount;
    b.hops = d_hops;
    b.fail = d_fail;
    if (!alloc) return HM_OK;  // the allocation-free path carves what is null
    HM_TRY(h->dc.ensure(o->nq), "alloc");
    HM_TRY(h->hops.ensure(o->nq), "alloc");
    HM_TRY(h->fail.ensure(o->nq), "alloc");
    if (!b.dc) b.dc = h->dc.p;
    if (!b.hops) b.hops = h->hops.p;
    if (!b.fail) b.fail = h->fail.p;
    b.learned = o->learned != 0;
    return HM_OK;
}

int check_dev(hm_index *h, const float *d_queries, const hm_search_opts *o, int exact_u8,
              int64_t *d_ids, double *d_dists) {
    if (!h || !d_queries || !o || !d_ids || !d_dists) return invalid("null argument");
    int st = check_opts(h, o);
    if (st) return st;
    if (exact_u8 && !h->exact_u8) return invalid("the index keeps no exact 8-bit copy");
    if (o->k || o->ef) return invalid("device batches take uniform k and ef");
    return HM_OK;
}

// Everything a captured batch bakes in: its arguments, the tuning and stop model, and the
// address and size of every scratch buffer its kernels use.
std::vector<int64_t> graph_key(const hm_index *h, const float *d_queries, const hm_search_opts *o,
                               const uint32_t *d_mask_bits, int exact_u8, const Batch &b) {
    auto P = [](const void *p) { return (int64_t)(intptr_t)p; };
    float logit = h->stop_logit;
    int32_t logit_bits;
    memcpy(&logit_bits, &logit, 4);
    return {P(d_queries), o->nq, o->k_all, o->ef_all, o->kstride, o->mask_all, P(o->mask_index),
            o->learned, P(d_mask_bits), exact_u8, P(b.ids), P(b.dists), P(b.dc), P(b.hops),
            P(b.fail), h->tune_fast_hlog, h->tune_hlog, h->tune_waves_per_cu, h->tune_probe_limit,
            h->tune_schedule, h->tune_descent_cap, h->stop_on, logit_bits, h->stop_tmin, h->stop_tmax,
            P(h->stop_w.p), P(h->bits.p), (int64_t)h->bits.cap, P(h->vlist.p),
            (int64_t)h->vlist.cap, P(h->spill_key.p), (int64_t)h->spill_key.cap,
            P(h->spill_id.p), (int64_t)h->spill_id.cap, P(h->queue.p), P(h->stats.p),
            P(h->pre_key.p), (int64_t)h->pre_key.cap, P(h->pre_id.p), (int64_t)h->pre_id.cap,
            P(h->pre_dc.p), (int64_t)h->pre_dc.cap, P(h->pre_hops.p), (int64_t)h->pre_hops.cap,
            P(h->order.p), (int64_t)h->order.cap};
}

}  // namespace

int hm_search_dev(hm_index *h, const float *d_queries, const hm_search_opts *o,
                  const uint32_t *d_mask_bits, int exact_u8, int64_t *d_ids, double *d_dists,
                  int32_t *d_dcount, int32_t *d_hops, int32_t *d_fail, void *stream) {
    int st = check_dev(h, d_queries, o, exact_u8, d_ids, d_dists);
    if (st || o->nq == 0) return st;
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    Batch b;
    bool filt = false;
    st = device_batch(h, d_queries, o, d_mask_bits, d_ids, d_dists, d_dcount, d_hops, d_fail, b,
                      filt);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    HM_TRY(order_before(h, s), "stream order");
    st = launch(h, exact_u8 ? kExactU8 : kSeqF32, filt, b, std::min<int64_t>(h->n, kDefaultSpill),
                0, s);
    if (st) return st;
    HM_TRY(order_after(h, s), "stream order");
    return HM_OK;
}

int hm_search_dev_noalloc(hm_index *h, const float *d_queries, const hm_search_opts *o,
                          const uint32_t *d_mask_bits, int exact_u8, int64_t *d_ids,
                          double *d_dists, int32_t *d_dcount, int32_t *d_hops, int32_t *d_fail,
                          void *stream) {
    int st = check_dev(h, d_queries, o, exact_u8, d_ids, d_dists);
    if (st || o->nq == 0) return st;
    if (!h->static_scratch)
        return invalid("the allocation-free scratch is unavailable: " + h->static_scratch_error);
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    Batch b;
    bool filt = false;
    st = device_batch(h, d_queries, o, d_mask_bits, d_ids, d_dists, d_dcount, d_hops, d_fail, b,
                      filt, false);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    HM_TRY(order_before(h, s), "stream order");
    Carver c;
    c.base = (uint8_t *)h->static_scratch;
    c.cap = kStaticScratchBytes;
    st = launch(h, exact_u8 ? kExactU8 : kSeqF32, filt, b, std::min<int64_t>(h->n, kDefaultSpill),
                0, s, 2, &c);
    if (st) return st;
    HM_TRY(order_after(h, s), "stream order");
    return HM_OK;
}

int hm_search_graph_dev(hm_index *h, const float *d_queries, const hm_search_opts *o,
                        const uint32_t *d_mask_bits, int exact_u8, int64_t *d_ids,
                        double *d_dists, int32_t *d_dcount, int32_t *d_hops, int32_t *d_fail,
                        void *stream) {
    int st = check_dev(h, d_queries, o, exact_u8, d_ids, d_dists);
    if (st || o->nq == 0) return st;
    if (!stream) return invalid("graph replay needs a stream other than the null stream");
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    Batch b;
    bool filt = false;
    st = device_batch(h, d_queries, o, d_mask_bits, d_ids, d_dists, d_dcount, d_hops, d_fail, b,
                      filt);
    if (st) return st;
    HM_TRY(order_before(h, s), "stream order");
    if (h->graph_exec && graph_key(h, d_queries, o, d_mask_bits, exact_u8, b) == h->graph_key) {
        HM_TRY(hipGraphLaunch(h->graph_exec, s), "graph launch");
        HM_TRY(order_after(h, s), "stream order");
        return HM_OK;
    }
    // A new shape: run it directly (this allocates whatever scratch it lacks), then capture the
    // same launches, which now allocate and clear nothing, for the next call.
    const int mode = exact_u8 ? kExactU8 : kSeqF32;
    const int64_t spill = std::min<int64_t>(h->n, kDefaultSpill);
    st = launch(h, mode, filt, b, spill, 0, s);
    if (st) return st;
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
    h->graph_key.clear();
    std::vector<int64_t> key = graph_key(h, d_queries, o, d_mask_bits, exact_u8, b);
    HM_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "graph capture");
    st = launch(h, mode, filt, b, spill, 0, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail_hip(e, "graph capture");
    if (graph_key(h, d_queries, o, d_mask_bits, exact_u8, b) != key) {  // cannot happen
        (void)hipGraphDestroy(g);
        return invalid("scratch changed during graph capture");
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail_hip(ei, "graph instantiate");
    h->graph_exec = x;
    h->graph_key = key;
    HM_TRY(order_after(h, s), "stream order");
    return HM_OK;
}

int hm_trace_dev(hm_index *h, const float *d_queries, int64_t nq, int32_t k, int32_t ef,
                 int exact_u8, const double *d_thr, float *d_trace, int32_t steps,
                 int32_t *d_len, int32_t *d_final, void *stream) {
    if (!h || !d_queries || !d_thr || !d_trace || !d_len || !d_final) return invalid("null argument");
    if (k < 1 || ef < 1 || std::max(ef, k) > 64 || steps < 1) return invalid("traces need k <= ef <= 64");
    if (exact_u8 && !h->exact_u8) return invalid("the index keeps no exact 8-bit copy");
    if (h->n == 0 || h->entry < 0) return invalid("cannot search an empty index");
    if (nq == 0) return HM_OK;
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    HM_TRY(h->ids.ensure((size_t)nq * k), "alloc");
    HM_TRY(h->dists.ensure((size_t)nq * k), "alloc");
    HM_TRY(h->dc.ensure(nq), "alloc");
    HM_TRY(h->hops.ensure(nq), "alloc");
    HM_TRY(h->fail.ensure(nq), "alloc");
    Batch b;
    b.nq = nq;
    b.q32 = d_queries;
    b.kstride = k;
    b.k_all = k;
    b.ef_all = std::max(ef, k);
    b.max_ef = b.ef_all;
    b.ids = h->ids.p;
    b.dists = h->dists.p;
    b.dc = h->dc.p;
    b.hops = h->hops.p;
    b.fail = h->fail.p;
    b.trace = d_trace;
    b.trace_steps = steps;
    b.trace_thr = d_thr;
    b.trace_len = d_len;
    b.trace_final = d_final;
    hipStream_t s = (hipStream_t)stream;
    HM_TRY(order_before(h, s), "stream order");
    int st = launch(h, exact_u8 ? kExactU8 : kSeqF32, false, b,
                    std::min<int64_t>(h->n, kDefaultSpill), 0, s);
    if (st) return st;
    HM_TRY(order_after(h, s), "stream order");
    return HM_OK;
}

int hm_trace(hm_index *h, const float *queries, int64_t nq, int32_t k, int32_t ef,
             const double *thr, float *trace, int32_t steps, int32_t *len, int32_t *final_hits) {
    if (!h || !queries || !thr || !trace || !len || !final_hits) return invalid("null argument");
    if (nq == 0) return HM_OK;
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = h->stream;
    const size_t qn = (size_t)nq * h->dim, tn = (size_t)nq * steps * kTraceCols;
    DevBuf<float> dq, dtr;
    DevBuf<double> dthr;
    DevBuf<int32_t> dlen, dfin;
    struct Free {
        DevBuf<float> *a, *b;
        DevBuf<double> *c;
        DevBuf<int32_t> *d, *e;
        ~Free() { a->release(); b->release(); c->release(); d->release(); e->release(); }
    } guard{&dq, &dtr, &dthr, &dlen, &dfin};
    HM_TRY(dq.ensure(qn), "alloc");
    HM_TRY(dtr.ensure(tn), "alloc");
    HM_TRY(dthr.ensure(nq), "alloc");
    HM_TRY(dlen.ensure(nq), "alloc");
    HM_TRY(dfin.ensure(nq), "alloc");
    HM_TRY(hipMemcpyAsync(dq.p, queries, qn * 4, hipMemcpyHostToDevice, s), "copy");
    HM_TRY(hipMemcpyAsync(dthr.p, thr, nq * 8, hipMemcpyHostToDevice, s), "copy");
    HM_TRY(hipMemsetAsync(dtr.p, 0, tn * 4, s), "memset");
    const int exact = h->exact_u8 && hm_queries_exact(queries, (int64_t)qn) ? 1 : 0;
    int st = hm_trace_dev(h, dq.p, nq, k, ef, exact, dthr.p, dtr.p, steps, dlen.p, dfin.p, s);
    if (st) return st;
    HM_TRY(hipMemcpyAsync(trace, dtr.p, tn * 4, hipMemcpyDeviceToHost, s), "copy");
    HM_TRY(hipMemcpyAsync(len, dlen.p, nq * 4, hipMemcpyDeviceToHost, s), "copy");
    HM_TRY(hipMemcpyAsync(final_hits, dfin.p, nq * 4, hipMemcpyDeviceToHost, s), "copy");
    HM_TRY(hipStreamSynchronize(s), "trace");
    return HM_OK;
}

int hm_attach_codes(hm_index *h, const uint8_t *codes) {
    if (!h || !codes) return invalid("null argument");
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    std::vector<uint8_t> u((size_t)h->n * h->dimp, 0);
    for (int64_t i = 0; i < h->n; ++i)
        memcpy(u.data() + (size_t)i * h->dimp, codes + (size_t)i * h->dim, h->dim);
    HM_TRY(h->codes.ensure(u.size()), "alloc");
    HM_TRY(hipMemcpy(h->codes.p, u.data(), u.size(), hipMemcpyHostToDevice), "upload codes");
    h->has_codes = true;
    return HM_OK;
}

int hm_search_codes(hm_index *h, const double *queries, const hm_search_opts *o, int64_t *ids,
                    double *dists, int64_t *dcount, int64_t *hops) {
    if (!h || !queries || !o || !ids || !dists) return invalid("null argument");
    if (!h->has_codes) return invalid("no codes attached");
    return search_host(h, kSeqU8, nullptr, queries, o, ids, dists, dcount, hops);
}

int hm_search_codes_rerank(hm_index *h, const double *code_queries, const float *queries,
                           const hm_search_opts *o, int32_t k, int64_t *ids, double *dists,
                           int64_t *dcount, int64_t *hops) {
    if (!h || !code_queries || !queries || !o || !ids || !dists) return invalid("null argument");
    if (!h->has_codes) return invalid("no codes attached");
    if (o->k || o->ef || o->mask_pool) return invalid("the quantized search takes uniform k and ef");
    if (k < 1 || k > o->kstride || o->k_all != o->kstride)
        return invalid("k must be in [1, fetch] and the traversal must keep fetch candidates");
    int p2 = 1;
    while (p2 < o->kstride) p2 <<= 1;
    if (p2 * 12 > 160 * 1024) return invalid("rerank fetch too large");
    return search_host(h, kSeqU8, nullptr, code_queries, o, ids, dists, dcount, hops, queries, k);
}

int hm_set_tuning(hm_index *h, int32_t fast_hash_log2, int32_t hash_log2, int32_t waves_per_cu,
                  int32_t probe_limit, int32_t one_launch, int32_t descent_cap) {
    if (!h) return invalid("null argument");
    if ((fast_hash_log2 && (fast_hash_log2 < 6 || fast_hash_log2 > 14)) ||
        (hash_log2 && (hash_log2 < 6 || hash_log2 > 14)) || waves_per_cu < 0 || probe_limit < 0 ||
        one_launch < 0 || one_launch > 1 || descent_cap < 0 || descent_cap > kDescentMaxCap)
        return invalid("tuning values out of range");
    h->tune_schedule = one_launch;
    h->tune_descent_cap = descent_cap;
    h->tune_fast_hlog = fast_hash_log2;
    h->tune_hlog = hash_log2;
    h->tune_waves_per_cu = waves_per_cu;
    h->tune_probe_limit = probe_limit;
    return HM_OK;
}

int hm_read_stats(hm_index *h, int32_t *out, int reset) {
    if (!h || !out) return invalid("null argument");
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    HM_TRY(h->stats.ensure(4, true, h->stream), "stats allocation");
    HM_TRY(hipStreamSynchronize(h->stream), "sync");
    HM_TRY(hipDeviceSynchronize(), "sync");
    HM_TRY(hipMemcpy(out, h->stats.p, 12, hipMemcpyDeviceToHost), "copy stats");
    out[3] = (int32_t)h->reruns;
    if (reset) {
        HM_TRY(hipMemset(h->stats.p, 0, 16), "reset stats");
        h->reruns = 0;
    }
    return HM_OK;
}

const char *hm_last_error(void) { return g_error.c_str(); }

}  // extern "C"
