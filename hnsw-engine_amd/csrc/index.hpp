// The index handle behind the C ABI: the graph in HBM plus scratch kept between calls.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "common.hpp"
#include "hostcopy.hpp"

struct hm_index {
    int device = 0;
    int64_t n = 0, rows = 0, entry = -1;
    int dim = 0, M = 0, metric = 0, max_level = -1;
    int w0 = 0, w1 = 0, dimp = 0;  // repacked row widths, padded dimension of uint8 rows
    int cus = 1;
    bool exact_u8 = false, has_codes = false;
    hipStream_t stream = nullptr;
    // Recorded after each call's work; the next call's stream waits on it (calls share scratch).
    hipEvent_t last_done = nullptr;
    // This device's copy of the allocation-free entry point's scratch array (null: unavailable,
    // static_scratch_error says why).
    void *static_scratch = nullptr;
    std::string static_scratch_error;
    // graph: links0 (n, w0), links_up (rows, w1), upper (n,) first upper row, -1 beyond counts
    hm::DevBuf<int32_t> links0, links_up, upper;
    hm::DevBuf<float> vec;        // (n, dim)
    hm::DevBuf<uint8_t> u8, codes;  // (n, dimp)
    // exact search over u8 (set at open): per row sum (v - 128)^2, sum (v - 128), and the range
    // search's accumulator start for this index's metric
    hm::DevBuf<int32_t> ex_xsq, ex_xsum, ex_cinit;
    // search scratch
    hm::DevBuf<uint32_t> bits, vlist;
    hm::DevBuf<uint64_t> spill_key;
    hm::DevBuf<uint32_t> spill_id;
    hm::DevBuf<int32_t> queue;
    // two-phase schedule: per query layer-0 entry, counters after the descent, and the order
    hm::DevBuf<uint64_t> pre_key;
    hm::DevBuf<uint32_t> pre_id;
    hm::DevBuf<int32_t> pre_dc, pre_hops, order;
    // staging of the host-buffer entry points
    hm::DevBuf<float> q32;
    hm::DevBuf<double> q64;
    hm::DevBuf<double> qwide;  // float32 queries as float64 for the float64 sum over u8 rows
    hm::DevBuf<int32_t> kq, efq, midx, dc, hops, fail, subset;
    hm::DevBuf<uint8_t> boost;  // per position of a mixed batch: order boosts
    // 8-bit queries of hm_search (exact mode): packed on the host into pinned memory, then copied
    hm::PinBuf<uint8_t> pin_q8;
    hm::DevBuf<uint8_t> q8;
    hm::DevBuf<uint32_t> maskbits;
    hm::DevBuf<int64_t> ids, cand, rids;
    hm::DevBuf<double> dists, rdists;
    // learned stop (hm_set_stop_model)
    hm::DevBuf<float> stop_w;
    float stop_logit = 0.f;
    int stop_tmin = 0, stop_tmax = 0;
    bool stop_on = false;
    // tuning (0: automatic): visited-hash size of the register-beam and general kernels, waves
    // per compute unit
    int tune_fast_hlog = 0, tune_hlog = 0, tune_waves_per_cu = 0, tune_probe_limit = 0;
    int tune_schedule = 0;  // 1: one launch per batch (no descents-first schedule)
    int tune_descent_cap = 0;  // > 0: descend_kernel visited entries per layer (else the maximum)
    hm::DevBuf<int32_t> stats;  // event counters, see SearchArgs::stats
    int64_t reruns = 0;         // queries hm_search ran again after a spill or tie overflow
    // hm_search_graph_dev: the captured launches of the last batch shape and what they bake in
    hipGraphExec_t graph_exec = nullptr;
    std::vector<int64_t> graph_key;
    // large host results: pinned staging and copy threads (made on first use)
    hm::Staging staging;

    void release_all() {
        hm::DevBuf<int32_t> *i32[] = {&links0, &links_up, &upper, &queue,    &kq,   &efq,
                                      &midx,   &dc,       &hops,  &fail,     &pre_dc,
                                      &pre_hops, &order, &subset};
        for (auto *b : i32) b->release();
        vec.release();
        u8.release();
        ex_xsq.release();
        ex_xsum.release();
        ex_cinit.release();
        codes.release();
        bits.release();
        vlist.release();
        spill_key.release();
        pre_key.release();
        pre_id.release();
        spill_id.release();
        q32.release();
        q64.release();
        qwide.release();
        q8.release();
        pin_q8.release();
        boost.release();
        maskbits.release();
        ids.release();
        cand.release();
        rids.release();
        dists.release();
        rdists.release();
        stop_w.release();
        stats.release();
        staging.release();
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
        graph_key.clear();
    }
};

namespace hm {
int invalid(const std::string &msg);
}
