#!/usr/bin/env python3
"""Benchmark: QPS at tie-aware recall@10 >= 0.95 on 1M x 128 (one MI355X).

Workload (``hnsw_amd.workloads.clustered_1m``): 1,000,000 x 128 float32 vectors with integer
values 0..255 from 1,000 Gaussian clusters, 10,000 queries, L2, M = 16, ef_construction = 200;
the graph is built by sequential insertion on the host (cached under ``.graph_cache/``).

Steps: exact top-10 ground truth with ``exact_search`` (GPU). Plain beam: a sweep of ef picks
the smallest ef whose recall@10 >= 0.95. Learned stop: traces of the plain search on 10,000
training queries (disjoint from the measured ones) label each layer-0 expansion as final or not,
the stop MLP is trained on the GPU (<= 200,000 rows), and a sweep of (ef, threshold) picks the
fastest setting with recall@10 >= 0.95. Then W untimed and K timed steps of the faster variant
(the learned stop, unless it cannot reach the recall), a step being one batch search of all
10,000 queries with queries and outputs already in HBM. The JSON line carries the roofline of
one batch search (its kernels: the batched upper-layer descent, the longest-first order, the
layer-0 beam; algorithmic bytes per batch over the batch's average duration, HIP events on the
launch stream), the same batches replayed from a captured HIP graph (``config.graph_replay``)
and the CPU baseline (the C oracle at the plain beam's ef, 1 thread and all
cores, on a bounded query sample, ids checked equal to the GPU's plain results).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "hnsw-engine_amd"), str(ROOT)]

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def exact_l2(data: np.ndarray, queries: np.ndarray, ids: np.ndarray) -> np.ndarray:
    """Exact squared distances of the returned ids (integer data: exact in int64)."""
    out = np.full(ids.shape, np.inf)
    valid = ids >= 0
    rows = data[np.where(valid, ids, 0)].astype(np.int64)
    diff = rows - queries.astype(np.int64)[:, None, :]
    d = (diff * diff).sum(axis=2).astype(np.float64)
    out[valid] = d[valid]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="clustered_1m")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--sweep", default="16,24,32,40,42,44,46,48,56,64,96,128,256,512")
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="target CPU time of each CPU-baseline leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-queries", type=int, default=10_000)
    ap.add_argument("--learned-ef", default="48,56,64")
    ap.add_argument("--thresholds", default="0.5,0.6,0.7,0.8,0.85,0.9,0.95")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB on the device three times (run under rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE: 1,048,576 KiB over the counter's KiB per "
                         "dispatch is the factor that turns the counter into bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (asked for {args.gpus}); the search path shards by "
            "replicating the index and splitting queries, which needs no collective.",
            file=sys.stderr,
        )
        return 2

    import torch

    from hnsw_amd import _native
    from hnsw_amd.exact import exact_search
    from hnsw_amd.recall import recall_at_k
    from hnsw_amd.workloads import WORKLOADS, load_or_build

    w = WORKLOADS[args.workload]
    eng, data, queries = load_or_build(w, log=log)
    nq, k = queries.shape[0], args.k
    handle = eng._device()
    lib = _native.lib()

    exact_u8 = bool(lib.hm_has_exact_u8(handle)) and bool(
        lib.hm_queries_exact(_native.ptr(queries), queries.size))
    dev = torch.device("cuda:0")
    q_dev = torch.from_numpy(queries).to(dev)
    stream = torch.cuda.current_stream()

    log("[bench] exact ground truth on the GPU (hm_index_exact_dev)")
    oracle_info = exact_ground_truth(lib, handle, data.shape[0], data.shape[1], q_dev, k, exact_u8)
    true_ids, true_d = oracle_info.pop("ids"), oracle_info.pop("dists")
    log(f"[bench] exact: {oracle_info}")
    ids_dev = torch.empty((nq, k), dtype=torch.int64, device=dev)
    d_dev = torch.empty((nq, k), dtype=torch.float64, device=dev)
    dc_dev = torch.empty(nq, dtype=torch.int32, device=dev)
    hops_dev = torch.empty(nq, dtype=torch.int32, device=dev)
    fail_dev = torch.zeros(nq, dtype=torch.int32, device=dev)

    def launch(ef: int, learned: bool = False) -> None:
        opts = _native.SearchOpts()
        opts.nq, opts.k_all, opts.ef_all, opts.kstride, opts.mask_all = nq, k, ef, k, -1
        opts.learned = int(learned)
        _native.check(lib.hm_search_dev(
            handle, ctypes.c_void_p(q_dev.data_ptr()), ctypes.byref(opts), None, int(exact_u8),
            ctypes.c_void_p(ids_dev.data_ptr()), ctypes.c_void_p(d_dev.data_ptr()),
            ctypes.c_void_p(dc_dev.data_ptr()), ctypes.c_void_p(hops_dev.data_ptr()),
            ctypes.c_void_p(fail_dev.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))

    def measure(ef: int, learned: bool = False) -> dict:
        launch(ef, learned)
        torch.cuda.synchronize()
        t = time.perf_counter()
        launch(ef, learned)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        if int(fail_dev.sum()) != 0:
            raise RuntimeError("a query overflowed its candidate spill")
        ids = ids_dev.cpu().numpy()
        rec = recall_at_k(ids, exact_l2(data, queries, ids), true_ids, true_d, k)
        dc = dc_dev.cpu().numpy().astype(np.int64)
        hops = hops_dev.cpu().numpy().astype(np.int64)
        return {"ef": ef, "recall": rec["recall"], "recall_ids": rec["recall_ids"],
                "qps": nq / dt, "dcount_mean": float(dc.mean()), "hops_mean": float(hops.mean()),
                "ids": ids, "dc": dc, "hops": hops, "learned": learned}

    sweep = []
    for ef in [int(v) for v in args.sweep.split(",")]:
        sweep.append(measure(ef))
        r = sweep[-1]
        log(f"[bench] plain ef={ef} recall={r['recall']:.4f} qps={r['qps']:,.0f} "
            f"dist/q={r['dcount_mean']:.0f} hops/q={r['hops_mean']:.0f}")
    plain = next((s for s in sweep if s["recall"] >= args.target_recall), sweep[-1])
    chosen = plain

    learned_info = None
    if not args.no_learned:
        from hnsw_amd import learned as lrn

        t0 = time.perf_counter()
        train_q = w.train_queries(args.train_queries)
        _, t_d = exact_search(data, train_q, k)
        thr = lrn.recall_thresholds(t_d, k)
        xs, ys = [], []
        efs_l = [int(v) for v in args.learned_ef.split(",")]
        for ef_l in efs_l:
            trace, length, final = eng.trace(train_q, k, ef_l, thr, steps=256)
            # Rows of long queries weigh more: they set the batch time.
            x, y = lrn.dataset_from_traces(trace, length, final, max_rows=200_000 // len(efs_l),
                                           seed=ef_l, long_weight=2.0)
            xs.append(x)
            ys.append(y)
        x = np.concatenate(xs)
        y = np.concatenate(ys)
        params, losses = lrn.train(x, y, epochs=12, seed=3)
        t_train = time.perf_counter() - t0
        log(f"[bench] stop model: {x.shape[0]} rows ({y.mean():.2f} positive), loss "
            f"{losses[0]:.4f} -> {losses[-1]:.4f}, {t_train:.1f}s")
        # Expansion caps: the tail-aware oracle's cap on the training traces (largest ef), and
        # looser ones; 0 is no cap.
        slots = 16 * torch.cuda.get_device_properties(0).multi_processor_count
        cap_plan = lrn.oracle_stop_tail(trace, length, k, args.target_recall, slots, length)
        c0 = int(cap_plan["cap"]) if cap_plan else 0
        caps = sorted({0, c0, int(1.25 * c0), int(1.5 * c0)} if c0 else {0})
        log(f"[bench] stop caps from the training traces: {caps}")
        lsweep = []
        for cap in caps:
            eng.set_stop_cap(cap)
            for ef_l in [int(v) for v in args.learned_ef.split(",")]:
                for tau in [float(v) for v in args.thresholds.split(",")]:
                    eng.set_stop_model(params, tau)
                    r = measure(ef_l, learned=True)
                    r["threshold"] = tau
                    r["cap"] = cap
                    lsweep.append(r)
                    log(f"[bench] learned ef={ef_l} threshold={tau} cap={cap} "
                        f"recall={r['recall']:.4f} qps={r['qps']:,.0f} "
                        f"dist/q={r['dcount_mean']:.0f} hops/q={r['hops_mean']:.0f}")
        # Ceiling of any stop rule on this beam: the oracle stopping points of the measured
        # queries (known future) at the target recall, against the plain beam's expansions.
        thr_m = lrn.recall_thresholds(true_d, k)
        tr_p, len_p, _ = eng.trace(queries, k, plain["ef"], thr_m, steps=256)
        tr_o, len_o, _ = eng.trace(queries, k, max(int(v) for v in args.learned_ef.split(",")),
                                   thr_m, steps=256)
        ceiling = lrn.oracle_stop(tr_o, len_o, k, args.target_recall)
        ceiling_info = None
        if ceiling is not None:
            ceiling_info = {"layer0_expansions": ceiling[0], "recall": ceiling[1],
                           "plain_layer0_expansions": float(np.minimum(len_p, 256).mean()),
                           "max_speedup_in_expansions":
                               float(np.minimum(len_p, 256).mean()) / ceiling[0]}
            log(f"[bench] stop ceiling: {ceiling_info}")
        # The same ceiling for the batch time: list scheduling of the per-query layer-0
        # expansions on the launch's wave slots (16 per compute unit), with stopping points
        # chosen to minimise that makespan.
        slots = 16 * torch.cuda.get_device_properties(0).multi_processor_count
        tail = lrn.oracle_stop_tail(tr_o, len_o, k, args.target_recall, slots, len_p)
        log(f"[bench] tail-aware stop ceiling: {tail}")
        ok = [r for r in lsweep if r["recall"] >= args.target_recall]
        best = max(ok, key=lambda r: r["qps"]) if ok else None
        learned_best = best
        learned_info = {
            "train_rows": int(x.shape[0]), "positive_share": float(y.mean()),
            "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
            "train_seconds": t_train,
            "oracle_stop_ceiling": ceiling_info,
            "oracle_stop_ceiling_tail_aware": tail,
            "sweep": [{kk: r[kk] for kk in ("ef", "threshold", "cap", "recall", "qps",
                                             "dcount_mean", "hops_mean")} for r in lsweep],
        }
        if best is not None and best["qps"] > plain["qps"]:
            chosen = best
            eng.set_stop_model(params, best["threshold"])
            eng.set_stop_cap(best["cap"])
        learned_info["best"] = None if learned_best is None else {
            kk: learned_best[kk] for kk in ("ef", "threshold", "cap", "recall", "qps",
                                            "dcount_mean")}
    ef = chosen["ef"]
    use_learned = chosen.get("learned", False)

    for _ in range(args.warmup):
        launch(ef, use_learned)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t_begin = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    t_begin.record(stream)
    for i in range(args.steps):
        starts[i].record(stream)
        launch(ef, use_learned)
        ends[i].record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_begin.elapsed_time(t_end)
    launch_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    qps = nq * args.steps / (total_ms / 1e3)
    ids_final = ids_dev.cpu().numpy()
    assert np.array_equal(ids_final, chosen["ids"]), "non-deterministic search results"

    # The same batches replayed from a captured HIP graph (hm_search_graph_dev, on a stream of
    # its own), reported beside the direct launches.
    gstream = torch.cuda.Stream()
    with torch.cuda.stream(gstream):
        opts = _native.SearchOpts()
        opts.nq, opts.k_all, opts.ef_all, opts.kstride, opts.mask_all = nq, k, ef, k, -1
        opts.learned = int(use_learned)
        gargs = (handle, ctypes.c_void_p(q_dev.data_ptr()), ctypes.byref(opts), None,
                 int(exact_u8), ctypes.c_void_p(ids_dev.data_ptr()),
                 ctypes.c_void_p(d_dev.data_ptr()), ctypes.c_void_p(dc_dev.data_ptr()),
                 ctypes.c_void_p(hops_dev.data_ptr()), ctypes.c_void_p(fail_dev.data_ptr()),
                 ctypes.c_void_p(gstream.cuda_stream))
        for _ in range(max(2, args.warmup)):  # the first call captures
            _native.check(lib.hm_search_graph_dev(*gargs))
        gstream.synchronize()
        g0 = torch.cuda.Event(enable_timing=True)
        g1 = torch.cuda.Event(enable_timing=True)
        g0.record(gstream)
        for _ in range(args.steps):
            _native.check(lib.hm_search_graph_dev(*gargs))
        g1.record(gstream)
        gstream.synchronize()
    graph_ms = g0.elapsed_time(g1) / args.steps
    assert np.array_equal(ids_dev.cpu().numpy(), chosen["ids"]), "graph replay differs"
    graph_info = {"ms_per_step": graph_ms, "qps": nq / (graph_ms / 1e3)}

    # Host call to host results: the same batch through hm_search (queries copied in over PCIe,
    # ids and distances copied out), median of 3 calls.
    h_ids = np.empty((nq, k), dtype=np.int64)
    h_d = np.empty((nq, k), dtype=np.float64)
    hopts = _native.SearchOpts()
    hopts.nq, hopts.k_all, hopts.ef_all, hopts.kstride, hopts.mask_all = nq, k, ef, k, -1
    hopts.learned = int(use_learned)

    def host_call() -> float:
        t = time.perf_counter()
        _native.check(lib.hm_search(handle, _native.ptr(queries), ctypes.byref(hopts),
                                    _native.ptr(h_ids), _native.ptr(h_d), None, None))
        return time.perf_counter() - t

    host_call()
    host_times = sorted(host_call() for _ in range(3))
    host_info = {"qps": nq / host_times[1], "ms_per_batch": host_times[1] * 1e3,
                 "ids_match_device_path": bool(np.array_equal(h_ids, chosen["ids"])),
                 "note": "hm_search with host buffers: H2D queries + search + D2H ids/dists, "
                         "median of 3"}
    log(f"[bench] host-buffer (PCIe-inclusive) batch: {host_info}")

    # Algorithmic bytes of one launch: every distance reads one stored row, every expansion one
    # links row (upper-layer rows counted at the layer-0 width), as the reference counts them.
    row_bytes = eng.dim if exact_u8 else eng.dim * 4
    if exact_u8:
        row_bytes = (eng.dim + 15) // 16 * 16
    link_bytes = ((2 * eng.M + 15) // 16 * 16) * 4
    algo_bytes = float(chosen["dc"].sum() * row_bytes + chosen["hops"].sum() * link_bytes)
    achieved = algo_bytes / (launch_ms / 1e3) / 1e9
    # HBM traffic comes from rocprofv3 counters, which this process cannot read: the committed
    # profile of the same launch (same ef and variant) supplies it, labelled with its commit.
    traffic, traffic_source = None, None
    tfile = ROOT / "profiles" / f"search_{args.workload}_traffic.json"
    if tfile.is_file():
        tj = json.loads(tfile.read_text())
        if tj.get("ef") == ef and bool(tj.get("learned", False)) == bool(use_learned):
            traffic = tj.get("hbm_bytes_per_launch")
            traffic_source = (f"profiles/{tfile.name}: rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                              f"dispatch x copy calibration, measured at commit "
                              f"{tj.get('commit', 'unrecorded')}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(eng, queries, k, plain["ef"], plain["ids"], args.cpu_seconds)
        cpu["ids_match_gpu"] = bool(cpu.pop("match"))

    line = {
        "metric": "QPS at recall@10>=0.95, 1M x 128 fp32, learned-stop vs plain beam, 1 GPU",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8-exact" if exact_u8 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {w.n} x {w.dim} float32, integer values 0..255, "
            f"{w.n_clusters} clusters, {nq} queries, L2, M={w.M}, efC={w.ef_construction}",
            "k": k, "ef": ef, "recall": chosen["recall"],
            "variant": (f"learned stop (threshold {chosen['threshold']}, cap {chosen['cap']})"
                        if use_learned else "plain beam"),
            "plain": {"ef": plain["ef"], "recall": plain["recall"], "qps": plain["qps"]},
            "learned_speedup_vs_plain": (
                learned_info["best"]["qps"] / plain["qps"]
                if learned_info and learned_info.get("best") else None),
            "sweep": [{kk: s[kk] for kk in ("ef", "recall", "qps", "dcount_mean", "hops_mean")}
                      for s in sweep],
            "learned": learned_info,
            "graph_replay": graph_info,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source,
            "kernel": "one batch search: descend_kernel, order_by_entry_kernel, layer-0 kernel",
            "launch_ms": launch_ms, "algo_bytes": algo_bytes,
            "bytes_unit": (f"{row_bytes} B per distance (the "
                           + ("uint8 copy of the row, exact for these integer values"
                              if exact_u8 else "float32 row") + f"), {link_bytes} B per expansion "
                           "(one layer-0 link row)"),
        },
        "host_buffers": host_info,
        "oracle": oracle_info,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def copy_calibration() -> int:
    import torch

    src = torch.ones(1 << 28, device="cuda")  # float32: 1 GiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied 3 x {src.numel() * 4} bytes")
    return 0


def exact_ground_truth(lib, handle, n: int, dim: int, q_dev, k: int, exact_u8: bool) -> dict:
    """Exact top-k of every query over the index's rows (``hm_index_exact_dev``: rows, queries
    and outputs in HBM; integer rows read in place as bytes), timed.

    The range search reports its kernel times (HIP events on the launch stream) through
    ``hm_exact_timing``; the roofline is the algorithmic int8 work, 2 nq n dim operations, over the
    sum of its kernels (bound pass, threshold, range pass, collect), against the dense int8 peak,
    and over the whole call (host entry to results in HBM, median of 3).
    """
    import torch

    from hnsw_amd import _native

    nq = q_dev.shape[0]
    ids = torch.empty((nq, k), dtype=torch.int64, device=q_dev.device)
    dists = torch.empty((nq, k), dtype=torch.float64, device=q_dev.device)

    def call() -> float:
        torch.cuda.synchronize()
        t = time.perf_counter()
        _native.check(lib.hm_index_exact_dev(
            handle, ctypes.c_void_p(q_dev.data_ptr()), nq, k, None, int(exact_u8),
            ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(dists.data_ptr())))
        torch.cuda.synchronize()
        return time.perf_counter() - t

    call()  # first call loads the kernels and sizes the scratch
    walls = sorted(call() for _ in range(3))
    wall = walls[1]
    tm = np.zeros(4, dtype=np.float64)
    _native.check(lib.hm_exact_timing(_native.ptr(tm)))
    ops = 2.0 * nq * n * dim
    kernel_ms = float(tm[0] + tm[1] + tm[2])
    info = {
        "ids": ids.cpu().numpy(), "dists": dists.cpu().numpy(),
        "path": "int8 range search" if kernel_ms > 0 else ("int8 chunked" if exact_u8 else "float"),
        "call_ms": wall * 1e3,
        "bound_ms": float(tm[0]), "range_ms": float(tm[1]), "collect_ms": float(tm[2]),
        "requeried": int(tm[3]),
    }
    if kernel_ms > 0:
        tops = ops / (kernel_ms / 1e3) / 1e12
        info["roofline"] = {"bound": "mfma", "achieved": tops, "peak": INT8_PEAK_TOPS,
                            "unit": "TOPS (int8)", "frac": tops / INT8_PEAK_TOPS,
                            "kernel_ms": kernel_ms, "ops": ops}
        info["range_pass_frac"] = ops / (tm[1] / 1e3) / 1e12 / INT8_PEAK_TOPS
        info["call_frac"] = ops / wall / 1e12 / INT8_PEAK_TOPS
    return info


def cpu_baseline(eng, queries, k: int, ef: int, gpu_ids: np.ndarray, seconds: float) -> dict:
    """The C oracle on the host: 1 thread and all cores, bounded samples, ids checked."""
    import oracle

    arrays = eng.native_arrays()
    cores = oracle.threads_available()
    log(f"[bench] CPU baseline (C oracle), {cores} cores")
    probe = 100
    t = time.perf_counter()
    ids, _, _, _ = oracle.search_batch(arrays, eng.entry_point, eng.max_level, eng.metric,
                                       queries[:probe], k, ef, threads=1)
    per_q = (time.perf_counter() - t) / probe
    match = bool(np.array_equal(ids, gpu_ids[:probe]))
    n1 = int(min(queries.shape[0], max(probe, seconds / per_q)))
    t = time.perf_counter()
    ids1, _, _, _ = oracle.search_batch(arrays, eng.entry_point, eng.max_level, eng.metric,
                                        queries[:n1], k, ef, threads=1)
    qps1 = n1 / (time.perf_counter() - t)
    match &= bool(np.array_equal(ids1, gpu_ids[:n1]))
    nall = queries.shape[0]
    reps = max(1, int(seconds / (nall / (qps1 * cores))))
    t = time.perf_counter()
    for _ in range(reps):
        idsa, _, _, _ = oracle.search_batch(arrays, eng.entry_point, eng.max_level, eng.metric,
                                            queries, k, ef, threads=cores)
    qpsa = nall * reps / (time.perf_counter() - t)
    match &= bool(np.array_equal(idsa, gpu_ids))
    log(f"[bench] CPU 1-thread {qps1:,.0f} q/s ({n1} queries); {cores} threads {qpsa:,.0f} q/s")
    return {
        "value": qpsa, "unit": "queries/s", "cores": cores, "kind": "port",
        "sample": f"all {nall} queries x {reps} at ef={ef} on {cores} threads; "
                  f"1 thread: {qps1:.0f} q/s on the first {n1} queries",
        "single_thread_qps": qps1, "match": match,
    }


if __name__ == "__main__":
    sys.exit(main())
