"""Measurements of the secondary workloads (the headline one is ``bench.py`` at the repository
root). ``python -m hnsw_amd.bench_configs filtered|quantized|mixed|codes [ef ...]`` prints one
JSON line.

``filtered``: 100,000 x 784 8-bit vectors (about half exact zeros, 10 prototypes plus uniform
noise), 5,000 queries, k = 10, ef = 128, L2, M = 16, ef_construction = 200; masks of
selectivity 1 %, 5 %, 10 % and 50 %, i.i.d. random and one contiguous id range. For each mask:
the in-graph filtered search (time and tie-aware recall@10 against the exact top-10 over the
allowed rows), the masked brute force (time), and the planner's choice after calibrating its
threshold on this index. Times are host wall times of whole batches (queries uploaded, results
downloaded), best of 3. (Parity of both paths with the C oracle at 784 dimensions is a test:
tests/test_filtered784_gpu.py.)
"""

from __future__ import annotations

import json
import sys
import time

import numpy as np

# MI355X dense float32 matrix-core peak (TFLOP/s).
FP32_PEAK_TFLOPS = 157.3


def _best(fn, repeats: int = 3):
    out = fn()
    best = float("inf")
    for _ in range(repeats):
        t = time.perf_counter()
        out = fn()
        best = min(best, time.perf_counter() - t)
    return best, out


def _exact_l2(data: np.ndarray, queries: np.ndarray, ids: np.ndarray) -> np.ndarray:
    out = np.full(ids.shape, np.inf)
    for i in range(ids.shape[0]):
        v = ids[i] >= 0
        diff = data[ids[i, v]].astype(np.int64) - queries[i].astype(np.int64)
        out[i, v] = (diff * diff).sum(axis=1)
    return out


def filtered(log=print) -> dict:
    from .planner import FilterPlanner
    from .recall import recall_at_k
    from .synthetic import random_mask, range_mask
    from .workloads import WORKLOADS, load_or_build

    w = WORKLOADS["sparse784_100k"]
    eng, data, queries = load_or_build(w, log=log)
    n, k, ef = data.shape[0], 10, 128
    planner = FilterPlanner()
    # Calibrated on the batch it plans for: both paths' costs depend on the batch size (the
    # in-graph search needs thousands of queries to fill the GPU's wave slots).
    calib = planner.calibrate(eng, queries, k, ef,
                              selectivities=(0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5))
    log(f"[filtered] planner thresholds: random {planner.random_threshold}, "
        f"correlated {planner.correlated_threshold}; timings {calib}")
    rows = []
    for kind in ("random", "range"):
        for i, s in enumerate((0.01, 0.05, 0.1, 0.5)):
            mask = random_mask(n, s, 100 + i) if kind == "random" else range_mask(n, s, 200 + i)
            kk = min(k, int(mask.sum()))
            t_exact, (e_ids, e_d) = _best(lambda: eng.exact_search(queries, kk, mask))
            t_graph, got = _best(lambda: eng.search_many(queries, k, ef, mask_pool=mask,
                                                         mask_index=0))
            g_ids = got[0]
            rec = recall_at_k(g_ids[:, :kk], _exact_l2(data, queries, g_ids[:, :kk]), e_ids,
                              e_d, kk)
            plan = planner.plan(eng, mask)
            row = {
                "mask": kind, "selectivity": s, "allowed": int(mask.sum()),
                "graph_qps": queries.shape[0] / t_graph, "graph_recall": rec["recall"],
                "graph_dist_per_query": float(got[2].mean()),
                "exact_qps": queries.shape[0] / t_exact,
                "plan": plan.path, "plan_correlation": plan.correlation,
            }
            row["planned_qps"] = row["exact_qps"] if plan.path == "exact" else row["graph_qps"]
            rows.append(row)
            log(f"[filtered] {json.dumps(row)}")
    return {
        "workload": f"{w.name}: {n} x {w.dim}, {queries.shape[0]} queries, k={k}, ef={ef}",
        "planner": {"random_threshold": planner.random_threshold,
                    "correlated_threshold": planner.correlated_threshold,
                    "calibration": [{"selectivity": s, "graph_s": g, "exact_s": e}
                                    for s, g, e in calib]},
        "rows": rows,
    }


def quantized(log=print) -> dict:
    """The QuantizedIndex path on angular data: the graph searched with uint8-code distances
    (float64 code-space queries, the reference's batch_search_codes), then the ef candidates
    re-ranked by exact float64 distance (rerank = ef); recall against the exact top-10 (float64
    brute force over the float32 vectors); ef from 32 to 512."""
    from . import _native
    from .exact import exact_distances, last_range_search
    from .quantization import ScalarQuantizer
    from .recall import recall_at_k
    from .workloads import WORKLOADS, load_or_build

    w = WORKLOADS["angular_400k"]
    eng, data, queries = load_or_build(w, log=log)
    k = 10
    t_exact, (true_ids, true_d) = _best(lambda: eng.exact_search(queries, k), 2)
    tiles = last_range_search()
    # The same ground truth on the float64 chunked kernels, for comparison (identical results).
    _native.check(_native.lib().hm_set_exact_tuning(-1, 0))
    try:
        t_f64, (f_ids, f_d) = _best(lambda: eng.exact_search(queries, k), 1)
    finally:
        _native.check(_native.lib().hm_set_exact_tuning(0, 0))
    if not (np.array_equal(f_ids, true_ids) and np.array_equal(f_d, true_d)):
        raise RuntimeError("float32 range search and float64 kernels disagree")
    kernel_ms = tiles["bound_ms"] + tiles["range_ms"] + tiles["collect_ms"]
    flops = 2.0 * queries.shape[0] * data.shape[0] * data.shape[1]
    exact = {"seconds": t_exact, "float64_kernels_seconds": t_f64, **tiles,
             "roofline": {"bound": "mfma", "unit": "TFLOP/s (fp32)", "peak": FP32_PEAK_TFLOPS,
                          "achieved": flops / (kernel_ms / 1e3) / 1e12 if kernel_ms else None,
                          "range_pass": flops / (tiles["range_ms"] / 1e3) / 1e12
                          if tiles["range_ms"] else None}}
    log(f"[quantized] exact {json.dumps(exact)}")
    quant = ScalarQuantizer.fit(data)
    codes = quant.encode(data)
    qcode = quant.transform_queries(queries, w.metric)
    rows = []
    for ef in (32, 64, 128, 256, 512):
        # QuantizedIndex.batch_search(rerank=ef): traversal and re-rank in one device call, the
        # ef candidates kept in HBM (hm_search_codes_rerank)
        t, (ids, dists) = _best(lambda: eng.quantized_search(codes, qcode, queries, k, ef, ef))

        def two_calls():  # the candidates downloaded, then uploaded again for the re-rank
            cand, _ = eng.batch_search_codes(codes, qcode, ef, ef)
            return eng.rerank(queries, cand, k)
        t_two, (ids2, dists2) = _best(two_calls)
        if not (np.array_equal(ids, ids2) and np.array_equal(dists, dists2)):
            raise RuntimeError("one-call and two-call quantized searches disagree")
        t_codes, _ = _best(lambda: eng.batch_search_codes(codes, qcode, ef, ef))
        # scored as the reference scores recall: float64 distances with normalised rows
        rec = recall_at_k(ids, exact_distances(data, queries, ids, w.metric), true_ids, true_d, k)
        row = {"ef": ef, "rerank": ef, "qps": queries.shape[0] / t,
               "two_call_qps": queries.shape[0] / t_two,
               "codes_search_qps": queries.shape[0] / t_codes, "recall": rec["recall"]}
        rows.append(row)
        log(f"[quantized] {json.dumps(row)}")
    return {"workload": f"{w.name}: {w.n} x {w.dim} cosine, {queries.shape[0]} queries, k={k}",
            "quantizer": {"step": float(quant.step)}, "exact_seconds": t_exact, "exact": exact,
            "rows": rows}


HBM_PEAK_GBS = 8000.0


def _hbm(bytes_per_query: float, qps: float) -> dict:
    """Algorithmic bytes per query (rows gathered for each distance, one layer-0 links row per
    expansion) at the measured rate, against the 8 TB/s HBM peak."""
    gbs = bytes_per_query * qps / 1e9
    return {"bytes_per_query": bytes_per_query, "achieved_gbs": gbs,
            "frac": gbs / HBM_PEAK_GBS}


def codes(log=print, efs=(128,)) -> dict:
    """The codes traversal alone (``batch_search_codes``: uint8 codes, float64 code-space
    queries, the reference's ``numba_engine.py:674-701``) on angular_400k at each ef, k = ef as
    the quantized search asks for its re-rank candidates; host wall time of the whole batch,
    best of 3."""
    from .quantization import ScalarQuantizer
    from .workloads import WORKLOADS, load_or_build

    w = WORKLOADS["angular_400k"]
    eng, data, queries = load_or_build(w, log=log)
    quant = ScalarQuantizer.fit(data)
    cb = quant.encode(data)
    qcode = quant.transform_queries(queries, w.metric)
    nq = queries.shape[0]
    rows = []
    for ef in efs:
        d0, h0 = eng.distance_computations, eng.hops
        t, _ = _best(lambda: eng.batch_search_codes(cb, qcode, ef, ef))
        runs = 4  # _best runs the batch once untimed and three times timed
        dq = (eng.distance_computations - d0) / runs / nq
        hq = (eng.hops - h0) / runs / nq
        qps = nq / t
        row = {"ef": ef, "qps": qps, "seconds": t, "mean_distances": dq, "mean_hops": hq,
               "roofline": _hbm(dq * w.dim + hq * 4 * 2 * w.M, qps)}
        rows.append(row)
        log(f"[codes] {json.dumps(row)}")
    return {"workload": f"{w.name}: {w.n} x {w.dim} uint8 codes, {nq} queries, k = ef",
            "rows": rows}


def fractional(log=print, efs=(16, 46, 64, 128)) -> dict:
    """Fractional queries (float64 distances, the reference's sequential sum) on clustered_1m:
    the 10,000 queries each moved by uniform offsets in (-0.5, 0.5) (seed 1), k = 10, queries
    and outputs in HBM (``hm_search_dev``), median of 5 batches. On this index of 8-bit integers
    the sum reads the rows' one-byte copy, so the algorithmic bytes are 128 per distance and 128
    per expansion (one layer-0 links row)."""
    import ctypes

    import torch

    from . import _native
    from .workloads import WORKLOADS, load_or_build

    w = WORKLOADS["clustered_1m"]
    eng, data, queries = load_or_build(w, log=log)
    lib, h = _native.lib(), eng._device()
    dev = torch.device("cuda:0")
    nq, k = queries.shape[0], 10
    qf = queries + np.random.default_rng(1).uniform(-0.5, 0.5, queries.shape).astype(np.float32)
    q = torch.from_numpy(qf).to(dev)
    ids = torch.empty((nq, k), dtype=torch.int64, device=dev)
    dists = torch.empty((nq, k), dtype=torch.float64, device=dev)
    dc = torch.empty(nq, dtype=torch.int32, device=dev)
    hops = torch.empty(nq, dtype=torch.int32, device=dev)
    fail = torch.zeros(nq, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rows = []
    for ef in efs:
        opts = _native.SearchOpts()
        opts.nq, opts.k_all, opts.ef_all, opts.kstride, opts.mask_all = nq, k, ef, k, -1

        def go():
            _native.check(lib.hm_search_dev(h, p(q), ctypes.byref(opts), None, 0, p(ids),
                                            p(dists), p(dc), p(hops), p(fail),
                                            ctypes.c_void_p(stream.cuda_stream)))

        go()
        torch.cuda.synchronize()
        ts = []
        for _ in range(5):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            go()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        t = sorted(ts)[2]
        dq = dc.double().mean().item()
        hq = hops.double().mean().item()
        row = {"ef": ef, "qps": nq / t, "ms": t * 1e3, "mean_distances": dq, "mean_hops": hq,
               "roofline": _hbm(dq * 128 + hq * 128, nq / t)}
        rows.append(row)
        log(f"[fractional] {json.dumps(row)}")
    return {"workload": f"{w.name}: {w.n} x {w.dim}, {nq} fractional queries, k = 10",
            "rows": rows}


def mixed_batch(w, nq: int = 100_000):
    """The mixed batch on workload ``w`` (fixed seeds): ``(queries, k, ef, pool, mask_index)``
    with ef drawn from {16, 32, 64, 128, 256, 512} (raised to k), k from {1, 10, 100}, and half
    of the queries carrying one of 8 i.i.d. masks of selectivity 5-50 % (-1: unfiltered)."""
    from .synthetic import clustered_u8_queries, random_mask

    queries = clustered_u8_queries(nq, w.dim, n_clusters=w.n_clusters, sigma=w.sigma,
                                   seed=w.seed, query_seed=w.seed + 3)
    rng = np.random.default_rng(5)
    k = rng.choice([1, 10, 100], size=nq)
    ef = rng.choice([16, 32, 64, 128, 256, 512], size=nq)
    sels = np.linspace(0.05, 0.5, 8)
    pool = np.stack([random_mask(w.n, s, 300 + i) for i, s in enumerate(sels)])
    midx = np.where(rng.random(nq) < 0.5, rng.integers(0, len(sels), size=nq), -1)
    return queries, k, np.maximum(ef, k), pool, midx


def mixed(log=print) -> dict:
    """One persistent launch over 100,000 mixed queries on the 1M x 128 index
    (:func:`mixed_batch`). Each query follows exactly the reference semantics of its own
    arguments."""
    from .workloads import WORKLOADS, load_or_build

    w = WORKLOADS["clustered_1m"]
    eng, data, _ = load_or_build(w, log=log)
    queries, k, ef, pool, midx = mixed_batch(w)
    nq = queries.shape[0]
    t, got = _best(lambda: eng.search_many(queries, k, ef, mask_pool=pool, mask_index=midx), 2)
    plain = midx < 0
    dq, hq = float(got[2].mean()), float(got[3].mean())
    return {"workload": f"{w.name}: {nq} mixed queries (ef 16-512, k 1/10/100, "
                        f"{int((~plain).sum())} filtered at 5-50 %)",
            "qps": nq / t, "seconds": t, "mean_distances": dq, "mean_hops": hq,
            "roofline": _hbm(dq * w.dim + hq * 4 * 2 * w.M, nq / t)}


def main(argv: list[str]) -> int:
    which = argv[1] if len(argv) > 1 else "filtered"
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    runs = {"filtered": filtered, "quantized": quantized, "mixed": mixed, "codes": codes,
            "fractional": fractional}
    if which not in runs:
        print(f"unknown workload {which!r}; one of {sorted(runs)}", file=sys.stderr)
        return 2
    extra = ({"efs": tuple(int(v) for v in argv[2:])}
             if which in ("codes", "fractional") and argv[2:] else {})
    print(json.dumps(runs[which](log=log, **extra)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
