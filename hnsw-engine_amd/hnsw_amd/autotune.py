"""Autotuner: times a sample batch under each candidate setting of the library's tuning knobs and
keeps the fastest (an extension; the reference has no launch parameters to tune).

Search (``tune_search``), through ``hm_set_tuning`` on one index:

* waves per compute unit: how many queries each CU keeps in flight (0: as many as the LDS and
  registers allow);
* log2 of the register beam's LDS visited-table slots: a larger table keeps more layer
  searches in LDS before they spill to the per-wave global bitset, but leaves room for fewer
  waves;
* one launch per batch, or the batched descent with layer 0 longest-first.

Exact search (``tune_exact``), through ``hm_set_exact_tuning`` (process-wide): the tile shape of
the int8 range search, query blocks of 32 per wave and the bound pass's superblock sampling.

No knob changes results, and both tuners assert it: every setting's ids (and, for the search,
distances and counters) must equal the first setting's. Timings are HIP events on a stream of
the tuner's own around ``repeats`` launches with device-resident inputs; the best setting stays
applied and is returned with the table of timings.
"""

from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import _native

SEARCH_GRID = {
    "waves_per_cu": (0, 20, 16),
    "fast_hash_log2": (0, 12),
    "one_launch": (0, 1),
}
EXACT_GRID = {"query_blocks": (0, 2, 4), "sample_every": (0, 2, 8)}


def _settings(grid: dict) -> list[dict]:
    keys = list(grid)
    return [dict(zip(keys, vals)) for vals in itertools.product(*(grid[k] for k in keys))]


def _set_search(handle, s: dict) -> None:
    _native.check(_native.lib().hm_set_tuning(handle, s.get("fast_hash_log2", 0),
                                              s.get("hash_log2", 0), s.get("waves_per_cu", 0),
                                              s.get("probe_limit", 0), s.get("one_launch", 0),
                                              s.get("descent_cap", 0)))


def tune_search(engine, queries: np.ndarray, k: int, ef: int, *, grid: dict | None = None,
                repeats: int = 3) -> dict:
    """Fastest ``hm_set_tuning`` setting for batches like ``queries`` (float32 (nq, dim)) at
    (k, ef) on ``engine``'s index; applies it and returns ``{"best", "ms", "table"}``."""
    import torch

    lib, handle = _native.lib(), engine._device()
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    nq = queries.shape[0]
    exact_u8 = bool(lib.hm_has_exact_u8(handle)) and bool(
        lib.hm_queries_exact(_native.ptr(queries), queries.size))
    dev = torch.device("cuda:0")
    q = torch.from_numpy(queries).to(dev)
    ids = torch.empty((nq, k), dtype=torch.int64, device=dev)
    dists = torch.empty((nq, k), dtype=torch.float64, device=dev)
    dc = torch.empty(nq, dtype=torch.int32, device=dev)
    hops = torch.empty(nq, dtype=torch.int32, device=dev)
    fail = torch.zeros(nq, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream()
    opts = _native.SearchOpts()
    opts.nq, opts.k_all, opts.ef_all, opts.kstride, opts.mask_all = nq, k, ef, k, -1
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    args = (handle, p(q), ctypes.byref(opts), None, int(exact_u8), p(ids), p(dists), p(dc),
            p(hops), p(fail), ctypes.c_void_p(stream.cuda_stream))
    table, ref = [], None
    try:
        for s in _settings(grid or SEARCH_GRID):
            _set_search(handle, s)
            with torch.cuda.stream(stream):
                _native.check(lib.hm_search_dev(*args))  # warm-up (allocations)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(repeats):
                    _native.check(lib.hm_search_dev(*args))
                e1.record(stream)
                stream.synchronize()
            out = tuple(t.cpu().numpy() for t in (ids, dists, dc, hops))
            if ref is None:
                ref = out
            elif not all(np.array_equal(a, b) for a, b in zip(out, ref)):
                raise AssertionError(f"tuning {s} changed the search results")
            table.append({**s, "ms": e0.elapsed_time(e1) / repeats})
    except BaseException:
        _set_search(handle, {})
        raise
    best = min(table, key=lambda r: r["ms"])
    setting = {kk: best[kk] for kk in best if kk != "ms"}
    _set_search(handle, setting)
    return {"best": setting, "ms": best["ms"], "table": table}


def tune_exact(data: np.ndarray, queries: np.ndarray, k: int, metric: str = "l2", *,
               grid: dict | None = None, repeats: int = 2) -> dict:
    """Fastest ``hm_set_exact_tuning`` tile shape for the int8 range search of ``queries`` over
    ``data`` (both 8-bit integer valued float32); applies it and returns
    ``{"best", "ms", "table"}``. Each timing is the kernels' own (``hm_exact_timing``)."""
    import torch

    lib = _native.lib()
    data = np.ascontiguousarray(data, dtype=np.float32)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    if not (lib.hm_queries_exact(_native.ptr(data), data.size)
            and lib.hm_queries_exact(_native.ptr(queries), queries.size)):
        raise ValueError("the int8 range search needs 8-bit integer data and queries")
    n, dim = data.shape
    nq = queries.shape[0]
    dev = torch.device("cuda:0")
    x = torch.from_numpy(data).to(dev)
    q = torch.from_numpy(queries).to(dev)
    ids = torch.empty((nq, k), dtype=torch.int64, device=dev)
    dists = torch.empty((nq, k), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    code = _native.METRIC_CODES[metric]
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    tm = (ctypes.c_double * 4)()
    table, ref = [], None
    try:
        for s in _settings(grid or EXACT_GRID):
            _native.check(lib.hm_set_exact_tuning(s["query_blocks"], s["sample_every"]))
            best = float("inf")
            for _ in range(repeats):
                _native.check(lib.hm_exact_dev(p(x), n, dim, code, p(q), nq, k, None, 1, p(ids),
                                               p(dists), ctypes.c_void_p(stream.cuda_stream)))
                _native.check(lib.hm_exact_timing(tm))
                best = min(best, tm[0] + tm[1] + tm[2])
            out = (ids.cpu().numpy(), dists.cpu().numpy())
            if ref is None:
                ref = out
            elif not all(np.array_equal(a, b) for a, b in zip(out, ref)):
                raise AssertionError(f"tile shape {s} changed the exact results")
            table.append({**s, "ms": best})
    except BaseException:
        _native.check(lib.hm_set_exact_tuning(0, 0))
        raise
    best = min(table, key=lambda r: r["ms"])
    setting = {kk: best[kk] for kk in best if kk != "ms"}
    _native.check(lib.hm_set_exact_tuning(setting["query_blocks"], setting["sample_every"]))
    return {"best": setting, "ms": best["ms"], "table": table}
