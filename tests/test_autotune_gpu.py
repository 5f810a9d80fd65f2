"""GPU tests of the autotuner: every candidate setting returns the same results (the tuners
assert it), the fastest stays applied, and the search and the exact search still equal the
oracle under it."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from hnsw_amd import _native, autotune
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.exact import exact_search
from hnsw_amd.synthetic import clustered_u8

pytestmark = pytest.mark.gpu

pytest.importorskip("torch")


def test_tune_search_keeps_results(capsys) -> None:
    n, dim, M = 8_000, 128, 16
    data, queries = clustered_u8(n, dim, 397, n_clusters=40, sigma=30.0, seed=71)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(72).random(n))
                      / np.log(M)).astype(np.int64)
    eng = MI355XEngine(dim, "l2", M, 100, "heuristic", False, False)
    eng.insert(data, levels)
    print("case autotune search ef=48", flush=True)
    grid = {"waves_per_cu": (0, 1), "fast_hash_log2": (0, 12), "one_launch": (0, 1)}
    try:
        res = autotune.tune_search(eng, queries, 10, 48, grid=grid, repeats=2)
        assert len(res["table"]) == 8 and res["best"] in [
            {kk: r[kk] for kk in r if kk != "ms"} for r in res["table"]]
        got = eng.search_many(queries, 10, 48)
    finally:
        autotune._set_search(eng._device(), {})
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries, 10, 48, threads=8)
    for g, w, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w, err_msg=name)


def test_tune_exact_keeps_results() -> None:
    data, queries = clustered_u8(60_000, 128, 300, n_clusters=50, sigma=30.0, seed=73)
    print("case autotune exact", flush=True)
    grid = {"query_blocks": (0, 1, 4), "sample_every": (0, 1, 8)}
    try:
        res = autotune.tune_exact(data, queries, 10, grid=grid, repeats=1)
        assert len(res["table"]) == 9
        ids, dists = oracle.exact_search(data, queries, 10, "l2", threads=8)
        g_ids, g_d = exact_search(data, queries, 10, "l2")
    finally:
        _native.check(_native.lib().hm_set_exact_tuning(0, 0))
    np.testing.assert_array_equal(g_ids, ids)
    np.testing.assert_array_equal(g_d, dists)
