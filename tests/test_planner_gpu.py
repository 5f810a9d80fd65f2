"""Planned filtered search on the GPU: the path the planner picks, and its results.

The in-graph path must equal the oracle's filtered search bit for bit (the reference's
_search_layer0_filtered); the brute-force path must equal the oracle's exact top-k over the
allowed rows (the reference's exact_subset_search semantics).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from hnsw_amd import HNSWIndex
from hnsw_amd.planner import FilterPlanner
from hnsw_amd.synthetic import clustered_u8, random_mask

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def index_and_data():
    data, queries = clustered_u8(20_000, 32, 200, n_clusters=40, seed=21)
    index = HNSWIndex(32, "l2", M=16, ef_construction=100, seed=5)
    index.add(data)
    return index, data, queries


@pytest.mark.parametrize("sel", [0.5, 0.01])
def test_planned_search_takes_the_cheaper_exact_path(index_and_data, sel: float) -> None:
    index, data, queries = index_and_data
    mask = random_mask(data.shape[0], sel, 31)
    planner = FilterPlanner(random_threshold=0.05, correlated_threshold=0.1)
    print(f"case planned sel={sel}", flush=True)
    ids, dists, plan = index.batch_search_planned(queries, 10, 64, mask, planner)
    if sel >= 0.05:
        assert plan.path == "graph"
        eng = index._engine
        want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                                   queries, 10, 64, mask=mask, threads=8)
        np.testing.assert_array_equal(ids, want[0])
        np.testing.assert_array_equal(dists, want[1])
    else:
        assert plan.path == "exact"
        want_ids, want_d = oracle.exact_search(data, queries, 10, allowed=mask, threads=8)
        np.testing.assert_array_equal(ids, want_ids)
        np.testing.assert_array_equal(dists, want_d)
    assert np.all(mask[ids[ids >= 0]])


def test_calibration_sets_a_threshold(index_and_data) -> None:
    index, data, queries = index_and_data
    planner = FilterPlanner()
    print("case calibrate", flush=True)
    rows = planner.calibrate(index._engine, queries[:100], 10, 64,
                             selectivities=(0.01, 0.1, 0.5), repeats=1)
    assert len(rows) == 3
    assert 0.0 < planner.random_threshold <= 1.0
    assert planner.correlated_threshold >= planner.random_threshold or planner.random_threshold == 1.0
