"""The reference's own special cases of the search, run through the HIP kernels.

Each case has an answer fixed without the reference's code:

* a complete graph (simple selection, M >= n: every node links to every earlier one) searched
  with ef = n visits everything, so it equals brute force, ties included
  (``tests/test_index.py:157-170, 458-472``, ``tests/test_properties.py:92-104, 145-170`` of the
  reference). With n up to 150 the layer-0 rows hold up to 149 links, which the kernels take 64
  at a time (a path no other test reaches), on the register beam (ef <= 64) and on the general
  kernel (ef > 64);
* a hand-built graph whose node 2 has no incoming link pads with -1 / inf
  (``tests/test_index.py:341-352``);
* an all-true filter takes exactly the unfiltered steps: same ids, distances and counters
  (``tests/test_index.py:427-440``);
* an all-false filter returns padding without searching (``tests/test_index.py:489-499``);
* fewer allowed nodes than k pad (``tests/test_index.py:475-486``);
* random small integer data (values 0-3: ties and duplicates everywhere, exact sums): filtered
  results are valid (allowed, distinct, sorted, distances equal to a rescoring) and complete
  graphs equal brute force over the allowed nodes (hypothesis,
  ``tests/test_properties.py:116-170``).
"""

from __future__ import annotations

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import oracle
from hnsw_amd import HNSWIndex, exact_distances, exact_search
from hnsw_amd.synthetic import clustered_u8

pytestmark = pytest.mark.gpu


def _small_ints(n: int, dim: int, seed: int, high: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(0, high, size=(n, dim)).astype(np.float32)


@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("n", [60, 150])
def test_complete_graph_with_full_ef_is_brute_force(metric: str, n: int) -> None:
    data = _small_ints(n, 4, 8 + n, 5)
    queries = _small_ints(25, 4, 9 + n, 5)
    index = HNSWIndex(4, metric=metric, M=n, ef_construction=n, selection="simple", seed=1)
    index.add(data)
    counts = index._engine.native_arrays()["count0"]
    assert counts.max() == n - 1  # complete: every row holds every other node
    for ef in (n, 64) if n > 64 else (n,):
        print(f"case complete graph n={n} {metric} ef={ef}", flush=True)
        if ef < n:  # register beam over wide rows: equal to the oracle
            got = index._engine.search_many(queries, 10, ef)
            want = oracle.search_batch(index._engine.native_arrays(), index.entry_point,
                                       index.max_level, metric, queries, 10, ef)
            for g, w in zip(got, want):
                np.testing.assert_array_equal(g, w)
            continue
        ids, dists = index.batch_search(queries, k=10, ef=ef)
        t_ids, t_d = oracle.exact_search(data, queries, 10, metric)
        np.testing.assert_array_equal(ids, t_ids)
        np.testing.assert_array_equal(dists, t_d)
        g_ids, g_d = exact_search(data, queries, 10, metric)
        np.testing.assert_array_equal(g_ids, t_ids)
        np.testing.assert_array_equal(g_d, t_d)


def test_complete_graph_filtered_is_exact_over_allowed() -> None:
    n = 60
    data = _small_ints(n, 4, 12, 5)
    queries = _small_ints(20, 4, 13, 5)
    index = HNSWIndex(4, M=n, ef_construction=n, selection="simple", seed=1)
    index.add(data)
    allowed = np.random.default_rng(0).random(n) < 0.4
    subset = np.flatnonzero(allowed)
    print("case complete graph filtered", flush=True)
    ids, dists = index.batch_search(queries, k=8, ef=n, filter=allowed)
    t_ids, t_d = oracle.exact_search(data[subset], queries, 8)
    np.testing.assert_array_equal(ids, subset[t_ids])
    np.testing.assert_array_equal(dists, t_d)


def test_unreachable_nodes_are_padded() -> None:
    index = HNSWIndex(2, M=2, seed=0)
    vectors = np.array([[0, 0], [1, 0], [5, 5]], dtype=np.float32)
    links0 = np.full((3, 5), -1, np.int32)
    links0[0, 0], links0[1, 0], links0[2, 0] = 1, 0, 0
    arrays = {
        "vectors": vectors, "levels": np.zeros(3, np.int64), "links0": links0,
        "count0": np.ones(3, np.int32), "upper_offset": np.full(3, -1, np.int64),
        "links_up": np.zeros((0, 3), np.int32), "count_up": np.zeros(0, np.int32),
    }
    index._engine.attach_native(arrays, 0, 0)
    print("case unreachable node", flush=True)
    ids, dist = index.search(np.array([5.0, 5.0]), k=3, ef=10)
    np.testing.assert_array_equal(ids, [1, 0, -1])
    np.testing.assert_array_equal(dist, [41.0, 50.0, np.inf])


@pytest.mark.parametrize("dim", [6, 128])
def test_all_true_filter_equals_unfiltered_search(dim: int) -> None:
    data, queries = clustered_u8(3000, dim, 60, n_clusters=12, seed=1)
    index = HNSWIndex(dim, M=8 if dim < 128 else 16, ef_construction=40, seed=0)
    index.add(data)
    for ef in (30, 200):
        print(f"case all-true filter dim={dim} ef={ef}", flush=True)
        index.reset_counters()
        plain = index.batch_search(queries, k=10, ef=ef)
        counters = (index.distance_computations, index.hops)
        index.reset_counters()
        filtered = index.batch_search(queries, k=10, ef=ef, filter=np.ones(3000, dtype=bool))
        for a, b in zip(plain, filtered):
            np.testing.assert_array_equal(a, b)
        assert (index.distance_computations, index.hops) == counters


def test_all_false_filter_returns_padding_without_searching() -> None:
    data, queries = clustered_u8(500, 8, 4, n_clusters=5, seed=2)
    index = HNSWIndex(8, M=4, seed=0)
    index.add(data)
    index.reset_counters()
    nothing = np.zeros(500, dtype=bool)
    print("case all-false filter", flush=True)
    ids, dist = index.batch_search(queries, k=3, filter=nothing)
    assert (ids == -1).all() and np.isinf(dist).all()
    one, _ = index.search(queries[0], k=2, filter=nothing)
    assert (one == -1).all()
    # per-query masks: the all-false row pads, the other rows search
    pool = np.stack([nothing, np.ones(500, dtype=bool)])
    m_ids, m_d = index.batch_search_mixed(queries, 3, 20, filters=pool, filter_index=[0, 1, -1, 0])
    assert (m_ids[[0, 3]] == -1).all() and np.isinf(m_d[[0, 3]]).all()
    assert (m_ids[[1, 2]] >= 0).all()
    assert index.distance_computations > 0  # only the two searched rows count
    index.reset_counters()
    index.batch_search(queries, k=3, filter=nothing)
    assert (index.distance_computations, index.hops) == (0, 0)


def test_fewer_allowed_nodes_than_k_pads() -> None:
    data, _ = clustered_u8(300, 4, 1, n_clusters=4, seed=6)
    index = HNSWIndex(4, M=6, seed=0)
    index.add(data)
    allowed = np.zeros(300, dtype=bool)
    allowed[[7, 70, 170]] = True
    print("case fewer allowed than k", flush=True)
    ids, dist = index.search(data[70], k=5, ef=10, filter=allowed)
    assert ids[0] == 70 and dist[0] == 0
    assert sorted(ids[:3].tolist()) == [7, 70, 170]
    np.testing.assert_array_equal(ids[3:], [-1, -1])
    assert np.isinf(dist[3:]).all()


@st.composite
def _datasets(draw):
    n = draw(st.integers(1, 40))
    dim = draw(st.integers(1, 5))
    values = draw(st.lists(st.integers(0, 3), min_size=n * dim, max_size=n * dim))
    return np.array(values, dtype=np.float32).reshape(n, dim)


@st.composite
def _params(draw):
    selection = draw(st.sampled_from(["simple", "heuristic"]))
    heuristic = selection == "heuristic"
    return {
        "M": draw(st.integers(2, 6)),
        "ef_construction": draw(st.integers(1, 16)),
        "selection": selection,
        "extend_candidates": heuristic and draw(st.booleans()),
        "keep_pruned_connections": heuristic and draw(st.booleans()),
        "seed": draw(st.integers(0, 2**16)),
    }


@settings(max_examples=30)
@given(data=_datasets(), params=_params(), metric=st.sampled_from(["l2", "ip"]),
       k=st.integers(1, 40), ef=st.integers(1, 50), mask_seed=st.integers(0, 2**16))
def test_filtered_results_are_valid(data, params, metric, k, ef, mask_seed) -> None:
    n = data.shape[0]
    k = min(k, n)
    allowed = np.random.default_rng(mask_seed).random(n) < 0.5
    index = HNSWIndex(data.shape[1], metric=metric, **params)
    index.add(data)
    ids, dist = index.search(data[0], k=k, ef=ef, filter=allowed)
    found = ids >= 0
    assert found[: found.sum()].all()
    real = ids[found]
    assert allowed[real].all()
    assert len(set(real.tolist())) == real.size <= allowed.sum()
    assert (np.diff(dist[found]) >= 0).all()
    np.testing.assert_array_equal(dist[found],
                                  exact_distances(data, data[:1], real[None, :], metric)[0])
    assert (dist[~found] == np.inf).all()
    want = oracle.search_batch(index._engine.native_arrays(), index.entry_point,
                               index.max_level, metric, data[:1], k, max(ef, k), mask=allowed)
    np.testing.assert_array_equal(ids, want[0][0])
    np.testing.assert_array_equal(dist, want[1][0])


@settings(max_examples=30)
@given(data=_datasets(), metric=st.sampled_from(["l2", "ip"]), mask_seed=st.integers(0, 2**16),
       filtered=st.booleans())
def test_complete_graph_search_equals_brute_force(data, metric, mask_seed, filtered) -> None:
    n = data.shape[0]
    allowed = np.random.default_rng(mask_seed).random(n) < 0.5
    subset = np.flatnonzero(allowed) if filtered else np.arange(n)
    index = HNSWIndex(data.shape[1], metric=metric, M=max(n, 2), ef_construction=n,
                      selection="simple", seed=mask_seed)
    index.add(data)
    k = min(5, n)
    ids, dist = index.batch_search(data[:3], k=k, ef=n, filter=allowed if filtered else None)
    if subset.size == 0:
        assert (ids == -1).all()
        return
    kk = min(k, subset.size)
    t_ids, t_d = oracle.exact_search(data[subset], data[:3], kk, metric)
    np.testing.assert_array_equal(ids[:, :kk], subset[t_ids])
    np.testing.assert_array_equal(dist[:, :kk], t_d)
    assert (ids[:, kk:] == -1).all()


def test_native_round_trip_searches_identically(tmp_path) -> None:
    """A saved index reopened from read-only memory maps searches exactly as the original
    (``tests/test_numba_engine.py:237-253``), with and without a filter."""
    data, queries = clustered_u8(2500, 8, 40, n_clusters=10, seed=1)
    index = HNSWIndex(8, M=6, ef_construction=40, seed=3)
    index.add(data)
    index.save_native(tmp_path / "idx")
    opened = HNSWIndex.load_native(tmp_path / "idx")
    assert not opened._engine.native_arrays()["links0"].flags.writeable
    allowed = np.random.default_rng(0).random(2500) < 0.2
    for mask in (None, allowed):
        print(f"case native round trip masked={mask is not None}", flush=True)
        for a, b in zip(index.batch_search(queries, 10, ef=40, filter=mask),
                        opened.batch_search(queries, 10, ef=40, filter=mask)):
            np.testing.assert_array_equal(a, b)


def test_ef_below_k_is_raised_to_k() -> None:
    """``ef = max(ef, k)`` (``index.py:294-299``): a search with ef < k is the search with
    ef = k, counters included."""
    data, queries = clustered_u8(3000, 16, 50, n_clusters=10, seed=5)
    index = HNSWIndex(16, M=8, ef_construction=60, seed=2)
    index.add(data)
    print("case ef below k", flush=True)
    index.reset_counters()
    low = index.batch_search(queries, k=30, ef=5)
    c_low = (index.distance_computations, index.hops)
    index.reset_counters()
    same = index.batch_search(queries, k=30, ef=30)
    assert c_low == (index.distance_computations, index.hops)
    for a, b in zip(low, same):
        np.testing.assert_array_equal(a, b)
    want = oracle.search_batch(index._engine.native_arrays(), index.entry_point, index.max_level,
                               "l2", queries, 30, 30)
    np.testing.assert_array_equal(low[0], want[0])
