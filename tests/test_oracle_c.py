"""Pins the CPU oracle (oracle/hnsw_oracle.c) before anything is checked against it.

Which check fixes which oracle function:

* ``search_batch`` (plain, filtered) -- the committed fixtures written from the reference's
  engine (``tests/golden/make_golden.py``): ids, distances and per-query distance/hop counters
  bit for bit on integer data; on continuous data ids exactly and distances to 1e-12, the
  reference's own cross-engine bar (``tests/test_numba_engine.py:52-63``). Special cases fixed by
  mathematics: a complete graph searched with ``ef = n`` is brute force
  (``tests/test_index.py:157-170, 458-472``); the hand-built disconnected graph pads
  (``tests/test_index.py:341-352``); an all-true mask equals the plain search including counters
  (``tests/test_index.py:427-440``).
* ``search_batch`` over uint8 codes -- on lossless codes it equals the float32 search exactly
  (``tests/test_quantization.py:69-82``).
* ``rerank_batch`` -- re-ranking every id of the dataset is brute force; distances equal the
  oracle's exact distances.
* ``exact_search`` -- the reference's hand-checkable cases (``tests/test_oracle.py:9-46,
  109-116``); the fixtures from ``oracle.py`` (integer data bit for bit; float data ids and 1e-12,
  cosine up to last-bit near-ties); a brute-force pure-Python k-NN on tie-heavy integer data.

On continuous data the reference's float64 bits depend on numpy's reduction order (its
``einsum``) or on Numba's code generation; those bits stay unpinned and the tolerance is stated
in each test.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import golden_io
import oracle

INT_FIXTURES = ["search_u8_l2", "search_ties_l2", "search_ties_ip"]
FLOAT_FIXTURES = ["search_float_l2", "search_float_cosine"]


@pytest.mark.parametrize("name", INT_FIXTURES + FLOAT_FIXTURES)
def test_search_matches_reference_engine_fixtures(name: str) -> None:
    fx, arrays, entry, top, metric, _ = golden_io.graph(name)
    queries = golden_io.prepared_queries(fx)
    exact = name in INT_FIXTURES
    for tag in golden_io.tags(fx):
        k, ef = (int(v) for v in fx[f"{tag}_k_ef"])
        mask = fx.get(f"{tag}_mask")
        ids, dists, dc, hops = oracle.search_batch(
            arrays, entry, top, metric, queries, k, ef, mask=mask, threads=2
        )
        np.testing.assert_array_equal(ids, fx[f"{tag}_ids"], err_msg=f"{name}/{tag}")
        np.testing.assert_array_equal(dc, fx[f"{tag}_dc"], err_msg=f"{name}/{tag}")
        np.testing.assert_array_equal(hops, fx[f"{tag}_hops"], err_msg=f"{name}/{tag}")
        if exact:
            np.testing.assert_array_equal(dists, fx[f"{tag}_dists"], err_msg=f"{name}/{tag}")
        else:
            np.testing.assert_allclose(dists, fx[f"{tag}_dists"], rtol=0, atol=1e-12)


def _complete_graph(data: np.ndarray) -> tuple[dict, int, int]:
    n = data.shape[0]
    links0 = np.full((n, 2 * n + 1), -1, np.int32)
    for i in range(n):
        others = [j for j in range(n) if j != i]
        links0[i, : n - 1] = others
    arrays = {
        "vectors": data.astype(np.float32),
        "levels": np.zeros(n, np.int64),
        "links0": links0,
        "count0": np.full(n, n - 1, np.int32),
        "upper_offset": np.full(n, -1, np.int64),
        "links_up": np.zeros((0, n + 1), np.int32),
        "count_up": np.zeros(0, np.int32),
    }
    return arrays, 0, 0


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_complete_graph_with_full_ef_is_brute_force(metric: str) -> None:
    rng = np.random.default_rng(8)
    data = rng.integers(0, 5, size=(60, 4)).astype(np.float32)
    queries = rng.integers(0, 5, size=(25, 4)).astype(np.float32)
    arrays, entry, top = _complete_graph(data)
    ids, dists, _, _ = oracle.search_batch(arrays, entry, top, metric, queries, 10, 60)
    t_ids, t_d = _pure_python_knn(data, queries, 10, metric)
    np.testing.assert_array_equal(ids, t_ids)
    np.testing.assert_array_equal(dists, t_d)
    allowed = rng.random(60) < 0.4
    ids, dists, _, _ = oracle.search_batch(arrays, entry, top, metric, queries, 8, 60, mask=allowed)
    sub = np.flatnonzero(allowed)
    s_ids, s_d = _pure_python_knn(data[sub], queries, 8, metric)
    np.testing.assert_array_equal(ids, sub[s_ids])
    np.testing.assert_array_equal(dists, s_d)


def test_disconnected_hand_graph_pads() -> None:
    data = np.array([[0, 0], [1, 0], [5, 5]], dtype=np.float32)
    arrays = {
        "vectors": data,
        "levels": np.zeros(3, np.int64),
        "links0": np.array([[1, -1, -1, -1, -1], [0, -1, -1, -1, -1], [0, -1, -1, -1, -1]], np.int32),
        "count0": np.array([1, 1, 1], np.int32),
        "upper_offset": np.full(3, -1, np.int64),
        "links_up": np.zeros((0, 3), np.int32),
        "count_up": np.zeros(0, np.int32),
    }
    ids, dists, dc, hops = oracle.search_batch(arrays, 0, 0, "l2", np.array([[5.0, 5.0]]), 3, 10)
    np.testing.assert_array_equal(ids, [[1, 0, -1]])
    assert dists[0, 2] == np.inf
    assert dists[0, 0] == 16 + 25 and dists[0, 1] == 50
    assert dc[0] == 2 and hops[0] == 2


def test_all_true_mask_equals_plain_search_including_counters() -> None:
    fx, arrays, entry, top, metric, _ = golden_io.graph("search_u8_l2")
    q = golden_io.prepared_queries(fx)
    plain = oracle.search_batch(arrays, entry, top, metric, q, 10, 40)
    masked = oracle.search_batch(arrays, entry, top, metric, q, 10, 40,
                                 mask=np.ones(arrays["levels"].shape[0], bool))
    for a, b in zip(plain, masked):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_lossless_codes_search_equals_float_search(metric: str) -> None:
    name = "search_ties_l2" if metric == "l2" else "search_ties_ip"
    fx, arrays, entry, top, _, _ = golden_io.graph(name)
    q = golden_io.prepared_queries(fx)
    codes = arrays["vectors"].astype(np.uint8)  # values 0..4, offset 0, step 1
    a = oracle.search_batch(arrays, entry, top, metric, q, 10, 40)
    b = oracle.search_batch(arrays, entry, top, metric, q.astype(np.float64), 10, 40, codes=codes)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_rerank_of_every_id_is_brute_force() -> None:
    rng = np.random.default_rng(3)
    data = rng.integers(0, 7, size=(90, 5)).astype(np.float32)
    q = rng.integers(0, 7, size=(12, 5)).astype(np.float32)
    cand = np.tile(rng.permutation(90), (12, 1)).astype(np.int64)
    cand[:, ::7] = -1  # skipped slots
    kept = np.sort(np.where(cand >= 0, cand, 10**9), axis=1)[:, : (cand >= 0).sum(1).min()]
    ids, d = oracle.rerank_batch(data, "l2", q, cand, 9)
    for r in range(12):
        allowed = np.zeros(90, bool)
        allowed[kept[r]] = True
        sub = np.flatnonzero(allowed)
        t_ids, t_d = _pure_python_knn(data[sub], q[r : r + 1], 9, "l2")
        np.testing.assert_array_equal(ids[r], sub[t_ids[0]])
        np.testing.assert_array_equal(d[r], t_d[0])


def test_exact_search_hand_cases() -> None:
    data = np.array([[0, 0], [1, 0], [0, 2], [3, 3]], dtype=np.float32)
    ids, d = oracle.exact_search(data, np.array([[0.0, 0.0]]), 3)
    np.testing.assert_array_equal(ids, [[0, 1, 2]])
    np.testing.assert_array_equal(d, [[0.0, 1.0, 4.0]])
    data = np.array([[1, 0], [0, 3], [2, 2]], dtype=np.float32)
    ids, d = oracle.exact_search(data, np.array([[1.0, 1.0]]), 3, "ip")
    np.testing.assert_array_equal(ids, [[2, 1, 0]])
    np.testing.assert_array_equal(d, [[-4.0, -3.0, -1.0]])
    ids, d = oracle.exact_search(data, np.array([[1.0, 1.0]]), 3, "cosine")
    np.testing.assert_array_equal(ids, [[2, 0, 1]])
    np.testing.assert_allclose(d[0], [0.0, 1 - 2**-0.5, 1 - 2**-0.5], atol=1e-12)
    data = np.array([[9, 9], [0, -1], [1, 0], [-1, 0], [0, 1]], dtype=np.float32)
    ids, _ = oracle.exact_search(data, np.zeros((1, 2)), 4)
    np.testing.assert_array_equal(ids, [[1, 2, 3, 4]])
    data = np.array([[2, 0], [0, 0], [-2, 0], [0, 2]], dtype=np.float32)
    ids, _ = oracle.exact_search(data, np.zeros((1, 2)), 2)
    np.testing.assert_array_equal(ids, [[1, 0]])
    ids, d = oracle.exact_search(np.array([[0, 0], [3, 4], [1, 1]]), np.array([[0, 0]]), 3)
    np.testing.assert_array_equal(ids, [[0, 2, 1]])
    np.testing.assert_array_equal(d, [[0.0, 2.0, 25.0]])


def test_exact_search_matches_reference_oracle_fixture() -> None:
    fx = golden_io.load("oracle_exact")
    for metric in ("l2", "ip"):
        ids, d = oracle.exact_search(fx["int_data"], fx["int_queries"], 12, metric)
        np.testing.assert_array_equal(ids, fx[f"int_{metric}_ids"])
        np.testing.assert_array_equal(d, fx[f"int_{metric}_dists"])
    for metric in ("l2", "ip", "cosine"):
        ids, d = oracle.exact_search(fx["float_data"], fx["float_queries"], 10, metric)
        np.testing.assert_allclose(d, fx[f"float_{metric}_dists"], rtol=0, atol=1e-12)
        np.testing.assert_array_equal(ids, fx[f"float_{metric}_ids"])


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_exact_search_matches_pure_python_on_tie_heavy_data(metric: str) -> None:
    rng = np.random.default_rng(7)
    data = rng.integers(1, 5, size=(80, 4)).astype(np.float32)
    q = rng.integers(1, 5, size=(15, 4)).astype(np.float32)
    ids, d = oracle.exact_search(data, q, 10, metric)
    t_ids, t_d = _pure_python_knn(data, q, 10, metric)
    np.testing.assert_array_equal(ids, t_ids)
    np.testing.assert_array_equal(d, t_d)
    allowed = rng.random(80) < 0.1
    ids, d = oracle.exact_search(data, q, 10, metric, allowed=allowed)
    sub = np.flatnonzero(allowed)
    kk = min(10, sub.size)
    t_ids, t_d = _pure_python_knn(data[sub], q, kk, metric)
    np.testing.assert_array_equal(ids[:, :kk], sub[t_ids])
    assert (ids[:, kk:] == -1).all() and np.isinf(d[:, kk:]).all()


def _pure_python_knn(data, queries, k, metric):
    """Textbook k-NN in Python floats (no numpy arithmetic), ties to the lower id."""
    out_ids, out_d = [], []
    rows = data.astype(np.float64).tolist()
    for q in queries.astype(np.float64).tolist():
        scored = []
        for i, row in enumerate(rows):
            if metric == "l2":
                s = math.fsum((a - b) ** 2 for a, b in zip(row, q))
            else:
                s = -math.fsum(a * b for a, b in zip(row, q))
            scored.append((s, i))
        scored.sort()
        out_ids.append([i for _, i in scored[:k]])
        out_d.append([s for s, _ in scored[:k]])
    return np.array(out_ids, np.int64), np.array(out_d, np.float64)
