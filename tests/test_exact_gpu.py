"""GPU exact search (hnsw_amd.exact_search -> hm_exact) against the oracle and the reference.

Integer data runs the int8 matrix-core path: ids and float64 distances must equal the oracle and
the reference oracle.py fixture bit for bit. Continuous data runs the float64 path, whose
left-to-right sums equal the C oracle's bit for bit; against oracle.py (numpy reduction order)
distances are held to 1e-12.
"""

from __future__ import annotations

import numpy as np
import pytest

import golden_io
import oracle
from hnsw_amd import exact_search
from hnsw_amd.synthetic import clustered_u8, sparse_prototypes_u8

pytestmark = pytest.mark.gpu


def test_reference_hand_cases() -> None:
    data = np.array([[0, 0], [1, 0], [0, 2], [3, 3]], dtype=np.float32)
    ids, d = exact_search(data, np.array([[0.0, 0.0]]), 3)
    np.testing.assert_array_equal(ids, [[0, 1, 2]])
    np.testing.assert_array_equal(d, [[0.0, 1.0, 4.0]])
    data = np.array([[1, 0], [0, 3], [2, 2]], dtype=np.float32)
    ids, d = exact_search(data, np.array([[1.0, 1.0]]), 3, "ip")
    np.testing.assert_array_equal(ids, [[2, 1, 0]])
    np.testing.assert_array_equal(d, [[-4.0, -3.0, -1.0]])
    ids, d = exact_search(data, np.array([[1.0, 1.0]]), 3, "cosine")
    np.testing.assert_array_equal(ids, [[2, 0, 1]])
    np.testing.assert_allclose(d[0], [0.0, 1 - 2**-0.5, 1 - 2**-0.5], atol=1e-12)
    data = np.array([[5, 5], [1, 1], [5, 5], [1, 1], [9, 9]], dtype=np.float32)
    ids, d = exact_search(data, np.array([[1.0, 1.0], [5.0, 5.0]]), 4)
    np.testing.assert_array_equal(ids, [[1, 3, 0, 2], [0, 2, 1, 3]])
    data = np.array([[2, 0], [0, 0], [-2, 0], [0, 2]], dtype=np.float32)
    ids, _ = exact_search(data, np.array([[0.0, 0.0]]), 2)
    np.testing.assert_array_equal(ids, [[1, 0]])


def test_reference_oracle_fixture() -> None:
    fx = golden_io.load("oracle_exact")
    for metric in ("l2", "ip"):
        ids, d = exact_search(fx["int_data"], fx["int_queries"], 12, metric)
        np.testing.assert_array_equal(ids, fx[f"int_{metric}_ids"])
        np.testing.assert_array_equal(d, fx[f"int_{metric}_dists"])
    for metric in ("l2", "ip", "cosine"):
        ids, d = exact_search(fx["float_data"], fx["float_queries"], 10, metric)
        np.testing.assert_array_equal(ids, fx[f"float_{metric}_ids"])
        np.testing.assert_allclose(d, fx[f"float_{metric}_dists"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_int8_path_matches_oracle(metric: str, k: int) -> None:
    data, queries = clustered_u8(30_000, 128, 200, n_clusters=30, seed=5)
    print(f"case int8 {metric} k={k}", flush=True)
    ids, d = exact_search(data, queries, k, metric)
    t_ids, t_d = oracle.exact_search(data, queries, k, metric, threads=8)
    np.testing.assert_array_equal(ids, t_ids)
    np.testing.assert_array_equal(d, t_d)


def test_int8_path_784d_masked_matches_oracle() -> None:
    data, queries = sparse_prototypes_u8(8_000, 64, seed=6)
    for sel in (0.5, 0.05, 0.001):
        allowed = np.random.default_rng(7).random(8000) < sel
        print(f"case int8 784 masked sel={sel}", flush=True)
        ids, d = exact_search(data, queries, 10, allowed=allowed)
        t_ids, t_d = oracle.exact_search(data, queries, 10, allowed=allowed, threads=8)
        np.testing.assert_array_equal(ids, t_ids)
        np.testing.assert_array_equal(d, t_d)


@pytest.mark.parametrize("metric", ["l2", "ip", "cosine"])
def test_float64_path_matches_oracle(metric: str) -> None:
    rng = np.random.default_rng(8)
    data = rng.standard_normal((20_000, 100)).astype(np.float32)
    queries = rng.standard_normal((100, 100)).astype(np.float32)
    print(f"case float64 {metric}", flush=True)
    ids, d = exact_search(data, queries, 10, metric)
    t_ids, t_d = oracle.exact_search(data, queries, 10, metric, threads=8)
    np.testing.assert_array_equal(ids, t_ids)
    np.testing.assert_array_equal(d, t_d)


def test_ties_and_duplicates_go_to_the_lower_id() -> None:
    rng = np.random.default_rng(9)
    data = rng.integers(0, 3, size=(5000, 4)).astype(np.float32)  # massive ties
    queries = rng.integers(0, 3, size=(50, 4)).astype(np.float32)
    ids, d = exact_search(data, queries, 40)
    t_ids, t_d = oracle.exact_search(data, queries, 40, threads=8)
    np.testing.assert_array_equal(ids, t_ids)
    np.testing.assert_array_equal(d, t_d)
