"""The int8 matrix-core range search of the exact oracle (exact_tile.hip) against the C oracle.

The range search runs for 8-bit integer data with at least 8192 (allowed) rows, dim <= 256 and
k <= 128. Bar: ids and float64 distances bit for bit. The cases cover every tile shape the
launcher picks (dim 16 / 40 / 128 / 200 -> 1, 2, 4, 8 K-steps of 32; 100, 300 and 3000 queries
-> 1, 2 and 4 query blocks per wave), both metrics, masks before and after compaction, and
tie-heavy data whose records overflow, so that the flagged queries are re-run on the chunked
kernels; and an index's own byte rows read in place (hm_index_exact / hm_index_exact_dev).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from hnsw_amd import _native, exact_search
from hnsw_amd.exact import last_range_search
from hnsw_amd.synthetic import clustered_u8

pytestmark = pytest.mark.gpu


def _check(data, queries, k, metric="l2", allowed=None, requeried=False, tile=True):
    ids, d = exact_search(data, queries, k, metric, allowed=allowed)
    info = last_range_search()
    print(info, flush=True)
    if not tile:  # below the size where the range search applies: the chunked kernels ran
        assert info["range_ms"] == 0
    else:
        assert info["range_ms"] > 0, "the range search did not run"
        if requeried:
            assert info["requeried"] > 0, "no query was re-run on the chunked kernels"
        else:  # an unlucky sample may overflow a record list: rare, and the re-run is exact
            assert info["requeried"] <= max(2, queries.shape[0] // 100)
    t_ids, t_d = oracle.exact_search(data, queries, k, metric, allowed=allowed, threads=8)
    np.testing.assert_array_equal(ids, t_ids)
    np.testing.assert_array_equal(d, t_d)
    return ids, d


@pytest.mark.parametrize("dim,nq", [(16, 100), (40, 300), (128, 3000), (200, 300)])
def test_tile_shapes_match_oracle(dim: int, nq: int) -> None:
    data, queries = clustered_u8(60_000, dim, nq, n_clusters=60, seed=dim)
    print(f"case tile dim={dim} nq={nq}", flush=True)
    _check(data, queries, 10)


@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("k", [1, 10, 64])
def test_tile_metrics_and_k_match_oracle(metric: str, k: int) -> None:
    data, queries = clustered_u8(100_000, 128, 600, n_clusters=100, seed=11)
    print(f"case tile {metric} k={k}", flush=True)
    _check(data, queries, k, metric)


@pytest.mark.parametrize("sel", [0.5, 0.1, 0.02])
@pytest.mark.parametrize("kind", ["random", "range"])
def test_tile_masked_matches_oracle(sel: float, kind: str) -> None:
    n = 100_000
    data, queries = clustered_u8(n, 64, 400, n_clusters=80, seed=12)
    if kind == "random":
        allowed = np.random.default_rng(13).random(n) < sel
    else:
        allowed = np.zeros(n, dtype=bool)
        start = int(n * 0.3)
        allowed[start : start + int(n * sel)] = True
    # 100,000 x 0.5 allowed rows make 12 chunks; 10,000 or 2,000 rows are too few for the
    # sampled bound (fewer than 2k block scores) and run on the chunked kernels.
    print(f"case tile masked {kind} sel={sel}", flush=True)
    ids, _ = _check(data, queries, 10, allowed=allowed, tile=allowed.sum() >= 16384)
    assert np.all(allowed[ids[ids >= 0]])


def test_tile_tie_overflow_reruns_on_chunked_kernels() -> None:
    rng = np.random.default_rng(14)
    data = rng.integers(0, 3, size=(40_000, 8)).astype(np.float32)
    queries = rng.integers(0, 3, size=(200, 8)).astype(np.float32)
    print("case tile ties", flush=True)
    _check(data, queries, 20, requeried=True)


def test_exact_dev_matches_host_entry_point() -> None:
    import torch

    data, queries = clustered_u8(50_000, 128, 700, n_clusters=50, seed=15)
    k = 10
    want_ids, want_d = exact_search(data, queries, k)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(data).to(dev)
    q = torch.from_numpy(queries).to(dev)
    ids = torch.empty((queries.shape[0], k), dtype=torch.int64, device=dev)
    d = torch.empty((queries.shape[0], k), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    print("case exact_dev", flush=True)
    _native.check(_native.lib().hm_exact_dev(
        ctypes.c_void_p(x.data_ptr()), data.shape[0], data.shape[1], 0,
        ctypes.c_void_p(q.data_ptr()), queries.shape[0], k, None, 1,
        ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(d.data_ptr()),
        ctypes.c_void_p(stream.cuda_stream)))
    np.testing.assert_array_equal(ids.cpu().numpy(), want_ids)
    np.testing.assert_array_equal(d.cpu().numpy(), want_d)


def _u8_engine(data, metric="l2"):
    """An index over the rows (graph degree kept small: only its exact search is used)."""
    from hnsw_amd.engine import MI355XEngine

    n = data.shape[0]
    levels = np.floor(-np.log(1.0 - np.random.default_rng(1).random(n)) / np.log(4)).astype(np.int64)
    eng = MI355XEngine(data.shape[1], metric, 4, 8, "heuristic", False, False)
    eng.insert(data, levels)
    return eng


@pytest.mark.parametrize("dim,metric", [(128, "l2"), (100, "l2"), (100, "ip"), (40, "l2")])
def test_index_rows_read_in_place_match_oracle(dim: int, metric: str) -> None:
    """The index keeps raw 0..255 bytes (rows padded to 16 bytes); the range search reads them
    in place (byte ^ 0x80 = the shifted value) with row terms computed at open. dim 100: the
    rows are 112 bytes and the tile reads 128, the last 16 as zeros."""
    data, queries = clustered_u8(40_000, dim, 500, n_clusters=40, seed=16 + dim)
    eng = _u8_engine(data, metric)
    for allowed in (None, np.random.default_rng(17).random(40_000) < 0.6):
        print(f"case index exact dim={dim} {metric} masked={allowed is not None}", flush=True)
        ids, d = eng.exact_search(queries, 10, allowed)
        info = last_range_search()
        print(info, flush=True)
        assert info["range_ms"] > 0
        t_ids, t_d = oracle.exact_search(data, queries, 10, metric, allowed=allowed, threads=8)
        np.testing.assert_array_equal(ids, t_ids)
        np.testing.assert_array_equal(d, t_d)


def test_index_exact_dev_matches_host_entry_point() -> None:
    import torch

    data, queries = clustered_u8(30_000, 128, 300, n_clusters=30, seed=18)
    eng = _u8_engine(data)
    k = 10
    want_ids, want_d = eng.exact_search(queries, k)
    dev = torch.device("cuda:0")
    q = torch.from_numpy(queries).to(dev)
    ids = torch.empty((queries.shape[0], k), dtype=torch.int64, device=dev)
    d = torch.empty((queries.shape[0], k), dtype=torch.float64, device=dev)
    print("case index_exact_dev", flush=True)
    for u8 in (1, 0):  # the int8 path, then the float path on the same index
        _native.check(_native.lib().hm_index_exact_dev(
            eng._device(), ctypes.c_void_p(q.data_ptr()), queries.shape[0], k, None, u8,
            ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(d.data_ptr())))
        np.testing.assert_array_equal(ids.cpu().numpy(), want_ids)
        np.testing.assert_array_equal(d.cpu().numpy(), want_d)
