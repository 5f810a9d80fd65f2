"""The float32 matrix-core range search of the exact oracle (exact_ftile.hip) against the C oracle.

Float data with at least 8192 (allowed) rows, dim <= 256 and k <= 128 takes this path: a float32
shortlist whose threshold carries a derived rounding margin, then float64 distances of the
shortlist summed left to right. Bar: ids and float64 distances bit for bit against the C oracle
(``oracle.exact_search``, every row in float64) and against the float64 chunked kernels run on
the same input. Cases: every metric and k in {1, 10, 100}; padded widths from 8 to 256 floats
(dims not a multiple of 4 take the padded copy); masks before and after compaction; data far
from the origin (a wide margin); tie-heavy data whose records overflow, so that flagged queries
re-run on the chunked kernels.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from hnsw_amd import _native, exact_search
from hnsw_amd.exact import last_range_search
from hnsw_amd.synthetic import angular_heavy_tailed

pytestmark = pytest.mark.gpu


def _chunked(data, queries, k, metric, allowed=None):
    """The same search on the float64 chunked kernels only."""
    _native.check(_native.lib().hm_set_exact_tuning(-1, 0))
    try:
        return exact_search(data, queries, k, metric, allowed=allowed)
    finally:
        _native.check(_native.lib().hm_set_exact_tuning(0, 0))


def _check(data, queries, k, metric="l2", allowed=None, requeried=False, vs_chunked=False,
           tile=True):
    ids, d = exact_search(data, queries, k, metric, allowed=allowed)
    info = last_range_search()
    print(info, flush=True)
    if not tile:  # too few allowed rows for the sampled bound: the chunked kernels ran
        assert info["range_ms"] == 0
    else:
        assert info["range_ms"] > 0, "the float range search did not run"
    if requeried:
        assert info["requeried"] > 0, "no query was re-run on the chunked kernels"
    t_ids, t_d = oracle.exact_search(data, queries, k, metric, allowed=allowed, threads=8)
    np.testing.assert_array_equal(ids, t_ids)
    np.testing.assert_array_equal(d, t_d)
    if vs_chunked:
        c_ids, c_d = _chunked(data, queries, k, metric, allowed)
        assert last_range_search()["range_ms"] == 0
        np.testing.assert_array_equal(ids, c_ids)
        np.testing.assert_array_equal(d, c_d)
    return ids, d


@pytest.mark.parametrize("metric", ["l2", "ip", "cosine"])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_angular_metrics_and_k(metric: str, k: int) -> None:
    data, queries = angular_heavy_tailed(60_000, 100, 400, n_clusters=60, seed=21)
    print(f"case ftile angular {metric} k={k}", flush=True)
    _check(data, queries, k, metric, vs_chunked=(k == 10))


@pytest.mark.parametrize("dim,nq", [(4, 200), (30, 300), (64, 700), (128, 1100), (200, 300),
                                    (256, 150)])
def test_padded_widths(dim: int, nq: int) -> None:
    rng = np.random.default_rng(dim)
    centres = rng.standard_normal((50, dim)).astype(np.float32)
    lab = rng.integers(0, 50, size=30_000)
    data = (centres[lab] + 0.3 * rng.standard_normal((30_000, dim))).astype(np.float32)
    queries = (centres[rng.integers(0, 50, size=nq)]
               + 0.3 * rng.standard_normal((nq, dim))).astype(np.float32)
    print(f"case ftile dim={dim} nq={nq}", flush=True)
    _check(data, queries, 10, "l2")


@pytest.mark.parametrize("sel", [0.5, 0.2, 0.1])
@pytest.mark.parametrize("kind", ["random", "range"])
def test_masked(sel: float, kind: str) -> None:
    n = 100_000
    data, queries = angular_heavy_tailed(n, 48, 300, n_clusters=40, seed=22)
    if kind == "random":
        allowed = np.random.default_rng(23).random(n) < sel
    else:
        allowed = np.zeros(n, dtype=bool)
        allowed[int(n * 0.4): int(n * 0.4) + int(n * sel)] = True
    # 50,000 allowed rows make 12 chunks; 10,000 rows are too few for the sampled bound (fewer
    # than 2k block scores) and run on the chunked kernels.
    print(f"case ftile masked {kind} sel={sel}", flush=True)
    ids, _ = _check(data, queries, 10, "cosine", allowed=allowed, tile=allowed.sum() >= 16384)
    assert np.all(allowed[ids[ids >= 0]])


def test_far_from_origin() -> None:
    """Rows near (1000, ..., 1000): |x|^2 dwarfs the distances, so the margin is wide; the
    threshold still admits every true neighbour."""
    rng = np.random.default_rng(24)
    data = (1000.0 + rng.standard_normal((20_000, 40))).astype(np.float32)
    queries = (1000.0 + rng.standard_normal((100, 40))).astype(np.float32)
    print("case ftile far", flush=True)
    _check(data, queries, 10, "l2")
    _check(data, queries, 10, "ip")


def test_tie_overflow_reruns_on_chunked_kernels() -> None:
    rng = np.random.default_rng(25)
    data = (rng.integers(0, 3, size=(40_000, 8)) * 0.5).astype(np.float32)
    queries = (rng.integers(0, 3, size=(200, 8)) * 0.5).astype(np.float32)
    print("case ftile ties", flush=True)
    _check(data, queries, 20, "l2", requeried=True, vs_chunked=True)
