"""GPU parity on the measured workloads' own shapes, against the C oracle.

* Clustered 100K (100,000 x 128 8-bit vectors, 100 clusters, M = 16, ef_construction = 200,
  1,000 queries): k = 10 at every ef in {16, 32, 64, 128, 256}; ids, float64 distances and the
  per-query distance / hop counters bit-identical (``numba_engine.py:422-503``).
* The quantized search on angular data (20,000 x 100 L2-normalised vectors with heavy-tailed
  per-dimension scales, cosine, 200 queries) at ef in {128, 256, 512}: the codes traversal
  (the general kernel over uint8 codes with float64 code-space queries,
  ``numba_engine.py:674-701``) bit-identical to the oracle's codes search, then with rerank = ef
  the exact re-rank (``:506-530``), and with rerank = 0 the code distances mapped back by the
  quantizer (pinned on its own by the reference's fixture in tests/test_host_rules.py). The
  code-space queries are computed here from the quantizer's offset and step, not by its
  ``transform_queries``.
* The same codes traversal with the l2 metric on continuous data (fractional code-space
  queries) at ef = 128.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from hnsw_amd import HNSWIndex, QuantizedIndex
from hnsw_amd.metrics import prepare
from hnsw_amd.synthetic import angular_heavy_tailed
from hnsw_amd.workloads import WORKLOADS, load_or_build

pytestmark = pytest.mark.gpu

THREADS = oracle.threads_available()


def _same(got, want, msg=""):
    for g, w, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w, err_msg=f"{msg} {name}")


@pytest.fixture(scope="module")
def clustered_100k():
    eng, data, queries = load_or_build(WORKLOADS["clustered_100k"], log=lambda m: None)
    return eng, data, queries


@pytest.mark.parametrize("ef", [16, 32, 64, 128, 256])
def test_clustered_100k_matches_oracle(clustered_100k, ef: int) -> None:
    eng, data, queries = clustered_100k
    assert queries.shape == (1000, 128) and data.shape == (100_000, 128)
    print(f"case clustered_100k ef={ef}", flush=True)
    got = eng.search_many(queries, 10, ef)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries, 10, ef, threads=THREADS)
    _same(got, want, f"ef={ef}")


@pytest.fixture(scope="module")
def angular():
    data, queries = angular_heavy_tailed(20_000, 100, 200, n_clusters=40, seed=61)
    index = HNSWIndex(100, "cosine", M=16, ef_construction=100, seed=62)
    index.add(data)
    return index, QuantizedIndex(index), queries, prepare(queries, "cosine")


@pytest.mark.parametrize("ef", [128, 256, 512])
def test_quantized_general_kernel_matches_oracle(angular, ef: int) -> None:
    index, qidx, queries, prepared = angular
    eng = index._engine
    arrays = eng.native_arrays()
    # cosine: the code-space query is the prepared query itself
    print(f"case quantized cosine ef={ef} codes", flush=True)
    ids, cd = eng.batch_search_codes(qidx.codes, prepared.astype(np.float64), ef, ef)
    want = oracle.search_batch(arrays, eng.entry_point, eng.max_level, "cosine",
                               prepared.astype(np.float64), ef, ef, codes=qidx.codes,
                               threads=THREADS)
    np.testing.assert_array_equal(ids, want[0])
    np.testing.assert_array_equal(cd, want[1])
    print(f"case quantized cosine ef={ef} rerank={ef}", flush=True)
    r_ids, r_d = qidx.batch_search(queries, 10, ef=ef, rerank=ef)
    w_ids, w_d = oracle.rerank_batch(arrays["vectors"], "cosine", prepared, want[0], 10,
                                     threads=THREADS)
    np.testing.assert_array_equal(r_ids, w_ids)
    np.testing.assert_array_equal(r_d, w_d)
    print(f"case quantized cosine ef={ef} rerank=0", flush=True)
    p_ids, p_d = qidx.batch_search(queries, 10, ef=ef, rerank=0)
    w10 = oracle.search_batch(arrays, eng.entry_point, eng.max_level, "cosine",
                              prepared.astype(np.float64), 10, ef, codes=qidx.codes,
                              threads=THREADS)
    np.testing.assert_array_equal(p_ids, w10[0])
    np.testing.assert_array_equal(p_d, qidx.quantizer.to_distance(w10[1], prepared, "cosine"))


def test_quantized_l2_fractional_queries_match_oracle() -> None:
    rng = np.random.default_rng(71)
    data = (rng.standard_normal((8000, 48)) * rng.lognormal(size=48)).astype(np.float32)
    queries = (rng.standard_normal((150, 48)) * rng.lognormal(size=48)).astype(np.float32)
    index = HNSWIndex(48, "l2", M=12, ef_construction=80, seed=72)
    index.add(data)
    qidx = QuantizedIndex(index)
    q = qidx.quantizer
    qcode = (queries.astype(np.float64) - q.offset) / q.step
    eng = index._engine
    print("case quantized l2 ef=128", flush=True)
    ids, cd = eng.batch_search_codes(qidx.codes, qcode, 40, 128)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2", qcode,
                               40, 128, codes=qidx.codes, threads=THREADS)
    np.testing.assert_array_equal(ids, want[0])
    np.testing.assert_array_equal(cd, want[1])
    r_ids, r_d = qidx.batch_search(queries, 10, ef=128, rerank=40)
    w_ids, w_d = oracle.rerank_batch(data, "l2", queries, want[0], 10, threads=THREADS)
    np.testing.assert_array_equal(r_ids, w_ids)
    np.testing.assert_array_equal(r_d, w_d)
