"""QuantizedIndex on angular data (the quantized workload's kind) against the C oracle.

20,000 x 100 L2-normalised vectors from Gaussian clusters with log-normal per-dimension scales,
cosine, M = 16, ef_construction = 100; QuantizedIndex.batch_search(k = 10, ef = 64,
rerank = 64). Bar: the codes traversal's candidates and the re-ranked ids and float64 distances
bit-identical to the oracle's codes search (float64 code-space queries, left-to-right sums)
followed by its exact re-rank.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from hnsw_amd import HNSWIndex, QuantizedIndex
from hnsw_amd.metrics import prepare
from hnsw_amd.synthetic import angular_heavy_tailed

pytestmark = pytest.mark.gpu


def test_quantized_cosine_matches_oracle() -> None:
    data, queries = angular_heavy_tailed(20_000, 100, 150, n_clusters=40, seed=51)
    index = HNSWIndex(100, "cosine", M=16, ef_construction=100, seed=52)
    index.add(data)
    qidx = QuantizedIndex(index)
    print("case quantized cosine", flush=True)
    ids, dists = qidx.batch_search(queries, 10, ef=64, rerank=64)
    eng = index._engine
    arrays = eng.native_arrays()
    prepared = prepare(queries, "cosine")
    qcode = qidx.quantizer.transform_queries(prepared, "cosine")
    cand, _, _, _ = oracle.search_batch(arrays, eng.entry_point, eng.max_level, "cosine", qcode,
                                        64, 64, codes=qidx.codes, threads=8)
    want_ids, want_d = oracle.rerank_batch(arrays["vectors"], "cosine", prepared, cand, 10,
                                           threads=8)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dists, want_d)
