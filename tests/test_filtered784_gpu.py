"""Filtered search at 784 dimensions (the filtered workload's shape) against the C oracle.

10,000 x 784 8-bit vectors with about half exact zeros (10 prototypes plus uniform noise),
M = 16, ef_construction = 100; k = 10, ef = 128; i.i.d. and contiguous-range masks at 1 %, 5 %,
10 % and 50 %. Bar: the in-graph filtered search bit-identical to the oracle (ids, float64
distances, distance and hop counters), the masked brute force bit-identical to the oracle's
exact top-k over the allowed rows. At 784 dimensions partial sums exceed 2^24, so the exact
8-bit kernel's integer sums are what makes the float64 distances exact.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.synthetic import random_mask, range_mask, sparse_prototypes_u8

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def built784():
    data, queries = sparse_prototypes_u8(10_000, 120, seed=41)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(42).random(10_000)) / np.log(16))
    eng = MI355XEngine(784, "l2", 16, 100, "heuristic", False, False)
    eng.insert(data, levels.astype(np.int64))
    return eng, data, queries


@pytest.mark.parametrize("kind", ["random", "range"])
@pytest.mark.parametrize("sel", [0.01, 0.05, 0.1, 0.5])
def test_filtered_784_matches_oracle(built784, kind: str, sel: float) -> None:
    eng, data, queries = built784
    n = data.shape[0]
    mask = random_mask(n, sel, 43) if kind == "random" else range_mask(n, sel, 44)
    print(f"case 784 {kind} sel={sel}", flush=True)
    got = eng.search_many(queries, 10, 128, mask_pool=mask, mask_index=0)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries, 10, 128, mask=mask, threads=8)
    for g, w_, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w_, err_msg=name)
    kk = min(10, int(mask.sum()))
    e_ids, e_d = eng.exact_search(queries, kk, mask)
    t_ids, t_d = oracle.exact_search(data, queries, kk, allowed=mask, threads=8)
    np.testing.assert_array_equal(e_ids, t_ids)
    np.testing.assert_array_equal(e_d, t_d)
