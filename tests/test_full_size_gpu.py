"""Parity at the headline workload's full size: 1,000,000 x 128 clustered 8-bit vectors, the graph
built by sequential insertion (cached under .graph_cache/ by the bench; the test is skipped when
the cache is absent, since building it takes minutes).

* search: 200 of the 10,000 queries, k = 10, every ef in {16, 32, 64, 128, 256, 512}: ids,
  float64 distances and the per-query distance / hop counters bit-identical to the C oracle;
* exact oracle: 64 queries over all 1M rows, k = 10, ids and distances bit-identical to the C
  oracle's exact top-k; all 10,000 queries: results sorted by (distance, id), and every
  returned distance equal to an int64 recomputation of that row's distance.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

import oracle
from hnsw_amd.exact import exact_search
from hnsw_amd.workloads import CACHE, WORKLOADS, load_or_build

pytestmark = pytest.mark.gpu

W = WORKLOADS["clustered_1m"]


@pytest.fixture(scope="module")
def full():
    header = CACHE / W.name / "header.json"
    if not header.is_file():
        pytest.skip("the 1M graph is not cached (bench.py builds it)")
    eng, data, queries = load_or_build(W, log=lambda m: None)
    if json.loads(header.read_text()).get("workload", {}).get("n") != W.n:
        pytest.skip("cache holds another workload")
    return eng, data, queries


@pytest.mark.parametrize("ef", [16, 32, 64, 128, 256, 512])
def test_search_matches_oracle_at_full_size(full, ef: int) -> None:
    eng, data, queries = full
    q = queries[:200]
    print(f"case 1M search ef={ef}", flush=True)
    got = eng.search_many(q, 10, ef)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               q, 10, ef, threads=8)
    for g, w_, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w_, err_msg=f"ef={ef} {name}")


def test_exact_matches_oracle_at_full_size(full) -> None:
    eng, data, queries = full
    print("case 1M exact", flush=True)
    ids, d = exact_search(data, queries, 10)
    t_ids, t_d = oracle.exact_search(data, queries[:64], 10, threads=8)
    np.testing.assert_array_equal(ids[:64], t_ids)
    np.testing.assert_array_equal(d[:64], t_d)
    # size-independent properties over all queries
    order_ok = (d[:, 1:] > d[:, :-1]) | ((d[:, 1:] == d[:, :-1]) & (ids[:, 1:] > ids[:, :-1]))
    assert order_ok.all()
    rows = data[ids].astype(np.int64)
    diff = rows - queries.astype(np.int64)[:, None, :]
    np.testing.assert_array_equal(d, (diff * diff).sum(axis=2).astype(np.float64))
