"""Parity at the headline workload's full size: 1,000,000 x 128 clustered 8-bit vectors, the graph
built by sequential insertion and cached under .graph_cache/ (``python -m hnsw_amd.workloads
clustered_1m`` builds it in about ten minutes on the host; without it these tests fail rather than
skip, so a run without the cache cannot pass silently).

* search: 200 of the 10,000 queries, k = 10, every ef in {16, 32, 64, 128, 256, 512}: ids,
  float64 distances and the per-query distance / hop counters bit-identical to the C oracle;
* per-query arguments: 200 queries with ef drawn from 16..512, k from {1, 10, 100}, and half of
  them masked by one of four i.i.d. masks of 5-50 % (the filtered 128-dimension kernel), in one
  batch; ids, distances and counters bit-identical to the oracle;
* the mixed batch of the secondary measurement (100,000 queries, ef 16-512, k 1/10/100, half
  masked at 5-50 %) run as one batch; 2,000 sampled queries equal the oracle's results for
  their own arguments, counters included;
* fractional queries (float64 distances): all 10,000 queries moved by uniform offsets in
  (-0.5, 0.5), ef 16 and 64, in one batch (batched float64 descents, then layer 0 longest
  first) equal to the same batch in one launch, and 200 of them equal to the oracle;
* exact oracle: 64 queries over all 1M rows, k = 10, ids and distances bit-identical to the C
  oracle's exact top-k; all 10,000 queries: results sorted by (distance, id), and every
  returned distance equal to an int64 recomputation of that row's distance.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from hnsw_amd.bench_configs import mixed_batch
from hnsw_amd.exact import exact_search
from hnsw_amd.synthetic import random_mask
from hnsw_amd.workloads import WORKLOADS, is_cached, load_or_build

pytestmark = pytest.mark.gpu

W = WORKLOADS["clustered_1m"]


@pytest.fixture(scope="module")
def full():
    if not is_cached(W):
        pytest.fail("the 1M graph is not cached: run python -m hnsw_amd.workloads clustered_1m")
    return load_or_build(W, log=lambda m: None)


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


@pytest.mark.parametrize("ef", [16, 64])
def test_fractional_queries_at_full_size(full, ef: int) -> None:
    from hnsw_amd import _native

    eng, data, queries = full
    rng = np.random.default_rng(71)
    q = (queries + rng.uniform(-0.5, 0.5, queries.shape)).astype(np.float32)
    print(f"case 1M fractional ef={ef}", flush=True)
    got = eng.search_many(q, 10, ef)
    lib = _native.lib()
    try:
        _native.check(lib.hm_set_tuning(eng._device(), 0, 0, 0, 0, 1, 0))
        one = eng.search_many(q, 10, ef)
    finally:
        _native.check(lib.hm_set_tuning(eng._device(), 0, 0, 0, 0, 0, 0))
    for g, o, name in zip(got, one, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, o, err_msg=f"ef={ef} one launch {name}")
    pick = rng.choice(q.shape[0], size=200, replace=False)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               q[pick], 10, ef, threads=8)
    for g, w_, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g[pick], w_, err_msg=f"ef={ef} {name}")


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


def _same(got, want, msg=""):
    for g, w_, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w_, err_msg=f"{msg} {name}")


def test_per_query_arguments_and_masks_at_full_size(full) -> None:
    eng, data, queries = full
    rng = np.random.default_rng(31)
    q = queries[1000:1200]
    nq = q.shape[0]
    k = rng.choice([1, 10, 100], size=nq)
    ef = np.maximum(rng.integers(16, 513, size=nq), k)
    pool = np.stack([random_mask(W.n, s, 40 + i) for i, s in enumerate((0.05, 0.1, 0.25, 0.5))])
    midx = np.where(rng.random(nq) < 0.5, rng.integers(0, 4, size=nq), -1)
    print("case 1M per-query ef/k/masks", flush=True)
    got = eng.search_many(q, k, ef, mask_pool=pool, mask_index=midx)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2", q, k,
                               ef, mask_pool=pool, mask_index=midx,
                               threads=oracle.threads_available())
    _same(got, want, "per-query")
    masked = midx >= 0
    hit = got[0] >= 0
    rows, cols = np.nonzero(hit & masked[:, None])
    assert pool[midx[rows], got[0][rows, cols]].all()


def test_mixed_batch_sample_matches_oracle(full) -> None:
    eng, data, _ = full
    queries, k, ef, pool, midx = mixed_batch(W)
    print("case 1M mixed 100K batch", flush=True)
    got = eng.search_many(queries, k, ef, mask_pool=pool, mask_index=midx)
    pick = np.random.default_rng(8).choice(queries.shape[0], 2000, replace=False)
    print("case 1M mixed oracle sample", flush=True)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries[pick], k[pick], ef[pick], mask_pool=pool,
                               mask_index=midx[pick], threads=oracle.threads_available())
    kmax = want[0].shape[1]
    _same((got[0][pick, :kmax], got[1][pick, :kmax], got[2][pick], got[3][pick]), want, "mixed")
    assert (got[0][pick, kmax:] == -1).all()
