"""The general register beam's candidate overflow paths against the C oracle.

A filtered search whose results fill slowly holds many more candidates than its registers do
(the accepted nodes the filter rejects): they go to the per-wave spill in global memory and
come back by a radix select when the registers run dry. With a filter that allows almost
nothing, a query can hold more candidates than the first spill allows (16,384); it then runs
again, with other such queries, on room for every node. Each case is checked bit for bit
(ids, float64 distances, per-query distance / hop counters) against the oracle
(``numba_engine.py:151-198``), and the event counters show that the path ran.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from hnsw_amd import _native
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.synthetic import clustered_u8, random_mask

pytestmark = pytest.mark.gpu


def _stats(eng) -> list[int]:
    out = (ctypes.c_int32 * 4)()
    _native.check(_native.lib().hm_read_stats(eng._device(), out, 1))
    return [int(v) for v in out]


def _engine(n, dim, seed, M=12):
    data, queries = clustered_u8(n, dim, 80, n_clusters=30, sigma=30.0, seed=seed)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(seed + 1).random(n))
                      / np.log(M)).astype(np.int64)
    eng = MI355XEngine(dim, "l2", M, 60, "heuristic", False, False)
    eng.insert(data, levels)
    return eng, queries


def _check(eng, q, k, ef, mask, msg):
    got = eng.search_many(q, k, ef, mask_pool=mask, mask_index=0)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2", q, k,
                               ef, mask=mask, threads=oracle.threads_available())
    for g, w, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w, err_msg=f"{msg} {name}")
    return got


@pytest.mark.parametrize("ef", [40, 100])
def test_filtered_candidates_spill_and_refill(ef: int) -> None:
    eng, queries = _engine(30_000, 24, 3)
    _stats(eng)
    mask = random_mask(30_000, 0.01, 5)
    print(f"case beam spill ef={ef}", flush=True)
    got = _check(eng, queries, 10, ef, mask, f"ef={ef}")
    st = _stats(eng)
    assert st[2] > 0, st  # candidates went to the spill
    assert st[3] == 0, st  # and every query fitted it
    assert got[2].mean() > 5_000  # the filter made the searches long


def test_candidates_beyond_the_spill_run_again() -> None:
    eng, queries = _engine(80_000, 8, 7)
    _stats(eng)
    mask = np.zeros(80_000, dtype=bool)
    mask[np.random.default_rng(8).choice(80_000, 30, replace=False)] = True
    print("case beam spill overflow", flush=True)
    got = _check(eng, queries[:40], 10, 64, mask, "overflow")
    st = _stats(eng)
    assert st[3] > 0, st  # some queries held more than 16,384 candidates
    assert (got[0] >= 0).sum(axis=1).min() > 0
