"""GPU parity of the register-beam kernel's constant-shape build (M = 16, 128 one-byte
dimensions) and of every form of its visited set: 16-bit quotient slots, 32-bit whole-id slots
(ids with too many bits for the table), the move to the global bitset when the table fills or a
probe runs out of distance bits (forced with a probe limit of 1), and other table sizes. Every
variant must give the oracle's ids, distances and distance/hop counters bit for bit.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from hnsw_amd import _native, learned
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.synthetic import clustered_u8, random_mask

pytestmark = pytest.mark.gpu


def _same(got, want, msg=""):
    for g, w, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w, err_msg=f"{msg} {name}")


def _tune(eng, fast_hash_log2=0, probe_limit=0, waves_per_cu=0, one_launch=0,
          descent_cap=0) -> None:
    _native.check(_native.lib().hm_set_tuning(eng._device(), fast_hash_log2, 0, waves_per_cu,
                                              probe_limit, one_launch, descent_cap))


def _moves(eng) -> int:
    """Layer searches whose visited set moved to the global bitset since the last call."""
    out = (ctypes.c_int32 * 4)()
    _native.check(_native.lib().hm_read_stats(eng._device(), out, 1))
    return int(out[0])


def _build(n, dim, M, seed, metric="l2"):
    data, queries = clustered_u8(n, dim, 400, n_clusters=60, sigma=30.0, seed=seed)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(seed + 1).random(n))
                      / np.log(M)).astype(np.int64)
    eng = MI355XEngine(dim, metric, M, 100, "heuristic", False, False)
    eng.insert(data, levels)
    return eng, data, queries


@pytest.fixture(scope="module")
def m16d128():
    """20,000 x 128, M = 16: the constant-shape kernel; 15-bit ids, quotient slots."""
    return _build(20_000, 128, 16, 31)


@pytest.fixture(scope="module")
def m16d128_ip():
    """6,000 x 128, M = 16, inner product: the constant shape with the other exact metric."""
    return _build(6_000, 128, 16, 37, "ip")


@pytest.fixture(scope="module")
def m16d120():
    """6,000 x 120, M = 16: rows padded to 128 bytes, so the constant shape applies with zero
    padding in the last 8 bytes of every row and query."""
    return _build(6_000, 120, 16, 43)


@pytest.fixture(scope="module")
def wide_ids():
    """70,000 x 16, M = 8: 17-bit ids, so a 64-slot table needs 32-bit slots."""
    return _build(70_000, 16, 8, 41)


def _oracle(eng, queries, k, ef, metric="l2"):
    return oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, metric,
                               queries, k, ef, threads=8)


@pytest.mark.parametrize("ef,k", [(16, 10), (48, 10), (64, 64), (10, 1), (65, 10), (96, 10),
                                  (128, 100), (100, 128), (129, 10), (200, 100), (256, 256),
                                  (257, 10), (512, 100)])
def test_constant_shape_kernel_matches_oracle(m16d128, ef: int, k: int) -> None:
    """ef <= 64: one R entry per lane; 65..128: two; 129..256: four; 257..512: eight (the wide
    register beams)."""
    eng, _, queries = m16d128
    print(f"case m16d128 ef={ef} k={k}", flush=True)
    _same(eng.search_many(queries, k, ef), _oracle(eng, queries, k, ef), f"ef={ef}")


def test_wide_beam_mixed_batch_matches_oracle(m16d128) -> None:
    """Per-query ef in 16..128 and k in {1, 10, 100} in one launch of the wide beam."""
    eng, _, queries = m16d128
    rng = np.random.default_rng(17)
    nq = queries.shape[0]
    k = rng.choice([1, 10, 100], size=nq)
    ef = np.maximum(rng.choice([16, 40, 64, 80, 128], size=nq), k)
    print("case m16d128 mixed ef/k", flush=True)
    got = eng.search_many(queries, k, ef)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries, k, ef, threads=8)
    _same(got, want, "mixed")


@pytest.mark.parametrize("fast_hash_log2,probe_limit", [(0, 1), (0, 3), (6, 0), (8, 0), (12, 0)])
def test_visited_variants_match_oracle(m16d128, fast_hash_log2: int, probe_limit: int) -> None:
    eng, _, queries = m16d128
    want = _oracle(eng, queries, 10, 48)
    print(f"case m16d128 hash_log2={fast_hash_log2} probe_limit={probe_limit}", flush=True)
    _moves(eng)
    _tune(eng, fast_hash_log2, probe_limit)
    try:
        got = eng.search_many(queries, 10, 48)
        moves = _moves(eng)
    finally:
        _tune(eng)
    _same(got, want, f"hash_log2={fast_hash_log2} probe_limit={probe_limit}")
    if probe_limit == 1 or fast_hash_log2 == 6:
        assert moves > 0  # the bitset path ran


def test_four_entry_beam_mixed_batch_matches_oracle(m16d128) -> None:
    """Per-query ef in 16..256 and k in {1, 10, 100} in one launch of the four-entry beam."""
    eng, _, queries = m16d128
    rng = np.random.default_rng(19)
    nq = queries.shape[0]
    k = rng.choice([1, 10, 100], size=nq)
    ef = np.maximum(rng.choice([16, 64, 128, 192, 256], size=nq), k)
    print("case m16d128 mixed ef/k up to 256", flush=True)
    got = eng.search_many(queries, k, ef)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries, k, ef, threads=8)
    _same(got, want, "mixed 256")


@pytest.mark.parametrize("ef,k", [(48, 10), (96, 100), (300, 10)])
def test_two_phase_schedule_matches_oracle(m16d128, ef: int, k: int) -> None:
    """With fewer wave slots than queries (one wave per CU) a batch on the register beams runs
    every descent first (16 lanes per query) and layer 0 longest-first. One launch gives the
    same results, and so do descent caps that move some (20) or all (1) of the descents into the
    layer-0 launch. 397 queries: the last block of the descent kernel holds 13 of its 16
    queries."""
    eng, _, queries = m16d128
    queries = queries[:397]
    want = _oracle(eng, queries, k, ef)
    runs = {}
    try:
        for name, kw in (("two", {}), ("cap20", {"descent_cap": 20}),
                         ("cap1", {"descent_cap": 1}), ("one", {"one_launch": 1})):
            print(f"case m16d128 two-phase ef={ef} k={k} {name}", flush=True)
            _tune(eng, waves_per_cu=1, **kw)
            runs[name] = eng.search_many(queries, k, ef)
    finally:
        _tune(eng)
    for name, got in runs.items():
        _same(got, want, f"{name} ef={ef}")


@pytest.mark.parametrize("ef", [32, 100])
def test_two_phase_inner_product_matches_oracle(m16d128_ip, ef: int) -> None:
    eng, _, queries = m16d128_ip
    queries = queries[:397]
    want = _oracle(eng, queries, 10, ef, "ip")
    print(f"case m16d128 ip ef={ef}", flush=True)
    try:
        _tune(eng, waves_per_cu=1)
        two = eng.search_many(queries, 10, ef)
    finally:
        _tune(eng)
    _same(two, want, f"ip two-phase ef={ef}")
    _same(eng.search_many(queries, 10, ef), want, f"ip ef={ef}")


@pytest.mark.parametrize("which,metric", [("m16d128", "l2"), ("m16d128_ip", "ip")])
def test_two_phase_batch_beyond_the_order_kernels_lds(request, which: str, metric: str) -> None:
    """17,000 queries: the order kernel recomputes each query's bucket instead of keeping it in
    LDS (more than 16,384 queries), and inner-product keys are bucketed relative to the batch's
    smallest key."""
    eng, data, _ = request.getfixturevalue(which)
    rng = np.random.default_rng(53)
    rows = data[rng.integers(0, data.shape[0], size=17_000)]
    queries = np.clip(rows + rng.integers(-20, 21, size=rows.shape), 0, 255).astype(np.float32)
    want = _oracle(eng, queries, 10, 32, metric)
    print(f"case {which} two-phase 17000 queries", flush=True)
    try:
        _tune(eng, waves_per_cu=1)
        got = eng.search_many(queries, 10, 32)
    finally:
        _tune(eng)
    _same(got, want, f"{which} 17000 queries")


@pytest.mark.parametrize("which,metric", [("m16d128", "l2"), ("m16d128_ip", "ip")])
@pytest.mark.parametrize("ef", [32, 100])
def test_two_phase_fractional_queries_match_oracle(request, which: str, metric: str,
                                                   ef: int) -> None:
    """Fractional queries take the float64 distance (sequential sums): with fewer wave slots
    than queries their descents run 16 lanes per query with the query in LDS, and the order
    buckets the float64 keys by their excess over the batch's smallest. 397 queries, and 17,000
    (the order kernel's recomputed buckets)."""
    eng, data, queries = request.getfixturevalue(which)
    rng = np.random.default_rng(61)
    q = (queries[:397] + rng.uniform(-0.5, 0.5, (397, 128))).astype(np.float32)
    rows = data[rng.integers(0, data.shape[0], size=17_000)]
    big = (rows + rng.uniform(-20.0, 20.0, size=rows.shape)).astype(np.float32)
    cases = [("397", q)] + ([("17000", big)] if ef == 32 else [])
    for name, qq in cases:
        want = _oracle(eng, qq, 10, ef, metric)
        print(f"case {which} fractional two-phase ef={ef} {name} queries", flush=True)
        try:
            _tune(eng, waves_per_cu=1)
            got = eng.search_many(qq, 10, ef)
        finally:
            _tune(eng)
        _same(got, want, f"{which} fractional ef={ef} {name}")


def test_fractional_query_routes_match_oracle(m16d128) -> None:
    """Fractional queries on the 8-bit index: unfiltered batches sum over the rows' one-byte copy
    with the queries widened to float64 (also a batch with per-query k and ef, one launch per
    kernel class over a subset of the widened rows); a masked batch sums over the float32 rows.
    All give the oracle's results."""
    eng, data, queries = m16d128
    rng = np.random.default_rng(73)
    q = (queries[:300] + rng.uniform(-0.5, 0.5, (300, 128))).astype(np.float32)
    mask = random_mask(data.shape[0], 0.3, 5)
    print("case m16d128 fractional unfiltered / masked", flush=True)
    for ef in (40, 200):
        _same(eng.search_many(q, 10, ef), _oracle(eng, q, 10, ef), f"fractional ef={ef}")
        got = eng.search_many(q, 10, ef, mask_pool=mask, mask_index=0)
        want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                                   q, 10, ef, mask=mask, threads=8)
        _same(got, want, f"fractional masked ef={ef}")
    print("case m16d128 fractional per-query k / ef", flush=True)
    k = rng.choice([1, 10, 50], size=300)
    ef = np.maximum(rng.choice([16, 48, 100, 300], size=300), k)
    _same(eng.search_many(q, k, ef), _oracle(eng, q, k, ef), "fractional per-query")


@pytest.mark.parametrize("ef", [16, 64])
def test_two_phase_codes_search_matches_oracle(m16d128, ef: int) -> None:
    """The codes traversal (uint8 codes, float64 code-space queries) with fewer wave slots than
    queries: batched float64 descents, then layer 0 in the order of their keys."""
    eng, data, queries = m16d128
    rng = np.random.default_rng(67)
    codes = data.astype(np.uint8)  # the vectors' own integer values as codes
    qc = queries[:397].astype(np.float64) + rng.uniform(-0.5, 0.5, (397, 128))
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2", qc,
                               ef, ef, codes=codes, threads=8)
    print(f"case m16d128 codes two-phase ef={ef}", flush=True)
    d0, h0 = eng.distance_computations, eng.hops
    try:
        _tune(eng, waves_per_cu=1)
        ids, d = eng.batch_search_codes(codes, qc, ef, ef)
    finally:
        _tune(eng)
    np.testing.assert_array_equal(ids, want[0])
    np.testing.assert_array_equal(d, want[1])
    assert eng.distance_computations - d0 == int(want[2].sum())
    assert eng.hops - h0 == int(want[3].sum())


@pytest.mark.parametrize("ef", [48, 100])
def test_padded_rows_match_oracle(m16d120, ef: int) -> None:
    """120 dimensions on the constant shape, in one launch and after the batched descent."""
    eng, _, queries = m16d120
    queries = queries[:397]
    want = _oracle(eng, queries, 10, ef)
    print(f"case m16d120 ef={ef}", flush=True)
    _same(eng.search_many(queries, 10, ef), want, f"d120 ef={ef}")
    try:
        _tune(eng, waves_per_cu=1)
        two = eng.search_many(queries, 10, ef)
    finally:
        _tune(eng)
    _same(two, want, f"d120 two-phase ef={ef}")


def test_wide_beam_on_the_bitset_matches_oracle(m16d128) -> None:
    """The two-entry beam with its visited set moved to the bitset after every probe step."""
    eng, _, queries = m16d128
    want = _oracle(eng, queries, 10, 96)
    print("case m16d128 ef=96 probe_limit=1", flush=True)
    _moves(eng)
    _tune(eng, 0, 1)
    try:
        got = eng.search_many(queries, 10, 96)
        moves = _moves(eng)
    finally:
        _tune(eng)
    _same(got, want, "ef=96 probe_limit=1")
    assert moves > 0


def test_whole_id_slots_match_oracle(wide_ids) -> None:
    eng, _, queries = wide_ids
    want = _oracle(eng, queries, 10, 32)
    for hl in (6, 0):  # 6: 17 - 6 = 11 remainder bits, whole-id slots; default: quotient slots
        print(f"case wide ids hash_log2={hl}", flush=True)
        _tune(eng, hl)
        try:
            got = eng.search_many(queries, 10, 32)
        finally:
            _tune(eng)
        _same(got, want, f"hash_log2={hl}")


def test_constant_shape_stop_kernel_that_never_fires_is_the_plain_search(m16d128) -> None:
    """In one launch and (one wave per CU) after the batched descent."""
    eng, _, queries = m16d128
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2000, learned.N_FEATURES)).astype(np.float32)
    p = learned.init_model(x, 3)
    p[learned.MODEL_B2] = -1e4
    eng.set_stop_model(p, 0.5)
    try:
        got = eng.search_many(queries, 10, 48, learned=True)
        _tune(eng, waves_per_cu=1)
        two = eng.search_many(queries, 10, 48, learned=True)
    finally:
        _tune(eng)
        eng.set_stop_model(None)
    want = _oracle(eng, queries, 10, 48)
    _same(got, want, "stop kernel")
    _same(two, want, "stop kernel after the batched descent")
