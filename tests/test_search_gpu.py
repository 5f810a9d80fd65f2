"""GPU search parity: the HIP kernels against the CPU oracle and the reference's fixtures.

Bar: bit-identical ids, float64 distances and per-query distance/hop counters on every case --
the exact 8-bit kernel on integer data, the sequential float64 kernel on continuous data (the
oracle computes the same unfused left-to-right sum), the codes kernel and the re-rank. Against
the reference's own fixtures on continuous data the distances are held to 1e-12, the reference's
cross-engine tolerance (its engines sum in different orders there).
"""

from __future__ import annotations

import numpy as np
import pytest

import golden_io
import oracle
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.synthetic import clustered_u8, random_mask, range_mask

pytestmark = pytest.mark.gpu

INT_FIXTURES = ["search_u8_l2", "search_ties_l2", "search_ties_ip"]
FLOAT_FIXTURES = ["search_float_l2", "search_float_cosine"]


def _engine(arrays, entry, top, metric, M) -> MI355XEngine:
    e = MI355XEngine(arrays["vectors"].shape[1], metric, M, 100, "heuristic", False, False)
    e.attach_native(arrays, entry, top)
    return e


def _same(got, want, msg=""):
    for g, w, name in zip(got, want, ("ids", "dists", "dcount", "hops")):
        np.testing.assert_array_equal(g, w, err_msg=f"{msg} {name}")


@pytest.mark.parametrize("name", INT_FIXTURES + FLOAT_FIXTURES)
def test_fixture_searches(name: str) -> None:
    fx, arrays, entry, top, metric, M = golden_io.graph(name)
    q = golden_io.prepared_queries(fx)
    eng = _engine(arrays, entry, top, metric, M)
    for tag in golden_io.tags(fx):
        k, ef = (int(v) for v in fx[f"{tag}_k_ef"])
        mask = fx.get(f"{tag}_mask")
        print(f"case {name}/{tag} k={k} ef={ef}", flush=True)
        got = eng.search_many(q, k, ef, mask_pool=mask, mask_index=None if mask is None else 0)
        want = oracle.search_batch(arrays, entry, top, metric, q, k, ef, mask=mask, threads=4)
        _same(got, want, f"{name}/{tag}")
        np.testing.assert_array_equal(got[0], fx[f"{tag}_ids"])
        np.testing.assert_array_equal(got[2], fx[f"{tag}_dc"])
        np.testing.assert_array_equal(got[3], fx[f"{tag}_hops"])
        if name in INT_FIXTURES:
            np.testing.assert_array_equal(got[1], fx[f"{tag}_dists"])
        else:
            np.testing.assert_allclose(got[1], fx[f"{tag}_dists"], rtol=0, atol=1e-12)


@pytest.fixture(scope="module")
def built_u8():
    """20,000 x 32 8-bit clustered vectors, M=16, efC=100, built by the host builder."""
    data, queries = clustered_u8(20_000, 32, 300, n_clusters=50, sigma=30.0, seed=3)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(9).random(20_000)) / np.log(16)).astype(np.int64)
    eng = MI355XEngine(32, "l2", 16, 100, "heuristic", False, False)
    eng.insert(data, levels)
    return eng, data, queries


@pytest.mark.parametrize("ef,k", [(16, 10), (64, 10), (200, 1), (512, 100), (10, 40)])
def test_plain_search_matches_oracle(built_u8, ef: int, k: int) -> None:
    eng, data, queries = built_u8
    print(f"case plain ef={ef} k={k}", flush=True)
    got = eng.search_many(queries, k, ef)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries, k, ef, threads=8)
    _same(got, want)


@pytest.mark.parametrize("kind", ["random", "range"])
@pytest.mark.parametrize("sel", [0.5, 0.1, 0.02, 0.005])
def test_filtered_search_matches_oracle(built_u8, kind: str, sel: float) -> None:
    eng, data, queries = built_u8
    n = data.shape[0]
    mask = random_mask(n, sel, 11) if kind == "random" else range_mask(n, sel, 12)
    q = queries[:100]
    print(f"case filtered {kind} sel={sel}", flush=True)
    got = eng.search_many(q, 10, 32, mask_pool=mask, mask_index=0)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               q, 10, 32, mask=mask, threads=8)
    _same(got, want, f"{kind} {sel}")
    assert np.all(mask[got[0][got[0] >= 0]])


def test_mixed_batch_per_query_ef_k_masks(built_u8) -> None:
    eng, data, queries = built_u8
    n = data.shape[0]
    rng = np.random.default_rng(5)
    nq = queries.shape[0]
    k = rng.choice([1, 10, 100], size=nq)
    ef = rng.choice([16, 32, 64, 128, 256, 512], size=nq)
    pool = np.stack([random_mask(n, s, 20 + i) for i, s in enumerate([0.05, 0.2, 0.5])])
    midx = np.where(rng.random(nq) < 0.5, rng.integers(0, 3, size=nq), -1)
    print("case mixed", flush=True)
    got = eng.search_many(queries, k, ef, mask_pool=pool, mask_index=midx)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               queries, k, ef, mask_pool=pool, mask_index=midx, threads=8)
    _same(got, want)


def test_continuous_data_sequential_kernel_matches_oracle() -> None:
    rng = np.random.default_rng(2)
    centres = rng.standard_normal((20, 24))
    data = (centres[rng.integers(0, 20, 6000)] + 0.3 * rng.standard_normal((6000, 24))).astype(np.float32)
    q = (centres[rng.integers(0, 20, 150)] + 0.3 * rng.standard_normal((150, 24))).astype(np.float32)
    levels = np.floor(-np.log(1.0 - rng.random(6000)) / np.log(12)).astype(np.int64)
    for metric in ("l2", "ip"):
        eng = MI355XEngine(24, metric, 12, 80, "heuristic", False, False)
        eng.insert(data, levels)
        for ef, mask in ((40, None), (100, random_mask(6000, 0.1, 4))):
            print(f"case continuous {metric} ef={ef} masked={mask is not None}", flush=True)
            got = eng.search_many(q, 10, ef, mask_pool=mask, mask_index=None if mask is None else 0)
            want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, metric,
                                       q, 10, ef, mask=mask, threads=8)
            _same(got, want, metric)


def test_codes_search_matches_oracle() -> None:
    rng = np.random.default_rng(4)
    data = (rng.standard_normal((5000, 20)) * rng.lognormal(size=20)).astype(np.float32)
    levels = np.floor(-np.log(1.0 - rng.random(5000)) / np.log(10)).astype(np.int64)
    eng = MI355XEngine(20, "l2", 10, 60, "heuristic", False, False)
    eng.insert(data, levels)
    low = data.astype(np.float64).min(axis=0)
    step = float((data.max(axis=0).astype(np.float64) - low).max()) / 255.0
    codes = np.clip(np.rint((data - low) / step), 0, 255).astype(np.uint8)
    q = (rng.standard_normal((120, 20)) * rng.lognormal(size=20)).astype(np.float32)
    qc = (q.astype(np.float64) - low) / step
    print("case codes", flush=True)
    ids, d = eng.batch_search_codes(codes, qc, 10, 64)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                               qc, 10, 64, codes=codes, threads=8)
    np.testing.assert_array_equal(ids, want[0])
    np.testing.assert_array_equal(d, want[1])


def test_rerank_matches_oracle(built_u8) -> None:
    eng, data, queries = built_u8
    cand, _, _, _ = eng.search_many(queries, 64, 64)
    cand[:, 5] = -1
    print("case rerank", flush=True)
    got = eng.rerank(queries, cand, 10)
    want = oracle.rerank_batch(data, "l2", queries, cand, 10, threads=8)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])


def test_massive_ties_overflow_the_register_beam_and_rerun() -> None:
    """16 distinct points repeated 200 times: the evicted ties at the worst distance outgrow
    the register-beam's tie list, so those queries rerun on the general kernel."""
    rng = np.random.default_rng(12)
    base = rng.integers(0, 2, size=(3200, 4)).astype(np.float32)
    q = rng.integers(0, 2, size=(80, 4)).astype(np.float32)
    levels = np.floor(-np.log(1.0 - rng.random(3200)) / np.log(8)).astype(np.int64)
    eng = MI355XEngine(4, "l2", 8, 40, "heuristic", False, False)
    eng.insert(base, levels)
    for ef, k in ((10, 10), (40, 5), (64, 64)):
        print(f"case ties ef={ef} k={k}", flush=True)
        got = eng.search_many(q, k, ef)
        want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                                   q, k, ef, threads=8)
        _same(got, want, f"ties ef={ef}")


def test_host_query_forms_on_an_8bit_index(built_u8) -> None:
    """hm_search on an index of 8-bit rows: integer queries 0..255 travel as bytes (packed on the
    host) and take the exact kernel; a batch with one fractional, negative or out-of-range value
    takes the float64 kernel for all its queries. Both against the oracle, which computes the
    float64 distance of whatever values it is given; also at 2^18 values, where the host packs
    with four threads."""
    eng, data, queries = built_u8
    base = queries[:300].copy()
    cases = {"integer": base}
    for tag, (row, col, v) in {"fraction": (5, 3, 17.5), "negative": (140, 0, -1.0),
                               "above_255": (299, 31, 256.0)}.items():
        q = base.copy()
        q[row, col] = v
        cases[tag] = q
    # A batch of 2^18 values is checked and packed by four host threads (a quarter each): all
    # integers, and one fraction in the last quarter.
    big = np.tile(base, (28, 1))[:8192]
    cases["parallel_integer"] = big
    q = big.copy()
    q[8000, 7] = 3.25
    cases["parallel_fraction_last_quarter"] = q
    for tag, q in cases.items():
        print(f"case host queries {tag}", flush=True)
        got = eng.search_many(q, 10, 48)
        want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2",
                                   q, 10, 48, threads=4)
        _same(got, want, tag)
