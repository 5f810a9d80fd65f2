"""The host side of the search path against the reference's own cases (no GPU needed).

* ``recall_at_k`` -- the reference's known answers (``tests/test_bench.py:23-48``) and, on random
  batches with repeats, misses and ties, a plain per-row count (valid, first occurrence, exact
  distance within the tie tolerance of the true k-th).
* ``ScalarQuantizer`` -- the fixture written from the reference's quantizer
  (``tests/golden/quantizer.npz``: offsets, steps, codes, code-space queries and the mapping of
  code distances back, bit for bit, on continuous and integer data), and its own test cases
  (``tests/test_quantization.py:13-66``): lossless 8-bit data, half-step error bound, scaled
  wide integers, clipping, fit errors; the mapped-back distances equal the distances to the
  decoded vectors for every metric.
* ``exact_distances`` -- against the reference oracle's fixture (integer data bit for bit;
  continuous data to 1e-12: the reference sums with numpy's ``einsum``, whose order is not
  pinned), the hand case of ``tests/test_oracle.py:154-161`` and its messages.
* argument rules and messages of the facade, ``exact_search`` and the checks
  (``tests/test_index.py:30-101, 415-425``, ``tests/test_metrics.py:10-87``,
  ``tests/test_oracle.py:121-151``).
* the native directory (``tests/test_numba_engine.py:237-372``): a saved index reopens with
  read-only memory maps used in place and the same graph; foreign, inconsistent and corrupted
  directories are refused with the reference's messages; ``validate=False`` skips the check.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

import golden_io
from hnsw_amd import HNSWIndex, ScalarQuantizer, exact_distances, exact_search, recall_at_k
from hnsw_amd import checks
from hnsw_amd.graph_io import diff_states
from hnsw_amd.metrics import prepare
from hnsw_amd.synthetic import clustered_u8

# -- recall -------------------------------------------------------------------------------------


def test_recall_counts_tied_alternatives() -> None:
    out = recall_at_k(np.array([[0, 1, 9]]), np.array([[1.0, 2.0, 3.0]]), np.array([[0, 1, 2]]),
                      np.array([[1.0, 2.0, 3.0]]), k=3)
    assert out["recall"] == 1.0 and out["recall_ids"] == pytest.approx(2 / 3)


def test_recall_ignores_missing_and_repeated_ids() -> None:
    true_ids = np.array([[0, 1, 2], [0, 1, 2]])
    true_d = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0]])
    out = recall_at_k(np.array([[0, 0, 0], [0, -1, 7]]),
                      np.array([[1.0, 1.0, 1.0], [1.0, np.inf, 5.0]]), true_ids, true_d, k=3)
    assert out["recall"] == pytest.approx(1 / 3) and out["recall_ids"] == pytest.approx(1 / 3)


def test_recall_uses_only_the_first_k_columns() -> None:
    out = recall_at_k(np.array([[0, 5, 1, 2]]), np.array([[1.0, 9.0, 2.0, 3.0]]),
                      np.array([[0, 1, 2, 3]]), np.array([[1.0, 2.0, 3.0, 4.0]]), k=2)
    assert out["recall"] == 0.5


def test_recall_equals_a_per_row_count() -> None:
    rng = np.random.default_rng(3)
    nq, k = 300, 10
    found = rng.integers(-1, 25, size=(nq, 14))
    dist = np.where(found >= 0, rng.integers(0, 30, size=found.shape).astype(float), np.inf)
    true_ids = np.sort(rng.integers(0, 25, size=(nq, 12)), axis=1)
    true_d = np.sort(rng.integers(0, 30, size=(nq, 12)).astype(float), axis=1)
    hits = id_hits = 0
    for r in range(nq):
        kth = true_d[r, k - 1]
        tol = kth + 1e-9 * max(1.0, abs(kth))
        seen = []
        for c in range(k):
            node = int(found[r, c])
            if node >= 0 and node not in seen:
                seen.append(node)
                hits += dist[r, c] <= tol
        id_hits += len(set(seen) & set(true_ids[r, :k].tolist()))
    out = recall_at_k(found, dist, true_ids, true_d, k)
    assert out["recall"] == pytest.approx(hits / (nq * k), abs=1e-15)
    assert out["recall_ids"] == pytest.approx(id_hits / (nq * k), abs=1e-15)


# -- quantizer ----------------------------------------------------------------------------------


@pytest.mark.parametrize("kind", ["cont", "int"])
def test_quantizer_matches_the_reference_fixture(kind: str) -> None:
    fx = golden_io.load("quantizer")
    q = ScalarQuantizer.fit(fx[f"{kind}_data"])
    np.testing.assert_array_equal(q.offset, fx[f"{kind}_offset"])
    assert q.step == float(fx[f"{kind}_step"])
    np.testing.assert_array_equal(q.encode(fx[f"{kind}_data"]), fx[f"{kind}_codes"])


@pytest.mark.parametrize("metric", ["l2", "ip", "cosine"])
def test_quantizer_queries_and_distances_match_the_fixture(metric: str) -> None:
    fx = golden_io.load("quantizer")
    q = ScalarQuantizer.fit(fx["cont_data"])
    queries = fx["cont_queries"]
    np.testing.assert_array_equal(q.transform_queries(queries, metric), fx[f"cont_tq_{metric}"])
    np.testing.assert_array_equal(q.to_distance(fx[f"cont_cd_{metric}"], queries, metric),
                                  fx[f"cont_todist_{metric}"])


def test_quantizer_reference_cases() -> None:
    rng = np.random.default_rng(1)
    ints = rng.integers(0, 256, size=(500, 6)).astype(np.float64) + 7.0
    qi = ScalarQuantizer.fit(ints)
    assert qi.step == 1.0
    np.testing.assert_array_equal(qi.decode(qi.encode(ints)), ints)
    cont = rng.standard_normal((1000, 8)) * 3.0
    qc = ScalarQuantizer.fit(cont)
    assert qc.step == pytest.approx(float((cont.max(0) - cont.min(0)).max()) / 255)
    assert np.abs(qc.decode(qc.encode(cont)) - cont).max() <= qc.step / 2 + 1e-9
    assert ScalarQuantizer.fit(np.array([[0.0], [300.0]])).step == pytest.approx(300 / 255)
    clip = ScalarQuantizer.fit(np.array([[0.0, 0.0], [255.0, 10.0]]))
    np.testing.assert_array_equal(clip.encode(np.array([[-5.0, 300.0]])), [[0, 255]])
    assert ScalarQuantizer.fit(np.full((4, 2), 0.5)).step == 1.0
    with pytest.raises(ValueError, match="zero vectors"):
        ScalarQuantizer.fit(np.zeros((0, 3)))
    with pytest.raises(ValueError, match="NaN"):
        ScalarQuantizer.fit(np.array([[np.nan, 1.0]]))


@pytest.mark.parametrize("metric", ["l2", "ip", "cosine"])
def test_code_distances_map_to_decoded_distances(metric: str) -> None:
    rng = np.random.default_rng(3)
    data, queries = rng.standard_normal((300, 5)), rng.standard_normal((20, 5))
    if metric == "cosine":
        data /= np.linalg.norm(data, axis=1, keepdims=True)
        queries /= np.linalg.norm(queries, axis=1, keepdims=True)
    q = ScalarQuantizer.fit(data)
    codes = q.encode(data).astype(np.float64)
    cq = q.transform_queries(queries, metric)
    dec = q.decode(q.encode(data))
    if metric == "l2":
        cd = ((cq[:, None] - codes[None]) ** 2).sum(-1)
        true = ((queries[:, None] - dec[None]) ** 2).sum(-1)
    elif metric == "ip":
        cd, true = -(cq @ codes.T), -(queries @ dec.T)
    else:
        cd, true = 1.0 - cq @ codes.T, 1.0 - queries @ dec.T
    np.testing.assert_allclose(q.to_distance(cd, queries, metric), true, rtol=1e-9, atol=1e-9)


# -- exact_distances ----------------------------------------------------------------------------


def test_exact_distances_against_the_oracle_fixture() -> None:
    fx = golden_io.load("oracle_exact")
    for metric in ("l2", "ip"):
        got = exact_distances(fx["int_data"], fx["int_queries"], fx[f"int_{metric}_ids"], metric)
        np.testing.assert_array_equal(got, fx[f"int_{metric}_dists"])
    for metric in ("l2", "ip", "cosine"):
        data, queries = fx["float_data"], fx["float_queries"]
        got = exact_distances(data, queries, fx[f"float_{metric}_ids"], metric)
        np.testing.assert_allclose(got, fx[f"float_{metric}_dists"], rtol=0, atol=1e-12)


def test_exact_distances_hand_case_and_messages() -> None:
    data = np.array([[0, 0], [3, 4]], dtype=np.float32)
    out = exact_distances(data, np.array([[0.0, 0.0]]), np.array([[1, -1, 0]]))
    np.testing.assert_array_equal(out, [[25.0, np.inf, 0.0]])
    with pytest.raises(ValueError, match="outside"):
        exact_distances(data, np.array([[0.0, 0.0]]), np.array([[2]]))
    with pytest.raises(ValueError, match="shape"):
        exact_distances(data, np.array([[0.0, 0.0]]), np.array([1]))


# -- argument rules and messages ----------------------------------------------------------------


@pytest.mark.parametrize("kwargs,error,message", [
    ({"dim": 0}, ValueError, "dim must be >= 1"),
    ({"dim": 2.0}, TypeError, "dim must be an integer"),
    ({"M": 1}, ValueError, "M must be >= 2"),
    ({"ef_construction": 0}, ValueError, "ef_construction must be >= 1"),
    ({"ef_search": 0}, ValueError, "ef_search must be >= 1"),
    ({"seed": -1}, ValueError, "seed must be >= 0"),
    ({"seed": 1.5}, TypeError, "seed must be an integer"),
    ({"metric": "hamming"}, ValueError, "unknown metric 'hamming'"),
    ({"selection": "random"}, ValueError, "unknown selection 'random'"),
    ({"selection": "simple", "extend_candidates": True}, ValueError, "apply to selection="),
    ({"selection": "simple", "keep_pruned_connections": True}, ValueError, "apply to selection="),
    ({"extend_candidates": 1}, TypeError, "extend_candidates must be a bool"),
    ({"m_L": -0.1}, ValueError, "m_L must be finite and >= 0"),
    ({"m_L": float("nan")}, ValueError, "m_L must be finite"),
    ({"m_L": float("inf")}, ValueError, "m_L must be finite"),
    ({"m_L": True}, TypeError, "m_L must be a number"),
    ({"engine": "gpu"}, ValueError, "unknown engine 'gpu'"),
])
def test_constructor_rules(kwargs, error, message) -> None:
    with pytest.raises(error, match=message):
        HNSWIndex(**{"dim": 4, **kwargs})


def test_defaults_and_insertion_rules() -> None:
    index = HNSWIndex(3, M=16)
    assert index.m_L == pytest.approx(1 / np.log(16)) and index.selection == "heuristic"
    assert len(index) == 0 and index.add(np.empty((0, 3))).shape == (0,)
    with pytest.raises(ValueError, match="dimension 2, expected 3"):
        index.add(np.ones((2, 2)))
    with pytest.raises(ValueError, match="2-D"):
        index.add(np.ones(3))
    with pytest.raises(ValueError, match="NaN"):
        index.add(np.array([[1.0, np.nan, 0.0]]))
    with pytest.raises(ValueError, match="levels"):
        index._add_with_levels(np.ones((2, 3)), [0])
    with pytest.raises(ValueError, match="levels"):
        index._add_with_levels(np.ones((1, 3)), [-1])
    cos = HNSWIndex(3, "cosine")
    with pytest.raises(ValueError, match="zero vectors"):
        cos.add(np.zeros((1, 3)))
    with pytest.raises(ValueError, match="overflow"):
        prepare(np.array([[1e39, 0.0]]), "l2")


def test_search_argument_rules() -> None:
    """Every refusal happens on the host before any device work."""
    index = HNSWIndex(2)
    with pytest.raises(ValueError, match="empty index"):
        index.search(np.zeros(2), k=1)
    index.add(np.eye(2))
    for bad_k, error, message in [(0, ValueError, ">= 1"), (3, ValueError, "exceeds"),
                                  (1.0, TypeError, "k must be an integer")]:
        with pytest.raises(error, match=message):
            index.search(np.zeros(2), k=bad_k)
    with pytest.raises(ValueError, match="ef must be >= 1"):
        index.search(np.zeros(2), k=1, ef=0)
    with pytest.raises(ValueError, match="dimension"):
        index.search(np.zeros(3), k=1)
    with pytest.raises(ValueError, match="1-D"):
        index.search(np.zeros((1, 2)), k=1)
    with pytest.raises(ValueError, match="NaN"):
        index.search(np.array([np.nan, 0.0]), k=1)
    with pytest.raises(ValueError, match=">= 1"):
        index.ef_search = 0
    with pytest.raises(TypeError, match="boolean"):
        index.search(np.zeros(2), k=1, filter=np.array([1, 0]))
    with pytest.raises(ValueError, match="one entry per indexed vector"):
        index.search(np.zeros(2), k=1, filter=np.array([True]))
    with pytest.raises(ValueError, match="one entry per indexed vector"):
        index.batch_search(np.zeros((1, 2)), k=1, filter=np.ones((2, 1), dtype=bool))
    index.reset_counters()  # insertion counts its distances too
    ids, dists = index.batch_search(np.zeros((3, 2)), k=2, filter=np.zeros(2, dtype=bool))
    assert (ids == -1).all() and np.isinf(dists).all() and ids.shape == (3, 2)
    assert (index.distance_computations, index.hops) == (0, 0)


def test_checks_reference_cases() -> None:
    with pytest.raises(ValueError, match=">= 1"):
        checks.integer(0, "x", minimum=1)
    for bad in (1.5, "2", True, None):
        with pytest.raises(TypeError):
            checks.integer(bad, "x")
    with pytest.raises(ValueError, match="2-D"):
        checks.real_array([1, 2], "m", 2)
    with pytest.raises(ValueError, match="dimension 2, expected 3"):
        checks.real_array([[1, 2]], "m", 2, 3)
    with pytest.raises(ValueError, match="NaN or infinite"):
        checks.real_array([[1.0, np.inf]], "m", 2)
    for bad in (np.array([[1 + 2j]]), np.array([[True]])):
        with pytest.raises(ValueError, match="real numbers"):
            checks.real_array(bad, "m", 2)
    with pytest.raises(ValueError, match="1-D"):
        checks.real_array([[1.0, 2.0]], "v", 1, 2)
    with pytest.raises(ValueError, match="exceeds"):
        checks.neighbour_count(4, 3)


@pytest.mark.parametrize("data,queries,k,metric,message", [
    (np.ones((3, 2)), np.ones((1, 2)), 0, "l2", ">= 1"),
    (np.ones((3, 2)), np.ones((1, 2)), 4, "l2", "exceeds"),
    (np.ones((3, 2)), np.ones((1, 2)), -1, "l2", ">= 1"),
    (np.ones((3, 2)), np.ones((1, 3)), 1, "l2", "dimension 3, expected 2"),
    (np.ones(3), np.ones((1, 3)), 1, "l2", "2-D"),
    (np.ones((3, 2)), np.ones(2), 1, "l2", "2-D"),
    (np.empty((0, 2)), np.ones((1, 2)), 1, "l2", "non-empty"),
    (np.array([[1.0, np.nan]]), np.ones((1, 2)), 1, "l2", "NaN"),
    (np.ones((3, 2)), np.array([[np.inf, 0.0]]), 1, "l2", "NaN or infinite"),
    (np.array([[0.0, 0.0], [1.0, 1.0]]), np.ones((1, 2)), 1, "cosine", "zero vectors"),
    (np.ones((3, 2)), np.zeros((1, 2)), 1, "cosine", "zero vectors"),
    (np.ones((3, 2)), np.ones((1, 2)), 1, "hamming", "unknown metric"),
])
def test_exact_search_argument_rules(data, queries, k, metric, message) -> None:
    with pytest.raises(ValueError, match=message):
        exact_search(data, queries, k, metric)


@pytest.mark.parametrize("k", [1.5, "2", True, None])
def test_exact_search_non_integer_k(k) -> None:
    with pytest.raises(TypeError):
        exact_search(np.ones((3, 2)), np.ones((1, 2)), k)


# -- the native directory -----------------------------------------------------------------------


def _saved(tmp_path, n=400, dim=4, M=4):
    data, _ = clustered_u8(n, dim, 1, n_clusters=8, seed=1)
    index = HNSWIndex(dim, M=M, ef_construction=40, seed=0)
    index.add(data)
    root = tmp_path / "idx"
    index.save_native(root)
    return index, root


def test_native_round_trip_is_memory_mapped(tmp_path) -> None:
    index, root = _saved(tmp_path, 2500, 8, 6)
    opened = HNSWIndex.load_native(root)
    assert opened.engine == "mi355x"
    arrays = opened._engine.native_arrays()
    assert not arrays["vectors"].flags.writeable and not arrays["links0"].flags.writeable
    assert diff_states(index.graph_state(), opened.graph_state()) == []
    header = json.loads((root / "header.json").read_text())
    assert header["format"] == "hnsw-engine-native" and header["format_version"] == 1
    assert set(header["arrays"]) == {"vectors", "levels", "links0", "count0", "upper_offset",
                                     "links_up", "count_up"}
    # adding vectors copies the maps and leaves the files as they were
    before = {p.name: p.read_bytes() for p in root.iterdir()}
    opened.add(np.ones((3, 8), dtype=np.float32))
    assert {p.name: p.read_bytes() for p in root.iterdir()} == before
    assert len(opened) == 2503


def test_native_refuses_foreign_and_inconsistent_directories(tmp_path) -> None:
    with pytest.raises(ValueError, match="not an hnsw-engine native index"):
        HNSWIndex.load_native(tmp_path)
    _, root = _saved(tmp_path, 50, 3, 4)
    header = json.loads((root / "header.json").read_text())
    cases = [({**header, "format": "other"}, "not an hnsw-engine native index"),
             ({**header, "format_version": 99}, "version"),
             ({**header, "params": {**header["params"], "M": 5}}, "links0"),
             ({**header, "entry_point": 10**6}, "entry point")]
    for changed, message in cases:
        (root / "header.json").write_text(json.dumps(changed))
        with pytest.raises(ValueError, match=message):
            HNSWIndex.load_native(root)
    (root / "header.json").write_text(json.dumps(header))
    np.save(root / "count0.npy", np.zeros(49, dtype=np.int32))
    with pytest.raises(ValueError, match="count0"):
        HNSWIndex.load_native(root)


def _corrupt(root, name, change) -> None:
    array = np.load(root / f"{name}.npy")
    change(array)
    np.save(root / f"{name}.npy", array)


@pytest.mark.parametrize("name,change,message", [
    ("links0", lambda a: a.__setitem__((0, 0), 10**6), "layer 0: a link points outside the graph"),
    ("links0", lambda a: a.__setitem__((0, 0), -5), "layer 0: a link points outside the graph"),
    ("count0", lambda a: a.__setitem__(0, 99), r"layer 0: link counts outside \[0, 8\]"),
    ("upper_offset", lambda a: a.__setitem__(np.flatnonzero(a >= 0)[0], 10**6),
     "upper-layer rows of the nodes overlap or leave gaps"),
    ("upper_offset", lambda a: a.__setitem__(np.flatnonzero(a >= 0)[-1], 0),
     "upper-layer rows of the nodes overlap or leave gaps"),
    ("count_up", lambda a: a.__setitem__(0, 5), r"upper layers: link counts outside \[0, 4\]"),
    ("links_up", lambda a: a.__setitem__((0, 0), 10**6),
     "upper layers: a link points outside the graph"),
    ("levels", None, "negative node level"),
])
def test_native_load_rejects_corrupted_arrays(tmp_path, name, change, message) -> None:
    _, root = _saved(tmp_path)
    if change is None:  # one node to -1, another to 1: the row count stays consistent
        def change(a):
            ground = np.flatnonzero(a == 0)
            a[ground[0]], a[ground[1]] = -1, 1
    _corrupt(root, name, change)
    with pytest.raises(ValueError, match=message):
        HNSWIndex.load_native(root)
    HNSWIndex.load_native(root, validate=False)  # trusted mode skips the check


def test_native_load_rejects_an_upper_link_to_a_node_absent_from_the_layer(tmp_path) -> None:
    _, root = _saved(tmp_path)
    levels = np.load(root / "levels.npy")
    ground = int(np.flatnonzero(levels == 0)[0])
    _corrupt(root, "links_up", lambda a: a.__setitem__((0, 0), ground))
    with pytest.raises(ValueError, match="absent from that layer"):
        HNSWIndex.load_native(root)


def test_native_load_of_an_empty_index(tmp_path) -> None:
    index = HNSWIndex(3)
    index.save_native(tmp_path / "e")
    assert len(HNSWIndex.load_native(tmp_path / "e")) == 0
