"""GPU checks of the learned stop: training kernels against the float64 reference, the
fallback property, and what the trained model gains over the plain fixed-ef beam."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import oracle.mlp as mlp
from hnsw_amd import exact_search, learned
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.recall import recall_at_k
from hnsw_amd.synthetic import clustered_u8, clustered_u8_queries

pytestmark = pytest.mark.gpu


def _model(seed=0):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((5000, 16)) * rng.lognormal(size=16) + rng.standard_normal(16)).astype(np.float32)
    y = (rng.random(5000) < 1 / (1 + np.exp(-x[:, 0] + x[:, 3]))).astype(np.float32)
    p = learned.init_model(x, seed)
    p[learned.MODEL_B1:learned.MODEL_W2] = rng.standard_normal(64).astype(np.float32) * 0.2
    return x, y, p


def test_gradient_matches_float64_reference() -> None:
    x, y, p = _model()
    g, loss = learned.gradient(x, y, p, pos_weight=1.5)
    ref_loss, ref_g = mlp.loss_and_grad(x, y, p, pos_weight=1.5)
    assert abs(loss - ref_loss) < 1e-4 * max(1.0, abs(ref_loss))
    np.testing.assert_allclose(g, ref_g, rtol=2e-3, atol=2e-6)


def test_forward_matches_float64_reference() -> None:
    x, _, p = _model(1)
    np.testing.assert_allclose(learned.logits(x, p), mlp.forward(x, p), rtol=1e-4, atol=1e-4)


def test_training_is_deterministic_and_lowers_the_loss() -> None:
    x, y, _ = _model(2)
    p1, l1 = learned.train(x, y, epochs=6, batch=256, seed=7)
    p2, l2 = learned.train(x, y, epochs=6, batch=256, seed=7)
    np.testing.assert_array_equal(p1, p2)
    assert l1[-1] < l1[0]


@pytest.fixture(scope="module")
def setup():
    data, queries = clustered_u8(40_000, 32, 1000, n_clusters=80, seed=21)
    train_q = clustered_u8_queries(3000, 32, n_clusters=80, seed=21, query_seed=23)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(22).random(40_000)) / np.log(16)).astype(np.int64)
    eng = MI355XEngine(32, "l2", 16, 100, "heuristic", False, False)
    eng.insert(data, levels)
    return eng, data, queries, train_q


def test_model_that_never_fires_is_the_plain_search(setup) -> None:
    eng, data, queries, _ = setup
    x, _, p = _model(3)
    p[learned.MODEL_B2] = -1e4
    eng.set_stop_model(p, 0.5)
    plain = eng.search_many(queries, 10, 48)
    with_model = eng.search_many(queries, 10, 48, learned=True)
    for a, b in zip(plain, with_model):
        np.testing.assert_array_equal(a, b)
    want = oracle.search_batch(eng.native_arrays(), eng.entry_point, eng.max_level, "l2", queries,
                               10, 48, threads=8)
    np.testing.assert_array_equal(with_model[0], want[0])
    eng.set_stop_model(None)


def test_traces_record_the_plain_search(setup) -> None:
    eng, data, queries, _ = setup
    ids, d = exact_search(data, queries, 10)
    thr = learned.recall_thresholds(d, 10)
    trace, length, final = eng.trace(queries, 10, 48, thr, steps=256)
    got = eng.search_many(queries, 10, 48)
    # hops = layer-0 expansions + upper-layer expansions; the final hits are the recall hits
    assert (length <= got[3]).all() and (length > 0).all()
    found_d = got[1]
    hits = (found_d <= thr[:, None]).sum(axis=1)
    np.testing.assert_array_equal(final, hits)
    last = trace[np.arange(len(length)), np.minimum(length, 256) - 1, 16]
    np.testing.assert_array_equal(last[length <= 256], final[length <= 256])


def test_trained_stop_beats_plain_beam_at_equal_recall(setup) -> None:
    """What the trained model gains, in distance computations (hardware-independent): for some
    threshold the learned stop at ef = 64 reaches a recall that the plain beam reaches only with
    more distance computations (the plain cost is the cheapest ef in the sweep whose recall is at
    least as high)."""
    eng, data, queries, train_q = setup
    t_ids, t_d = exact_search(data, train_q, 10)
    trace, length, final = eng.trace(train_q, 10, 64, learned.recall_thresholds(t_d, 10), steps=256)
    x, y = learned.dataset_from_traces(trace, length, final)
    params, losses = learned.train(x, y, epochs=10, seed=5)
    assert losses[-1] < losses[0]
    q_ids, q_d = exact_search(data, queries, 10)

    def rec(ids):
        exact = np.array([[np.sum((data[j] - q) ** 2) if j >= 0 else np.inf for j in row]
                          for row, q in zip(ids, queries.astype(np.float64))])
        return recall_at_k(ids, exact, q_ids, q_d, 10)["recall"]

    plain = []
    for ef in (10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 56, 64):
        ids, _, dc, _ = eng.search_many(queries, 10, ef)
        plain.append((rec(ids), dc.mean(), ef))
    gains = []
    for thr in (0.5, 0.7, 0.8, 0.9, 0.95, 0.98):
        eng.set_stop_model(params, thr)
        ids, _, dc, _ = eng.search_many(queries, 10, 64, learned=True)
        r = rec(ids)
        match = [cost for pr, cost, _ in plain if pr >= r]
        if match:
            gains.append(min(match) / dc.mean())
        print(f"threshold {thr}: recall {r:.4f} dist/q {dc.mean():.0f}; plain at that recall: "
              f"{min(match) if match else float('nan'):.0f} dist/q", flush=True)
    eng.set_stop_model(None)
    print("plain curve:", [(round(r, 4), round(c), ef) for r, c, ef in plain], flush=True)
    assert gains and max(gains) > 1.05
