"""CPU checks of the learned-stop pieces: the float64 MLP reference and the trace labelling."""

from __future__ import annotations

import numpy as np

import oracle.mlp as mlp
from hnsw_amd.learned import MODEL_SIZE, dataset_from_traces, init_model


def test_reference_gradient_matches_finite_differences() -> None:
    rng = np.random.default_rng(0)
    x = rng.standard_normal((64, 16)) * 3 + 1
    y = (rng.random(64) < 0.4).astype(np.float64)
    p = init_model(x.astype(np.float32), seed=1).astype(np.float64)
    p[mlp.B1:mlp.W2] = rng.standard_normal(64) * 0.1
    loss, g = mlp.loss_and_grad(x, y, p, pos_weight=2.0, round_weights=False)
    eps = 1e-6
    for i in list(range(mlp.W1, mlp.W1 + 40)) + list(range(mlp.B1, mlp.SIZE)):
        q = p.copy()
        q[i] += eps
        lp, _ = mlp.loss_and_grad(x, y, q, pos_weight=2.0, round_weights=False)
        q[i] -= 2 * eps
        lm, _ = mlp.loss_and_grad(x, y, q, pos_weight=2.0, round_weights=False)
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-6 * max(1.0, abs(g[i])), i
    assert g[:mlp.W1].sum() == 0.0
    assert mlp.SIZE == MODEL_SIZE


def test_bf16_rounding_is_round_to_nearest_even() -> None:
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 1.0 + 2**-9], dtype=np.float32)
    got = mlp.bf16(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7 (bfloat16 keeps 7 fraction bits): ties go to
    # the even neighbour 1.0; 1 + 3*2^-9 rounds up; 1 + 2^-9 rounds down.
    np.testing.assert_array_equal(got, np.array([1.0, 1.0, 1.0 + 2**-7, -2.5, 1.0],
                                                dtype=np.float32))


def test_labels_mark_steps_whose_hits_equal_the_final_hits() -> None:
    trace = np.zeros((2, 5, 18), dtype=np.float32)
    trace[0, :, 16] = [1, 3, 3, 5, 9]
    trace[1, :, 16] = [0, 2, 2, 0, 0]
    trace[..., 0] = np.arange(5)
    length = np.array([5, 3], dtype=np.int32)
    final = np.array([5, 2], dtype=np.int32)
    x, y = dataset_from_traces(trace, length, final)
    assert x.shape == (8, 16)
    np.testing.assert_array_equal(y, [0, 0, 0, 1, 1, 0, 1, 1])
    np.testing.assert_array_equal(x[:, 0], [0, 1, 2, 3, 4, 0, 1, 2])


def test_oracle_stop_on_hand_traces() -> None:
    """Two queries, k = 2: query 0 reaches both hits at step 1, query 1 its first hit at step 2
    and the second at step 5. Full recall costs (1 + 5) / 2 = 3 expansions; recall 0.75 lets
    query 1 stop at step 2: (1 + 2) / 2 = 1.5."""
    from hnsw_amd.learned import N_FEATURES, oracle_stop

    trace = np.zeros((2, 6, N_FEATURES + 2), dtype=np.float32)
    trace[0, :, N_FEATURES] = [2, 2, 2, 2, 2, 2]
    trace[1, :, N_FEATURES] = [0, 1, 1, 1, 2, 2]
    length = np.array([6, 6])
    exp_full, rec_full = oracle_stop(trace, length, 2, 1.0)
    assert rec_full == 1.0 and exp_full == 3.0
    exp, rec = oracle_stop(trace, length, 2, 0.75)
    assert rec == 0.75 and exp == 1.5
    assert oracle_stop(trace, np.array([1, 1]), 2, 0.8) is None


def test_makespan_of_list_scheduling() -> None:
    from hnsw_amd.learned import makespan

    rng = np.random.default_rng(3)
    ln = rng.integers(1, 100, size=500).astype(np.float64)
    assert makespan(ln, 1) == ln.sum()  # one slot runs everything in turn
    assert makespan(ln, 1000) == ln.max()  # a slot per query: the longest sets the time
    m = makespan(ln, 16)
    # longest-first list scheduling is within 4/3 of the optimum, itself at least the larger of
    # the mean load and the longest query
    lower = max(ln.sum() / 16, ln.max())
    assert lower <= m <= 4.0 / 3.0 * lower + 1e-9
    assert makespan([5, 5, 4, 4, 3, 3, 3], 3) == 11  # the textbook case: 5+4, 5+4, 3+3+3 -> 11


def test_tail_aware_stop_ceiling() -> None:
    from hnsw_amd.learned import N_FEATURES, oracle_stop_tail

    # 40 queries, hits reach k at step s_q and stay; the full searches run 2 s_q steps.
    k, steps = 10, 64
    rng = np.random.default_rng(4)
    s_q = rng.integers(4, 30, size=40)
    trace = np.zeros((40, steps, N_FEATURES + 2), dtype=np.float32)
    for i, s in enumerate(s_q):
        trace[i, :, N_FEATURES] = np.minimum(np.arange(1, steps + 1) * k // s, k)
    length = 2 * s_q
    out = oracle_stop_tail(trace, length, k, 1.0, 4, length)
    assert out is not None and out["recall"] >= 1.0
    # stopping at s_q is free: the batch time can halve at most
    assert 1.0 < out["max_speedup_in_batch_time"] <= 2.0 + 1e-9
    assert oracle_stop_tail(trace, np.minimum(length, 3), k, 1.0, 4, length) is None
