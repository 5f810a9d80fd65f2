"""Learned early stopping of the layer-0 beam (no counterpart in the reference).

After every layer-0 expansion the register-beam kernel evaluates a 16 -> 64 -> 1 MLP
(bfloat16 weights, float32 accumulation) on features of the beam state and stops when the
predicted probability that the current top-k is already final exceeds a threshold. The plain
fixed-ef search is the fallback: a model that never fires reproduces it exactly, and the stop can
only shorten a search.

Pipeline (all on the GPU): traces of the plain search on training queries (features and the
tie-aware hits of the current top-k against the exact oracle, per expansion), labels
``hits_t == hits_final`` (stopping here loses nothing against running the full beam), training
by the library's own forward/backward and Adam kernels, then a threshold chosen on held-out
queries for the recall target.

Feature order (``features`` in search.hip): expansions so far, log2(1 + distances computed),
ef, k, |R|/ef, unexpanded entries/ef, signed log2 of the best / k-th / next candidate's
distance, (next - k-th)/(|k-th - best| + 1), (k-th - best)/(|best| + 1),
(worst - k-th)/(|k-th - best| + 1), expansions since the k-th
changed, accepted and new neighbours of the last expansion, entries of the current top-k not
yet expanded. The search evaluates the model after every eighth expansion.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .recall import RELATIVE_TIE_TOLERANCE

N_FEATURES = 16
HIDDEN = 64
MODEL_W1 = 32
MODEL_B1 = MODEL_W1 + N_FEATURES * HIDDEN
MODEL_W2 = MODEL_B1 + HIDDEN
MODEL_B2 = MODEL_W2 + HIDDEN
MODEL_SIZE = MODEL_B2 + 1


def recall_thresholds(true_dist: np.ndarray, k: int) -> np.ndarray:
    """Per query: exact k-th distance plus the tie tolerance of the recall definition."""
    kth = true_dist[:, k - 1]
    return kth + RELATIVE_TIE_TOLERANCE * np.maximum(1.0, np.abs(kth))


def dataset_from_traces(trace: np.ndarray, length: np.ndarray, final: np.ndarray,
                        max_rows: int | None = None, seed: int = 0, long_weight: float = 0.0):
    """``(x, y)``: one row per recorded expansion; label 1 when the hits already equal the
    final hits. With ``max_rows`` a sample (fixed seed) of that many rows: uniform, or with
    ``long_weight`` > 0 drawn with probability proportional to (query length / mean
    length) ** long_weight, so that the long queries, which set a batch's time under the
    longest-first schedule, weigh more in training."""
    steps = trace.shape[1]
    ln = np.minimum(length, steps)
    t_idx = np.arange(steps)[None, :]
    valid = t_idx < ln[:, None]
    x = trace[..., :N_FEATURES][valid]
    hits = trace[..., N_FEATURES][valid]
    fin = np.broadcast_to(final[:, None], valid.shape)[valid]
    y = (hits >= fin).astype(np.float32)
    if max_rows is not None and x.shape[0] > max_rows:
        rng = np.random.default_rng(seed)
        if long_weight > 0:
            w = np.broadcast_to((ln / max(1.0, ln.mean()))[:, None] ** long_weight, valid.shape)
            p = w[valid].astype(np.float64)
            pick = np.sort(rng.choice(x.shape[0], max_rows, replace=False, p=p / p.sum()))
        else:
            pick = np.sort(rng.choice(x.shape[0], max_rows, replace=False))
        x, y = x[pick], y[pick]
    return np.ascontiguousarray(x, dtype=np.float32), np.ascontiguousarray(y, dtype=np.float32)


def makespan(lengths: np.ndarray, slots: int) -> float:
    """Batch time, in expansion times, of the layer-0 launch's schedule modelled as list
    scheduling: queries in decreasing length (the kernel orders by entry distance, which
    correlates with length) onto ``slots`` identical wave slots, each query holding its slot for
    its length."""
    import heapq

    lengths = np.asarray(lengths, dtype=np.float64)
    free = [0.0] * max(1, min(int(slots), lengths.size))
    heapq.heapify(free)
    end = 0.0
    for ln in np.sort(lengths)[::-1]:
        t = heapq.heappop(free) + float(ln)
        end = max(end, t)
        heapq.heappush(free, t)
    return end


def oracle_stop(trace: np.ndarray, length: np.ndarray, k: int, target_recall: float):
    """The cheapest per-query stopping points (in layer-0 expansions) that keep the mean
    recall of the traced queries at ``target_recall``, knowing every query's future: for a
    price ``lam`` per expansion each query stops where ``hits_t - lam t`` is largest, and the
    price is bisected until the mean hits reach the target. Returns ``(mean expansions, mean
    recall)``, or ``None`` when even the full searches stay below the target. A ceiling for
    any stop rule on the same beam."""
    steps = trace.shape[1]
    ln = np.minimum(length, steps)
    hits = trace[..., N_FEATURES].astype(np.float64)
    t = np.arange(1, steps + 1, dtype=np.float64)[None, :]
    valid = t <= ln[:, None]

    def pick(lam: float):
        score = np.where(valid, hits - lam * t, -np.inf)
        j = np.argmax(score, axis=1)
        rows = np.arange(trace.shape[0])
        return float(t[0, j].mean()), float(hits[rows, j].mean() / k)

    full = pick(0.0)
    if full[1] < target_recall:
        return None
    lo, hi = 0.0, float(k)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if pick(mid)[1] >= target_recall:
            lo = mid
        else:
            hi = mid
    return pick(lo)


def oracle_stop_tail(trace: np.ndarray, length: np.ndarray, k: int, target_recall: float,
                     slots: int, plain_length: np.ndarray):
    """Tail-aware ceiling of any stop rule on this beam: per-query stopping points that keep the
    mean recall at ``target_recall`` (knowing every query's future) and minimise the batch's
    makespan (``makespan`` over ``slots`` wave slots) rather than the mean: every cap C on a
    query's expansions is tried, with the per-expansion price of ``oracle_stop`` bisected under
    it. Returns a dict with the best makespan, the plain beam's (``plain_length``), their
    ratio, and the cap and mean expansions that reach it; ``None`` when even the full searches
    stay below the target."""
    steps = trace.shape[1]
    ln = np.minimum(length, steps)
    hits = trace[..., N_FEATURES].astype(np.float64)
    t = np.arange(1, steps + 1, dtype=np.float64)[None, :]
    rows = np.arange(trace.shape[0])

    def pick(lam: float, cap: float):
        valid = (t <= ln[:, None]) & (t <= cap)
        score = np.where(valid, hits - lam * t, -np.inf)
        j = np.argmax(score, axis=1)
        return t[0, j], hits[rows, j].mean() / k

    plain_span = makespan(np.minimum(plain_length, steps), slots)
    best = None
    for cap in sorted(set(int(c) for c in np.linspace(8, int(ln.max()), 40))):
        stops, rec = pick(0.0, cap)
        if rec < target_recall:
            continue
        lo, hi = 0.0, float(k)
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if pick(mid, cap)[1] >= target_recall:
                lo = mid
            else:
                hi = mid
        stops, rec = pick(lo, cap)
        span = makespan(stops, slots)
        if best is None or span < best["makespan"]:
            best = {"makespan": span, "cap": cap, "mean_expansions": float(stops.mean()),
                    "recall": float(rec)}
    if best is None:
        return None
    best["plain_makespan"] = plain_span
    best["max_speedup_in_batch_time"] = plain_span / best["makespan"]
    return best


def init_model(x: np.ndarray, seed: int = 0) -> np.ndarray:
    """Feature normalisation from ``x`` and He-initialised weights."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MODEL_SIZE, dtype=np.float32)
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    p[:N_FEATURES] = mean
    p[N_FEATURES:MODEL_W1] = 1.0 / np.where(std > 1e-6, std, 1.0)
    p[MODEL_W1:MODEL_B1] = (rng.standard_normal(N_FEATURES * HIDDEN) * np.sqrt(2.0 / N_FEATURES))
    p[MODEL_W2:MODEL_B2] = rng.standard_normal(HIDDEN) * np.sqrt(1.0 / HIDDEN)
    return p


def train(x: np.ndarray, y: np.ndarray, *, epochs: int = 12, batch: int = 1024,
          lr: float = 3e-3, pos_weight: float = 1.0, seed: int = 0):
    """GPU training (``hm_mlp_train``); returns ``(params, epoch_losses)``."""
    params = init_model(x, seed)
    losses = np.zeros(epochs, dtype=np.float32)
    xc = np.ascontiguousarray(x, dtype=np.float32)
    yc = np.ascontiguousarray(y, dtype=np.float32)
    _native.check(_native.lib().hm_mlp_train(
        _native.ptr(xc), _native.ptr(yc), xc.shape[0], _native.ptr(params), epochs, batch,
        ctypes.c_float(lr), ctypes.c_float(pos_weight), ctypes.c_uint64(seed), _native.ptr(losses)))
    return params, losses


def gradient(x: np.ndarray, y: np.ndarray, params: np.ndarray, pos_weight: float = 1.0):
    """``(grad, loss)`` of the mean loss over ``x`` (``hm_mlp_grad``)."""
    xc = np.ascontiguousarray(x, dtype=np.float32)
    yc = np.ascontiguousarray(y, dtype=np.float32)
    pc = np.ascontiguousarray(params, dtype=np.float32)
    g = np.zeros(MODEL_SIZE, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _native.check(_native.lib().hm_mlp_grad(_native.ptr(xc), _native.ptr(yc), xc.shape[0],
                                            _native.ptr(pc), ctypes.c_float(pos_weight),
                                            _native.ptr(g), _native.ptr(loss)))
    return g, float(loss[0])


def logits(x: np.ndarray, params: np.ndarray) -> np.ndarray:
    xc = np.ascontiguousarray(x, dtype=np.float32)
    pc = np.ascontiguousarray(params, dtype=np.float32)
    out = np.zeros(xc.shape[0], dtype=np.float32)
    _native.check(_native.lib().hm_mlp_forward(_native.ptr(xc), xc.shape[0], _native.ptr(pc),
                                               _native.ptr(out)))
    return out
