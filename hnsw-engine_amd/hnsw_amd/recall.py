"""Tie-aware recall@k, the definition behind "recall@10 >= 0.95" (``bench/metrics.py:13-51``).

A returned id counts if it is valid, not a repeat, and its exact distance is at most the exact
k-th neighbour distance plus a relative tolerance of 1e-9 (equally close points are
interchangeable). ``recall_ids`` is the plain id overlap with the exact top-k.
"""

from __future__ import annotations

import numpy as np

RELATIVE_TIE_TOLERANCE = 1e-9


def recall_at_k(found_ids, found_exact_dist, true_ids, true_dist, k: int) -> dict[str, float]:
    found = np.asarray(found_ids)[:, :k]
    dist = np.asarray(found_exact_dist)[:, :k]
    kth = np.asarray(true_dist)[:, k - 1 : k]
    threshold = kth + RELATIVE_TIE_TOLERANCE * np.maximum(1.0, np.abs(kth))
    nq = found.shape[0]
    # vectorised form of the per-row loop: first occurrences of valid ids only
    order = np.argsort(found, axis=1, kind="stable")
    sorted_ids = np.take_along_axis(found, order, axis=1)
    first = np.ones_like(sorted_ids, dtype=bool)
    first[:, 1:] = sorted_ids[:, 1:] != sorted_ids[:, :-1]
    unique = np.zeros_like(first)
    np.put_along_axis(unique, order, first, axis=1)
    valid = unique & (found >= 0)
    hits = (valid & (dist <= threshold)).sum(axis=1)
    truth = np.asarray(true_ids)[:, :k]
    id_hits = np.array([np.intersect1d(found[r][valid[r]], truth[r]).size for r in range(nq)])
    return {"recall": float(hits.mean() / k), "recall_ids": float(id_hits.mean() / k)}
