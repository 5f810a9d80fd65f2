"""``exact_search`` and ``exact_distances`` with the reference oracle's API (``oracle.py:35-150``).

The k nearest neighbours by (distance, id) are computed on the GPU (``hm_exact``): int8 matrix
cores for 8-bit integer data (exact integers, so the float64 distances are the reference's bit
for bit), float64 left-to-right sums otherwise. ``query_block`` / ``data_block`` are accepted for
API compatibility; the kernel chooses its own tiles (results never depend on them).
"""

from __future__ import annotations

import numpy as np

from . import _native
from . import checks
from .metrics import validate_metric


def exact_search(data, queries, k: int, metric: str = "l2", *, query_block: int = 256,
                 data_block: int = 65_536, allowed=None, device: int = 0):
    """Exact ``(ids, distances)`` of shape ``(nq, k)``, ascending, ties to the lower id.

    ``allowed`` (boolean per data row) restricts the search to those rows (the masked brute
    force of ``benchmarks/run_filtered.py:76-106``); fewer allowed rows than ``k`` pad with
    ``-1`` / ``inf``.
    """
    metric_name = validate_metric(metric)
    x = checks.real_array(data, "data", 2)
    n, dim = x.shape
    if n == 0 or dim == 0:
        raise ValueError(f"data must be non-empty with at least one dimension, got {x.shape}")
    q = checks.real_array(queries, "queries", 2, dim)
    k = checks.neighbour_count(k, n)
    checks.integer(query_block, "query_block", minimum=1)
    checks.integer(data_block, "data_block", minimum=1)
    if metric_name == "cosine":
        for arr, name in ((x, "data"), (q, "queries")):
            zero = np.flatnonzero(np.linalg.norm(np.asarray(arr, np.float64), axis=1) == 0.0)
            if zero.size:
                raise ValueError(
                    f"cosine distance is undefined for zero vectors ({name} rows {zero[:5].tolist()})"
                )
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    q32 = np.ascontiguousarray(q, dtype=np.float32)
    ids = np.empty((q32.shape[0], k), dtype=np.int64)
    dists = np.empty((q32.shape[0], k), dtype=np.float64)
    mask = None
    if allowed is not None:
        mask = np.ascontiguousarray(np.asarray(allowed, dtype=bool), dtype=np.uint8)
        if mask.shape != (n,):
            raise ValueError(f"allowed must have one entry per data row ({n})")
    if q32.shape[0]:
        _native.check(
            _native.lib().hm_exact(
                _native.ptr(x32), n, dim, _native.METRIC_CODES[metric_name], _native.ptr(q32),
                q32.shape[0], k, _native.ptr(mask), device, _native.ptr(ids), _native.ptr(dists),
            )
        )
    return ids, dists


def last_range_search() -> dict:
    """Kernel times (ms) and re-run count of the last ``exact_search`` on this thread that took
    a range-search path, int8 or float32 (``hm_exact_timing``); zeros when it took another."""
    out = np.zeros(4, dtype=np.float64)
    _native.check(_native.lib().hm_exact_timing(_native.ptr(out)))
    return {"bound_ms": float(out[0]), "range_ms": float(out[1]), "collect_ms": float(out[2]),
            "requeried": int(out[3])}


def exact_distances(data, queries, ids, metric: str = "l2") -> np.ndarray:
    """Float64 distances of each query to the rows named in ``ids`` (``-1`` -> ``inf``).

    Scoring helper (``oracle.py:129-150``): rows are gathered and summed in float64.
    """
    metric_name = validate_metric(metric)
    x = checks.real_array(data, "data", 2)
    q = checks.real_array(queries, "queries", 2, x.shape[1])
    id_array = np.asarray(ids, dtype=np.int64)
    if id_array.ndim != 2 or id_array.shape[0] != q.shape[0]:
        raise ValueError(f"ids must have shape (n_queries, m), got {id_array.shape}")
    if ((id_array < -1) | (id_array >= x.shape[0])).any():
        raise ValueError("ids contains values outside [-1, n)")
    out = np.full(id_array.shape, np.inf, dtype=np.float64)
    for row in range(q.shape[0]):
        valid = id_array[row] >= 0
        if not valid.any():
            continue
        rows = np.asarray(x[id_array[row, valid]], dtype=np.float64)
        q64 = np.asarray(q[row], dtype=np.float64)
        if metric_name == "l2":
            diff = rows - q64
            out[row, valid] = np.einsum("ij,ij->i", diff, diff)
        elif metric_name == "ip":
            out[row, valid] = -(rows @ q64)
        else:
            rows = rows / np.linalg.norm(rows, axis=1)[:, None]
            out[row, valid] = 1.0 - rows @ (q64 / np.linalg.norm(q64))
    return out
