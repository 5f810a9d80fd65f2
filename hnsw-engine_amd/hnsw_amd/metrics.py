"""Metric names and the stored representation (``metrics.py:35-68`` of the reference).

Distances are "smaller is closer": ``l2`` = sum of squared differences, ``ip`` = -x.q,
``cosine`` = 1 - x.q on vectors normalised once (in float64, then cast to float32).
"""

from __future__ import annotations

import numpy as np

METRICS = ("l2", "cosine", "ip")


def validate_metric(metric: object) -> str:
    from .checks import one_of

    return one_of(metric, METRICS, "metric")


def prepare(vectors: np.ndarray, metric: str) -> np.ndarray:
    """Contiguous float32 rows; cosine rows divided by their float64 norm first."""
    if metric != "cosine":
        with np.errstate(over="ignore"):
            direct = np.ascontiguousarray(vectors, dtype=np.float32)
        if not np.isfinite(direct).all():
            raise ValueError("vectors overflow float32 (values beyond about 3.4e38)")
        return direct
    as64 = np.asarray(vectors, dtype=np.float64)
    norms = np.linalg.norm(as64, axis=1)
    zero_rows = np.flatnonzero(norms == 0.0)
    if zero_rows.size:
        raise ValueError(
            f"cosine distance is undefined for zero vectors (rows {zero_rows[:5].tolist()})"
        )
    with np.errstate(over="ignore"):
        out = np.ascontiguousarray(as64 / norms[:, None], dtype=np.float32)
    if not np.isfinite(out).all():
        raise ValueError("vectors overflow float32 (values beyond about 3.4e38)")
    return out
