"""Input checks of every public entry point, with the reference's messages
(``_validation.py:16-66``): ``ValueError`` for bad values, ``TypeError`` for wrong types."""

from __future__ import annotations

import numbers

import numpy as np


def check_int(value: object, name: str, *, minimum: int | None = None) -> int:
    if isinstance(value, bool) or not isinstance(value, numbers.Integral):
        raise TypeError(f"{name} must be an integer, got {type(value).__name__}")
    result = int(value)
    if minimum is not None and result < minimum:
        raise ValueError(f"{name} must be >= {minimum}, got {result}")
    return result


def as_matrix(values, name: str, *, dim: int | None = None) -> np.ndarray:
    array = np.asarray(values)
    if array.ndim != 2:
        raise ValueError(f"{name} must be a 2-D array, got shape {array.shape}")
    _real_finite(array, name)
    if dim is not None and array.shape[1] != dim:
        raise ValueError(f"{name} has dimension {array.shape[1]}, expected {dim}")
    return array


def as_vector(values, name: str, *, dim: int) -> np.ndarray:
    array = np.asarray(values)
    if array.ndim != 1:
        raise ValueError(f"{name} must be a 1-D array, got shape {array.shape}")
    _real_finite(array, name)
    if array.shape[0] != dim:
        raise ValueError(f"{name} has dimension {array.shape[0]}, expected {dim}")
    return array


def check_k(k: object, n: int) -> int:
    k_int = check_int(k, "k", minimum=1)
    if k_int > n:
        raise ValueError(f"k={k_int} exceeds the number of indexed vectors ({n})")
    return k_int


def _real_finite(array: np.ndarray, name: str) -> None:
    if not (np.issubdtype(array.dtype, np.integer) or np.issubdtype(array.dtype, np.floating)):
        raise ValueError(f"{name} must contain real numbers, got dtype {array.dtype}")
    if np.issubdtype(array.dtype, np.floating) and not np.isfinite(array).all():
        raise ValueError(f"{name} contains NaN or infinite values")
