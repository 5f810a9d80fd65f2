"""Argument checks of the public API.

The exception types and message texts are the reference's (``_validation.py:16-66``,
``metrics.py:35-40``, ``index.py:133-147, 294-312, 474-481``): its callers and tests match on
them. ``TypeError`` for a wrong Python type, ``ValueError`` for a bad value.
"""

from __future__ import annotations

import math
import numbers

import numpy as np


def integer(value: object, name: str, minimum: int | None = None) -> int:
    """``value`` as an int; bools are refused although they are ints."""
    if not isinstance(value, numbers.Integral) or isinstance(value, bool):
        raise TypeError(f"{name} must be an integer, got {type(value).__name__}")
    out = int(value)
    if minimum is not None and out < minimum:
        raise ValueError(f"{name} must be >= {minimum}, got {out}")
    return out


def flag(value: object, name: str) -> bool:
    if isinstance(value, bool):
        return value
    raise TypeError(f"{name} must be a bool, got {type(value).__name__}")


def level_factor(value: object) -> float:
    """The level normalisation ``m_L``: a finite number >= 0."""
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise TypeError(f"m_L must be a number, got {type(value).__name__}")
    out = float(value)
    if out >= 0 and math.isfinite(out):
        return out
    raise ValueError(f"m_L must be finite and >= 0, got {value}")


def one_of(value: object, options: tuple[str, ...], what: str) -> str:
    if value in options:
        return value  # type: ignore[return-value]
    raise ValueError(f"unknown {what} {value!r}; expected one of {options}")


def real_array(values, name: str, ndim: int, dim: int | None = None) -> np.ndarray:
    """``values`` as an array of ``ndim`` dimensions (1 or 2) of finite integers or floats,
    dtype unchanged; ``dim`` fixes the length of the last axis."""
    out = np.asarray(values)
    if out.ndim != ndim:
        raise ValueError(f"{name} must be a {ndim}-D array, got shape {out.shape}")
    kind = out.dtype.kind
    if kind not in "iuf":
        raise ValueError(f"{name} must contain real numbers, got dtype {out.dtype}")
    if kind == "f" and not np.isfinite(out).all():
        raise ValueError(f"{name} contains NaN or infinite values")
    if dim is not None and out.shape[-1] != dim:
        raise ValueError(f"{name} has dimension {out.shape[-1]}, expected {dim}")
    return out


def neighbour_count(k: object, size: int) -> int:
    """``k`` in ``[1, size]``."""
    out = integer(k, "k", minimum=1)
    if out <= size:
        return out
    raise ValueError(f"k={out} exceeds the number of indexed vectors ({size})")


def node_mask(mask, size: int) -> np.ndarray | None:
    """A filter: None, or a boolean array with one entry per indexed vector."""
    if mask is None:
        return None
    out = np.asarray(mask)
    if out.dtype != np.bool_:
        raise TypeError(f"filter must be a boolean array, got dtype {out.dtype}")
    if out.shape == (size,):
        return np.ascontiguousarray(out)
    raise ValueError(
        f"filter must have one entry per indexed vector ({size}), got shape {out.shape}"
    )
