"""Deterministic synthetic workloads for the search path (all seeds explicit).

The reference ships float generators for its own tests (``synthetic.py:14-68``); the workloads
measured here are integer-valued 8-bit data, where every float64 distance sum is exact in any
order, plus one continuous angular set. Generators:

* :func:`clustered_u8` -- isotropic Gaussian clusters, centres uniform in ``[0, 255]``, values
  rounded and clipped to the integers ``0..255`` (stored as float32). Queries come from the same
  centres with ``seed + 1``.
* :func:`sparse_prototypes_u8` -- 784-d vectors with about half exact zeros; the non-zero
  coordinates mix one of 10 prototypes with uniform noise in ``[0, 255]``.
* :func:`angular_heavy_tailed` -- Gaussian clusters with log-normal per-dimension scales,
  L2-normalised (cosine workload).
* :func:`random_mask`, :func:`range_mask` -- i.i.d. and position-correlated (one contiguous id
  range) filter masks of a given selectivity.
"""

from __future__ import annotations

import numpy as np


def clustered_u8(
    n: int, dim: int, n_queries: int, *, n_clusters: int, sigma: float = 30.0, seed: int = 0
) -> tuple[np.ndarray, np.ndarray]:
    """``(data, queries)`` float32 arrays with integer values in ``0..255``."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 255.0, size=(n_clusters, dim))
    data = _around(rng, centres, n, sigma)
    qrng = np.random.default_rng(seed + 1)
    queries = _around(qrng, centres, n_queries, sigma)
    return data, queries


def clustered_u8_queries(
    n_queries: int, dim: int, *, n_clusters: int, sigma: float = 30.0, seed: int = 0,
    query_seed: int,
) -> np.ndarray:
    """More queries around the centres of ``clustered_u8(..., seed=seed)`` (the centres are the
    first draw of that generator), from ``query_seed`` -- e.g. training queries disjoint from
    the measured ones."""
    centres = np.random.default_rng(seed).uniform(0.0, 255.0, size=(n_clusters, dim))
    return _around(np.random.default_rng(query_seed), centres, n_queries, sigma)


def _around(rng: np.random.Generator, centres: np.ndarray, n: int, sigma: float) -> np.ndarray:
    out = np.empty((n, centres.shape[1]), dtype=np.float32)
    labels = rng.integers(0, centres.shape[0], size=n)
    block = 65_536
    for start in range(0, n, block):
        stop = min(n, start + block)
        noise = rng.standard_normal((stop - start, centres.shape[1]))
        vals = np.rint(centres[labels[start:stop]] + sigma * noise)
        out[start:stop] = np.clip(vals, 0.0, 255.0)
    return out


def sparse_prototypes_u8(
    n: int, n_queries: int, *, dim: int = 784, n_prototypes: int = 10, seed: int = 0
) -> tuple[np.ndarray, np.ndarray]:
    """``(data, queries)``: each prototype fixes which half of the coordinates are zero; the
    other coordinates are ``rint((prototype + U[0, 255]) / 2)``, so about 50 % of the values are
    exact zeros and all values are 8-bit integers."""
    rng = np.random.default_rng(seed)
    support = rng.random((n_prototypes, dim)) < 0.5
    level = rng.uniform(0.0, 255.0, size=(n_prototypes, dim))

    def draw(r: np.random.Generator, m: int) -> np.ndarray:
        labels = r.integers(0, n_prototypes, size=m)
        noise = r.uniform(0.0, 255.0, size=(m, dim))
        vals = np.rint((level[labels] + noise) / 2.0)
        vals[~support[labels]] = 0.0
        return np.clip(vals, 0.0, 255.0).astype(np.float32)

    return draw(rng, n), draw(np.random.default_rng(seed + 1), n_queries)


def angular_heavy_tailed(
    n: int, dim: int, n_queries: int, *, n_clusters: int = 200, seed: int = 0
) -> tuple[np.ndarray, np.ndarray]:
    """``(data, queries)`` float32, L2-normalised rows (continuous values)."""
    rng = np.random.default_rng(seed)
    scales = rng.lognormal(mean=0.0, sigma=1.0, size=dim)
    centres = rng.standard_normal((n_clusters, dim)) * scales

    def draw(r: np.random.Generator, m: int) -> np.ndarray:
        labels = r.integers(0, n_clusters, size=m)
        x = centres[labels] + 0.35 * r.standard_normal((m, dim)) * scales
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        return x.astype(np.float32)

    return draw(rng, n), draw(np.random.default_rng(seed + 1), n_queries)


def random_mask(n: int, selectivity: float, seed: int) -> np.ndarray:
    """I.i.d. mask: each id allowed with probability ``selectivity``."""
    return np.random.default_rng(seed).random(n) < selectivity


def range_mask(n: int, selectivity: float, seed: int) -> np.ndarray:
    """Position-correlated mask: one contiguous id range holding ``selectivity`` of the ids."""
    width = max(1, int(round(selectivity * n)))
    start = int(np.random.default_rng(seed).integers(0, n - width + 1))
    mask = np.zeros(n, dtype=bool)
    mask[start : start + width] = True
    return mask
