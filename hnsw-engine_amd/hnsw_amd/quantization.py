"""Scalar quantization and the quantized search (``quantization.py:29-138`` of the reference).

``ScalarQuantizer``: per-dimension offset (the minimum) and ONE shared step (the widest range
over 255; 1 when every value is an integer and every range spans at most 256 values, i.e.
lossless 8-bit data). Codes are ``clip(rint((x - offset) / step), 0, 255)`` as uint8.

``QuantizedIndex.batch_search`` traverses the index's graph with distances to the codes
(float64 code-space queries, the HIP codes kernel) and optionally re-ranks the best ``R`` with
exact float64 distances to the float32 vectors (the HIP re-rank kernel).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._validation import as_matrix, check_int

CODE_LEVELS = 256


@dataclass(frozen=True)
class ScalarQuantizer:
    offset: np.ndarray
    step: float

    @classmethod
    def fit(cls, vectors) -> "ScalarQuantizer":
        x = np.asarray(as_matrix(vectors, "vectors"), dtype=np.float64)
        if x.shape[0] == 0:
            raise ValueError("cannot fit a quantizer on zero vectors")
        low = x.min(axis=0)
        widest = float((x.max(axis=0) - low).max())
        integer_valued = bool(np.array_equal(x, np.rint(x)))
        if (integer_valued and widest <= CODE_LEVELS - 1) or widest == 0.0:
            step = 1.0
        else:
            step = widest / (CODE_LEVELS - 1)
        return cls(low, step)

    def encode(self, vectors) -> np.ndarray:
        x = np.asarray(vectors, dtype=np.float64)
        codes = np.rint((x - self.offset) / self.step)
        return np.ascontiguousarray(np.clip(codes, 0, CODE_LEVELS - 1), dtype=np.uint8)

    def decode(self, codes: np.ndarray) -> np.ndarray:
        return self.offset + self.step * codes.astype(np.float64)

    def transform_queries(self, queries, metric: str) -> np.ndarray:
        q = np.asarray(queries, dtype=np.float64)
        if metric == "l2":
            return np.ascontiguousarray((q - self.offset) / self.step)
        return np.ascontiguousarray(q)

    def to_distance(self, code_distances: np.ndarray, queries, metric: str) -> np.ndarray:
        if metric == "l2":
            return code_distances * self.step**2
        shift = np.asarray(queries, dtype=np.float64) @ self.offset
        if metric == "ip":
            return self.step * code_distances - shift[:, None]
        return 1.0 - shift[:, None] - self.step + self.step * code_distances


class QuantizedIndex:
    """Searches an :class:`HNSWIndex`'s graph with distances to uint8 codes."""

    def __init__(self, index) -> None:
        if len(index) == 0:
            raise ValueError("cannot quantize an empty index")
        self._index = index
        self._engine = index._engine
        vectors = self._engine.native_arrays()["vectors"]
        self.quantizer = ScalarQuantizer.fit(vectors)
        self.codes = self.quantizer.encode(vectors)

    @property
    def code_bytes(self) -> int:
        return int(self.codes.nbytes)

    @property
    def vector_bytes(self) -> int:
        return len(self._index) * self._index.dim * 4

    def batch_search(self, queries, k: int = 10, ef: int | None = None, rerank: int = 0):
        """``max(R, k)`` candidates from the codes traversal (``ef`` raised to at least that),
        then, with ``rerank = R``, the best ``k`` by exact distance, ties to the lower id."""
        if len(self._index) != self.codes.shape[0]:
            raise ValueError("the index changed after quantization; build a new QuantizedIndex")
        k_valid, ef_valid = self._index._check_search_args(k, ef)
        r = check_int(rerank, "rerank", minimum=0)
        metric = self._index.metric
        prepared = self._index._prepare(queries)
        fetch = min(max(r, k_valid), len(self._index))
        ids, code_dist = self._engine.batch_search_codes(
            self.codes, self.quantizer.transform_queries(prepared, metric), fetch,
            max(ef_valid, fetch),
        )
        if r == 0:
            return ids, self.quantizer.to_distance(code_dist, prepared, metric)
        return self._engine.rerank(prepared, ids, k_valid)
