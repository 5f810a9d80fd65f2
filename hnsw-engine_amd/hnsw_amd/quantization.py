"""8-bit scalar quantization and the quantized search (the reference's ``ScalarQuantizer`` and
``QuantizedIndex``, ``quantization.py:29-138``).

Codes: one byte per dimension, ``clip(rint((x - offset) / step), 0, 255)`` with ``offset`` the
per-dimension minimum and ONE step shared by all dimensions (the widest range over 255, or 1
when the data are integers spanning at most 256 values, which makes 8-bit data lossless). With a
shared step the code-space distance is an increasing affine function of the distance to the
decoded vectors, so a traversal over the codes ranks candidates as the decoded vectors would:
for l2 the code-space query is ``(q - offset) / step`` and the distance scales by ``step**2``;
for ip and cosine the query is unchanged and the distance maps through
``step * d_code + bias(q)``.

``QuantizedIndex.batch_search`` runs on the device in one call: the traversal over the codes
keeps its ``fetch`` candidates in HBM and, with ``rerank > 0``, the exact float64 re-rank scores
them there (``hm_search_codes_rerank``); only the final ``(nq, k)`` results cross PCIe.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import checks

LEVELS = 256


@dataclass(frozen=True)
class ScalarQuantizer:
    offset: np.ndarray  # float64 per-dimension minimum
    step: float

    @classmethod
    def fit(cls, vectors) -> "ScalarQuantizer":
        x = np.asarray(checks.real_array(vectors, "vectors", 2), dtype=np.float64)
        if x.shape[0] == 0:
            raise ValueError("cannot fit a quantizer on zero vectors")
        low = x.min(axis=0)
        span = float(np.max(x.max(axis=0) - low))
        lossless = span == 0.0 or (span <= LEVELS - 1 and bool(np.all(np.rint(x) == x)))
        return cls(low, 1.0 if lossless else span / (LEVELS - 1))

    def encode(self, vectors) -> np.ndarray:
        """uint8 codes; values outside the fitted range clip to 0 / 255."""
        scaled = (np.asarray(vectors, dtype=np.float64) - self.offset) / self.step
        return np.ascontiguousarray(np.clip(np.rint(scaled), 0, LEVELS - 1).astype(np.uint8))

    def decode(self, codes: np.ndarray) -> np.ndarray:
        return self.offset + self.step * codes.astype(np.float64)

    def transform_queries(self, queries, metric: str) -> np.ndarray:
        """float64 code-space queries: shifted and scaled for l2, unchanged otherwise."""
        q = np.asarray(queries, dtype=np.float64)
        return np.ascontiguousarray((q - self.offset) / self.step if metric == "l2" else q)

    def to_distance(self, code_distances, queries, metric: str) -> np.ndarray:
        """Code-space distances (one row per query) as distances to the decoded vectors."""
        d = np.asarray(code_distances, dtype=np.float64)
        if metric == "l2":
            return d * self.step**2
        # ip: -(offset.q) - step c.q; cosine: 1 - offset.q - step c.q, with d = -c.q / 1 - c.q
        shift = (np.asarray(queries, dtype=np.float64) @ self.offset)[:, None]
        bias = -shift if metric == "ip" else (1.0 - shift) - self.step
        return self.step * d + bias


class QuantizedIndex:
    """Searches an :class:`HNSWIndex`'s graph (built from its float32 vectors) with distances
    to 8-bit codes; the codes are a snapshot of the index at construction."""

    def __init__(self, index) -> None:
        if len(index) == 0:
            raise ValueError("cannot quantize an empty index")
        self._index = index
        vectors = index._engine.native_arrays()["vectors"]
        self.quantizer = ScalarQuantizer.fit(vectors)
        self.codes = self.quantizer.encode(vectors)

    @property
    def code_bytes(self) -> int:
        return int(self.codes.nbytes)

    @property
    def vector_bytes(self) -> int:
        return len(self._index) * self._index.dim * 4

    def batch_search(self, queries, k: int = 10, ef: int | None = None, rerank: int = 0):
        """``k`` nearest neighbours per query. The traversal collects ``fetch = max(rerank, k)``
        candidates (capped at the index size) with ef raised to at least ``fetch``; with
        ``rerank > 0`` the best ``k`` of them by exact float64 distance are returned (ties to
        the lower id), else the first ``k`` with their code distances mapped back."""
        if self.codes.shape[0] != len(self._index):
            raise ValueError("the index changed after quantization; build a new QuantizedIndex")
        req = self._index._request(k, ef, None)
        r = checks.integer(rerank, "rerank", minimum=0)
        rows = self._index._prepare(queries)
        metric = self._index.metric
        fetch = min(len(self._index), max(r, req.k))
        code_q = self.quantizer.transform_queries(rows, metric)
        engine = self._index._engine
        if r:
            return engine.quantized_search(self.codes, code_q, rows, req.k, fetch,
                                           max(req.ef, fetch))
        ids, d = engine.batch_search_codes(self.codes, code_q, fetch, max(req.ef, fetch))
        return ids, self.quantizer.to_distance(d, rows, metric)
