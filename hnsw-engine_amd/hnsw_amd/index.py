"""``HNSWIndex``: the reference's index API (``index.py:96-466``) over the MI355X engine.

The product boundary is :class:`hnsw_amd.engine.MI355XEngine`, the reference's ``Engine``
protocol; this facade gives it the reference's constructor, search and persistence surface so
that a caller of ``hnsw_engine.HNSWIndex`` finds the same arguments, results, padding and
errors. Construction is sequential insertion on the host (the same graphs as the reference);
every search runs the HIP kernels.

Search rules kept from the reference (``index.py:244-312``): ``1 <= k <= len(index)``; ``ef``
defaults to ``ef_search`` and is raised to ``k``; a filter is a boolean array with one entry
per indexed vector and applies to every query; a filter that allows nothing returns padding
(-1 / inf) without searching. Extensions: per-query ``k`` / ``ef`` / filters in one batch
(:meth:`batch_search_mixed`) and the filter planner (:meth:`batch_search_planned`).
"""

from __future__ import annotations

import math
from dataclasses import asdict, dataclass, replace
from typing import Any, NamedTuple

import numpy as np

from . import checks
from .engine import MI355XEngine
from .graph_io import GraphState, load_native, load_state, save_native, save_state
from .metrics import METRICS, prepare

ENGINES = ("mi355x",)
# Engine names a file written by the reference may carry; they open on this engine.
FOREIGN_ENGINES = ("numba", "reference")
SELECTIONS = ("simple", "heuristic")


@dataclass(frozen=True)
class IndexParams:
    """Construction parameters, checked in the reference's order and with its messages."""

    dim: int
    metric: str = "l2"
    M: int = 16
    ef_construction: int = 200
    ef_search: int = 50
    seed: int | None = None
    selection: str = "heuristic"
    extend_candidates: bool = False
    keep_pruned_connections: bool = False
    m_L: float | None = None
    engine: str = "mi355x"

    def __post_init__(self) -> None:
        fix = lambda field, value: object.__setattr__(self, field, value)  # noqa: E731
        fix("dim", checks.integer(self.dim, "dim", minimum=1))
        checks.one_of(self.metric, METRICS, "metric")
        fix("M", checks.integer(self.M, "M", minimum=2))
        fix("ef_construction", checks.integer(self.ef_construction, "ef_construction", minimum=1))
        fix("ef_search", checks.integer(self.ef_search, "ef_search", minimum=1))
        if self.seed is not None:
            fix("seed", checks.integer(self.seed, "seed", minimum=0))
        checks.one_of(self.selection, SELECTIONS, "selection")
        checks.flag(self.extend_candidates, "extend_candidates")
        checks.flag(self.keep_pruned_connections, "keep_pruned_connections")
        if self.selection == "simple" and (self.extend_candidates or self.keep_pruned_connections):
            raise ValueError(
                "extend_candidates and keep_pruned_connections apply to selection='heuristic'")
        fix("m_L", 1.0 / math.log(self.M) if self.m_L is None else checks.level_factor(self.m_L))
        checks.one_of(self.engine, ENGINES, "engine")


class _Request(NamedTuple):
    k: int
    ef: int
    mask: np.ndarray | None
    nothing_allowed: bool


class HNSWIndex:
    """Hierarchical Navigable Small World index whose searches run on an MI355X.

    Arguments are the reference's (``dim``, ``metric``, ``M``, ``ef_construction``,
    ``ef_search``, ``seed``, ``selection``, ``extend_candidates``, ``keep_pruned_connections``,
    ``m_L``, ``engine``), with ``engine="mi355x"``, plus ``device``, the HIP device ordinal.
    """

    def __init__(self, dim: int, *args, device: int = 0, **kwargs) -> None:
        self._params = IndexParams(dim, *args, **kwargs)
        self._rng = np.random.default_rng(self._params.seed)
        p = self._params
        self._engine = MI355XEngine(p.dim, p.metric, p.M, p.ef_construction, p.selection,
                                    p.extend_candidates, p.keep_pruned_connections, device=device)

    # -- parameters and state ---------------------------------------------------------------

    dim = property(lambda self: self._params.dim)
    metric = property(lambda self: self._params.metric)
    M = property(lambda self: self._params.M)
    ef_construction = property(lambda self: self._params.ef_construction)
    m_L = property(lambda self: self._params.m_L)
    selection = property(lambda self: self._params.selection)
    engine = property(lambda self: self._params.engine)
    entry_point = property(lambda self: self._engine.entry_point)
    max_level = property(lambda self: self._engine.max_level)
    distance_computations = property(lambda self: self._engine.distance_computations)
    hops = property(lambda self: self._engine.hops)

    @property
    def ef_search(self) -> int:
        return self._params.ef_search

    @ef_search.setter
    def ef_search(self, value: int) -> None:
        self._params = replace(self._params, ef_search=value)

    def __len__(self) -> int:
        return self._engine.size

    def __repr__(self) -> str:
        p = self._params
        return (f"HNSWIndex(dim={p.dim}, metric={p.metric!r}, M={p.M}, "
                f"ef_construction={p.ef_construction}, ef_search={p.ef_search}, "
                f"selection={p.selection!r}, engine={p.engine!r}, size={len(self)})")

    def reset_counters(self) -> None:
        """Zero the cumulative distance-computation and hop counters."""
        self._engine.distance_computations = 0
        self._engine.hops = 0

    def graph_state(self) -> GraphState:
        return self._engine.export_state()

    # -- building ------------------------------------------------------------------------------

    def add(self, vectors) -> np.ndarray:
        """Insert ``(n, dim)`` vectors by sequential insertion on the host; returns their ids.
        Levels are ``floor(-ln(1 - U) * m_L)`` from the index's generator (``levels.py:22-31``)."""
        rows = self._prepare(vectors)
        u = 1.0 - self._rng.random(rows.shape[0])
        return self._insert(rows, np.floor(-np.log(u) * self._params.m_L).astype(np.int64))

    def _add_with_levels(self, vectors, levels) -> np.ndarray:
        rows = self._prepare(vectors)
        lv = np.asarray(levels, dtype=np.int64)
        if lv.shape != (rows.shape[0],) or (lv < 0).any():
            raise ValueError("levels must be a non-negative integer array, one per vector")
        return self._insert(rows, lv)

    def _insert(self, rows: np.ndarray, levels: np.ndarray) -> np.ndarray:
        first = len(self)
        if rows.shape[0]:
            self._engine.insert(rows, levels)
        return np.arange(first, first + rows.shape[0], dtype=np.int64)

    def _prepare(self, vectors) -> np.ndarray:
        return prepare(checks.real_array(vectors, "vectors", 2, self.dim), self.metric)

    # -- search -------------------------------------------------------------------------------

    def _request(self, k, ef, filter) -> _Request:
        if len(self) == 0:
            raise ValueError("cannot search an empty index")
        kk = checks.neighbour_count(k, len(self))
        width = self.ef_search if ef is None else checks.integer(ef, "ef", minimum=1)
        mask = checks.node_mask(filter, len(self))
        return _Request(kk, max(width, kk), mask, mask is not None and not mask.any())

    def _run(self, req: _Request, rows: np.ndarray):
        if req.nothing_allowed:
            shape = (rows.shape[0], req.k)
            return np.full(shape, -1, dtype=np.int64), np.full(shape, np.inf)
        return self._engine.batch_search(rows, req.k, req.ef, req.mask)

    def search(self, query, k: int = 10, ef: int | None = None, filter=None):
        """Approximate ``k`` nearest neighbours of one ``(dim,)`` query: ``(ids, distances)``
        of length ``k``, closest first, padded with -1 / inf."""
        req = self._request(k, ef, filter)
        row = self._prepare(checks.real_array(query, "query", 1, self.dim)[None, :])
        ids, dists = self._run(req, row)
        return ids[0], dists[0]

    def batch_search(self, queries, k: int = 10, ef: int | None = None, filter=None):
        """:meth:`search` for every row of ``(nq, dim)`` queries; ``(nq, k)`` results."""
        req = self._request(k, ef, filter)
        return self._run(req, self._prepare(queries))

    def batch_search_planned(self, queries, k: int = 10, ef: int | None = None, filter=None,
                             planner=None):
        """Extension: :meth:`batch_search` through a filter planner (``planner.FilterPlanner``,
        default thresholds when None). Returns ``(ids, dists, plan)``: the in-graph filtered
        search (the reference's results) when ``plan.path == "graph"``, the exact top-k over
        the allowed vectors when ``"exact"``."""
        from .planner import FilterPlan, FilterPlanner

        req = self._request(k, ef, filter)
        rows = self._prepare(queries)
        if req.mask is None:
            return (*self._run(req, rows), FilterPlan("graph", 1.0, 1.0, 0.0))
        if req.nothing_allowed:
            return (*self._run(req, rows), FilterPlan("exact", 0.0, 1.0, 0.0))
        plan = (planner or FilterPlanner()).plan(self._engine, req.mask)
        if plan.path == "graph":
            return (*self._run(req, rows), plan)
        got = min(req.k, int(req.mask.sum()))
        ids = np.full((rows.shape[0], req.k), -1, dtype=np.int64)
        dists = np.full(ids.shape, np.inf)
        ids[:, :got], dists[:, :got] = self._engine.exact_search(rows, got, req.mask)
        return ids, dists, plan

    def batch_search_mixed(self, queries, k, ef, filters=None, filter_index=None):
        """Extension: one batch with per-query ``k`` and ``ef`` (ints or arrays) and per-query
        filters (``filters``: a pool of boolean rows; ``filter_index[i]`` picks query i's row,
        -1 for none). Each query follows the reference's rules for its own arguments. Returns
        ``(ids, dists)`` of shape ``(nq, max k)``; a row is padded past its query's k."""
        if len(self) == 0:
            raise ValueError("cannot search an empty index")
        rows = self._prepare(queries)
        nq = rows.shape[0]
        kq = np.broadcast_to(np.asarray(k, dtype=np.int64), (nq,))
        efq = np.broadcast_to(np.asarray(ef, dtype=np.int64), (nq,))
        if nq and not (kq.min() >= 1 and kq.max() <= len(self)):
            raise ValueError("every k must be in [1, len(index)]")
        if nq and efq.min() < 1:
            raise ValueError("every ef must be >= 1")
        run = np.ones(nq, dtype=bool)
        pool = which = None
        if filters is not None:
            pool = np.atleast_2d(np.asarray(filters))
            if pool.dtype != np.bool_:
                raise TypeError(f"filter must be a boolean array, got dtype {pool.dtype}")
            if pool.shape[1] != len(self):
                raise ValueError("every filter must have one entry per indexed vector")
            which = np.broadcast_to(np.asarray(filter_index, dtype=np.int64), (nq,))
            if (which >= pool.shape[0]).any():
                raise ValueError("filter_index out of range")
            # a query whose filter allows nothing gets padding and no search
            run = (which < 0) | pool.any(axis=1)[np.maximum(which, 0)]
        width = int(kq.max()) if nq else 1
        ids = np.full((nq, width), -1, dtype=np.int64)
        dists = np.full(ids.shape, np.inf)
        sel = np.flatnonzero(run)
        if sel.size:
            got = self._engine.search_many(
                rows[sel], kq[sel], np.maximum(efq[sel], kq[sel]), mask_pool=pool,
                mask_index=None if which is None else which[sel], kstride=width)
            ids[sel], dists[sel] = got[0], got[1]
        return ids, dists

    # -- persistence (the reference's formats) ----------------------------------------------------

    def _meta(self) -> dict[str, Any]:
        return {"params": asdict(self._params), "rng_state": self._rng.bit_generator.state}

    def save(self, path) -> None:
        """The parameters, generator state and graph as a ``.npz`` snapshot at ``path`` (the
        reference's ``HNSWIndex.save`` format, ``index.py:369-372``)."""
        save_state(path, self._meta(), self.graph_state())

    @classmethod
    def _reopen(cls, meta: dict[str, Any], device: int = 0, engine: str | None = None):
        params = dict(meta["params"])
        named = params.get("engine") if engine is None else engine
        params["engine"] = "mi355x" if named in FOREIGN_ENGINES else named
        index = cls(device=device, **params)
        index._rng.bit_generator.state = meta["rng_state"]
        return index

    @classmethod
    def load(cls, path, engine: str | None = None, device: int = 0) -> "HNSWIndex":
        """An index saved by :meth:`save` here or by the reference (its engine names open on
        this engine)."""
        meta, state = load_state(path)
        index = cls._reopen(meta, device, engine)
        if state.vectors.shape[1:] != (index.dim,):
            raise ValueError(
                f"stored vectors have shape {state.vectors.shape}, expected dim {index.dim}")
        index._engine.load_state(state)
        return index

    def save_native(self, directory) -> None:
        """The engine's arrays as ``.npy`` files plus ``header.json`` (``index.py:390-403``)."""
        meta = {**self._meta(), "entry_point": self.entry_point, "max_level": self.max_level}
        save_native(directory, meta, self._engine.native_arrays())

    @classmethod
    def load_native(cls, directory, mmap: bool = True, validate: bool = True,
                    device: int = 0) -> "HNSWIndex":
        """Open a directory written by ``save_native`` here or by the reference
        (``index.py:405-425``). With ``mmap`` the arrays are read-only maps used in place
        (adding vectors copies them first); ``validate`` checks every link before any search
        reads it (skip it only for files you wrote yourself)."""
        meta, arrays = load_native(directory, mmap=mmap)
        index = cls._reopen(meta, device)
        index._engine.attach_native(arrays, int(meta["entry_point"]), int(meta["max_level"]),
                                    validate=validate)
        return index
