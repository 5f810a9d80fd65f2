"""``HNSWIndex``: the reference's public index API (``index.py:96-466``) on the MI355X engine.

Validation, metric preparation, the seeded level draw, the search argument rules
(``ef = max(ef, k)``; an all-false filter returns padding without searching) and the native
file format are the reference's; graph construction runs the host builder and every search runs
the HIP kernels. ``engine="mi355x"`` is the only engine here.
"""

from __future__ import annotations

import math
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Any

import numpy as np

from ._validation import as_matrix, as_vector, check_int, check_k
from .engine import MI355XEngine
from .metrics import prepare, validate_metric
from .persistence import (
    GraphState,
    check_native_shapes,
    load_native,
    load_state,
    save_native,
    save_state,
    validate_native,
)

ENGINES = ("mi355x",)
SELECTIONS = ("simple", "heuristic")


@dataclass(frozen=True)
class IndexStats:
    """Sizes, layer structure, degrees, compact memory estimate and cumulative counters
    (``index.py:63-93``)."""

    size: int
    dim: int
    metric: str
    engine: str
    M: int
    max_links_layer0: int
    ef_construction: int
    ef_search: int
    m_L: float
    selection: str
    extend_candidates: bool
    keep_pruned_connections: bool
    max_level: int
    entry_point: int
    nodes_per_layer: tuple[int, ...]
    links_per_layer: tuple[int, ...]
    mean_degree_per_layer: tuple[float, ...]
    memory_bytes: int
    distance_computations: int
    hops: int

    def as_dict(self) -> dict[str, Any]:
        return asdict(self)


def sample_levels(rng: np.random.Generator, n: int, m_l: float) -> np.ndarray:
    """``floor(-ln(U) * m_L)`` with ``U = 1 - rng.random()`` (``levels.py:22-31``)."""
    u = 1.0 - rng.random(n)
    return np.floor(-np.log(u) * m_l).astype(np.int64)


class HNSWIndex:
    """Hierarchical Navigable Small World index searched on an MI355X.

    Arguments as in the reference: ``dim``, ``metric`` (``"l2"``, ``"cosine"``, ``"ip"``),
    ``M``, ``ef_construction``, ``ef_search``, ``seed``, ``selection``,
    ``extend_candidates``, ``keep_pruned_connections``, ``m_L``, ``engine`` (``"mi355x"``),
    plus ``device`` (HIP device ordinal).
    """

    def __init__(
        self,
        dim: int,
        metric: str = "l2",
        M: int = 16,
        ef_construction: int = 200,
        ef_search: int = 50,
        seed: int | None = None,
        selection: str = "heuristic",
        extend_candidates: bool = False,
        keep_pruned_connections: bool = False,
        m_L: float | None = None,
        engine: str = "mi355x",
        device: int = 0,
    ) -> None:
        self._dim = check_int(dim, "dim", minimum=1)
        self._metric = validate_metric(metric)
        self._M = check_int(M, "M", minimum=2)
        self._ef_construction = check_int(ef_construction, "ef_construction", minimum=1)
        self._ef_search = check_int(ef_search, "ef_search", minimum=1)
        self._seed = None if seed is None else check_int(seed, "seed", minimum=0)
        if selection not in SELECTIONS:
            raise ValueError(f"unknown selection {selection!r}; expected one of {SELECTIONS}")
        self._selection = selection
        self._extend = _check_bool(extend_candidates, "extend_candidates")
        self._keep_pruned = _check_bool(keep_pruned_connections, "keep_pruned_connections")
        if selection == "simple" and (self._extend or self._keep_pruned):
            raise ValueError(
                "extend_candidates and keep_pruned_connections apply to selection='heuristic'"
            )
        self._m_L = 1.0 / math.log(self._M) if m_L is None else _check_m_l(m_L)
        if engine not in ENGINES:
            raise ValueError(f"unknown engine {engine!r}; expected one of {ENGINES}")
        self._engine_name = engine
        self._rng = np.random.default_rng(self._seed)
        self._engine = MI355XEngine(
            self._dim, self._metric, self._M, self._ef_construction, self._selection,
            self._extend, self._keep_pruned, device=device,
        )

    # -- properties ---------------------------------------------------------------------

    @property
    def dim(self) -> int:
        return self._dim

    @property
    def metric(self) -> str:
        return self._metric

    @property
    def M(self) -> int:
        return self._M

    @property
    def ef_construction(self) -> int:
        return self._ef_construction

    @property
    def ef_search(self) -> int:
        return self._ef_search

    @ef_search.setter
    def ef_search(self, value: int) -> None:
        self._ef_search = check_int(value, "ef_search", minimum=1)

    @property
    def m_L(self) -> float:
        return self._m_L

    @property
    def selection(self) -> str:
        return self._selection

    @property
    def engine(self) -> str:
        return self._engine_name

    @property
    def entry_point(self) -> int:
        return self._engine.entry_point

    @property
    def max_level(self) -> int:
        return self._engine.max_level

    def __len__(self) -> int:
        return self._engine.size

    def __repr__(self) -> str:
        return (
            f"HNSWIndex(dim={self._dim}, metric={self._metric!r}, M={self._M}, "
            f"ef_construction={self._ef_construction}, ef_search={self._ef_search}, "
            f"selection={self._selection!r}, engine={self._engine_name!r}, size={len(self)})"
        )

    # -- building -------------------------------------------------------------------------

    def add(self, vectors) -> np.ndarray:
        """Insert ``(n, dim)`` vectors (sequential insertion on the host); returns their ids."""
        prepared = self._prepare(vectors)
        levels = sample_levels(self._rng, prepared.shape[0], self._m_L)
        return self._insert(prepared, levels)

    def _add_with_levels(self, vectors, levels) -> np.ndarray:
        prepared = self._prepare(vectors)
        level_array = np.asarray(levels, dtype=np.int64)
        if level_array.shape != (prepared.shape[0],) or (level_array < 0).any():
            raise ValueError("levels must be a non-negative integer array, one per vector")
        return self._insert(prepared, level_array)

    def _insert(self, prepared: np.ndarray, levels: np.ndarray) -> np.ndarray:
        start = len(self)
        if prepared.shape[0]:
            self._engine.insert(prepared, levels)
        return np.arange(start, start + prepared.shape[0], dtype=np.int64)

    def _prepare(self, vectors) -> np.ndarray:
        return prepare(as_matrix(vectors, "vectors", dim=self._dim), self._metric)

    # -- search ---------------------------------------------------------------------------

    def search(self, query, k: int = 10, ef: int | None = None, filter=None):
        """Approximate ``k`` nearest neighbours of one query; ``(ids, distances)`` of length
        ``k``, padded with ``-1`` / ``inf``."""
        k_valid, ef_valid = self._check_search_args(k, ef)
        allowed = self._check_filter(filter)
        prepared = self._prepare(as_vector(query, "query", dim=self._dim)[None, :])[0]
        if allowed is not None and not allowed.any():
            return np.full(k_valid, -1, dtype=np.int64), np.full(k_valid, np.inf)
        return self._engine.search(prepared, k_valid, ef_valid, allowed)

    def batch_search(self, queries, k: int = 10, ef: int | None = None, filter=None):
        """:meth:`search` for every row of ``queries``; one filter for every query."""
        k_valid, ef_valid = self._check_search_args(k, ef)
        allowed = self._check_filter(filter)
        prepared = self._prepare(queries)
        if allowed is not None and not allowed.any():
            shape = (prepared.shape[0], k_valid)
            return np.full(shape, -1, dtype=np.int64), np.full(shape, np.inf)
        return self._engine.batch_search(prepared, k_valid, ef_valid, allowed)

    def batch_search_planned(self, queries, k: int = 10, ef: int | None = None, filter=None,
                             planner=None):
        """Extension: :meth:`batch_search` with a filter planner (``planner.FilterPlanner``,
        default thresholds when None). Returns ``(ids, dists, plan)``: the in-graph filtered
        search (the reference's results) when ``plan.path == "graph"``, the exact top-k over
        the allowed vectors when ``"exact"``. Without a filter this is :meth:`batch_search`."""
        from .planner import FilterPlan, FilterPlanner

        k_valid, ef_valid = self._check_search_args(k, ef)
        allowed = self._check_filter(filter)
        prepared = self._prepare(queries)
        shape = (prepared.shape[0], k_valid)
        if allowed is None:
            ids, dists = self._engine.batch_search(prepared, k_valid, ef_valid, None)
            return ids, dists, FilterPlan("graph", 1.0, 1.0, 0.0)
        if not allowed.any():
            return (np.full(shape, -1, dtype=np.int64), np.full(shape, np.inf),
                    FilterPlan("exact", 0.0, 1.0, 0.0))
        plan = (planner or FilterPlanner()).plan(self._engine, allowed)
        if plan.path == "graph":
            ids, dists = self._engine.batch_search(prepared, k_valid, ef_valid, allowed)
            return ids, dists, plan
        kk = min(k_valid, int(allowed.sum()))
        got_ids, got_d = self._engine.exact_search(prepared, kk, allowed)
        ids = np.full(shape, -1, dtype=np.int64)
        dists = np.full(shape, np.inf)
        ids[:, :kk] = got_ids
        dists[:, :kk] = got_d
        return ids, dists, plan

    def batch_search_mixed(self, queries, k, ef, filters=None, filter_index=None):
        """Extension: one batch with per-query ``k`` and ``ef`` (arrays) and per-query filters
        (``filters``: a pool of boolean rows; ``filter_index[i]`` picks query i's row, -1 for
        none). Each query follows exactly the reference semantics of its own arguments.
        Returns ``(ids, dists)`` of shape ``(nq, max k)``; a query's row is padded past its k."""
        if len(self) == 0:
            raise ValueError("cannot search an empty index")
        prepared = self._prepare(queries)
        nq = prepared.shape[0]
        k_arr = np.broadcast_to(np.asarray(k, dtype=np.int64), (nq,))
        ef_arr = np.broadcast_to(np.asarray(ef, dtype=np.int64), (nq,))
        if nq and (k_arr.min() < 1 or k_arr.max() > len(self)):
            raise ValueError("every k must be in [1, len(index)]")
        if nq and ef_arr.min() < 1:
            raise ValueError("every ef must be >= 1")
        pool = None
        index = None
        if filters is not None:
            pool = np.asarray(filters)
            if pool.dtype != np.bool_:
                raise TypeError(f"filter must be a boolean array, got dtype {pool.dtype}")
            pool = np.atleast_2d(pool)
            if pool.shape[1] != len(self):
                raise ValueError("every filter must have one entry per indexed vector")
            index = np.broadcast_to(np.asarray(filter_index, dtype=np.int64), (nq,)).copy()
            if (index >= pool.shape[0]).any():
                raise ValueError("filter_index out of range")
            empty = ~pool.any(axis=1)
            index[(index >= 0) & empty[np.maximum(index, 0)]] = -2  # all-false: padding only
        ids = np.full((nq, int(k_arr.max()) if nq else 1), -1, dtype=np.int64)
        dists = np.full(ids.shape, np.inf)
        run = np.ones(nq, dtype=bool) if index is None else index != -2
        if run.any():
            sel = np.flatnonzero(run)
            got_ids, got_d, _, _ = self._engine.search_many(
                prepared[sel], k_arr[sel], np.maximum(ef_arr[sel], k_arr[sel]),
                mask_pool=pool, mask_index=None if index is None else index[sel],
                kstride=ids.shape[1],
            )
            ids[sel] = got_ids
            dists[sel] = got_d
        return ids, dists

    def _check_search_args(self, k: int, ef: int | None) -> tuple[int, int]:
        if len(self) == 0:
            raise ValueError("cannot search an empty index")
        k_valid = check_k(k, len(self))
        ef_valid = self._ef_search if ef is None else check_int(ef, "ef", minimum=1)
        return k_valid, max(ef_valid, k_valid)

    def _check_filter(self, filter) -> np.ndarray | None:
        if filter is None:
            return None
        mask = np.asarray(filter)
        if mask.dtype != np.bool_:
            raise TypeError(f"filter must be a boolean array, got dtype {mask.dtype}")
        if mask.shape != (len(self),):
            raise ValueError(
                f"filter must have one entry per indexed vector ({len(self)}), got shape "
                f"{mask.shape}"
            )
        return np.ascontiguousarray(mask)

    # -- inspection -----------------------------------------------------------------------

    def graph_state(self) -> GraphState:
        return self._engine.export_state()

    def neighbors(self, node: int, layer: int = 0) -> list[int]:
        state = self.graph_state()
        node = check_int(node, "node", minimum=0)
        layer = check_int(layer, "layer", minimum=0)
        if node >= state.size or layer > int(state.levels[node]):
            raise ValueError(f"node {node} is not present on layer {layer}")
        return state.adjacency(layer)[node]

    def stats(self) -> IndexStats:
        state = self.graph_state()
        nodes = tuple(int((state.levels >= layer).sum()) for layer in range(state.num_layers))
        links = tuple(int(state.indices[layer].shape[0]) for layer in range(state.num_layers))
        degrees = tuple(lk / nd if nd else 0.0 for lk, nd in zip(links, nodes))
        link_slots = sum(
            nd * ((2 * self._M if layer == 0 else self._M) + 1) for layer, nd in enumerate(nodes)
        )
        memory = state.size * self._dim * 4 + link_slots * 4 + state.size * 4
        return IndexStats(
            size=state.size, dim=self._dim, metric=self._metric, engine=self._engine_name,
            M=self._M, max_links_layer0=2 * self._M, ef_construction=self._ef_construction,
            ef_search=self._ef_search, m_L=self._m_L, selection=self._selection,
            extend_candidates=self._extend, keep_pruned_connections=self._keep_pruned,
            max_level=state.max_level, entry_point=state.entry_point, nodes_per_layer=nodes,
            links_per_layer=links, mean_degree_per_layer=degrees, memory_bytes=memory,
            distance_computations=self._engine.distance_computations, hops=self._engine.hops,
        )

    @property
    def distance_computations(self) -> int:
        return self._engine.distance_computations

    @property
    def hops(self) -> int:
        return self._engine.hops

    def reset_counters(self) -> None:
        self._engine.distance_computations = 0
        self._engine.hops = 0

    # -- persistence (the reference's native format) ---------------------------------------

    def save(self, path: str | Path) -> None:
        """The index (parameters, generator state, graph) as a ``.npz`` snapshot at ``path``
        (the reference's ``HNSWIndex.save``, ``index.py:369-372``; same file format)."""
        meta = {"params": self._params(), "rng_state": self._rng.bit_generator.state}
        save_state(path, meta, self.graph_state())

    @classmethod
    def load(cls, path: str | Path, engine: str | None = None) -> "HNSWIndex":
        """An index written by :meth:`save` (or by the reference), validated; ``engine``
        defaults to ``"mi355x"`` (the reference's engine names are accepted and mapped to it)."""
        meta, state = load_state(path)
        params = dict(meta["params"])
        params["engine"] = "mi355x" if engine is None else engine
        index = cls(**params)
        if state.vectors.shape[1:] != (index.dim,):
            raise ValueError(
                f"stored vectors have shape {state.vectors.shape}, expected dim {index.dim}"
            )
        index._rng.bit_generator.state = meta["rng_state"]
        index._engine.load_state(state)
        return index

    def save_native(self, directory: str | Path) -> None:
        """Write the native arrays and ``header.json`` (``index.py:390-403``)."""
        meta = {
            "params": self._params(),
            "rng_state": self._rng.bit_generator.state,
            "entry_point": self._engine.entry_point,
            "max_level": self._engine.max_level,
        }
        save_native(directory, meta, self._engine.native_arrays())

    @classmethod
    def load_native(cls, directory: str | Path, mmap: bool = True, validate: bool = True,
                    device: int = 0) -> "HNSWIndex":
        """Open a directory written by the reference's (or this) ``save_native``
        (``index.py:405-425``); ``validate`` checks every link before the kernels use it."""
        meta, arrays = load_native(directory, mmap=mmap)
        params = {**meta["params"], "engine": "mi355x"}
        index = cls(**params, device=device)
        index._rng.bit_generator.state = meta["rng_state"]
        entry, top = int(meta["entry_point"]), int(meta["max_level"])
        check_native_shapes(arrays, index.dim, index.M)
        if validate:
            validate_native(arrays, entry, top, index.M)
        index._engine.attach_native(arrays, entry, top)
        return index

    def _params(self) -> dict[str, Any]:
        return {
            "dim": self._dim, "metric": self._metric, "M": self._M,
            "ef_construction": self._ef_construction, "ef_search": self._ef_search,
            "seed": self._seed, "selection": self._selection,
            "extend_candidates": self._extend, "keep_pruned_connections": self._keep_pruned,
            "m_L": self._m_L, "engine": self._engine_name,
        }


def _check_bool(value: object, name: str) -> bool:
    if not isinstance(value, bool):
        raise TypeError(f"{name} must be a bool, got {type(value).__name__}")
    return value


def _check_m_l(value: object) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise TypeError(f"m_L must be a number, got {type(value).__name__}")
    result = float(value)
    if not math.isfinite(result) or result < 0:
        raise ValueError(f"m_L must be finite and >= 0, got {value}")
    return result
