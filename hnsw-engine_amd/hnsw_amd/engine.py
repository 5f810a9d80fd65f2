"""``MI355XEngine``: the reference's ``Engine`` protocol (``index.py:32-60``) on the GPU.

Besides the protocol (``insert``, ``search``, ``batch_search``, counters, ``export_state`` /
``load_state``) it provides what the reference's Numba engine adds and its facades call:
``native_arrays`` / ``attach_native`` (``numba_engine.py:814-868``), ``batch_search_codes``
(``:674-701``), ``rerank`` (``:703-713``) and ``memory_bytes`` (``:737-749``).

The graph lives twice: as host arrays in the reference's native layout (so ``save_native``
writes the reference's files and insertion runs the host builder on them) and in HBM, uploaded
on first search after any change. Searches always run the HIP kernels.
"""

from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from . import _native
from .graph_io import (
    STORAGE,
    GraphState,
    check_native_graph,
    check_native_layout,
    native_from_csr,
    state_from_native,
)

_INITIAL_CAPACITY = 1024


def _grow(array: np.ndarray, rows: int, fill: float) -> np.ndarray:
    out = np.full((rows, *array.shape[1:]), fill, dtype=array.dtype)
    out[: array.shape[0]] = array
    return out


class MI355XEngine:
    """Batched HNSW search on one MI355X; graph construction on the host."""

    name = "mi355x"

    def __init__(
        self,
        dim: int,
        metric: str,
        M: int,
        ef_construction: int,
        selection: str,
        extend_candidates: bool,
        keep_pruned_connections: bool,
        device: int = 0,
    ) -> None:
        self.dim = dim
        self.metric = metric
        self.M = M
        self.M0 = 2 * M
        self.ef_construction = ef_construction
        self.selection = selection
        self.extend_candidates = extend_candidates
        self.keep_pruned_connections = keep_pruned_connections
        self.device = device
        self._size = 0
        self._upper_rows = 0
        self._vectors = np.empty((_INITIAL_CAPACITY, dim), dtype=np.float32)
        self._levels = np.zeros(_INITIAL_CAPACITY, dtype=np.int64)
        self._links0 = np.full((_INITIAL_CAPACITY, self.M0 + 1), -1, dtype=np.int32)
        self._count0 = np.zeros(_INITIAL_CAPACITY, dtype=np.int32)
        self._upper_offset = np.full(_INITIAL_CAPACITY, -1, dtype=np.int64)
        self._links_up = np.full((max(_INITIAL_CAPACITY // 8, 1), M + 1), -1, dtype=np.int32)
        self._count_up = np.zeros(self._links_up.shape[0], dtype=np.int32)
        self._state = np.array([-1, -1], dtype=np.int64)  # entry point, max level
        self._counters = np.zeros(2, dtype=np.int64)  # distance computations, hops
        self._handle: ctypes.c_void_p | None = None
        self._codes: np.ndarray | None = None

    # -- protocol ------------------------------------------------------------------------

    @property
    def size(self) -> int:
        return self._size

    @property
    def entry_point(self) -> int:
        return int(self._state[0])

    @property
    def max_level(self) -> int:
        return int(self._state[1])

    @property
    def distance_computations(self) -> int:
        return int(self._counters[0])

    @distance_computations.setter
    def distance_computations(self, value: int) -> None:
        self._counters[0] = value

    @property
    def hops(self) -> int:
        return int(self._counters[1])

    @hops.setter
    def hops(self, value: int) -> None:
        self._counters[1] = value

    def insert(self, vectors: np.ndarray, levels: np.ndarray) -> None:
        """Sequential insertion on the host (the reference's ``_insert_batch``)."""
        self._ensure_writable()
        count = vectors.shape[0]
        start, stop = self._size, self._size + count
        self._reserve_nodes(stop)
        self._vectors[start:stop] = vectors
        self._levels[start:stop] = levels
        self._reserve_upper_rows(self._upper_rows + int(levels.sum()))
        for offset, level in enumerate(levels.tolist()):
            if level > 0:
                self._upper_offset[start + offset] = self._upper_rows
                self._upper_rows += level
        dc = np.zeros(1, dtype=np.int64)
        sel = 0 if self.selection == "simple" else 1
        _native.check(
            _native.lib().hm_build(
                start, stop, self.dim, _native.METRIC_CODES[self.metric], self.M,
                self.ef_construction, sel, int(self.extend_candidates),
                int(self.keep_pruned_connections),
                _native.ptr(self._vectors), _native.ptr(self._levels), _native.ptr(self._links0),
                _native.ptr(self._count0), _native.ptr(self._upper_offset),
                _native.ptr(self._links_up), _native.ptr(self._count_up),
                _native.ptr(self._state), _native.ptr(dc),
            )
        )
        self._counters[0] += int(dc[0])
        self._size = stop
        self._invalidate()

    def search(self, query: np.ndarray, k: int, ef: int, allowed: np.ndarray | None = None):
        ids, dists = self.batch_search(query[None, :], k, ef, allowed)
        return ids[0], dists[0]

    def batch_search(self, queries: np.ndarray, k: int, ef: int, allowed: np.ndarray | None = None):
        """Algorithm 5 for every row on the GPU; ``(nq, k)`` ids (int64) and dists (float64)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        masks = None if allowed is None else np.ascontiguousarray(allowed, dtype=np.bool_)[None, :]
        ids, dists, dc, hops = self.search_many(q, k, ef, mask_pool=masks,
                                                mask_index=None if masks is None else 0)
        return ids, dists

    def search_many(
        self,
        queries: np.ndarray,
        k,
        ef,
        *,
        mask_pool: np.ndarray | None = None,
        mask_index=None,
        kstride: int | None = None,
        learned: bool = False,
    ):
        """Batch search with per-query ``k``/``ef`` (ints or arrays) and per-query masks
        (``mask_pool`` rows, ``mask_index`` per query or one int; -1: unfiltered). Returns
        ``(ids, dists, dcount, hops)`` with per-query counters; the cumulative counters grow
        as the reference's do."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        opts, keep, kmax = self._opts(nq, k, ef, mask_pool, mask_index, kstride)
        opts.learned = int(bool(learned))
        ids = np.empty((nq, kmax), dtype=np.int64)
        dists = np.empty((nq, kmax), dtype=np.float64)
        dc = np.zeros(nq, dtype=np.int64)
        hops = np.zeros(nq, dtype=np.int64)
        if nq:
            handle = self._device()
            _native.check(
                _native.lib().hm_search(
                    handle, _native.ptr(q), ctypes.byref(opts), _native.ptr(ids),
                    _native.ptr(dists), _native.ptr(dc), _native.ptr(hops),
                )
            )
        self._count(dc, hops)
        return ids, dists, dc, hops

    def _attach_codes(self, codes: np.ndarray) -> ctypes.c_void_p:
        if codes.shape != (self._size, self.dim) or codes.dtype != np.uint8:
            raise ValueError(f"codes must be uint8 with shape {(self._size, self.dim)}")
        handle = self._device()
        if self._codes is not codes:
            c = np.ascontiguousarray(codes)
            _native.check(_native.lib().hm_attach_codes(handle, _native.ptr(c)))
            self._codes = codes
        return handle

    def _count(self, dc: np.ndarray, hops: np.ndarray) -> None:
        self._counters[0] += int(dc.sum())
        self._counters[1] += int(hops.sum())

    def batch_search_codes(self, codes: np.ndarray, queries: np.ndarray, k: int, ef: int):
        """Algorithm 5 over uint8 ``codes`` with float64 code-space queries (no filter)."""
        handle = self._attach_codes(codes)
        q = np.ascontiguousarray(queries, dtype=np.float64)
        nq = q.shape[0]
        opts, keep, kmax = self._opts(nq, k, ef, None, None, None)
        ids = np.empty((nq, kmax), dtype=np.int64)
        dists = np.empty((nq, kmax), dtype=np.float64)
        dc = np.zeros(nq, dtype=np.int64)
        hops = np.zeros(nq, dtype=np.int64)
        if nq:
            _native.check(
                _native.lib().hm_search_codes(
                    handle, _native.ptr(q), ctypes.byref(opts), _native.ptr(ids),
                    _native.ptr(dists), _native.ptr(dc), _native.ptr(hops),
                )
            )
        self._count(dc, hops)
        return ids, dists

    def quantized_search(self, codes: np.ndarray, code_queries: np.ndarray, queries: np.ndarray,
                         k: int, fetch: int, ef: int):
        """The codes traversal keeping ``fetch`` candidates, then their exact re-rank, in one
        device call (``hm_search_codes_rerank``): the best ``k`` ids and float64 distances."""
        handle = self._attach_codes(codes)
        cq = np.ascontiguousarray(code_queries, dtype=np.float64)
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        opts, keep, _ = self._opts(nq, fetch, ef, None, None, None)
        ids = np.empty((nq, k), dtype=np.int64)
        dists = np.empty((nq, k), dtype=np.float64)
        dc = np.zeros(nq, dtype=np.int64)
        hops = np.zeros(nq, dtype=np.int64)
        if nq:
            _native.check(_native.lib().hm_search_codes_rerank(
                handle, _native.ptr(cq), _native.ptr(q), ctypes.byref(opts), int(k),
                _native.ptr(ids), _native.ptr(dists), _native.ptr(dc), _native.ptr(hops)))
        self._count(dc, hops)
        return ids, dists

    def rerank(self, queries: np.ndarray, candidates: np.ndarray, k: int):
        """Exact float64 re-score of each row's candidates, best ``k`` by (distance, id)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        c = np.ascontiguousarray(candidates, dtype=np.int64)
        nq, fetch = c.shape
        ids = np.empty((nq, k), dtype=np.int64)
        dists = np.empty((nq, k), dtype=np.float64)
        if nq:
            _native.check(
                _native.lib().hm_rerank(
                    self._device(), _native.ptr(q), nq, _native.ptr(c), fetch, k,
                    _native.ptr(ids), _native.ptr(dists),
                )
            )
        return ids, dists

    def set_stop_model(self, params: np.ndarray | None, threshold: float = 0.5,
                       min_steps: int = 0) -> None:
        """Install (or remove, with ``None``) the learned-stop model used by ``learned=True``
        searches."""
        handle = self._device()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        _native.check(_native.lib().hm_set_stop_model(handle, _native.ptr(p), float(threshold),
                                                      int(min_steps)))
        self._stop = (p, float(threshold), int(min_steps))

    def set_stop_cap(self, max_steps: int) -> None:
        """Cap stopped beams at ``max_steps`` layer-0 expansions (0: no cap)."""
        _native.check(_native.lib().hm_set_stop_cap(self._device(), int(max_steps)))

    def exact_search(self, queries: np.ndarray, k: int, allowed: np.ndarray | None = None):
        """Exact k-NN (ids, float64 distances) over the vectors in HBM, optionally restricted to
        the allowed rows (the masked brute force, ``benchmarks/run_filtered.py:76-106``)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), dtype=np.int64)
        dists = np.empty((nq, k), dtype=np.float64)
        mask = None if allowed is None else np.ascontiguousarray(allowed, dtype=np.uint8)
        _native.check(_native.lib().hm_index_exact(self._device(), _native.ptr(q), nq, int(k),
                                                   _native.ptr(mask), _native.ptr(ids),
                                                   _native.ptr(dists)))
        return ids, dists

    def trace(self, queries: np.ndarray, k: int, ef: int, thresholds: np.ndarray, steps: int = 128):
        """Plain register-beam search recording the stop features per layer-0 expansion.

        Returns ``(trace (nq, steps, 18), length (nq,), final_hits (nq,))``; ``thresholds`` is
        each query's exact k-th distance plus the recall tolerance.
        """
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        thr = np.ascontiguousarray(thresholds, dtype=np.float64)
        tr = np.zeros((nq, steps, 18), dtype=np.float32)
        ln = np.zeros(nq, dtype=np.int32)
        fin = np.zeros(nq, dtype=np.int32)
        _native.check(_native.lib().hm_trace(
            self._device(), _native.ptr(q), nq, int(k), int(ef), _native.ptr(thr),
            _native.ptr(tr), int(steps), _native.ptr(ln), _native.ptr(fin)))
        return tr, ln, fin

    # -- state ------------------------------------------------------------------------------

    def native_arrays(self) -> dict[str, np.ndarray]:
        n, rows = self._size, self._upper_rows
        return {
            "vectors": self._vectors[:n],
            "levels": self._levels[:n],
            "links0": self._links0[:n],
            "count0": self._count0[:n],
            "upper_offset": self._upper_offset[:n],
            "links_up": self._links_up[:rows],
            "count_up": self._count_up[:rows],
        }

    def attach_native(self, arrays: dict[str, np.ndarray], entry_point: int, max_level: int,
                      *, validate: bool = True) -> None:
        """Use native arrays (possibly read-only maps) in place on the host side, as the
        reference's ``attach_native`` does (``numba_engine.py:827-868``): dtypes and shapes are
        always checked, the links too unless ``validate`` is False (files you wrote yourself);
        the first insertion copies them, so the arrays are never written."""
        check_native_layout(arrays, self.dim, self.M)
        if validate:
            check_native_graph(arrays, entry_point, max_level, self.M, self.metric)
        for name in STORAGE:
            setattr(self, f"_{name}", arrays[name])
        self._size = arrays["vectors"].shape[0]
        self._upper_rows = arrays["links_up"].shape[0]
        self._state[:] = (entry_point, max_level)
        self._invalidate()

    def export_state(self) -> GraphState:
        return state_from_native(self.native_arrays(), self.entry_point, self.max_level)

    def load_state(self, state: GraphState) -> None:
        arrays = native_from_csr(state.vectors, state.levels, list(state.indptr),
                                 list(state.indices), self.M)
        for name in STORAGE:
            setattr(self, f"_{name}", arrays[name])
        self._size = state.size
        self._upper_rows = arrays["links_up"].shape[0]
        self._state[:] = (state.entry_point, state.max_level)
        self._invalidate()

    def memory_bytes(self) -> int:
        arrays = (self._vectors, self._levels, self._links0, self._count0, self._upper_offset,
                  self._links_up, self._count_up)
        return int(sum(a.nbytes for a in arrays))

    def close(self) -> None:
        if self._handle is not None:
            _native.lib().hm_close(self._handle)
            self._handle = None

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass

    # -- internals --------------------------------------------------------------------------

    def _opts(self, nq, k, ef, mask_pool, mask_index, kstride):
        keep: dict[str, Any] = {}
        k_arr = np.asarray(k)
        ef_arr = np.asarray(ef)
        per_query = k_arr.ndim > 0 or ef_arr.ndim > 0
        kmax = int(k_arr.max()) if k_arr.size else 1
        if kstride is not None:
            kmax = max(kmax, int(kstride))
        opts = _native.SearchOpts()
        opts.nq = nq
        opts.kstride = kmax
        opts.mask_all = -1
        if per_query:
            keep["k"] = np.ascontiguousarray(np.broadcast_to(k_arr, (nq,)), dtype=np.int32)
            keep["ef"] = np.ascontiguousarray(np.broadcast_to(ef_arr, (nq,)), dtype=np.int32)
            opts.k = keep["k"].ctypes.data
            opts.ef = keep["ef"].ctypes.data
        else:
            opts.k_all = int(k_arr)
            opts.ef_all = int(ef_arr)
        if mask_pool is not None:
            pool = np.ascontiguousarray(mask_pool, dtype=np.uint8)
            if pool.ndim == 1:
                pool = pool[None, :]
            keep["pool"] = pool
            opts.mask_pool = pool.ctypes.data
            opts.n_masks = pool.shape[0]
            mi = np.asarray(mask_index if mask_index is not None else 0)
            if mi.ndim == 0:
                opts.mask_all = int(mi)
            else:
                keep["mi"] = np.ascontiguousarray(mi, dtype=np.int32)
                opts.mask_index = keep["mi"].ctypes.data
        opts._keep = keep  # keep the arrays alive as long as the struct
        return opts, keep, kmax

    def _device(self) -> ctypes.c_void_p:
        if self._handle is None:
            if self._size == 0:
                raise ValueError("cannot search an empty index")
            g, keep = _native.graph_struct(
                self.native_arrays(), self.entry_point, self.max_level, self.M, self.metric
            )
            handle = ctypes.c_void_p()
            _native.check(_native.lib().hm_open(ctypes.byref(g), self.device, ctypes.byref(handle)))
            self._handle = handle
            self._codes = None
            stop = getattr(self, "_stop", None)
            if stop is not None and stop[0] is not None:
                _native.check(_native.lib().hm_set_stop_model(
                    handle, _native.ptr(stop[0]), stop[1], stop[2]))
        return self._handle

    def _invalidate(self) -> None:
        self.close()
        self._codes = None

    def _reserve_nodes(self, capacity: int) -> None:
        current = self._vectors.shape[0]
        if capacity <= current:
            return
        new = max(capacity, 2 * current)
        self._vectors = _grow(self._vectors, new, 0.0)
        self._levels = _grow(self._levels, new, 0)
        self._links0 = _grow(self._links0, new, -1)
        self._count0 = _grow(self._count0, new, 0)
        self._upper_offset = _grow(self._upper_offset, new, -1)

    def _reserve_upper_rows(self, rows: int) -> None:
        current = self._links_up.shape[0]
        if rows <= current:
            return
        new = max(rows, 2 * current)
        self._links_up = _grow(self._links_up, new, -1)
        self._count_up = _grow(self._count_up, new, 0)

    def _ensure_writable(self) -> None:
        for name in STORAGE:
            array = getattr(self, f"_{name}")
            if not array.flags.writeable:
                setattr(self, f"_{name}", np.array(array))
