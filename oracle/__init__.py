"""CPU oracle for the search path: ctypes bindings to ``libhnsw_oracle.so`` (``hnsw_oracle.c``).

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` may import this package; the product library never does. Every function
restates the reference's algorithm (file:line cited in ``hnsw_oracle.c``):

* :func:`search_batch`   -- ``numba_engine.py:422-503`` (+ ``:151-198`` filtered, ``:674-701``
  codes), with per-query ``k``/``ef`` and masks;
* :func:`rerank_batch`   -- ``numba_engine.py:506-530``;
* :func:`exact_search`   -- ``oracle.py:35-126`` / ``benchmarks/run_filtered.py:76-106``.

Which check pins each function is listed in ``tests/test_oracle_c.py``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libhnsw_oracle.so"
_METRICS = {"l2": 0, "ip": 1, "cosine": 2}
_lib: ctypes.CDLL | None = None


class _Graph(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("dim", ctypes.c_int),
        ("entry_point", ctypes.c_int64),
        ("max_level", ctypes.c_int),
        ("links0", ctypes.c_void_p),
        ("count0", ctypes.c_void_p),
        ("upper_offset", ctypes.c_void_p),
        ("links_up", ctypes.c_void_p),
        ("count_up", ctypes.c_void_p),
        ("stride0", ctypes.c_int),
        ("stride_up", ctypes.c_int),
    ]


def build() -> None:
    """Compile ``libhnsw_oracle.so`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", str(_HERE)], check=True)


def _load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
    return _lib


def _ptr(a: np.ndarray | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if a is None else a.ctypes.data)


def threads_available() -> int:
    """Threads the oracle uses when asked for 'all cores'."""
    return len(os.sched_getaffinity(0))


def search_batch(
    arrays: dict,
    entry_point: int,
    max_level: int,
    metric: str,
    queries: np.ndarray,
    k,
    ef,
    *,
    mask: np.ndarray | None = None,
    mask_pool: np.ndarray | None = None,
    mask_index: np.ndarray | None = None,
    codes: np.ndarray | None = None,
    threads: int = 1,
):
    """Algorithm 5 for every query row on the CPU; returns ``(ids, dists, dcount, hops)``.

    ``k`` and ``ef`` are ints or per-query arrays. ``mask`` is one boolean mask for every query
    (the reference's ``filter``); ``mask_pool`` + ``mask_index`` give each query its own mask
    (``mask_index < 0``: unfiltered). With ``codes`` (uint8, one row per node) the traversal uses
    distances to the codes and ``queries`` must be float64 code-space queries.
    """
    lib = _load()
    vectors = np.ascontiguousarray(arrays["vectors"], dtype=np.float32)
    n, dim = vectors.shape
    links0 = np.ascontiguousarray(arrays["links0"], dtype=np.int32)
    count0 = np.ascontiguousarray(arrays["count0"], dtype=np.int32)
    upper = np.ascontiguousarray(arrays["upper_offset"], dtype=np.int64)
    links_up = np.ascontiguousarray(arrays["links_up"], dtype=np.int32)
    count_up = np.ascontiguousarray(arrays["count_up"], dtype=np.int32)
    if links_up.ndim != 2:
        links_up = links_up.reshape(0, 1)
    g = _Graph(
        n,
        dim,
        entry_point,
        max_level,
        links0.ctypes.data,
        count0.ctypes.data,
        upper.ctypes.data,
        links_up.ctypes.data if links_up.size else 0,
        count_up.ctypes.data if count_up.size else 0,
        links0.shape[1],
        links_up.shape[1],
    )
    if codes is not None:
        table = np.ascontiguousarray(codes, dtype=np.uint8)
        q = np.ascontiguousarray(queries, dtype=np.float64)
        vec_u8, q_f64 = 1, 1
    else:
        table = vectors
        q = np.ascontiguousarray(queries, dtype=np.float32)
        vec_u8, q_f64 = 0, 0
    nq = q.shape[0]
    k_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.int64), (nq,)))
    ef_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(ef, dtype=np.int64), (nq,)))
    kmax = int(k_arr.max()) if nq else int(np.max(k))
    if mask is not None:
        mask_pool = np.ascontiguousarray(mask, dtype=np.uint8)[None, :]
        mask_index = np.zeros(nq, dtype=np.int64)
    if mask_pool is not None:
        mask_pool = np.ascontiguousarray(mask_pool, dtype=np.uint8)
        mask_index = np.ascontiguousarray(mask_index, dtype=np.int64)
    ids = np.full((nq, kmax), -1, dtype=np.int64)
    dists = np.full((nq, kmax), np.inf, dtype=np.float64)
    dcount = np.zeros(nq, dtype=np.int64)
    hops = np.zeros(nq, dtype=np.int64)
    lib.ho_search_batch.restype = ctypes.c_int
    status = lib.ho_search_batch(
        ctypes.byref(g),
        _ptr(table),
        ctypes.c_int(vec_u8),
        ctypes.c_int(_METRICS[metric]),
        _ptr(q),
        ctypes.c_int(q_f64),
        ctypes.c_int64(nq),
        _ptr(k_arr),
        _ptr(ef_arr),
        ctypes.c_int64(kmax),
        _ptr(mask_pool),
        _ptr(mask_index),
        _ptr(ids),
        _ptr(dists),
        _ptr(dcount),
        _ptr(hops),
        ctypes.c_int(threads),
    )
    if status != 0:
        raise MemoryError("oracle search ran out of memory")
    return ids, dists, dcount, hops


def rerank_batch(vectors, metric: str, queries, candidates, k: int, threads: int = 1):
    """Exact re-score of candidate ids, best ``k`` by (distance, id)."""
    lib = _load()
    v = np.ascontiguousarray(vectors, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    c = np.ascontiguousarray(candidates, dtype=np.int64)
    nq, fetch = c.shape
    ids = np.empty((nq, k), dtype=np.int64)
    dists = np.empty((nq, k), dtype=np.float64)
    lib.ho_rerank_batch(
        _ptr(v), ctypes.c_int(v.shape[1]), ctypes.c_int(_METRICS[metric]), _ptr(q),
        ctypes.c_int64(nq), _ptr(c), ctypes.c_int64(fetch), ctypes.c_int64(k), _ptr(ids),
        _ptr(dists), ctypes.c_int(threads),
    )
    return ids, dists


def exact_search(data, queries, k: int, metric: str = "l2", allowed=None, threads: int = 0):
    """Exact top-``k`` by (distance, id) over all (or the allowed) rows."""
    lib = _load()
    x = np.ascontiguousarray(data, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    mask = None if allowed is None else np.ascontiguousarray(allowed, dtype=np.uint8)
    nq = q.shape[0]
    ids = np.empty((nq, k), dtype=np.int64)
    dists = np.empty((nq, k), dtype=np.float64)
    lib.ho_exact_search(
        _ptr(x), ctypes.c_int64(x.shape[0]), ctypes.c_int(x.shape[1]),
        ctypes.c_int(_METRICS[metric]), _ptr(q), ctypes.c_int64(nq), ctypes.c_int64(k),
        _ptr(mask), _ptr(ids), _ptr(dists), ctypes.c_int(threads),
    )
    return ids, dists
