"""ctypes bindings to ``libhnsw_mi355x.so`` (declared in ``include/hnsw_mi355x.h``).

There is no CPU fallback: if the library is missing every call raises, so a result can only
come from the HIP kernels.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libhnsw_mi355x.so"

# Every entry point the header declares.
SYMBOLS = (
    "hm_open",
    "hm_close",
    "hm_check_graph",
    "hm_has_exact_u8",
    "hm_search",
    "hm_search_dev",
    "hm_search_graph_dev",
    "hm_search_dev_noalloc",
    "hm_queries_exact",
    "hm_attach_codes",
    "hm_search_codes",
    "hm_search_codes_rerank",
    "hm_rerank",
    "hm_exact",
    "hm_exact_dev",
    "hm_exact_timing",
    "hm_set_exact_tuning",
    "hm_index_exact",
    "hm_index_exact_dev",
    "hm_build",
    "hm_model_size",
    "hm_set_stop_model",
    "hm_set_stop_cap",
    "hm_trace_dev",
    "hm_trace",
    "hm_mlp_grad",
    "hm_mlp_forward",
    "hm_mlp_train",
    "hm_set_tuning",
    "hm_read_stats",
    "hm_last_error",
)

HM_EINVAL, HM_EDEVICE, HM_ENOMEM = -1, -2, -3
METRIC_CODES = {"l2": 0, "ip": 1, "cosine": 2}


class NativeGraph(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("dim", ctypes.c_int32),
        ("M", ctypes.c_int32),
        ("metric", ctypes.c_int32),
        ("max_level", ctypes.c_int32),
        ("entry_point", ctypes.c_int64),
        ("rows", ctypes.c_int64),
        ("vectors", ctypes.c_void_p),
        ("levels", ctypes.c_void_p),
        ("links0", ctypes.c_void_p),
        ("count0", ctypes.c_void_p),
        ("upper_offset", ctypes.c_void_p),
        ("links_up", ctypes.c_void_p),
        ("count_up", ctypes.c_void_p),
    ]


class SearchOpts(ctypes.Structure):
    _fields_ = [
        ("nq", ctypes.c_int64),
        ("k", ctypes.c_void_p),
        ("ef", ctypes.c_void_p),
        ("k_all", ctypes.c_int32),
        ("ef_all", ctypes.c_int32),
        ("kstride", ctypes.c_int32),
        ("mask_pool", ctypes.c_void_p),
        ("n_masks", ctypes.c_int64),
        ("mask_index", ctypes.c_void_p),
        ("mask_all", ctypes.c_int32),
        ("learned", ctypes.c_int32),
    ]


_lib: ctypes.CDLL | None = None


class DeviceError(RuntimeError):
    """A HIP error reported by the library."""


def lib() -> ctypes.CDLL:
    """The loaded library (raises if it was not built: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH.name} is not built; run `python hnsw-engine_amd/build.py` "
                "(the search path has no CPU fallback)"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        sig = {
            "hm_open": ([ctypes.POINTER(NativeGraph), ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
            "hm_close": ([vp], None),
            "hm_check_graph": ([ctypes.POINTER(NativeGraph)], ctypes.c_int),
            "hm_has_exact_u8": ([vp], ctypes.c_int),
            "hm_search": ([vp, vp, ctypes.POINTER(SearchOpts), vp, vp, vp, vp], ctypes.c_int),
            "hm_search_dev": (
                [vp, vp, ctypes.POINTER(SearchOpts), vp, ctypes.c_int, vp, vp, vp, vp, vp, vp],
                ctypes.c_int,
            ),
            "hm_search_graph_dev": (
                [vp, vp, ctypes.POINTER(SearchOpts), vp, ctypes.c_int, vp, vp, vp, vp, vp, vp],
                ctypes.c_int,
            ),
            "hm_search_dev_noalloc": (
                [vp, vp, ctypes.POINTER(SearchOpts), vp, ctypes.c_int, vp, vp, vp, vp, vp, vp],
                ctypes.c_int,
            ),
            "hm_queries_exact": ([vp, i64], ctypes.c_int),
            "hm_attach_codes": ([vp, vp], ctypes.c_int),
            "hm_search_codes": ([vp, vp, ctypes.POINTER(SearchOpts), vp, vp, vp, vp], ctypes.c_int),
            "hm_search_codes_rerank": ([vp, vp, vp, ctypes.POINTER(SearchOpts), i32, vp, vp, vp,
                                        vp], ctypes.c_int),
            "hm_rerank": ([vp, vp, i64, vp, i32, i32, vp, vp], ctypes.c_int),
            "hm_exact": ([vp, i64, i32, i32, vp, i64, i32, vp, ctypes.c_int, vp, vp], ctypes.c_int),
            "hm_exact_dev": ([vp, i64, i32, i32, vp, i64, i32, vp, ctypes.c_int, vp, vp, vp],
                             ctypes.c_int),
            "hm_exact_timing": ([vp], ctypes.c_int),
            "hm_set_exact_tuning": ([i32, i32], ctypes.c_int),
            "hm_index_exact": ([vp, vp, i64, i32, vp, vp, vp], ctypes.c_int),
            "hm_index_exact_dev": ([vp, vp, i64, i32, vp, ctypes.c_int, vp, vp], ctypes.c_int),
            "hm_build": (
                [i64, i64, i32, i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp],
                ctypes.c_int,
            ),
            "hm_model_size": ([], ctypes.c_int),
            "hm_set_stop_model": ([vp, vp, ctypes.c_float, i32], ctypes.c_int),
            "hm_set_stop_cap": ([vp, i32], ctypes.c_int),
            "hm_trace_dev": ([vp, vp, i64, i32, i32, ctypes.c_int, vp, vp, i32, vp, vp, vp], ctypes.c_int),
            "hm_trace": ([vp, vp, i64, i32, i32, vp, vp, i32, vp, vp], ctypes.c_int),
            "hm_mlp_grad": ([vp, vp, i64, vp, ctypes.c_float, vp, vp], ctypes.c_int),
            "hm_mlp_forward": ([vp, i64, vp, vp], ctypes.c_int),
            "hm_mlp_train": ([vp, vp, i64, vp, i32, i32, ctypes.c_float, ctypes.c_float,
                              ctypes.c_uint64, vp], ctypes.c_int),
            "hm_set_tuning": ([vp, i32, i32, i32, i32, i32, i32], ctypes.c_int),
            "hm_read_stats": ([vp, vp, ctypes.c_int], ctypes.c_int),
            "hm_last_error": ([], ctypes.c_char_p),
        }
        for name, (args, res) in sig.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def check(status: int) -> None:
    """Raise the reference's exception type for a failed call."""
    if status == 0:
        return
    msg = (lib().hm_last_error() or b"").decode()
    if status == HM_EINVAL:
        raise ValueError(msg)
    if status == HM_ENOMEM:
        raise MemoryError(msg)
    raise DeviceError(msg)


def ptr(a: np.ndarray | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if a is None else a.ctypes.data)


def graph_struct(arrays: dict, entry_point: int, max_level: int, M: int, metric: str) -> tuple:
    """``(NativeGraph, keep-alive arrays)`` for contiguous copies of the native arrays."""
    keep = {
        "vectors": np.ascontiguousarray(arrays["vectors"], dtype=np.float32),
        "levels": np.ascontiguousarray(arrays["levels"], dtype=np.int64),
        "links0": np.ascontiguousarray(arrays["links0"], dtype=np.int32),
        "count0": np.ascontiguousarray(arrays["count0"], dtype=np.int32),
        "upper_offset": np.ascontiguousarray(arrays["upper_offset"], dtype=np.int64),
        "links_up": np.ascontiguousarray(arrays["links_up"], dtype=np.int32),
        "count_up": np.ascontiguousarray(arrays["count_up"], dtype=np.int32),
    }
    n, dim = keep["vectors"].shape
    g = NativeGraph(
        n, dim, M, METRIC_CODES[metric], max_level, entry_point, keep["links_up"].shape[0],
        *(keep[k].ctypes.data if keep[k].size else None for k in (
            "vectors", "levels", "links0", "count0", "upper_offset", "links_up", "count_up")),
    )
    return g, keep
