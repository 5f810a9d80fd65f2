"""Loading the committed golden fixtures (written by tests/golden/make_golden.py)."""

from __future__ import annotations

from pathlib import Path

import numpy as np

from hnsw_amd.graph_io import native_from_csr

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def csr_layers(fx: dict, prefix: str = ""):
    levels = fx[prefix + "levels"]
    layer_ptr = fx[prefix + "layer_ptr"]
    indptr_all, indices_all = fx[prefix + "indptr"], fx[prefix + "indices"]
    indptr, indices = [], []
    pos = 0
    for layer in range(len(layer_ptr) - 1):
        count = int((levels >= layer).sum())
        indptr.append(indptr_all[pos : pos + count + 1])
        pos += count + 1
        indices.append(indices_all[layer_ptr[layer] : layer_ptr[layer + 1]].astype(np.int64))
    return levels, indptr, indices


def graph(name: str):
    """``(fx, arrays, entry_point, max_level, metric, M)`` for a search fixture."""
    fx = load(name)
    M = int(fx["params"][0])
    levels, indptr, indices = csr_layers(fx)
    metric = str(fx["metric"])
    data = fx["data"].astype(np.float32)
    if metric == "cosine":
        as64 = data.astype(np.float64)
        data = (as64 / np.linalg.norm(as64, axis=1)[:, None]).astype(np.float32)
    arrays = native_from_csr(data, levels, indptr, indices, M)
    entry, top = (int(v) for v in fx["entry"])
    return fx, arrays, entry, top, metric, M


def prepared_queries(fx: dict) -> np.ndarray:
    q = fx["queries"].astype(np.float32)
    if str(fx["metric"]) == "cosine":
        q64 = q.astype(np.float64)
        q = (q64 / np.linalg.norm(q64, axis=1)[:, None]).astype(np.float32)
    return q


def tags(fx: dict) -> list[str]:
    return sorted(k[: -len("_k_ef")] for k in fx if k.endswith("_k_ef"))
