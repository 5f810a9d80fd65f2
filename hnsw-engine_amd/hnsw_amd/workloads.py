"""The measured workloads: data, queries, levels and the graph, all from fixed seeds.

``clustered_1m`` is the headline workload: 1,000,000 x 128 float32 vectors with integer values
0..255 from 1,000 isotropic Gaussian clusters (centres uniform in [0, 255], sigma 30), 10,000
queries from the same centres, L2, M = 16, ef_construction = 200, heuristic selection, levels
``floor(-ln(1 - U) / ln M)`` from ``numpy.random.default_rng(level_seed)`` (the reference
facade's draw, ``levels.py:22-31``).

The graph is built by sequential insertion on the host (``hm_build``) and cached in the
reference's native layout under ``<repo>/.graph_cache/<name>/`` (build: several minutes for 1M
nodes); the vectors are regenerated from the seed, which is faster than reading them.
"""

from __future__ import annotations

import hashlib
import json
import math
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .engine import MI355XEngine
from .metrics import prepare
from .synthetic import (
    angular_heavy_tailed,
    clustered_u8,
    clustered_u8_queries,
    sparse_prototypes_u8,
)

CACHE = Path(__file__).resolve().parents[2] / ".graph_cache"


@dataclass(frozen=True)
class Workload:
    name: str
    n: int
    dim: int
    n_queries: int
    n_clusters: int
    sigma: float
    seed: int
    level_seed: int
    M: int = 16
    ef_construction: int = 200
    metric: str = "l2"
    kind: str = "clustered"  # or "sparse_prototypes" (n_clusters = prototypes), "angular"

    def data(self) -> tuple[np.ndarray, np.ndarray]:
        if self.kind == "sparse_prototypes":
            return sparse_prototypes_u8(self.n, self.n_queries, dim=self.dim,
                                        n_prototypes=self.n_clusters, seed=self.seed)
        if self.kind == "angular":  # prepared as the facade prepares cosine vectors
            data, queries = angular_heavy_tailed(self.n, self.dim, self.n_queries,
                                                 n_clusters=self.n_clusters, seed=self.seed)
            return prepare(data, self.metric), prepare(queries, self.metric)
        return clustered_u8(self.n, self.dim, self.n_queries, n_clusters=self.n_clusters,
                            sigma=self.sigma, seed=self.seed)

    def train_queries(self, count: int) -> np.ndarray:
        """Queries from the same clusters, disjoint from the measured ones (seed + 2)."""
        if self.kind != "clustered":
            raise ValueError(f"{self.name} has no training-query generator")
        return clustered_u8_queries(count, self.dim, n_clusters=self.n_clusters,
                                    sigma=self.sigma, seed=self.seed, query_seed=self.seed + 2)

    def levels(self) -> np.ndarray:
        rng = np.random.default_rng(self.level_seed)
        u = 1.0 - rng.random(self.n)
        return np.floor(-np.log(u) * (1.0 / math.log(self.M))).astype(np.int64)


WORKLOADS = {
    "clustered_1m": Workload("clustered_1m", 1_000_000, 128, 10_000, 1000, 30.0, 20260101, 7),
    "clustered_100k": Workload("clustered_100k", 100_000, 128, 1_000, 100, 30.0, 20260102, 7),
    "clustered_200k": Workload("clustered_200k", 200_000, 128, 2_000, 200, 30.0, 20260103, 7),
    # Filtered search: 100,000 x 784 8-bit vectors, about half exact zeros, 10 prototypes plus
    # uniform noise; 5,000 queries.
    "sparse784_100k": Workload("sparse784_100k", 100_000, 784, 5_000, 10, 0.0, 20260104, 7,
                               kind="sparse_prototypes"),
    # Quantized search: 400,000 x 100 L2-normalised vectors from 200 Gaussian clusters with
    # heavy-tailed (log-normal) per-dimension scales, cosine; 10,000 queries.
    "angular_400k": Workload("angular_400k", 400_000, 100, 10_000, 200, 0.0, 20260105, 7,
                             metric="cosine", kind="angular"),
}


def _digest(data: np.ndarray, levels: np.ndarray, w: Workload) -> str:
    h = hashlib.sha256()
    h.update(json.dumps(w.__dict__, sort_keys=True).encode())
    h.update(np.ascontiguousarray(data[:: max(1, data.shape[0] // 4096)]).tobytes())
    h.update(levels.tobytes())
    return h.hexdigest()[:16]


GRAPH_ARRAYS = ("levels", "links0", "count0", "upper_offset", "links_up", "count_up")


def is_cached(w: Workload) -> bool:
    """True when the graph of ``w`` is in the cache (it may still be stale: ``load_or_build``
    compares the digest)."""
    header = CACHE / w.name / "header.json"
    return header.is_file() and json.loads(header.read_text()).get("workload", {}).get("n") == w.n


def load_or_build(w: Workload, *, log=print) -> tuple[MI355XEngine, np.ndarray, np.ndarray]:
    """``(engine, data, queries)`` with the graph loaded from the cache or built (and cached)."""
    t0 = time.perf_counter()
    data, queries = w.data()
    levels = w.levels()
    digest = _digest(data, levels, w)
    root = CACHE / w.name
    header = root / "header.json"
    eng = MI355XEngine(w.dim, w.metric, w.M, w.ef_construction, "heuristic", False, False)
    if header.is_file() and json.loads(header.read_text()).get("digest") == digest:
        meta = json.loads(header.read_text())
        arrays = {name: np.load(root / f"{name}.npy", allow_pickle=False) for name in GRAPH_ARRAYS}
        arrays["vectors"] = data
        eng.attach_native(arrays, int(meta["entry_point"]), int(meta["max_level"]))
        log(f"[workload] {w.name}: graph loaded from cache in {time.perf_counter() - t0:.1f}s")
        return eng, data, queries
    log(f"[workload] {w.name}: building the graph on the host (sequential insertion)")
    chunk = max(1, w.n // 20)
    for start in range(0, w.n, chunk):
        eng.insert(data[start : start + chunk], levels[start : start + chunk])
        log(f"[workload] {w.name}: inserted {min(w.n, start + chunk)}/{w.n} "
            f"({time.perf_counter() - t0:.0f}s)")
    root.mkdir(parents=True, exist_ok=True)
    header.unlink(missing_ok=True)
    arrays = eng.native_arrays()
    for name in GRAPH_ARRAYS:
        np.save(root / f"{name}.npy", np.ascontiguousarray(arrays[name]), allow_pickle=False)
    header.write_text(json.dumps({"digest": digest, "entry_point": eng.entry_point,
                                  "max_level": eng.max_level, "workload": w.__dict__}, indent=2))
    return eng, data, queries


if __name__ == "__main__":  # python -m hnsw_amd.workloads <name> ...: build and cache graphs
    import sys

    for name in sys.argv[1:] or list(WORKLOADS):
        load_or_build(WORKLOADS[name], log=lambda m: print(m, flush=True))
