"""Filter planner: in-graph filtered search or masked brute force, per batch (an extension; the
reference measured both paths but ships no planner, ``docs/STUDIES.md:255-287``,
``docs/REPORT.md:153``).

The two paths answer different questions exactly: the in-graph path is the reference's
``_search_layer0_filtered`` (``numba_engine.py:151-198``), bit for bit, and the brute force is
the exact top-k over the allowed rows (``benchmarks/run_filtered.py:76-106``), so a planned
search returns whichever the cheaper path computes. In-graph work grows as the mask tightens
(the beam must expand about 1/selectivity nodes to meet k allowed ones) while brute-force work
shrinks with the allowed rows, so below a crossover selectivity the brute force is both faster
and exact.

The planner estimates two statistics of a mask:

* selectivity ``s``: the allowed fraction;
* correlation ``c``: among sampled allowed nodes, the allowed fraction of their layer-0
  neighbours, divided by ``s`` (1 for an i.i.d. mask; much more when allowed nodes cluster in
  the graph, where a query outside the cluster walks a long way through disallowed nodes).

and picks the brute force when ``s`` is below the threshold for the mask's kind (``c`` below or
above ``correlation_cut``). :func:`calibrate` sets both thresholds from timings of the two paths
on the index itself.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class FilterPlan:
    path: str  # "graph" or "exact"
    selectivity: float
    correlation: float
    threshold: float


@dataclass
class FilterPlanner:
    """Selectivity thresholds below which the masked brute force runs (i.i.d.-like masks and
    graph-correlated masks), and the sampling of the correlation estimate."""

    random_threshold: float = 0.05
    correlated_threshold: float = 0.10
    correlation_cut: float = 2.0
    sample: int = 512
    seed: int = 0
    timings: list = field(default_factory=list)

    def statistics(self, engine, mask: np.ndarray) -> tuple[float, float]:
        """``(selectivity, correlation)`` of a boolean mask on the engine's layer-0 graph."""
        mask = np.asarray(mask, dtype=bool)
        n = mask.shape[0]
        s = float(mask.mean()) if n else 0.0
        allowed = np.flatnonzero(mask)
        if allowed.size == 0 or s >= 1.0:
            return s, 1.0
        rng = np.random.default_rng(self.seed)
        pick = allowed if allowed.size <= self.sample else rng.choice(allowed, self.sample,
                                                                      replace=False)
        arrays = engine.native_arrays()
        links0, count0 = arrays["links0"], arrays["count0"]
        hits = total = 0
        for node in pick:
            row = links0[node, : count0[node]]
            hits += int(mask[row].sum())
            total += int(row.size)
        local = hits / total if total else s
        return s, (local / s if s > 0 else 1.0)

    def plan(self, engine, mask: np.ndarray) -> FilterPlan:
        s, c = self.statistics(engine, mask)
        threshold = self.correlated_threshold if c >= self.correlation_cut else self.random_threshold
        return FilterPlan("exact" if s < threshold else "graph", s, c, threshold)

    def calibrate(self, engine, queries: np.ndarray, k: int, ef: int,
                  selectivities=(0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.5), repeats: int = 2):
        """Times both paths on random masks of each selectivity (queries: a batch of the size
        the planner will serve, since both costs depend on it: the in-graph search needs
        thousands of queries to fill the GPU) and sets ``random_threshold`` to the smallest
        selectivity where the in-graph search is faster, checking larger ones; the correlated
        threshold keeps its ratio to it. Returns the timings ``[(selectivity, graph_s,
        exact_s)]``."""
        n = engine.size
        rows = []
        for i, s in enumerate(sorted(selectivities)):
            mask = np.random.default_rng(self.seed + 1 + i).random(n) < s
            if not mask.any():
                continue
            tg = _best(lambda: engine.search_many(queries, k, ef, mask_pool=mask, mask_index=0),
                       repeats)
            te = _best(lambda: engine.exact_search(queries, min(k, int(mask.sum())), mask),
                       repeats)
            rows.append((float(s), tg, te))
        self.timings = rows
        faster_graph = [s for s, tg, te in rows if tg < te]
        if faster_graph:
            ratio = self.correlated_threshold / self.random_threshold
            self.random_threshold = min(faster_graph)
            self.correlated_threshold = min(1.0, self.random_threshold * ratio)
        else:
            self.random_threshold = self.correlated_threshold = 1.0
        return rows


def _best(fn, repeats: int) -> float:
    fn()  # warm
    best = float("inf")
    for _ in range(repeats):
        t = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t)
    return best
