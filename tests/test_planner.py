"""Filter planner statistics and decisions (host logic, no GPU).

The correlation estimate is checked against masks whose answer the construction fixes: an
i.i.d. mask gives about 1, a mask equal to one connected component of a graph made of
disjoint cliques gives 1 / selectivity (every neighbour of an allowed node is allowed).
"""

from __future__ import annotations

import numpy as np
import pytest

from hnsw_amd.planner import FilterPlanner


class _Graph:
    """Layer-0 arrays only, as the planner reads them from an engine."""

    def __init__(self, links0: np.ndarray, count0: np.ndarray):
        self._arrays = {"links0": links0, "count0": count0}
        self.size = links0.shape[0]

    def native_arrays(self):
        return self._arrays


def _cliques(n_groups: int, size: int) -> _Graph:
    n = n_groups * size
    links0 = np.full((n, size), -1, dtype=np.int32)
    count0 = np.full(n, size - 1, dtype=np.int32)
    for g in range(n_groups):
        members = np.arange(g * size, (g + 1) * size)
        for i, node in enumerate(members):
            links0[node, : size - 1] = np.delete(members, i)
    return _Graph(links0, count0)


def _random_graph(n: int, degree: int, seed: int) -> _Graph:
    rng = np.random.default_rng(seed)
    links0 = rng.integers(0, n, size=(n, degree)).astype(np.int32)
    return _Graph(links0, np.full(n, degree, dtype=np.int32))


def test_iid_mask_has_correlation_near_one() -> None:
    g = _random_graph(20_000, 32, 1)
    mask = np.random.default_rng(2).random(20_000) < 0.1
    s, c = FilterPlanner(sample=2000).statistics(g, mask)
    assert abs(s - mask.mean()) < 1e-12
    assert 0.85 < c < 1.15


def test_component_mask_has_correlation_one_over_selectivity() -> None:
    g = _cliques(50, 20)
    mask = np.zeros(1000, dtype=bool)
    mask[200:220] = True  # one clique
    s, c = FilterPlanner().statistics(g, mask)
    assert s == pytest.approx(0.02)
    assert c == pytest.approx(50.0)


def test_plan_thresholds() -> None:
    g = _cliques(50, 20)
    p = FilterPlanner(random_threshold=0.05, correlated_threshold=0.3, correlation_cut=2.0)
    iid = np.zeros(1000, dtype=bool)
    iid[::25] = True  # 4 %, every 25th node: one per clique, correlation below 1
    assert p.plan(g, iid).path == "exact"
    iid[::10] = True  # 11.6 %
    assert p.plan(g, iid).path == "graph"
    block = np.zeros(1000, dtype=bool)
    block[:200] = True  # 20 %, ten whole cliques: correlated, below 0.3
    plan = p.plan(g, block)
    assert plan.correlation == pytest.approx(5.0) and plan.path == "exact"
    block[:400] = True  # 40 %
    assert p.plan(g, block).path == "graph"


def test_empty_and_full_masks() -> None:
    g = _random_graph(100, 4, 3)
    p = FilterPlanner()
    assert p.statistics(g, np.zeros(100, dtype=bool)) == (0.0, 1.0)
    assert p.statistics(g, np.ones(100, dtype=bool)) == (1.0, 1.0)
