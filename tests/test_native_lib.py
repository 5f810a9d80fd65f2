"""CPU checks of the C ABI library and the host graph builder (no GPU calls)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import golden_io
from hnsw_amd import _native
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.graph_io import GraphState, diff_states

HEADER = Path(__file__).resolve().parents[1] / "include" / "hnsw_mi355x.h"


def test_library_exports_every_declared_symbol() -> None:
    declared = set(re.findall(r"^\S.*?\b(hm_\w+)\(", HEADER.read_text(), flags=re.M))
    assert declared == set(_native.SYMBOLS)
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("tag", ["simple", "heuristic", "extend", "keep", "extend_keep"])
def test_builder_reproduces_reference_graphs(metric: str, tag: str) -> None:
    fx = golden_io.load("builder_small_int")
    cfg = {"simple": ("simple", False, False), "heuristic": ("heuristic", False, False),
           "extend": ("heuristic", True, False), "keep": ("heuristic", False, True),
           "extend_keep": ("heuristic", True, True)}[tag]
    pre = f"{metric}_{tag}_"
    levels, indptr, indices = golden_io.csr_layers(fx, pre)
    data = fx["data"].astype(np.float32)
    eng = MI355XEngine(5, metric, 5, 30, *cfg)
    eng.insert(data[:300], levels[:300])  # two calls: ids continue, as chunked adds do
    eng.insert(data[300:], levels[300:])
    entry, top = (int(v) for v in fx[pre + "entry"])
    ref = GraphState(data, levels, entry, top, tuple(indptr), tuple(indices))
    assert diff_states(eng.export_state(), ref) == []


@pytest.mark.parametrize("name", ["search_u8_l2", "search_ties_l2", "search_ties_ip",
                                  "search_float_l2", "search_float_cosine"])
def test_builder_reproduces_search_fixture_graphs(name: str) -> None:
    fx, arrays, entry, top, metric, M = golden_io.graph(name)
    eng = MI355XEngine(arrays["vectors"].shape[1], metric, M, int(fx["params"][1]),
                       "heuristic", False, False)
    eng.insert(arrays["vectors"], arrays["levels"])
    ref = GraphState(arrays["vectors"], arrays["levels"], entry, top,
                     *[tuple(x) for x in golden_io.csr_layers(fx)[1:]])
    assert diff_states(eng.export_state(), ref) == []


def test_exact_query_check_accepts_only_8bit_integers() -> None:
    """hm_queries_exact decides whether a batch takes the exact 8-bit kernels (host code only):
    every value must be an integer in 0..255 (-0.0 counts as 0)."""
    lib = _native.lib()
    q = np.random.default_rng(0).integers(0, 256, size=(300, 37)).astype(np.float32)
    q[0, 0], q[-1, -1] = 0.0, 255.0
    assert lib.hm_queries_exact(_native.ptr(q), q.size) == 1
    for bad in (np.nan, np.inf, -np.inf, -1.0, 255.5, 256.0, 3.25, 1e-30):
        for pos in ((0, 0), (150, 20), (299, 36)):
            q2 = q.copy()
            q2[pos] = bad
            assert lib.hm_queries_exact(_native.ptr(q2), q2.size) == 0, (bad, pos)
    q3 = q.copy()
    q3[7, 7] = -0.0
    assert lib.hm_queries_exact(_native.ptr(q3), q3.size) == 1
    assert lib.hm_queries_exact(_native.ptr(q), 0) == 1


def test_exact_tuning_checks_its_range() -> None:
    """Argument checks only (host code): out-of-range tile shapes are refused."""
    lib = _native.lib()
    for qb, se in ((3, 0), (8, 0), (0, -1), (0, 65)):
        with pytest.raises(ValueError, match="out of range"):
            _native.check(lib.hm_set_exact_tuning(qb, se))
    _native.check(lib.hm_set_exact_tuning(0, 0))


def test_autotune_grids_enumerate_every_setting() -> None:
    from hnsw_amd import autotune

    s = autotune._settings(autotune.SEARCH_GRID)
    assert len(s) == 3 * 2 * 2 and len({tuple(sorted(d.items())) for d in s}) == len(s)
    e = autotune._settings(autotune.EXACT_GRID)
    assert len(e) == 9 and all(set(d) == {"query_blocks", "sample_every"} for d in e)
