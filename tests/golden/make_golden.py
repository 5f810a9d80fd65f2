"""Write the golden fixtures in this directory from the reference's pure-Python engine.

Run once, in a container where the reference checkout exists (it does not travel to the GPU
box): ``python tests/golden/make_golden.py /path/to/reference``. Only numpy is needed: the
reference's ``engine="reference"`` (``reference.py``), ``oracle.py`` and ``ScalarQuantizer``
import without Numba.

Why an import at all: the traversal order, the per-query counters and the graph a sequential
insertion builds on a general graph are fixed by no closed form, so these fixtures are the
vectors that pin them. The reference engine equals its Numba engine bit for bit on
integer-valued data (``tests/test_numba_engine.py:37-49, 187-204``), which is why most fixtures
use integer data. Each ``.npz`` holds only inputs and expected outputs.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path.insert(0, str(REPO / "hnsw-engine_amd"))


def _csr(state) -> dict[str, np.ndarray]:
    """Per-layer CSR of a GraphState flattened into three arrays."""
    layer_ptr = [0]
    indptr, indices = [], []
    for layer in range(state.num_layers):
        indptr.append(state.indptr[layer])
        indices.append(state.indices[layer])
        layer_ptr.append(layer_ptr[-1] + state.indices[layer].shape[0])
    return {
        "levels": state.levels.astype(np.int64),
        "entry": np.array([state.entry_point, state.max_level], dtype=np.int64),
        "indptr": np.concatenate(indptr) if indptr else np.zeros(0, np.int64),
        "indices": np.concatenate(indices).astype(np.int32) if indices else np.zeros(0, np.int32),
        "layer_ptr": np.array(layer_ptr, dtype=np.int64),
    }


def _per_query(index, queries, k, ef, allowed=None):
    """Search row by row so that the cumulative counters give per-query figures."""
    eng = index._engine
    nq = queries.shape[0]
    ids = np.empty((nq, k), np.int64)
    dists = np.empty((nq, k), np.float64)
    dc = np.empty(nq, np.int64)
    hops = np.empty(nq, np.int64)
    for row in range(nq):
        d0, h0 = eng.distance_computations, eng.hops
        ids[row], dists[row] = index.search(queries[row], k=k, ef=ef, filter=allowed)
        dc[row] = eng.distance_computations - d0
        hops[row] = eng.hops - h0
    return ids, dists, dc, hops


def search_fixture(name, data, queries, metric, M, efc, seed, plan, masks) -> None:
    from hnsw_engine import HNSWIndex

    index = HNSWIndex(data.shape[1], metric=metric, M=M, ef_construction=efc, seed=seed,
                      engine="reference")
    index.add(data)
    out = {"data": data, "queries": queries, "params": np.array([M, efc, seed], np.int64)}
    out.update(_csr(index.graph_state()))
    for tag, (k, ef) in plan.items():
        ids, dists, dc, hops = _per_query(index, queries, k, ef)
        out[f"{tag}_ids"], out[f"{tag}_dists"] = ids, dists
        out[f"{tag}_dc"], out[f"{tag}_hops"] = dc, hops
        out[f"{tag}_k_ef"] = np.array([k, ef], np.int64)
    for tag, (mask, k, ef) in masks.items():
        ids, dists, dc, hops = _per_query(index, queries, k, ef, mask)
        out[f"{tag}_ids"], out[f"{tag}_dists"] = ids, dists
        out[f"{tag}_dc"], out[f"{tag}_hops"] = dc, hops
        out[f"{tag}_k_ef"] = np.array([k, ef], np.int64)
        out[f"{tag}_mask"] = mask
    out["metric"] = np.array(metric)
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(name, "max_level", index.max_level, "n", len(index))


def builder_fixture() -> None:
    from hnsw_engine import HNSWIndex
    from hnsw_engine.synthetic import small_integers

    configs = {
        "simple": {"selection": "simple"},
        "heuristic": {"selection": "heuristic"},
        "extend": {"selection": "heuristic", "extend_candidates": True},
        "keep": {"selection": "heuristic", "keep_pruned_connections": True},
        "extend_keep": {"selection": "heuristic", "extend_candidates": True,
                        "keep_pruned_connections": True},
    }
    data = small_integers(700, 5, seed=31, high=6)
    out = {"data": data}
    for metric in ("l2", "ip"):
        for tag, cfg in configs.items():
            index = HNSWIndex(5, metric=metric, M=5, ef_construction=30, seed=7,
                              engine="reference", **cfg)
            index.add(data)
            for key, val in _csr(index.graph_state()).items():
                out[f"{metric}_{tag}_{key}"] = val
    np.savez_compressed(HERE / "builder_small_int.npz", **out)
    print("builder_small_int")


def oracle_fixture() -> None:
    from hnsw_engine import exact_search

    rng = np.random.default_rng(41)
    out = {}
    fdata = rng.standard_normal((600, 9)).astype(np.float32)
    fq = rng.standard_normal((25, 9)).astype(np.float32)
    out["float_data"], out["float_queries"] = fdata, fq
    for metric in ("l2", "ip", "cosine"):
        ids, dists = exact_search(fdata, fq, 10, metric)
        out[f"float_{metric}_ids"], out[f"float_{metric}_dists"] = ids, dists
    idata = rng.integers(0, 256, size=(700, 16)).astype(np.float32)
    iq = rng.integers(0, 256, size=(30, 16)).astype(np.float32)
    out["int_data"], out["int_queries"] = idata, iq
    for metric in ("l2", "ip"):
        ids, dists = exact_search(idata, iq, 12, metric)
        out[f"int_{metric}_ids"], out[f"int_{metric}_dists"] = ids, dists
    np.savez_compressed(HERE / "oracle_exact.npz", **out)
    print("oracle_exact")


def quantizer_fixture() -> None:
    from hnsw_engine import ScalarQuantizer

    rng = np.random.default_rng(51)
    out = {}
    cont = (rng.standard_normal((400, 7)) * rng.lognormal(size=7)).astype(np.float32)
    ints = (rng.integers(0, 256, size=(300, 5)) + 3).astype(np.float32)
    q = (rng.standard_normal((20, 7)) * 2).astype(np.float32)
    for tag, data in (("cont", cont), ("int", ints)):
        sq = ScalarQuantizer.fit(data)
        out[f"{tag}_data"] = data
        out[f"{tag}_offset"] = sq.offset
        out[f"{tag}_step"] = np.array(sq.step)
        out[f"{tag}_codes"] = sq.encode(data)
    sq = ScalarQuantizer.fit(cont)
    out["cont_queries"] = q
    for metric in ("l2", "ip", "cosine"):
        out[f"cont_tq_{metric}"] = sq.transform_queries(q, metric)
        cd = rng.standard_normal((20, 4))
        out[f"cont_cd_{metric}"] = cd
        out[f"cont_todist_{metric}"] = sq.to_distance(cd, q, metric)
    np.savez_compressed(HERE / "quantizer.npz", **out)
    print("quantizer")


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    sys.path.insert(0, str(ref / "src"))
    from hnsw_engine.synthetic import clustered, small_integers

    from hnsw_amd.synthetic import clustered_u8, random_mask, range_mask

    data, queries = clustered_u8(3000, 16, 60, n_clusters=20, sigma=30.0, seed=11)
    n = data.shape[0]
    masks = {
        "rand50": (random_mask(n, 0.5, 1), 10, 32),
        "rand10": (random_mask(n, 0.1, 2), 10, 32),
        "rand2": (random_mask(n, 0.02, 3), 10, 32),
        "range10": (range_mask(n, 0.1, 4), 10, 32),
        "range2": (range_mask(n, 0.02, 5), 10, 48),
        "few": (np.isin(np.arange(n), [7, 700, 1700]), 5, 10),
    }
    plan = {"ef8": (10, 8), "ef32": (10, 32), "ef128": (10, 128), "k1": (1, 16), "k40": (40, 20)}
    search_fixture("search_u8_l2", data, queries, "l2", 8, 64, 5, plan, masks)

    tdata = small_integers(900, 6, seed=21, high=5)
    tq = small_integers(40, 6, seed=22, high=5)
    tmasks = {"rand50": (random_mask(900, 0.5, 24), 10, 10),
              "rand10": (random_mask(900, 0.1, 25), 10, 50)}
    for metric in ("l2", "ip"):
        search_fixture(f"search_ties_{metric}", tdata, tq, metric, 6, 40, 23,
                       {"ef10": (10, 10), "ef50": (10, 50)}, tmasks)

    fdata, fq = clustered(1500, 12, seed=4), clustered(40, 12, seed=5)
    for metric in ("l2", "cosine"):
        search_fixture(f"search_float_{metric}", fdata, fq, metric, 8, 50, 6,
                       {"ef50": (10, 50)}, {"rand30": (random_mask(1500, 0.3, 9), 10, 40)})

    builder_fixture()
    oracle_fixture()
    quantizer_fixture()


if __name__ == "__main__":
    main()
