"""The .npz GraphState snapshot (the reference's HNSWIndex.save / load format) on the host.

Round trips through the facade (graph built by the host builder, no GPU needed), the exact key
layout the reference's save_state writes (persistence.py:155-169), and validation errors for
damaged snapshots (persistence.py:124-152 messages).
"""

from __future__ import annotations

import json

import numpy as np
import pytest

from hnsw_amd import HNSWIndex
from hnsw_amd.graph_io import GraphState, diff_states, load_state, save_state, validate_state
from hnsw_amd.synthetic import clustered_u8


def _index() -> HNSWIndex:
    data, _ = clustered_u8(600, 8, 1, n_clusters=6, seed=61)
    index = HNSWIndex(8, "l2", M=6, ef_construction=40, seed=62)
    index.add(data)
    return index


def test_save_load_round_trip(tmp_path) -> None:
    index = _index()
    path = tmp_path / "index.npz"
    index.save(path)
    back = HNSWIndex.load(path)
    assert diff_states(index.graph_state(), back.graph_state()) == []
    assert back.M == index.M and back.metric == index.metric and len(back) == len(index)
    # the generator state travels: the next insertion draws the same levels
    more, _ = clustered_u8(50, 8, 1, n_clusters=6, seed=63)
    index.add(more)
    back.add(more)
    assert diff_states(index.graph_state(), back.graph_state()) == []


def test_file_layout_is_the_reference_layout(tmp_path) -> None:
    index = _index()
    path = tmp_path / "index.npz"
    index.save(path)
    with np.load(path, allow_pickle=False) as z:
        state = index.graph_state()
        expected = {"meta", "vectors", "levels", "entry"} | {
            f"{kind}_{l}" for l in range(state.num_layers) for kind in ("indptr", "indices")}
        assert set(z.files) == expected
        meta = json.loads(str(z["meta"]))
        assert meta["format_version"] == 1 and set(meta) == {"format_version", "params", "rng_state"}
        np.testing.assert_array_equal(z["entry"], [state.entry_point, state.max_level])


def test_reference_written_snapshot_loads(tmp_path) -> None:
    """A file written the way the reference writes it (numpy.savez of these keys, engine
    "numba") loads into the MI355X engine."""
    index = _index()
    state = index.graph_state()
    params = dict(index._meta()["params"], engine="numba")
    arrays = {"meta": np.array(json.dumps({"format_version": 1, "params": params,
                                           "rng_state": index._rng.bit_generator.state})),
              "vectors": state.vectors, "levels": state.levels,
              "entry": np.array([state.entry_point, state.max_level], dtype=np.int64)}
    for l in range(state.num_layers):
        arrays[f"indptr_{l}"] = state.indptr[l]
        arrays[f"indices_{l}"] = state.indices[l]
    path = tmp_path / "ref.npz"
    with open(path, "wb") as f:
        np.savez(f, **arrays)
    back = HNSWIndex.load(path)
    assert back.engine == "mi355x"
    assert diff_states(state, back.graph_state()) == []


def _tiny() -> GraphState:
    vectors = np.zeros((3, 2), dtype=np.float32)
    levels = np.array([0, 1, 0], dtype=np.int64)
    indptr = (np.array([0, 2, 4, 6]), np.array([0, 0]))
    indices = (np.array([1, 2, 0, 2, 0, 1]), np.array([], dtype=np.int64))
    return GraphState(vectors, levels, 1, 1, indptr, indices)


def test_validation_errors() -> None:
    validate_state(_tiny())
    s = _tiny()
    with pytest.raises(ValueError, match="entry point"):
        validate_state(GraphState(s.vectors, s.levels, 0, 1, s.indptr, s.indices))
    with pytest.raises(ValueError, match="outside the layer"):
        validate_state(GraphState(s.vectors, s.levels, 1, 1, s.indptr,
                                  (np.array([1, 2, 0, 2, 0, 7]), s.indices[1])))
    with pytest.raises(ValueError, match="malformed link offsets"):
        validate_state(GraphState(s.vectors, s.levels, 1, 1,
                                  (np.array([0, 2, 4]), s.indptr[1]), s.indices))
    with pytest.raises(ValueError, match="number of layers"):
        validate_state(GraphState(s.vectors, s.levels, 1, 2, s.indptr, s.indices))


def test_version_check(tmp_path) -> None:
    path = tmp_path / "bad.npz"
    save_state(path, {"params": {}}, _tiny())
    meta, state = load_state(path)
    assert diff_states(state, _tiny()) == []
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    arrays["meta"] = np.array(json.dumps({"format_version": 2}))
    with open(path, "wb") as f:
        np.savez(f, **arrays)
    with pytest.raises(ValueError, match="unsupported index format version"):
        load_state(path)
