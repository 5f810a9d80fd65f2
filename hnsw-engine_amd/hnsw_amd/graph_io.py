"""Graphs on disk and between engines, in the reference's two formats.

* **Native directory** -- one ``<name>.npy`` per engine array and a ``header.json`` written
  last (format ``"hnsw-engine-native"``, version 1, each array's dtype and shape, the index
  parameters, generator state, entry point and top level); ``persistence.py:192-243`` of the
  reference. Read with ``mmap`` the arrays are read-only maps used in place.
* **Snapshot** (``.npz``, no pickle) -- ``meta`` (a JSON string holding ``format_version`` 1),
  ``vectors``, ``levels``, ``entry = [entry_point, max_level]`` and per layer ``indptr_<l>`` /
  ``indices_<l>``, the CSR links of the nodes on that layer in ascending id order;
  ``persistence.py:155-189``.

The key names, layouts and error messages are the reference's so that files and callers move
between the two engines. The content of native arrays is checked by the library
(``hm_check_graph``) before the kernels, which do not bounds-check, ever read them.
"""

from __future__ import annotations

import itertools
import json
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

from . import _native

NATIVE_FORMAT = "hnsw-engine-native"
NATIVE_FORMAT_VERSION = 1
NATIVE_HEADER = "header.json"
SNAPSHOT_VERSION = 1

# The engine arrays (numba_engine.py:814-825 of the reference): dtype and shape for n nodes,
# `rows` upper-layer rows, `dim` dimensions and M links per upper row (2M on layer 0); a row's
# first `count` slots are its links.
NATIVE_LAYOUT: dict[str, tuple[type, Any]] = {
    "vectors": (np.float32, lambda n, rows, dim, M: (n, dim)),
    "levels": (np.int64, lambda n, rows, dim, M: (n,)),
    "links0": (np.int32, lambda n, rows, dim, M: (n, 2 * M + 1)),
    "count0": (np.int32, lambda n, rows, dim, M: (n,)),
    "upper_offset": (np.int64, lambda n, rows, dim, M: (n,)),
    "links_up": (np.int32, lambda n, rows, dim, M: (rows, M + 1)),
    "count_up": (np.int32, lambda n, rows, dim, M: (rows,)),
}
STORAGE = tuple(NATIVE_LAYOUT)


@dataclass(frozen=True)
class GraphState:
    """Engine-independent snapshot of a graph: the nodes of layer ``l`` are those with
    ``levels >= l``, ascending; node ``i`` of that list has the links
    ``indices[l][indptr[l][i]:indptr[l][i + 1]]`` in stored order (the reference's
    ``GraphState``, ``persistence.py:34-68``)."""

    vectors: np.ndarray
    levels: np.ndarray
    entry_point: int
    max_level: int
    indptr: tuple[np.ndarray, ...]
    indices: tuple[np.ndarray, ...]

    @property
    def size(self) -> int:
        return int(self.levels.shape[0])

    @property
    def num_layers(self) -> int:
        return len(self.indptr)

    def layer_nodes(self, layer: int) -> np.ndarray:
        return np.flatnonzero(self.levels >= layer).astype(np.int64)

    def adjacency(self, layer: int) -> dict[int, list[int]]:
        ptr, idx = self.indptr[layer], self.indices[layer]
        return {int(v): idx[ptr[i]:ptr[i + 1]].tolist()
                for i, v in enumerate(self.layer_nodes(layer))}

    @classmethod
    def from_adjacency(cls, vectors, levels, entry_point: int, max_level: int,
                       layers: list[dict[int, list[int]]]) -> "GraphState":
        """A snapshot from per-layer ``{node: [links]}`` dictionaries."""
        levels = np.asarray(levels, dtype=np.int64)
        ptrs, idxs = [], []
        for layer, links in enumerate(layers):
            rows = [links[int(v)] for v in np.flatnonzero(levels >= layer)]
            ptrs.append(np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64))
            idxs.append(np.fromiter(itertools.chain.from_iterable(rows), np.int64))
        return cls(vectors, levels, entry_point, max_level, tuple(ptrs), tuple(idxs))


def _require(ok: bool, message: str) -> None:
    if not ok:
        raise ValueError(message)


def validate_state(state: GraphState) -> None:
    """``ValueError`` (the reference's messages, ``persistence.py:124-152``) unless the snapshot
    is consistent and every link names a node present on its layer."""
    n = state.size
    _require(state.vectors.ndim == 2 and state.vectors.shape[0] == n,
             "vectors and levels disagree on the number of nodes")
    if n == 0:
        _require((state.entry_point, state.max_level, state.num_layers) == (-1, -1, 0),
                 "an empty graph has no entry point and no layers")
        return
    lv = state.levels
    _require(int(lv.min()) >= 0, "negative node level")
    top = int(lv.max())
    _require(state.max_level == top and state.num_layers == len(state.indices) == top + 1,
             "the number of layers does not match the node levels")
    _require(0 <= state.entry_point < n and int(lv[state.entry_point]) == top,
             "the entry point is not a node of the top layer")
    on_layer = np.cumsum(np.bincount(lv, minlength=top + 1)[::-1])[::-1]  # nodes with level >= l
    for layer in range(top + 1):
        ptr, idx = state.indptr[layer], state.indices[layer]
        _require(ptr.shape == (int(on_layer[layer]) + 1,) and ptr[0] == 0
                 and ptr[-1] == idx.shape[0] and bool(np.all(ptr[1:] >= ptr[:-1])),
                 f"layer {layer}: malformed link offsets")
        _require(idx.size == 0 or (int(idx.min()) >= 0 and int(idx.max()) < n
                                   and int(lv[idx].min()) >= layer),
                 f"layer {layer}: a link points to a node outside the layer")


def diff_states(a: GraphState, b: GraphState) -> list[str]:
    """What differs between two snapshots (an empty list when the graphs are identical)."""
    out = []
    same_vectors = a.vectors.shape == b.vectors.shape and np.array_equal(a.vectors, b.vectors)
    if not same_vectors:
        out.append("vectors differ")
    if not np.array_equal(a.levels, b.levels):
        out.append("levels differ")
    if (a.entry_point, a.max_level) != (b.entry_point, b.max_level):
        out.append(f"entry point / max level: {(a.entry_point, a.max_level)} against "
                   f"{(b.entry_point, b.max_level)}")
    if a.num_layers != b.num_layers:
        return out + [f"{a.num_layers} layers against {b.num_layers}"]
    for layer, (pa, ia, pb, ib) in enumerate(zip(a.indptr, a.indices, b.indptr, b.indices)):
        if not (np.array_equal(pa, pb) and np.array_equal(ia, ib)):
            out.append(f"layer {layer}: links differ")
    return out


def state_from_native(arrays: dict[str, np.ndarray], entry_point: int, max_level: int) -> GraphState:
    """The CSR snapshot of native arrays (``numba_engine.py:753-781``)."""
    levels = np.asarray(arrays["levels"], dtype=np.int64)
    n = levels.shape[0]
    top = int(levels.max()) if n else -1
    indptr, indices = [], []
    for layer in range(top + 1):
        nodes = np.flatnonzero(levels >= layer)
        if layer == 0:
            counts, table = arrays["count0"][nodes], arrays["links0"][nodes]
        else:
            rows = arrays["upper_offset"][nodes] + layer - 1
            counts, table = arrays["count_up"][rows], arrays["links_up"][rows]
        counts = counts.astype(np.int64)
        used = np.arange(table.shape[1])[None, :] < counts[:, None]
        ptr = np.zeros(nodes.size + 1, dtype=np.int64)
        np.cumsum(counts, out=ptr[1:])
        indptr.append(ptr)
        indices.append(table[used].astype(np.int64))
    return GraphState(np.asarray(arrays["vectors"])[:n].copy(), levels.copy(), entry_point,
                      max_level, tuple(indptr), tuple(indices))


def native_from_csr(vectors, levels, indptr, indices, M: int) -> dict[str, np.ndarray]:
    """Native arrays of a CSR graph; upper rows go to the nodes in id order, as the reference's
    ``load_state`` assigns them (``numba_engine.py:783-810``); unused slots hold -1."""
    levels = np.asarray(levels, dtype=np.int64)
    n = levels.shape[0]
    rows = int(levels.sum())
    shapes = {name: shape(n, rows, 0, M) for name, (_, shape) in NATIVE_LAYOUT.items()}
    out = {
        "vectors": np.ascontiguousarray(vectors, dtype=np.float32),
        "levels": levels,
        "links0": np.full(shapes["links0"], -1, dtype=np.int32),
        "count0": np.zeros(n, dtype=np.int32),
        "upper_offset": np.full(n, -1, dtype=np.int64),
        "links_up": np.full(shapes["links_up"], -1, dtype=np.int32),
        "count_up": np.zeros(rows, dtype=np.int32),
    }
    upper = np.flatnonzero(levels > 0)
    out["upper_offset"][upper] = np.cumsum(levels[upper]) - levels[upper]
    for layer in range(len(indptr)):
        nodes = np.flatnonzero(levels >= layer)
        ptr, idx = np.asarray(indptr[layer]), np.asarray(indices[layer])
        counts = np.diff(ptr).astype(np.int32)
        where = nodes if layer == 0 else out["upper_offset"][nodes] + layer - 1
        table, cnt = (out["links0"], out["count0"]) if layer == 0 else (out["links_up"],
                                                                          out["count_up"])
        cnt[where] = counts
        slot = np.arange(idx.size) - np.repeat(ptr[:-1], counts)
        table[np.repeat(where, counts), slot] = idx
    return out


# -- the .npz snapshot -------------------------------------------------------------------------


def save_state(path, meta: dict[str, Any], state: GraphState) -> None:
    """Write the snapshot to exactly ``path`` (no suffix added), without pickle."""
    payload: dict[str, np.ndarray] = {
        "meta": np.array(json.dumps({"format_version": SNAPSHOT_VERSION, **meta})),
        "vectors": state.vectors,
        "levels": state.levels,
        "entry": np.array([state.entry_point, state.max_level], dtype=np.int64),
    }
    for layer, (ptr, idx) in enumerate(zip(state.indptr, state.indices)):
        payload[f"indptr_{layer}"] = ptr
        payload[f"indices_{layer}"] = idx
    with open(path, "wb") as fh:
        np.savez(fh, **payload)


def load_state(path) -> tuple[dict[str, Any], GraphState]:
    """``(meta, state)`` of a snapshot written here or by the reference, validated."""
    with np.load(path, allow_pickle=False) as z:
        if "meta" not in z.files:
            raise ValueError(f"{path} is not an hnsw-engine index file")
        meta = json.loads(str(z["meta"]))
        version = meta.pop("format_version", None)
        if version != SNAPSHOT_VERSION:
            raise ValueError(f"unsupported index format version {version!r}")
        entry_point, max_level = (int(v) for v in z["entry"])
        layers = range(max_level + 1)
        state = GraphState(
            np.ascontiguousarray(z["vectors"], dtype=np.float32),
            z["levels"].astype(np.int64), entry_point, max_level,
            tuple(z[f"indptr_{l}"].astype(np.int64) for l in layers),
            tuple(z[f"indices_{l}"].astype(np.int64) for l in layers),
        )
    validate_state(state)
    return meta, state


# -- the native directory ----------------------------------------------------------------------


def save_native(directory, meta: dict[str, Any], arrays: dict[str, np.ndarray]) -> None:
    """Write each array as ``<name>.npy`` and then ``header.json``: a directory whose header is
    missing was not written completely."""
    root = Path(directory)
    root.mkdir(parents=True, exist_ok=True)
    header_path = root / NATIVE_HEADER
    header_path.unlink(missing_ok=True)
    described = {}
    for name, array in arrays.items():
        np.save(root / f"{name}.npy", np.ascontiguousarray(array), allow_pickle=False)
        described[name] = {"dtype": str(array.dtype), "shape": list(array.shape)}
    header = {"format": NATIVE_FORMAT, "format_version": NATIVE_FORMAT_VERSION,
              "arrays": described, **meta}
    header_path.write_text(json.dumps(header, indent=2))


def load_native(directory, *, mmap: bool = True) -> tuple[dict[str, Any], dict[str, np.ndarray]]:
    """``(meta, arrays)`` of a native directory; with ``mmap`` the arrays are read-only maps of
    the files (pages are read as searches touch them)."""
    root = Path(directory)
    header_path = root / NATIVE_HEADER
    if not header_path.is_file():
        raise ValueError(f"{root} is not an hnsw-engine native index (no {NATIVE_HEADER})")
    meta = json.loads(header_path.read_text())
    if meta.pop("format", None) != NATIVE_FORMAT:
        raise ValueError(f"{root} is not an hnsw-engine native index")
    version = meta.pop("format_version", None)
    if version != NATIVE_FORMAT_VERSION:
        raise ValueError(f"unsupported native format version {version!r}")
    arrays = {}
    for name, want in meta.pop("arrays").items():
        got = np.load(root / f"{name}.npy", mmap_mode="r" if mmap else None, allow_pickle=False)
        if (str(got.dtype), list(got.shape)) != (want["dtype"], want["shape"]):
            raise ValueError(f"{name}.npy has dtype {got.dtype} and shape {got.shape}, header "
                             f"says {want['dtype']} and {tuple(want['shape'])}")
        arrays[name] = np.asarray(got)  # a plain ndarray view of the map
    return meta, arrays


def check_native_layout(arrays: dict[str, np.ndarray], dim: int, M: int) -> tuple[int, int]:
    """``ValueError`` unless every array has its dtype and shape (the reference's
    ``attach_native`` table, ``numba_engine.py:841-860``); returns ``(n, rows)``."""
    n, rows = arrays["vectors"].shape[0], arrays["links_up"].shape[0]
    for name, (dtype, shape) in NATIVE_LAYOUT.items():
        want = shape(n, rows, dim, M)
        got = arrays[name]
        if got.dtype != dtype or got.shape != want:
            raise ValueError(f"{name} has dtype {got.dtype} and shape {got.shape}, expected "
                             f"{np.dtype(dtype)} and {want}")
    if int(arrays["levels"].sum()) != rows:
        raise ValueError("upper-layer rows do not match the node levels")
    return n, rows


def check_native_graph(arrays: dict[str, np.ndarray], entry_point: int, max_level: int, M: int,
                       metric: str = "l2") -> None:
    """``ValueError`` unless the arrays form a well-formed graph: the reference's
    ``_validate_native`` (``numba_engine.py:878-927``), run by the library in one pass."""
    g, keep = _native.graph_struct(arrays, entry_point, max_level, M, metric)
    _native.check(_native.lib().hm_check_graph(g))
    del keep
