"""The reference's on-disk formats -- the ``.npz`` GraphState snapshot (``save_state`` /
``load_state``) and the native ``.npy`` directory -- their validation, and CSR <-> native
conversion.

Format (``persistence.py:192-243`` of the reference): one ``<name>.npy`` per engine array and a
``header.json`` written last, holding ``format = "hnsw-engine-native"``, ``format_version = 1``,
each array's dtype and shape, the index parameters, the generator state, the entry point and the
top level. Arrays (``numba_engine.py:814-825``):

========== ======= ================== ====================================================
name       dtype   shape              meaning
========== ======= ================== ====================================================
vectors    float32 (n, dim)           prepared vectors
levels     int64   (n,)               top layer of each node
links0     int32   (n, 2M + 1)        layer-0 links; the first ``count0[i]`` slots are used
count0     int32   (n,)
upper_offset int64 (n,)               first row in ``links_up`` of a node with level > 0
links_up   int32   (rows, M + 1)      one row per (node, layer >= 1); rows = sum(levels)
count_up   int32   (rows,)
========== ======= ================== ====================================================

Validation restates ``numba_engine.py:827-927`` with the same ``ValueError`` messages, because
the device kernels do not bounds-check either.
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

NATIVE_FORMAT = "hnsw-engine-native"
NATIVE_FORMAT_VERSION = 1
NATIVE_HEADER = "header.json"
STORAGE = ("vectors", "levels", "links0", "count0", "upper_offset", "links_up", "count_up")


@dataclass(frozen=True)
class GraphState:
    """Engine-independent snapshot: per-layer CSR links of the nodes on that layer (ascending
    ids), as in ``persistence.py:34-68`` of the reference."""

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

    def adjacency(self, layer: int) -> dict[int, list[int]]:
        nodes = np.flatnonzero(self.levels >= layer)
        ptr, idx = self.indptr[layer], self.indices[layer]
        return {int(v): idx[ptr[i] : ptr[i + 1]].tolist() for i, v in enumerate(nodes)}


def diff_states(a: GraphState, b: GraphState) -> list[str]:
    """Differences between two snapshots (empty when identical)."""
    problems: list[str] = []
    if a.vectors.shape != b.vectors.shape or not np.array_equal(a.vectors, b.vectors):
        problems.append("vectors differ")
    if not np.array_equal(a.levels, b.levels):
        problems.append("levels differ")
    if (a.entry_point, a.max_level) != (b.entry_point, b.max_level):
        problems.append("entry point/max level differ")
    if a.num_layers != b.num_layers:
        problems.append("layer count differs")
        return problems
    for layer in range(a.num_layers):
        if not (
            np.array_equal(a.indptr[layer], b.indptr[layer])
            and np.array_equal(a.indices[layer], b.indices[layer])
        ):
            problems.append(f"layer {layer}: links differ")
    return problems


STATE_FORMAT_VERSION = 1


def validate_state(state: GraphState) -> None:
    """``ValueError`` unless the snapshot is consistent and every link is in range and on its
    layer (``persistence.py:124-152`` of the reference: the kernels do not bounds-check)."""
    n = state.size
    if state.vectors.ndim != 2 or state.vectors.shape[0] != n:
        raise ValueError("vectors and levels disagree on the number of nodes")
    if n == 0:
        if state.entry_point != -1 or state.max_level != -1 or state.num_layers != 0:
            raise ValueError("an empty graph has no entry point and no layers")
        return
    if (state.levels < 0).any():
        raise ValueError("negative node level")
    top = int(state.levels.max())
    if state.max_level != top or state.num_layers != top + 1 or len(state.indices) != top + 1:
        raise ValueError("the number of layers does not match the node levels")
    if not 0 <= state.entry_point < n or int(state.levels[state.entry_point]) != top:
        raise ValueError("the entry point is not a node of the top layer")
    for layer in range(top + 1):
        on_layer = int((state.levels >= layer).sum())
        ptr, idx = state.indptr[layer], state.indices[layer]
        if (ptr.shape != (on_layer + 1,) or ptr[0] != 0 or ptr[-1] != idx.shape[0]
                or (np.diff(ptr) < 0).any()):
            raise ValueError(f"layer {layer}: malformed link offsets")
        if idx.size and ((idx < 0).any() or (idx >= n).any() or (state.levels[idx] < layer).any()):
            raise ValueError(f"layer {layer}: a link points to a node outside the layer")


def save_state(path: str | Path, meta: dict[str, Any], state: GraphState) -> None:
    """``.npz`` snapshot at exactly ``path`` (the reference's ``save_state``,
    ``persistence.py:155-169``): ``meta`` as a JSON string with ``format_version``, vectors,
    levels, ``entry = [entry_point, max_level]``, then ``indptr_<l>`` / ``indices_<l>`` per
    layer. Written without pickle."""
    arrays: dict[str, np.ndarray] = {
        "meta": np.array(json.dumps({"format_version": STATE_FORMAT_VERSION, **meta})),
        "vectors": state.vectors,
        "levels": state.levels,
        "entry": np.array([state.entry_point, state.max_level], dtype=np.int64),
    }
    for layer in range(state.num_layers):
        arrays[f"indptr_{layer}"] = state.indptr[layer]
        arrays[f"indices_{layer}"] = state.indices[layer]
    with open(path, "wb") as handle:
        np.savez(handle, **arrays)


def load_state(path: str | Path) -> tuple[dict[str, Any], GraphState]:
    """``(meta, state)`` of a file written by :func:`save_state` (or the reference's), read
    without pickle and validated."""
    with np.load(path, allow_pickle=False) as data:
        if "meta" not in data.files:
            raise ValueError(f"{path} is not an hnsw-engine index file")
        meta: dict[str, Any] = json.loads(str(data["meta"]))
        version = meta.pop("format_version", None)
        if version != STATE_FORMAT_VERSION:
            raise ValueError(f"unsupported index format version {version!r}")
        entry_point, max_level = (int(v) for v in data["entry"])
        layers = max_level + 1
        state = GraphState(
            vectors=np.ascontiguousarray(data["vectors"], dtype=np.float32),
            levels=data["levels"].astype(np.int64),
            entry_point=entry_point,
            max_level=max_level,
            indptr=tuple(data[f"indptr_{i}"].astype(np.int64) for i in range(layers)),
            indices=tuple(data[f"indices_{i}"].astype(np.int64) for i in range(layers)),
        )
    validate_state(state)
    return meta, state


def state_from_native(arrays: dict[str, np.ndarray], entry_point: int, max_level: int) -> GraphState:
    """CSR snapshot of native arrays (``numba_engine.py:753-781``)."""
    levels = np.asarray(arrays["levels"], dtype=np.int64)
    n = levels.shape[0]
    top = int(levels.max()) if n else -1
    indptr, indices = [], []
    for layer in range(top + 1):
        nodes = np.flatnonzero(levels >= layer)
        if layer == 0:
            counts = arrays["count0"][nodes].astype(np.int64)
            table = arrays["links0"][nodes]
        else:
            rows = arrays["upper_offset"][nodes] + layer - 1
            counts = arrays["count_up"][rows].astype(np.int64)
            table = arrays["links_up"][rows]
        used = np.arange(table.shape[1])[None, :] < counts[:, None]
        ptr = np.zeros(nodes.size + 1, dtype=np.int64)
        np.cumsum(counts, out=ptr[1:])
        indptr.append(ptr)
        indices.append(table[used].astype(np.int64))
    return GraphState(
        np.asarray(arrays["vectors"])[:n].copy(), levels.copy(), entry_point, max_level,
        tuple(indptr), tuple(indices),
    )


def native_from_csr(
    vectors: np.ndarray,
    levels: np.ndarray,
    indptr: list[np.ndarray],
    indices: list[np.ndarray],
    M: int,
) -> dict[str, np.ndarray]:
    """Native arrays of a CSR graph; upper rows are assigned in id order, as the reference's
    ``load_state`` does (``numba_engine.py:783-810``). Unused slots hold -1."""
    levels = np.asarray(levels, dtype=np.int64)
    n = levels.shape[0]
    rows = int(levels.sum())
    links0 = np.full((n, 2 * M + 1), -1, dtype=np.int32)
    count0 = np.zeros(n, dtype=np.int32)
    upper_offset = np.full(n, -1, dtype=np.int64)
    links_up = np.full((rows, M + 1), -1, dtype=np.int32)
    count_up = np.zeros(rows, dtype=np.int32)
    upper = np.flatnonzero(levels > 0)
    if upper.size:
        starts = np.zeros(upper.size, dtype=np.int64)
        np.cumsum(levels[upper][:-1], out=starts[1:])
        upper_offset[upper] = starts
    for layer in range(len(indptr)):
        nodes = np.flatnonzero(levels >= layer)
        ptr, idx = np.asarray(indptr[layer]), np.asarray(indices[layer])
        counts = np.diff(ptr).astype(np.int32)
        row_ids = nodes if layer == 0 else upper_offset[nodes] + layer - 1
        table = links0 if layer == 0 else links_up
        cnt = count0 if layer == 0 else count_up
        cnt[row_ids] = counts
        pos = np.arange(idx.size) - np.repeat(ptr[:-1], counts)
        table[np.repeat(row_ids, counts), pos] = idx
    return {
        "vectors": np.ascontiguousarray(vectors, dtype=np.float32),
        "levels": levels,
        "links0": links0,
        "count0": count0,
        "upper_offset": upper_offset,
        "links_up": links_up,
        "count_up": count_up,
    }


def save_native(directory: str | Path, meta: dict[str, Any], arrays: dict[str, np.ndarray]) -> None:
    """Write ``arrays`` as ``.npy`` files and ``header.json`` last (``persistence.py:192-213``)."""
    root = Path(directory)
    root.mkdir(parents=True, exist_ok=True)
    (root / NATIVE_HEADER).unlink(missing_ok=True)
    layout = {}
    for name, array in arrays.items():
        np.save(root / f"{name}.npy", np.ascontiguousarray(array), allow_pickle=False)
        layout[name] = {"dtype": str(array.dtype), "shape": list(array.shape)}
    header = {"format": NATIVE_FORMAT, "format_version": NATIVE_FORMAT_VERSION, "arrays": layout,
              **meta}
    (root / NATIVE_HEADER).write_text(json.dumps(header, indent=2))


def load_native(directory: str | Path, *, mmap: bool = True) -> tuple[dict[str, Any], dict[str, np.ndarray]]:
    """Read a directory written by ``save_native`` (``persistence.py:216-243``)."""
    root = Path(directory)
    header_path = root / NATIVE_HEADER
    if not header_path.is_file():
        raise ValueError(f"{root} is not an hnsw-engine native index (no {NATIVE_HEADER})")
    meta: dict[str, Any] = json.loads(header_path.read_text())
    if meta.pop("format", None) != NATIVE_FORMAT:
        raise ValueError(f"{root} is not an hnsw-engine native index")
    version = meta.pop("format_version", None)
    if version != NATIVE_FORMAT_VERSION:
        raise ValueError(f"unsupported native format version {version!r}")
    arrays: dict[str, np.ndarray] = {}
    for name, expected in meta.pop("arrays").items():
        array = np.load(root / f"{name}.npy", mmap_mode="r" if mmap else None, allow_pickle=False)
        if str(array.dtype) != expected["dtype"] or list(array.shape) != expected["shape"]:
            raise ValueError(
                f"{name}.npy has dtype {array.dtype} and shape {array.shape}, header says "
                f"{expected['dtype']} and {tuple(expected['shape'])}"
            )
        arrays[name] = np.asarray(array)
    return meta, arrays


def check_native_shapes(arrays: dict[str, np.ndarray], dim: int, M: int) -> tuple[int, int]:
    """Dtype/shape table of ``numba_engine.py:841-860``; returns ``(n, rows)``."""
    n = arrays["vectors"].shape[0]
    rows = arrays["links_up"].shape[0]
    expected = {
        "vectors": (np.float32, (n, dim)),
        "levels": (np.int64, (n,)),
        "links0": (np.int32, (n, 2 * M + 1)),
        "count0": (np.int32, (n,)),
        "upper_offset": (np.int64, (n,)),
        "links_up": (np.int32, (rows, M + 1)),
        "count_up": (np.int32, (rows,)),
    }
    for name, (dtype, shape) in expected.items():
        array = arrays[name]
        if array.dtype != dtype or array.shape != shape:
            raise ValueError(
                f"{name} has dtype {array.dtype} and shape {array.shape}, expected "
                f"{np.dtype(dtype)} and {shape}"
            )
    if int(arrays["levels"].sum()) != rows:
        raise ValueError("upper-layer rows do not match the node levels")
    return n, rows


def validate_native(arrays: dict[str, np.ndarray], entry_point: int, max_level: int, M: int) -> None:
    """``ValueError`` unless the arrays form a well-formed graph (``numba_engine.py:878-927``)."""
    levels = arrays["levels"]
    n, rows = levels.shape[0], arrays["links_up"].shape[0]
    if n == 0:
        if rows or entry_point != -1 or max_level != -1:
            raise ValueError("an empty graph has no entry point and no layers")
        return
    if (levels < 0).any():
        raise ValueError("negative node level")
    top = int(levels.max())
    if max_level != top or not 0 <= entry_point < n or int(levels[entry_point]) != top:
        raise ValueError("the entry point is not a node of the top layer")
    _check_links(arrays["links0"], arrays["count0"], n, 2 * M, "layer 0")
    upper = np.flatnonzero(levels > 0)
    offsets = arrays["upper_offset"][upper]
    order = np.argsort(offsets, kind="stable")
    starts, sizes = offsets[order], levels[upper][order]
    ends = starts + sizes
    tiles = starts.size == 0 or (
        starts[0] == 0 and ends[-1] == rows and bool((starts[1:] == ends[:-1]).all())
    )
    if not tiles or (starts.size == 0 and rows):
        raise ValueError("upper-layer rows of the nodes overlap or leave gaps")
    count_up = arrays["count_up"]
    targets = _check_links(arrays["links_up"], count_up, n, M, "upper layers")
    layer_of_row = np.arange(rows) - np.repeat(starts, sizes) + 1
    if (levels[targets] < np.repeat(layer_of_row, count_up)).any():
        raise ValueError("an upper-layer link points to a node absent from that layer")


def _check_links(links: np.ndarray, counts: np.ndarray, n: int, cap: int, where: str) -> np.ndarray:
    if (counts < 0).any() or (counts > cap).any():
        raise ValueError(f"{where}: link counts outside [0, {cap}]")
    used = np.arange(links.shape[1])[None, :] < counts[:, None]
    values = links[used]
    if values.size and ((values < 0).any() or (values >= n).any()):
        raise ValueError(f"{where}: a link points outside the graph")
    return values


def has_duplicate_links(links: np.ndarray, counts: np.ndarray) -> bool:
    """True if any row repeats a link among its used slots (the reference's builder never
    produces one; ``invariants.py`` treats it as a defect)."""
    if links.shape[0] == 0:
        return False
    used = np.arange(links.shape[1])[None, :] < counts[:, None]
    big = np.int64(links.shape[1]) + np.iinfo(np.int32).max
    vals = np.where(used, links.astype(np.int64), big + np.arange(links.shape[1])[None, :])
    vals.sort(axis=1)
    return bool((vals[:, 1:] == vals[:, :-1]).any())
