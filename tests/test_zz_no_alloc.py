"""hm_search_dev_noalloc (scratch carved from the library's static device array) against
hm_search_dev (scratch the handle allocates): ids, distances, counters and fail flags identical,
on two calls in a row with the second batch smaller, for the constant-shape register beam (more
queries than the static scratch's wave slots, so the descent-first schedule runs), the wide
register beam, the general kernel with per-query masks, and with the caller's counter arrays
left null. Calls go through the C ABI with device buffers."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hnsw_amd import _native
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.synthetic import clustered_u8, random_mask

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def index():
    n, dim, M = 60_000, 128, 16
    data, queries = clustered_u8(n, dim, 3000, n_clusters=60, sigma=30.0, seed=81)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(82).random(n))
                      / np.log(M)).astype(np.int64)
    eng = MI355XEngine(dim, "l2", M, 100, "heuristic", False, False)
    eng.insert(data, levels)
    return eng, queries


def _bits(masks: np.ndarray) -> np.ndarray:
    n = masks.shape[1]
    words = (n + 31) // 32
    padded = np.zeros((masks.shape[0], words * 32), dtype=bool)
    padded[:, :n] = masks
    # bit i of word i // 32 (little-endian bytes, least significant bit first)
    return np.packbits(padded, axis=1, bitorder="little").view(np.uint32).copy()


def _run(fn, eng, q, k, ef, masks=None, midx=None, counters=True):
    dev = torch.device("cuda:0")
    nq = q.shape[0]
    tq = torch.from_numpy(np.ascontiguousarray(q)).to(dev)
    ids = torch.full((nq, k), -7, dtype=torch.int64, device=dev)
    d = torch.full((nq, k), -7.0, dtype=torch.float64, device=dev)
    dc = torch.full((nq,), -7, dtype=torch.int32, device=dev)
    hops = torch.full((nq,), -7, dtype=torch.int32, device=dev)
    fail = torch.full((nq,), -7, dtype=torch.int32, device=dev)
    opts = _native.SearchOpts()
    opts.nq, opts.k_all, opts.ef_all, opts.kstride, opts.mask_all = nq, k, ef, k, -1
    keep = []
    bits_ptr = None
    if masks is not None:
        tb = torch.from_numpy(_bits(masks)).to(dev)
        tm = torch.from_numpy(np.ascontiguousarray(midx, dtype=np.int32)).to(dev)
        keep += [tb, tm]
        bits_ptr = ctypes.c_void_p(tb.data_ptr())
        opts.mask_index = tm.data_ptr()
        opts.n_masks = masks.shape[0]
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        _native.check(fn(eng._device(), p(tq), ctypes.byref(opts), bits_ptr, 1, p(ids), p(d),
                         p(dc) if counters else None, p(hops) if counters else None,
                         p(fail) if counters else None, ctypes.c_void_p(stream.cuda_stream)))
        stream.synchronize()
    out = [ids.cpu().numpy(), d.cpu().numpy()]
    if counters:
        out += [dc.cpu().numpy(), hops.cpu().numpy(), fail.cpu().numpy()]
    return out


def _same(got, want, msg):
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w, err_msg=msg)


@pytest.mark.parametrize("ef,k", [(48, 10), (200, 10)])
def test_noalloc_equals_allocating_entry_point(index, ef: int, k: int) -> None:
    eng, queries = index
    lib = _native.lib()
    for nq in (3000, 700):  # second call smaller
        print(f"case noalloc ef={ef} nq={nq}", flush=True)
        want = _run(lib.hm_search_dev, eng, queries[:nq], k, ef)
        got = _run(lib.hm_search_dev_noalloc, eng, queries[:nq], k, ef)
        assert want[4].sum() == 0
        _same(got, want, f"ef={ef} nq={nq}")


def test_noalloc_masks_and_null_counters(index) -> None:
    eng, queries = index
    lib = _native.lib()
    n = eng.size
    masks = np.stack([random_mask(n, s, 90 + i) for i, s in enumerate((0.05, 0.3))])
    rng = np.random.default_rng(91)
    for nq in (1200, 300):
        midx = np.where(rng.random(nq) < 0.6, rng.integers(0, 2, size=nq), -1)
        print(f"case noalloc masked nq={nq}", flush=True)
        want = _run(lib.hm_search_dev, eng, queries[:nq], 10, 64, masks, midx)
        got = _run(lib.hm_search_dev_noalloc, eng, queries[:nq], 10, 64, masks, midx)
        _same(got, want, f"masked nq={nq}")
        print(f"case noalloc null counters nq={nq}", flush=True)
        got = _run(lib.hm_search_dev_noalloc, eng, queries[:nq], 10, 48, counters=False)
        want = _run(lib.hm_search_dev, eng, queries[:nq], 10, 48)
        _same(got, want[:2], f"null counters nq={nq}")
