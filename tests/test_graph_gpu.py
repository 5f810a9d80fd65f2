"""GPU tests of hm_search_graph_dev: a batch replayed from its captured HIP graph gives the
direct launch's ids, distances and counters on every replay, with the program's own GPU work and
synchronisations between replays; a change of arguments or tuning recaptures; the null stream is
refused. Covers the one-launch register beam, the batched descent with the longest-first
layer-0 launch (one wave per CU: more queries than wave slots) and the wide register beam."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hnsw_amd import _native
from hnsw_amd.engine import MI355XEngine
from hnsw_amd.synthetic import clustered_u8

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def index():
    n, dim, M = 8_000, 128, 16
    data, queries = clustered_u8(n, dim, 397, n_clusters=40, sigma=30.0, seed=61)
    levels = np.floor(-np.log(1.0 - np.random.default_rng(62).random(n))
                      / np.log(M)).astype(np.int64)
    eng = MI355XEngine(dim, "l2", M, 100, "heuristic", False, False)
    eng.insert(data, levels)
    return eng, queries


class _Bufs:
    def __init__(self, queries: np.ndarray, k: int):
        dev = torch.device("cuda:0")
        nq = queries.shape[0]
        self.q = torch.from_numpy(queries).to(dev)
        self.ids = torch.empty((nq, k), dtype=torch.int64, device=dev)
        self.d = torch.empty((nq, k), dtype=torch.float64, device=dev)
        self.dc = torch.empty(nq, dtype=torch.int32, device=dev)
        self.hops = torch.empty(nq, dtype=torch.int32, device=dev)
        self.fail = torch.zeros(nq, dtype=torch.int32, device=dev)

    def call(self, fn, handle, nq, k, ef, stream_ptr) -> int:
        opts = _native.SearchOpts()
        opts.nq, opts.k_all, opts.ef_all, opts.kstride, opts.mask_all = nq, k, ef, k, -1
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        return fn(handle, p(self.q), ctypes.byref(opts), None, 1, p(self.ids), p(self.d),
                  p(self.dc), p(self.hops), p(self.fail), ctypes.c_void_p(stream_ptr))

    def host(self):
        return (self.ids.cpu().numpy(), self.d.cpu().numpy(), self.dc.cpu().numpy(),
                self.hops.cpu().numpy(), self.fail.cpu().numpy())

    def scramble(self) -> None:
        for t in (self.ids, self.d, self.dc, self.hops):
            t.fill_(-7)


def _same(got, want, msg):
    for g, w, name in zip(got, want, ("ids", "dists", "dcount", "hops", "fail")):
        np.testing.assert_array_equal(g, w, err_msg=f"{msg} {name}")


def _tune(eng, waves_per_cu=0) -> None:
    _native.check(_native.lib().hm_set_tuning(eng._device(), 0, 0, waves_per_cu, 0, 0, 0))


@pytest.mark.parametrize("ef,waves_per_cu", [(48, 0), (48, 1), (96, 1)])
def test_graph_replays_match_direct_launch(index, ef: int, waves_per_cu: int) -> None:
    eng, queries = index
    lib, h, k = _native.lib(), eng._device(), 10
    nq = queries.shape[0]
    b = _Bufs(queries, k)
    stream = torch.cuda.Stream()
    print(f"case graph ef={ef} waves_per_cu={waves_per_cu}", flush=True)
    try:
        _tune(eng, waves_per_cu)
        with torch.cuda.stream(stream):
            _native.check(b.call(lib.hm_search_dev, h, nq, k, ef, stream.cuda_stream))
            stream.synchronize()
            want = b.host()
            assert want[4].sum() == 0
            for rep in range(4):  # the first call captures, the others replay
                b.scramble()
                _native.check(b.call(lib.hm_search_graph_dev, h, nq, k, ef, stream.cuda_stream))
                stream.synchronize()
                _same(b.host(), want, f"replay {rep}")
                torch.cuda.synchronize()
            # another ef recaptures and still matches its direct launch
            _native.check(b.call(lib.hm_search_dev, h, nq, k, 32, stream.cuda_stream))
            stream.synchronize()
            want32 = b.host()
            for rep in range(2):
                b.scramble()
                _native.check(b.call(lib.hm_search_graph_dev, h, nq, k, 32, stream.cuda_stream))
                stream.synchronize()
                _same(b.host(), want32, f"ef 32 replay {rep}")
    finally:
        _tune(eng)
    assert want[0].shape == (nq, k)


def test_graph_refuses_the_null_stream(index) -> None:
    eng, queries = index
    b = _Bufs(queries[:16], 10)
    with pytest.raises(ValueError, match="null stream"):
        _native.check(b.call(_native.lib().hm_search_graph_dev, eng._device(), 16, 10, 48, 0))
