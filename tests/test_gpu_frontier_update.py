"""kgmt_frontier_update — the planner's own append and expand kernels run on
caller buffers — against the oracle's frontier update (orc_frontier_update,
KGMT.cu:305-403) on identical inputs: mid-plan states of every model captured
from the numpy replay of whole plans, and edge cases built from them
(capacity clamp, several goal hits, no children, no re-activation
candidates). Bar: tree samples, parents, costs, frontier flags, the next
expansion list and the counts bit for bit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tests.gpu_util import call, dev, empty, host, ptr, stream  # noqa: E402
from tests.test_oracle_replay import compare_update, oracle_update, update_states  # noqa: E402
from kgmt_rules import run_update  # noqa: E402


def gpu_update(orc, cfg, itr, tree, samples, parents, costs, frontier, child, parent_u, fnext, valid, score, goal):
    S, Cd, D = cfg.dims
    N = cfg.capacity
    d_s = dev(np.ascontiguousarray(np.asarray(samples, np.float32).reshape(N, D).T))
    d_p = dev(np.ascontiguousarray(parents, np.int32))
    d_c = dev(np.ascontiguousarray(costs, np.float32))
    d_f = dev(np.ascontiguousarray(frontier, np.uint8))
    n = len(parent_u)
    d_ch = dev(np.ascontiguousarray(child, np.float32)) if n else None
    d_pu = dev(np.ascontiguousarray(parent_u, np.int32)) if n else None
    d_fn = dev(np.ascontiguousarray(fnext, np.uint8)) if n else None
    d_v = dev(np.ascontiguousarray(valid, np.int32))
    d_sc = dev(np.ascontiguousarray(score, np.float32))
    rep = empty(N, torch.int32)
    info = empty(4, torch.int32)
    g = np.zeros(3, np.float32)
    g[: cfg.wdim] = np.asarray(goal, np.float32)[: cfg.wdim]
    call("kgmt_frontier_update", C.byref(cfg), itr, tree, ptr(d_s), ptr(d_p), ptr(d_c), ptr(d_f), ptr(d_ch),
         ptr(d_pu), ptr(d_fn), n, ptr(d_v), ptr(d_sc), g.ctypes.data_as(C.c_void_p), ptr(rep), ptr(info), stream())
    inf = host(info)
    R = int(inf[1])
    return {"samples": np.ascontiguousarray(host(d_s).T), "parents": host(d_p), "costs": host(d_c),
            "frontier": host(d_f), "M": host(rep)[: min(R, N)].view(np.uint32), "R": R, "next": int(inf[0]),
            "goal": int(inf[2]), "react": int(inf[3])}


@pytest.fixture(scope="module")
def states(orc):
    return update_states(orc)


def test_frontier_update_bit_exact(states, orc):
    for name, st in states:
        print(f"frontier update {name}: tree {st['tree']}, children {len(st['parent_u'])}", flush=True)
        want = run_update(oracle_update, orc, st)
        got = run_update(gpu_update, orc, st)
        assert compare_update(want, got) is None, (name, compare_update(want, got))


def test_frontier_update_argument_errors():
    from kgmt_amd import _lib
    from kgmt_amd import workloads as W
    cfg = W.config("di6d", capacity=1 << 10)
    buf = empty(1 << 12, torch.float32)
    g = np.zeros(3, np.float32)
    L = _lib.lib()
    args = lambda itr, tree, n: (C.byref(cfg), itr, tree, ptr(buf), ptr(buf), ptr(buf), ptr(buf), ptr(buf), ptr(buf),
                                 ptr(buf), n, ptr(buf), ptr(buf), g.ctypes.data_as(C.c_void_p), ptr(buf), ptr(buf),
                                 stream())
    assert L.kgmt_frontier_update(*args(0, 1, 0)) == -1           # itr must be >= 1
    assert L.kgmt_frontier_update(*args(1, 0, 0)) == -1           # empty tree
    assert L.kgmt_frontier_update(*args(1, 1025, 0)) == -1        # tree larger than capacity
    assert L.kgmt_frontier_update(*args(1, 1, 1025)) == -4        # more children than capacity
