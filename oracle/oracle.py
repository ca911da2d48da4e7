"""ctypes wrapper of the CPU oracle (libkgmt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the kgmt_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from kgmt_amd.abi import KGMT_MAX_SAMPLE, KgmtAdjointParams, KgmtConfig, dims

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkgmt_oracle.so")


class OrcStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("tree_size", C.c_uint32),
        ("goal_index", C.c_int32),
        ("path_rows", C.c_uint32),
        ("path_cost", C.c_float),
        ("path_length", C.c_float),
        ("expanded", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_u32 = C.c_uint32
_cfgp = C.POINTER(KgmtConfig)
_SIG = {
    "orc_philox4x32_10": (None, [_P, _P, _P]),
    "orc_u01": (C.c_float, [_u32]),
    "orc_sincosf": (None, [C.c_float, _P, _P]),
    "orc_draw_child": (None, [_cfgp, _u32, _u32, _u32, _P, _P, _P]),
    "orc_draw_react": (C.c_float, [_cfgp, _u32, _u32, _u32]),
    "orc_segment_valid": (C.c_int, [_P, _P, _P, _u32]),
    "orc_segment_valid2": (C.c_int, [_P, _P, _P, _u32]),
    "orc_region": (C.c_int, [_cfgp, _P]),
    "orc_subregion": (C.c_int, [_cfgp, _P, C.c_int]),
    "orc_propagate": (C.c_int, [_cfgp, _P, _P, C.c_int32, _P, _u32, _P]),
    "orc_propagate_f64": (C.c_int, [_cfgp, _P, _P, C.c_int32, _P, _u32, _P]),
    "orc_quad_ode_f64": (None, [_cfgp, _P, _P, _P]),
    "orc_region_update": (None, [_cfgp, _P, _P, _P, _P, _u32, _P, _P, _P, _P, _P, _P]),
    "orc_score": (None, [_cfgp, _P, _P, _P, _P]),
    "orc_faccept": (C.c_float, [_cfgp, _u32, _u32, _u32]),
    "orc_frontier_update": (C.c_int, [_cfgp, _u32, _u32, _P, _P, _P, _P, _P, _P, _P, _u32, _P, _P, _P, _P, _P, _P, _P,
                                      _P, _P]),
    "orc_plan": (C.c_int, [_cfgp, _P, _P, _P, _u32, _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(OrcStats)]),
    "orc_plan_many": (C.c_int, [_cfgp, _u32, _P, _P, _P, _P, _P, C.c_int, C.POINTER(OrcStats), _P]),
    "orc_sincosf_n": (None, [_P, _u32, _P, _P]),
    "orc_philox_n": (None, [_P, _P, _u32, _P]),
    "orc_draw_children": (None, [_cfgp, _u32, _u32, _u32, _u32, _P, _P, _P]),
    "orc_propagate_children": (C.c_int, [_cfgp, _P, _P, _u32, _P, _P, _u32, _P, _u32, _P, _P, _P, _P]),
    "orc_propagate_children_f64": (C.c_int, [_cfgp, _P, _P, _P, _u32, _P, _u32, _P, _P]),
    "orc_segment_valid_n": (None, [_P, _P, _u32, _P, _u32, _P]),
    "orc_region_n": (None, [_cfgp, _P, _u32, _P, _P]),
    "orc_quad_vjp_f64": (None, [_cfgp, _P, _P, _P, _P, _P]),
    "orc_traj_cost_f64": (C.c_double, [_cfgp, C.POINTER(KgmtAdjointParams), C.c_int32, _P, _P, _P, _P, _u32, _P]),
    "orc_traj_grad_f64": (C.c_double, [_cfgp, C.POINTER(KgmtAdjointParams), C.c_int32, _P, _P, _P, _P, _u32, _P]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(path: str):
    """A build of the oracle at `path` with every function's signature set
    (the tests also load deliberately altered builds this way)."""
    h = C.CDLL(path)
    for name, (res, args) in _SIG.items():
        f = getattr(h, name)
        f.restype = res
        f.argtypes = args
    return h


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(np.asarray(a, dtype=dt))


# ---- scalar / vector helpers ------------------------------------------------
def philox(ctr, key):
    ctr = _c(ctr, np.uint32).reshape(-1, 4)
    key = _c(key, np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    lib().orc_philox_n(_p(ctr), _p(key), len(ctr), _p(out))
    return out


def u01(x: int) -> float:
    return lib().orc_u01(int(x) & 0xFFFFFFFF)


def sincos(x):
    x = _c(x, np.float32).ravel()
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().orc_sincosf_n(_p(x), len(x), _p(s), _p(c))
    return s, c


def draw_children(cfg: KgmtConfig, query: int, itr: int, u0: int, n: int):
    S, Cd, D = dims(cfg.model)
    ctrl = np.zeros((Cd, n), np.float32)
    steps = np.zeros(n, np.int32)
    uf = np.zeros(n, np.float32)
    lib().orc_draw_children(C.byref(cfg), query, itr, u0, n, _p(ctrl), _p(steps), _p(uf))
    return ctrl, steps, uf


def draw_react(cfg, query, itr, node) -> float:
    return lib().orc_draw_react(C.byref(cfg), query, itr, node)


def segment_valid(p0, p1, obs):
    p0 = _c(p0, np.float32).reshape(-1, 3)
    p1 = _c(p1, np.float32).reshape(-1, 3)
    ob = _c(obs, np.float32).reshape(-1)
    out = np.zeros(len(p0), np.uint8)
    lib().orc_segment_valid_n(_p(p0), _p(p1), len(p0), _p(ob) if ob.size else None, ob.size // 6, _p(out))
    return out


def segment_valid2(p0, p1, obs):
    """Planar segment test, rows xmin ymin xmax ymax (one segment)."""
    p0 = _c(p0, np.float32)
    p1 = _c(p1, np.float32)
    ob = _c(obs, np.float32).reshape(-1)
    return bool(lib().orc_segment_valid2(_p(p0), _p(p1), _p(ob) if ob.size else None, ob.size // 4))


def regions(cfg, pos):
    pos = _c(pos, np.float32).reshape(-1, 3)
    r1 = np.zeros(len(pos), np.int32)
    r2 = np.zeros(len(pos), np.int32)
    lib().orc_region_n(C.byref(cfg), _p(pos), len(pos), _p(r1), _p(r2))
    return r1, r2


def propagate_children(cfg, parent_idx, parent_states, controls, steps, obs):
    """SoA in/out like kgmt_propagate_collide: parent_states [S][np],
    controls [C][n] -> child [D][n], valid, r1, r2."""
    S, Cd, D = dims(cfg.model)
    pidx = _c(parent_idx, np.int32)
    ps = _c(parent_states, np.float32).reshape(S, -1)
    ctrl = _c(controls, np.float32).reshape(Cd, -1)
    st = _c(steps, np.int32)
    ob = _c(obs, np.float32).reshape(-1)
    n = len(pidx)
    child = np.zeros((D, n), np.float32)
    valid = np.zeros(n, np.uint8)
    r1 = np.zeros(n, np.int32)
    r2 = np.zeros(n, np.int32)
    rc = lib().orc_propagate_children(C.byref(cfg), _p(pidx), _p(ps), ps.shape[1], _p(ctrl), _p(st), n,
                                      _p(ob) if ob.size else None, ob.size // (2 * cfg.wdim), _p(child), _p(valid),
                                      _p(r1), _p(r2))
    assert rc == 0, rc
    return child, valid, r1, r2


def propagate_f64(cfg, x0s, ctrls, steps, obs):
    """FP64 twin: x0s (n, S), ctrls (n, C) -> x1 (n, D), valid."""
    S, Cd, D = dims(cfg.model)
    x0 = _c(x0s, np.float64).reshape(-1, S)
    u = _c(ctrls, np.float64).reshape(-1, Cd)
    st = _c(steps, np.int32)
    ob = _c(obs, np.float64).reshape(-1)
    n = len(x0)
    x1 = np.zeros((n, D), np.float64)
    valid = np.zeros(n, np.uint8)
    rc = lib().orc_propagate_children_f64(C.byref(cfg), _p(x0), _p(u), _p(st), n, _p(ob) if ob.size else None,
                                          ob.size // (2 * cfg.wdim), _p(x1), _p(valid))
    assert rc == 0, rc
    return x1, valid


def quad_ode_f64(cfg, x, u):
    x = _c(x, np.float64)
    u = _c(u, np.float64)
    out = np.zeros(12, np.float64)
    lib().orc_quad_ode_f64(C.byref(cfg), _p(x), _p(u), _p(out))
    return out


def region_update(cfg, r1, r2, valid, ufront, score, active_prev, active_next, total, valid_count):
    """In-place on active_next/total/valid_count (bytes per subregion); returns fnext."""
    n = len(r1)
    fnext = np.zeros(n, np.uint8)
    r1 = _c(r1, np.int32)
    r2 = _c(r2, np.int32)
    valid = _c(valid, np.uint8)
    ufront = _c(ufront, np.float32)
    score = _c(score, np.float32)
    ap = _c(active_prev, np.uint8)
    for a, dt in ((active_next, np.uint8), (total, np.int32), (valid_count, np.int32)):
        assert a.dtype == dt and a.flags.c_contiguous
    lib().orc_region_update(C.byref(cfg), _p(r1), _p(r2), _p(valid), _p(ufront), n, _p(score), _p(ap),
                            _p(active_next), _p(total), _p(valid_count), _p(fnext))
    return fnext


def score(cfg, total, valid_count, active_bytes):
    R1 = cfg.regions
    out = np.zeros(R1, np.float32)
    lib().orc_score(C.byref(cfg), _p(_c(total, np.int32)), _p(_c(valid_count, np.int32)),
                    _p(_c(active_bytes, np.uint8)), _p(out))
    return out


def faccept(cfg, itr, tree, nxt) -> float:
    return lib().orc_faccept(C.byref(cfg), itr, tree, nxt)


def frontier_update(cfg, itr, tree, samples, parents, costs, frontier, child, parent_u, fnext, valid, score, goal):
    """orc_frontier_update on copies: samples (capacity, D), parents, costs,
    frontier (capacity), child (D, n) as the stage entries take children,
    parent_u, fnext (n), region valid counts and scores, goal (>= W floats).
    Returns the updated arrays, the next expansion list (its first capacity
    entries), its size R, the appended count, the goal node (-1) and the
    re-activated count."""
    S, Cd, D = dims(cfg.model)
    N = cfg.capacity
    smp = _c(samples, np.float32).reshape(N, D).copy()
    par = _c(parents, np.int32).copy()
    cst = _c(costs, np.float32).copy()
    fr = _c(frontier, np.uint8).copy()
    ch = np.ascontiguousarray(np.asarray(child, np.float32).reshape(D, -1).T)
    pu = _c(parent_u, np.int32)
    fn = _c(fnext, np.uint8)
    g = np.zeros(KGMT_MAX_SAMPLE, np.float32)
    g[: min(len(goal), KGMT_MAX_SAMPLE)] = np.asarray(goal, np.float32)[:KGMT_MAX_SAMPLE]
    M = np.zeros(N, np.uint32)
    nM, nxt, nr = C.c_uint32(), C.c_uint32(), C.c_uint32()
    R = C.c_uint64()
    gi = C.c_int32()
    n = len(pu)
    rc = lib().orc_frontier_update(C.byref(cfg), itr, tree, _p(smp), _p(par), _p(cst), _p(fr), _p(ch) if n else None,
                                   _p(pu) if n else None, _p(fn) if n else None, n, _p(_c(valid, np.int32)),
                                   _p(_c(score, np.float32)), _p(g), _p(M), C.byref(nM), C.byref(R), C.byref(nxt),
                                   C.byref(gi), C.byref(nr))
    assert rc == 0, rc
    return {"samples": smp, "parents": par, "costs": cst, "frontier": fr, "M": M[: nM.value], "R": R.value,
            "next": nxt.value, "goal": gi.value, "react": nr.value}


def plan(cfg: KgmtConfig, start, goal, obs, log_iters: bool = False, lib_=None):
    """Full planner on the CPU. Returns dict with stats, tree, path, regions.
    lib_: another build of the oracle (from load()) to run instead."""
    S, Cd, D = dims(cfg.model)
    N = cfg.capacity
    samples = np.zeros((N, D), np.float32)
    parents = np.zeros(N, np.int32)
    costs = np.zeros(N, np.float32)
    path = np.zeros((cfg.max_iter + 2, D), np.float32)
    R1 = cfg.regions
    total = np.zeros(R1, np.int32)
    valid = np.zeros(R1, np.int32)
    score_ = np.zeros(R1, np.float32)
    log = np.zeros((cfg.max_iter, 4), np.uint32) if log_iters else None
    st = OrcStats()
    s = np.zeros(KGMT_MAX_SAMPLE, np.float32)
    s[:D] = np.asarray(start, np.float32)[:D]
    g = np.zeros(KGMT_MAX_SAMPLE, np.float32)
    g[:D] = np.asarray(goal, np.float32)[:D]
    ob = _c(obs, np.float32).reshape(-1)
    rc = (lib_ or lib()).orc_plan(C.byref(cfg), _p(s), _p(g), _p(ob) if ob.size else None, ob.size // (2 * cfg.wdim), _p(samples),
                        _p(parents), _p(costs), _p(path), _p(total), _p(valid), _p(score_),
                        _p(log) if log is not None else None, C.byref(st))
    assert rc == 0, rc
    n = st.tree_size
    out = {"stats": st.as_dict(), "samples": samples[:n].copy(), "parents": parents[:n].copy(),
           "costs": costs[:n].copy(), "path": path[: st.path_rows].copy(), "total": total, "valid": valid,
           "score": score_}
    if log is not None:
        out["iter_log"] = log[: st.iterations].copy()
    return out


def plan_many(cfgs, starts, goals, obstacles, offsets, counts, threads: int = 0):
    n = len(cfgs)
    arr = (KgmtConfig * n)(*cfgs)
    starts = _c(starts, np.float32).reshape(n, KGMT_MAX_SAMPLE)
    goals = _c(goals, np.float32).reshape(n, KGMT_MAX_SAMPLE)
    obs = _c(obstacles, np.float32).reshape(-1)
    if obs.size == 0:
        obs = np.zeros(6, np.float32)
    offs = _c(offsets, np.uint32)
    cnts = _c(counts, np.uint32)
    stats = (OrcStats * n)()
    ms = np.zeros(n, np.float64)
    rc = lib().orc_plan_many(arr, n, _p(starts), _p(goals), _p(obs), _p(offs), _p(cnts), int(threads), stats,
                             _p(ms))
    assert rc == 0, rc
    return [stats[i].as_dict() for i in range(n)], ms


# ---- trajectory refinement (FP64) -----------------------------------------
def quad_vjp_f64(cfg, s, c, a):
    s, c, a = (_c(v, np.float64) for v in (s, c, a))
    gs = np.zeros(12)
    gc = np.zeros(4)
    lib().orc_quad_vjp_f64(C.byref(cfg), _p(s), _p(c), _p(a), _p(gs), _p(gc))
    return gs, gc


def traj_cost_f64(cfg, params, x0, u, goal, obs):
    """u: (T, 4) controls; returns (J, x_T)."""
    x0, u, goal = _c(x0, np.float64), _c(u, np.float64), _c(goal, np.float64)
    ob = _c(obs, np.float64).reshape(-1)
    xT = np.zeros(12)
    J = lib().orc_traj_cost_f64(C.byref(cfg), C.byref(params), len(u), _p(x0), _p(u), _p(goal),
                                _p(ob) if ob.size else None, ob.size // 6, _p(xT))
    return J, xT


def traj_grad_f64(cfg, params, x0, u, goal, obs):
    """Cost and its gradient (T, 4) by the discrete adjoint, in FP64."""
    x0, u, goal = _c(x0, np.float64), _c(u, np.float64), _c(goal, np.float64)
    ob = _c(obs, np.float64).reshape(-1)
    g = np.zeros_like(u)
    J = lib().orc_traj_grad_f64(C.byref(cfg), C.byref(params), len(u), _p(x0), _p(u), _p(goal),
                                _p(ob) if ob.size else None, ob.size // 6, _p(g))
    return J, g
