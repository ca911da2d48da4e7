"""Python mirror of the reference planner interface over the C-ABI.

    Planner  include/planners/Planner.cuh:19-44   (abstract plan + result members)
    KGMT     include/planners/KGMT.cuh:5-36        (plan, planBench)

Member names follow the reference so that code written against it reads the
same: after plan(), h_treeSize_, h_itr_, h_pathToGoal_ (0 = not found, else
the goal node index) and h_controlPathToGoal_ (rows goal -> root, D floats
each) hold the results; tree arrays come from tree().
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .abi import KGMT_MAX_SAMPLE, KgmtConfig, KgmtDeviceTree, KgmtProfile, KgmtResult, dims, wdim
from .workloads import config as workload_config


def _f32(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if n is not None and a.size < n:
        b = np.zeros(n, np.float32)
        b[: a.size] = a.ravel()
        a = b
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _obstacle_arg(obstacles, n, w=3):
    """(pointer, on_device, keepalive) for host arrays, torch tensors or raw
    device addresses (ints, like the reference's d_obstacles_ptr); rows of
    2w floats."""
    if obstacles is None:
        return None, 0, None
    if isinstance(obstacles, int):
        return C.c_void_p(obstacles), 1, None
    mod = type(obstacles).__module__
    if mod.startswith("torch"):
        t = obstacles.contiguous().float()
        if t.numel() < 2 * w * n:
            raise ValueError(f"obstacle tensor shorter than {2 * w} * count")
        return C.c_void_p(t.data_ptr()), 1 if t.is_cuda else 0, t
    a = _f32(obstacles).reshape(-1)
    if a.size < 2 * w * n:
        raise ValueError(f"obstacle array shorter than {2 * w} * count")
    return _ptr(a), 0, a


class Planner:
    """include/planners/Planner.cuh:19-44"""

    def __init__(self):
        self.h_treeSize_ = 0
        self.h_itr_ = 0
        self.h_costToGoal_ = 0.0
        self.h_pathToGoal_ = 0
        self.h_controlPathToGoal_ = np.zeros((0, 0), np.float32)

    def plan(self, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount):  # pragma: no cover - abstract
        raise NotImplementedError


class KGMT(Planner):
    """include/planners/KGMT.cuh:5-36, backed by libkgmt (HIP, gfx950)."""

    def __init__(self, cfg: KgmtConfig | None = None):
        super().__init__()
        self.cfg = (cfg or workload_config("reference_quad")).copy()
        self.S, self.C, self.D = dims(self.cfg.model)
        L = _lib.lib()
        h = C.c_void_p()
        _lib.check(L.kgmt_create(C.byref(self.cfg), C.byref(h)))
        self._h = h
        self.result = KgmtResult()
        self.h_pathRows_ = 0
        # start / goal staging: kgmt_plan copies them before it returns, so
        # one pair of buffers (addresses taken once) serves every plan
        self._sg = np.zeros((2, self.D), np.float32)
        self._sg_ptr = (C.c_void_p(self._sg.ctypes.data), C.c_void_p(self._sg.ctypes.data + 4 * self.D))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().kgmt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, hip_stream: int | None):
        _lib.check(_lib.lib().kgmt_set_stream(self._h, C.c_void_p(hip_stream or 0)))

    def set_seed(self, seed: int):
        """Seed of the following plans (the reference seeds from the clock)."""
        self.cfg.seed = int(seed)
        _lib.check(_lib.lib().kgmt_set_seed(self._h, int(seed)))

    def set_profiling(self, on=True):
        """Event timing: False/0 off, True/1 propagate only, 2 every kernel."""
        _lib.check(_lib.lib().kgmt_set_profiling(self._h, int(on)))

    def profile(self) -> dict:
        p = KgmtProfile()
        _lib.check(_lib.lib().kgmt_get_profile(self._h, C.byref(p)))
        return p.as_dict()

    def synchronize(self):
        """Wait for every launch of this planner (a plan returns with its result
        while the launches queued past its last iteration still drain)."""
        _lib.check(_lib.lib().kgmt_synchronize(self._h))

    def set_graph(self, on: bool = True):
        """Replay iterations from a captured hipGraph (ignored while profiling)."""
        _lib.check(_lib.lib().kgmt_set_graph(self._h, 1 if on else 0))

    def set_launch(self, propagate_block=0, propagate_grid=0, chunk_grid=0, lanes_per_child=0):
        """Launch shape of the persistent kernels (0 keeps the current value);
        results do not depend on it."""
        _lib.check(_lib.lib().kgmt_set_launch(self._h, int(propagate_block), int(lanes_per_child),
                                              int(propagate_grid), int(chunk_grid)))

    # KGMT::plan, src/planners/KGMT.cu:41-105
    def _stage(self, row, v):
        a = np.asarray(v, dtype=np.float32).ravel()
        if a.size >= self.D:
            self._sg[row] = a[: self.D]
        else:
            self._sg[row] = 0.0
            self._sg[row, : a.size] = a

    def plan(self, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount) -> KgmtResult:
        self._stage(0, h_initial)
        self._stage(1, h_goal)
        n = int(h_obstaclesCount)
        p, on_dev, keep = _obstacle_arg(d_obstacles_ptr, n, wdim(self.cfg.model))
        res = KgmtResult()
        _lib.check(_lib.lib().kgmt_plan(self._h, self._sg_ptr[0], self._sg_ptr[1], p, n, on_dev, C.byref(res)))
        del keep
        self.result = res
        self._publish()
        return res

    def _publish(self):
        """Result members from self.result (Planner.cuh:31-34)."""
        res = self.result
        self.h_treeSize_ = res.tree_size
        self.h_itr_ = res.iterations
        self.h_pathToGoal_ = res.goal_index if res.success else 0
        self.h_costToGoal_ = res.path_cost if res.success else 0.0
        self.h_pathRows_ = res.path_rows
        self._pathRows = None  # h_controlPathToGoal_, read on first access

    @property
    def h_controlPathToGoal_(self) -> np.ndarray:
        """Planner.cuh:34: the goal path's rows (goal -> root, D floats each) of
        the last plan, read from the library on first access after it."""
        if getattr(self, "_pathRows", None) is None:
            ok = getattr(self, "_h", None) is not None and self.result.success
            self._pathRows = self.path() if ok else np.zeros((0, getattr(self, "D", 0)), np.float32)
        return self._pathRows

    @h_controlPathToGoal_.setter
    def h_controlPathToGoal_(self, rows):
        self._pathRows = rows

    # KGMT::planBench, src/planners/KGMT.cu:107-165: the plan's iterations one
    # at a time, the state after each written as CSV under out_dir like
    # writeDeviceVectorsToCSV (KGMT.cu:405-488)
    def planBench(self, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount, benchItr: int,
                  out_dir: str = "Data", verbose: bool = False) -> KgmtResult:
        from .export import plan_bench_csv
        return plan_bench_csv(self, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount, out_dir, benchItr,
                              verbose=verbose)

    def tree(self):
        """(samples (n, D), parents (n,), costs (n,)) of the last plan."""
        L = _lib.lib()
        n = C.c_uint32()
        _lib.check(L.kgmt_get_tree(self._h, None, None, None, 0, C.byref(n)))
        k = n.value
        s = np.zeros((k, self.D), np.float32)
        p = np.zeros(k, np.int32)
        c = np.zeros(k, np.float32)
        _lib.check(L.kgmt_get_tree(self._h, _ptr(s), _ptr(p), _ptr(c), k, C.byref(n)))
        return s, p, c

    # step-by-step planning (the iterations of KGMT::planBench, KGMT.cu:107-165)
    def plan_begin(self, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount):
        """Prepare a plan like plan() without running an iteration."""
        L = _lib.lib()
        S, Cd, D = dims(self.cfg.model)
        x0 = _f32(h_initial, KGMT_MAX_SAMPLE)
        xg = _f32(h_goal, KGMT_MAX_SAMPLE)
        n = int(h_obstaclesCount)
        optr, on_dev, keep = _obstacle_arg(d_obstacles_ptr, n, wdim(self.cfg.model))
        self._keep = keep
        _lib.check(L.kgmt_plan_begin(self._h, _ptr(x0), _ptr(xg), optr, n, on_dev))
        self.h_treeSize_, self.h_itr_, self.h_pathToGoal_, self.h_pathRows_ = 1, 0, 0, 0

    def plan_step(self) -> bool:
        """Run one iteration; returns True once the plan has ended. The
        result members and getters then describe the state after it."""
        done = C.c_int32()
        _lib.check(_lib.lib().kgmt_plan_step(self._h, C.byref(self.result), C.byref(done)))
        self._publish()
        return bool(done.value)

    def frontier(self):
        """Frontier flags and repeat counts of the tree nodes after the last
        iteration (d_frontier_, d_activeFrontierRepeatCount_)."""
        L = _lib.lib()
        n = C.c_uint32()
        _lib.check(L.kgmt_get_frontier(self._h, None, None, 0, C.byref(n)))
        f = np.zeros(n.value, np.uint8)
        r = np.zeros(n.value, np.uint32)
        _lib.check(L.kgmt_get_frontier(self._h, _ptr(f), _ptr(r), n.value, C.byref(n)))
        return f, r

    def tree_device(self):
        """The tree in the reference's device layout (d_treeSamples_ rows of D
        floats, d_treeSamplesParentIdxs_, d_treeSampleCosts_): a KgmtDeviceTree
        of device addresses into library-owned buffers of `capacity` entries,
        valid until the next plan."""
        t = KgmtDeviceTree()
        _lib.check(_lib.lib().kgmt_get_tree_device(self._h, C.byref(t)))
        return t

    def path(self) -> np.ndarray:
        L = _lib.lib()
        cap = self.cfg.max_iter + 2
        rows = np.zeros((cap, self.D), np.float32)
        n = C.c_uint32()
        _lib.check(L.kgmt_get_path(self._h, _ptr(rows), cap, C.byref(n)))
        return rows[: n.value].copy()

    def regions(self):
        L = _lib.lib()
        n = C.c_uint32()
        _lib.check(L.kgmt_get_regions(self._h, None, None, None, 0, C.byref(n)))
        k = n.value
        t = np.zeros(k, np.int32)
        v = np.zeros(k, np.int32)
        s = np.zeros(k, np.float32)
        _lib.check(L.kgmt_get_regions(self._h, _ptr(t), _ptr(v), _ptr(s), k, C.byref(n)))
        return t, v, s


class BatchKGMT:
    """Many independent queries served by one launch sequence per iteration."""

    def __init__(self, cfgs):
        self.cfgs = [c.copy() for c in cfgs]
        arr = (KgmtConfig * len(self.cfgs))(*self.cfgs)
        h = C.c_void_p()
        _lib.check(_lib.lib().kgmt_batch_create(arr, len(self.cfgs), C.byref(h)))
        self._h = h
        self.n = len(self.cfgs)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().kgmt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, on=True):
        """As KGMT.set_profiling."""
        _lib.check(_lib.lib().kgmt_batch_set_profiling(self._h, int(on)))

    def profile(self) -> dict:
        p = KgmtProfile()
        _lib.check(_lib.lib().kgmt_batch_get_profile(self._h, C.byref(p)))
        return p.as_dict()

    def set_launch(self, propagate_block=0, propagate_grid=0, chunk_grid=0, lanes_per_child=0):
        """As KGMT.set_launch."""
        _lib.check(_lib.lib().kgmt_batch_set_launch(self._h, int(propagate_block), int(lanes_per_child),
                                                    int(propagate_grid), int(chunk_grid)))

    def plan(self, starts, goals, obstacles, offsets, counts, use_graph=True):
        starts = _f32(starts).reshape(self.n, KGMT_MAX_SAMPLE)
        goals = _f32(goals).reshape(self.n, KGMT_MAX_SAMPLE)
        obs = _f32(obstacles).reshape(-1)
        if obs.size == 0:
            obs = np.zeros(6, np.float32)
        offs = np.ascontiguousarray(offsets, np.uint32)
        cnts = np.ascontiguousarray(counts, np.uint32)
        w = wdim(self.cfgs[0].model)
        end = int((offs.astype(np.int64) + cnts).max()) if self.n else 0
        if obs.size < 2 * w * end and end:
            raise ValueError(f"obstacle array shorter than {2 * w} * (offset + count)")
        res = (KgmtResult * self.n)()
        fin = np.zeros(self.n, np.float32)
        _lib.check(_lib.lib().kgmt_batch_plan(self._h, _ptr(starts), _ptr(goals), _ptr(obs), _ptr(offs),
                                              _ptr(cnts), 1 if use_graph else 0, res, _ptr(fin)))
        return [res[i] for i in range(self.n)], fin

    def path(self, q: int) -> np.ndarray:
        D = dims(self.cfgs[q].model)[2]
        cap = self.cfgs[q].max_iter + 2
        rows = np.zeros((cap, D), np.float32)
        n = C.c_uint32()
        _lib.check(_lib.lib().kgmt_batch_get_path(self._h, q, _ptr(rows), cap, C.byref(n)))
        return rows[: n.value].copy()

    def tree(self, q: int):
        L = _lib.lib()
        D = dims(self.cfgs[q].model)[2]
        n = C.c_uint32()
        _lib.check(L.kgmt_batch_get_tree(self._h, q, None, None, None, 0, C.byref(n)))
        k = n.value
        s = np.zeros((k, D), np.float32)
        p = np.zeros(k, np.int32)
        c = np.zeros(k, np.float32)
        _lib.check(L.kgmt_batch_get_tree(self._h, q, _ptr(s), _ptr(p), _ptr(c), k, C.byref(n)))
        return s, p, c
