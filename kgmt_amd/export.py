"""CSV dump of a plan in the reference's Data/ layout, after every iteration.

Mirrors KGMT::planBench (src/planners/KGMT.cu:107-165), which writes the
planner's state after each iteration with writeDeviceVectorsToCSV
(KGMT.cu:405-488) through the writers of include/helper/helper.cuh:56-109,
so the reference's MATLAB viewers (viz/KGMT_3D*.m) read our trees:

  Samples/Samples<b>/samples<itr>.csv        capacity rows of D floats (zeros past the tree)
  Parents/Parents<b>/parents<itr>.csv        capacity rows of one int (-1 past the tree)
  TotalCountPerVertex<b>/totalCountPerVertex.csv, ValidCountPerVertex<b>/...,
  VertexScores<b>/vertexScores.csv           one row of R1 values per iteration
  Frontier<b>/frontier.csv, FrontierRepeatCount<b>/frontierRepeatCount.csv
                                             one row of capacity values per iteration
  FrontierSize<b>, TreeSize<b>, ExpandedNodes<b>  one value per iteration
  ControlPathToGoal<b>/controlPathToGoal.csv the goal -> root rows, when found
  ExecutionTime/executionTime.csv            seconds per plan

Formats are the reference's: floats fixed with six decimals in the array
files and ten in the single-value files, integers and flags as integers,
comma separated; every file is appended to, and the sample and parent files
are one per iteration.
"""
from __future__ import annotations

import io
import os

import numpy as np

SUBDIRS = ("Samples", "Parents", "TotalCountPerVertex", "ValidCountPerVertex", "Frontier", "FrontierRepeatCount",
           "VertexScores", "FrontierSize", "TreeSize", "ExpandedNodes", "ControlPathToGoal")


def _rows_text(a, fmt) -> str:
    a = np.atleast_2d(a)
    buf = io.StringIO()
    np.savetxt(buf, a, fmt=fmt, delimiter=",")
    return buf.getvalue()


def _write(path, text, append):
    with open(path, "a" if append else "w") as f:
        f.write(text)


def _value(path, v):
    """writeValueToCSV: integers as they are, floats with ten decimals."""
    text = f"{v:.10f}" if isinstance(v, (float, np.floating)) else f"{int(v)}"
    _write(path, text + "\n", True)


def _pad(a, n, fill):
    out = np.full((n,) + a.shape[1:], fill, a.dtype)
    out[: len(a)] = a
    return out


def write_iteration(planner, out_dir: str, bench_itr: int, frontier_size: int, expanded: int):
    """The state of `planner` (a kgmt_amd.planner.KGMT) after its last
    iteration, appended to out_dir the way writeDeviceVectorsToCSV does."""
    b, itr, cap = bench_itr, planner.h_itr_, planner.cfg.capacity
    d = lambda sub, name: os.path.join(out_dir, sub, f"{sub}{b}", name)
    for sub in SUBDIRS:
        os.makedirs(os.path.join(out_dir, sub, f"{sub}{b}"), exist_ok=True)
    s, p, _ = planner.tree()
    total, valid, score = planner.regions()
    fr, rep = planner.frontier()
    _write(d("Samples", f"samples{itr}.csv"), _rows_text(_pad(s, cap, np.float32(0)), "%.6f"), True)
    _write(d("Parents", f"parents{itr}.csv"), _rows_text(_pad(p, cap, np.int32(-1))[:, None], "%d"), True)
    _write(d("TotalCountPerVertex", "totalCountPerVertex.csv"), _rows_text(total[None], "%d"), True)
    _write(d("ValidCountPerVertex", "validCountPerVertex.csv"), _rows_text(valid[None], "%d"), True)
    _write(d("Frontier", "frontier.csv"), _rows_text(_pad(fr, cap, np.uint8(0))[None], "%d"), True)
    _write(d("FrontierRepeatCount", "frontierRepeatCount.csv"), _rows_text(_pad(rep, cap, np.uint32(0))[None], "%d"),
           True)
    _write(d("VertexScores", "vertexScores.csv"), _rows_text(score[None], "%.6f"), True)
    _value(d("FrontierSize", "frontierSize.csv"), frontier_size)
    _value(d("TreeSize", "treeSize.csv"), planner.h_treeSize_)
    _value(d("ExpandedNodes", "expandedNodes.csv"), expanded)
    if planner.h_pathToGoal_:
        _write(d("ControlPathToGoal", "controlPathToGoal.csv"), _rows_text(planner.h_controlPathToGoal_, "%.6f"),
               True)


def plan_bench_csv(planner, h_initial, h_goal, obstacles, n_obstacles, out_dir: str, bench_itr: int,
                   verbose: bool = False):
    """KGMT::planBench: plan one iteration at a time and write the state
    after each under out_dir. Returns the plan's result."""
    import time
    t0 = time.perf_counter()
    planner.plan_begin(h_initial, h_goal, obstacles, n_obstacles)
    frontier_size, printed_size, expanded_before, done = 1, 0, 0, False
    while not done:
        if verbose:  # the frontier size the reference prints lags one iteration
            print(f"Iteration: {planner.h_itr_ + 1}, Tree Size: {planner.h_treeSize_}, "
                  f"Frontier Size: {printed_size}", flush=True)
        printed_size = frontier_size
        done = planner.plan_step()
        expanded = int(planner.result.expanded) - expanded_before
        expanded_before = int(planner.result.expanded)
        write_iteration(planner, out_dir, bench_itr, frontier_size, expanded)
        frontier_size = int(planner.frontier()[0].sum())
    os.makedirs(os.path.join(out_dir, "ExecutionTime"), exist_ok=True)
    _value(os.path.join(out_dir, "ExecutionTime", "executionTime.csv"), float(time.perf_counter() - t0))
    return planner.result


def read_rows(path) -> np.ndarray:
    """Read back a file written above (numbers only)."""
    return np.loadtxt(path, delimiter=",", ndmin=2)
