"""CSV export of a finished plan in the reference's Data/ layout.

Mirrors KGMT::writeDeviceVectorsToCSV (src/planners/KGMT.cu:405-488) and the
writers of include/helper/helper.cuh:56-109 so that the reference's MATLAB
viewers (viz/KGMT_3D*.m) can read our trees: fixed six decimals, one row per
tree node, comma separated; scalar files are appended one value per line.
The reference writes every iteration; this writes the state after the plan.
"""
from __future__ import annotations

import os

import numpy as np


def _write_rows(path, rows, append=False):
    rows = np.atleast_2d(rows)
    with open(path, "a" if append else "w") as f:
        for r in rows:
            f.write(",".join(f"{float(x):.6f}" for x in r) + "\n")


def _append_value(path, v, fmt="{}"):
    with open(path, "a") as f:
        f.write(fmt.format(v) + "\n")


def write_plan_csv(planner, out_dir: str, bench_itr: int) -> str:
    """Write samples, parents, region counters/scores, tree size, iteration
    count and the goal->root path of `planner` (a kgmt_amd.planner.KGMT after
    plan()) under out_dir. Returns out_dir."""
    it = planner.h_itr_
    s, p, c = planner.tree()
    total, valid, score = planner.regions()
    d = lambda *parts: os.path.join(out_dir, *parts)
    for sub in ("Samples", "Parents", "TotalCountPerVertex", "ValidCountPerVertex", "VertexScores", "TreeSize",
                "ControlPathToGoal", "ExecutionTime"):
        os.makedirs(d(sub, f"{sub}{bench_itr}") if sub != "ExecutionTime" else d(sub), exist_ok=True)
    _write_rows(d("Samples", f"Samples{bench_itr}", f"samples{it}.csv"), s)
    _write_rows(d("Parents", f"Parents{bench_itr}", f"parents{it}.csv"), p[:, None].astype(np.float64))
    _write_rows(d("TotalCountPerVertex", f"TotalCountPerVertex{bench_itr}", "totalCountPerVertex.csv"),
                total[None].astype(np.float64))
    _write_rows(d("ValidCountPerVertex", f"ValidCountPerVertex{bench_itr}", "validCountPerVertex.csv"),
                valid[None].astype(np.float64))
    _write_rows(d("VertexScores", f"VertexScores{bench_itr}", "vertexScores.csv"), score[None])
    _append_value(d("TreeSize", f"TreeSize{bench_itr}", "treeSize.csv"), planner.h_treeSize_)
    if planner.h_pathToGoal_:
        _write_rows(d("ControlPathToGoal", f"ControlPathToGoal{bench_itr}", "controlPathToGoal.csv"),
                    planner.h_controlPathToGoal_)
    _append_value(d("ExecutionTime", "executionTime.csv"), planner.result.time_ms / 1000.0, "{:.10f}")
    return out_dir


def read_rows(path) -> np.ndarray:
    """Read back a file written above (numbers only)."""
    return np.loadtxt(path, delimiter=",", ndmin=2)
