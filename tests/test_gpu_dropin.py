"""The C++ drop-in class (include/kgmt_planner.hpp) driven like the
reference's examples/gpu/executionTimeMain.cu — readObstaclesFromCSV with
W_DIM, SAMPLE_DIM start/goal arrays, the reference's quadTrees obstacle file
and start/goal, device obstacle pointer — checked against the CPU oracle on
the same inputs: result members, h_controlPathToGoal_ rows, and the device
tree members (rows, parents, costs; entries past the tree 0, -1, 0) read
back with hipMemcpy and compared with the host getters and the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from kgmt_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_main")
CSV = os.path.join(ROOT, "tests", "golden", "obstacles", "quadTrees.csv")


def _bit_sum(a):
    return int(np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64).sum())


@pytest.mark.parametrize("cap,iters,seed", [(1 << 14, 30, 7), (1 << 13, 12, 3)])
def test_cpp_dropin_matches_oracle(cap, iters, seed, orc):
    assert os.path.exists(EXE), "tests/cpp/dropin_main is built by python -m kgmt_amd.build"
    out = subprocess.run([EXE, CSV, str(cap), str(iters), str(seed), "2"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    runs = [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]
    assert len(runs) == 2
    cfg = W.config("reference_quad", capacity=cap, max_iter=iters, seed=seed)
    obs = W.read_obstacles_csv(CSV)
    start = np.zeros(17, np.float32)
    start[:3] = [10.0, 8.0, 5.0]
    goal = np.zeros(17, np.float32)
    goal[:3] = [80.0, 95.0, 90.0]
    ref = orc.plan(cfg, start, goal, obs)
    st = ref["stats"]
    for r in runs:  # the second plan on the same object starts from scratch
        assert r["device_tree_matches_host"] and r["device_tree_tail_ok"]
        assert r["tree_size"] == st["tree_size"]
        assert r["iterations"] == st["iterations"]
        assert r["goal"] == (st["goal_index"] if st["goal_index"] >= 0 else 0)
        assert r["path_rows"] == st["path_rows"]
        assert r["parent_sum"] == int(ref["parents"].astype(np.int64).sum())
        assert r["sample_bits"] == _bit_sum(ref["samples"])
        assert r["cost_bits"] == _bit_sum(ref["costs"])
        assert r["path_bits"] == _bit_sum(ref["path"])
