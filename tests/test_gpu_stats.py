"""Planning outcomes over many seeds on the GPU against the oracle's records
(tests/golden/plan_stats_<workload>.json, written by make_plan_stats.py):
50 seeds of the 6D double integrator, 200 of the 5D Dubins airplane, 200 of
the 12D quadcopter and 200 of the planar unicycle at their full sizes, and
the full mixed batch workload (512 queries, double integrator / quadcopter,
capacities log-uniform in [2^12, 2^17], 10..500 boxes each).

Every seed must match exactly (iterations, tree size, goal node, path rows,
path cost and length bits, expanded children); on top of that the two
distributions are compared the way a statistical check would: success rates
within 2 percentage points and a two-sample Kolmogorov-Smirnov test on the
path costs of solved queries."""
import json
import os
import struct

import numpy as np
import pytest
import torch

from kgmt_amd import workloads as W
from kgmt_amd.planner import KGMT, BatchKGMT

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _f32(bits):
    return struct.unpack("<f", struct.pack("<I", bits))[0]


def _load(name):
    path = os.path.join(HERE, f"plan_stats_{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated")
    with open(path) as f:
        return json.load(f)


def _ks(a, b):
    """Two-sample KS statistic and its asymptotic p-value."""
    a, b = np.sort(a), np.sort(b)
    x = np.concatenate([a, b])
    d = np.max(np.abs(np.searchsorted(a, x, side="right") / len(a) - np.searchsorted(b, x, side="right") / len(b)))
    en = np.sqrt(len(a) * len(b) / (len(a) + len(b)))
    lam = (en + 0.12 + 0.11 / en) * d
    k = np.arange(1, 101)
    p = float(np.clip(2 * np.sum((-1) ** (k - 1) * np.exp(-2 * (k * lam) ** 2)), 0, 1)) if d > 0 else 1.0
    return d, p


def _records(results):
    return [[s, r.iterations, r.tree_size, r.goal_index if r.success else -1, r.path_rows if r.success else 0,
             struct.unpack("<I", struct.pack("<f", r.path_cost))[0] if r.success else 0,
             struct.unpack("<I", struct.pack("<f", r.path_length))[0] if r.success else 0, r.expanded]
            for s, r in enumerate(results)]


def _compare(name, got, ref):
    bad = [(g, r) for g, r in zip(got, ref["records"]) if g != r]
    assert not bad, f"{name}: {len(bad)} seeds differ, first {bad[0]}"
    sg = np.array([g[3] >= 0 for g in got])
    sr = np.array([r[3] >= 0 for r in ref["records"]])
    assert abs(sg.mean() - sr.mean()) <= 0.02
    cg = [_f32(g[5]) for g in got if g[3] >= 0]
    cr = [_f32(r[5]) for r in ref["records"] if r[3] >= 0]
    d, p = _ks(np.array(cg), np.array(cr))
    print(f"{name}: {len(got)} seeds, success {sg.mean():.3f} (oracle {sr.mean():.3f}), "
          f"median path cost {np.median(cg):.4f}, KS D={d:.4f} p={p:.3f}", flush=True)
    assert p > 0.05


@pytest.mark.parametrize("name", ["di6d", "dubins5d", "quad12d", "unicycle2d"])
def test_plain_planner_over_seeds(name):
    ref = _load(name)
    n = ref["seeds"]
    qs = [W.make_query(name, s) for s in range(n)]
    assert qs[0].cfg.capacity == ref["capacity"] and qs[0].cfg.max_iter == ref["max_iter"]
    print(f"plain {name}: {n} seeds", flush=True)
    g = KGMT(qs[0].cfg)
    results = []
    for q in qs:
        d = torch.from_numpy(q.obstacles.reshape(-1)).cuda()
        g.set_seed(q.cfg.seed)
        results.append(g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles)))
    _compare(name, _records(results), ref)


@pytest.mark.parametrize("name", ["di6d", "dubins5d", "quad12d", "unicycle2d"])
def test_batched_planner_over_seeds(name):
    ref = _load(name)
    n = ref["seeds"]
    qs = [W.make_query(name, s) for s in range(n)]
    print(f"batched {name}: {n} queries in one batch", flush=True)
    b = BatchKGMT([q.cfg for q in qs])
    results, _ = b.plan(*W.pack_batch(qs)[1:], use_graph=True)
    _compare(name, _records(results), ref)


@pytest.mark.parametrize("use_graph", [True, False])
def test_full_mixed_batch(use_graph):
    """All 512 queries of the mixed batch workload in one batch, every query
    identical to the oracle's record of it."""
    ref = _load("batch512")
    qs = W.make_batch(ref["seeds"], seed=0)
    assert [q.cfg.capacity for q in qs] == ref["capacity"]
    print(f"mixed batch: {len(qs)} queries, graph={use_graph}", flush=True)
    b = BatchKGMT([q.cfg for q in qs])
    results, _ = b.plan(*W.pack_batch(qs)[1:], use_graph=use_graph)
    _compare("batch512", _records(results), ref)
