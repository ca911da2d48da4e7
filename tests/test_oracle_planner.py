"""Planner-level checks of the CPU oracle (no GPU): invariants every KGMT
tree must satisfy, the reference's launch-regime arithmetic, and a
brute-force replay of one iteration's bookkeeping with numpy."""
import numpy as np
import pytest

from kgmt_amd import workloads as W
from kgmt_amd.abi import dims


def _check_tree(q, out):
    cfg = q.cfg
    S, Cd, D = dims(cfg.model)
    st = out["stats"]
    n = st["tree_size"]
    assert 1 <= n <= cfg.capacity
    s, p, c = out["samples"], out["parents"], out["costs"]
    assert p[0] == -1 and np.array_equal(s[0], q.start[:D])
    # parents are older nodes
    assert np.all(p[1:] >= 0) and np.all(p[1:] < np.arange(1, n))
    # durations: whole steps within the configured range
    k = np.rint(s[1:, D - 1] / np.float32(cfg.dt))
    assert np.all(k >= cfg.steps_min) and np.all(k <= cfg.steps_max)
    assert np.allclose(s[1:, D - 1], np.float32(cfg.dt) * k.astype(np.float32), rtol=0, atol=0)
    # controls inside their draw ranges
    for i in range(Cd):
        assert np.all(s[1:, S + i] >= cfg.ctrl_min[i]) and np.all(s[1:, S + i] <= cfg.ctrl_max[i])
    # cost = distance of the node position (W coordinates) to the goal
    w = cfg.wdim
    d = np.sqrt(((s[1:, :w].astype(np.float64) - q.goal[:w]) ** 2).sum(1))
    assert np.allclose(c[1:], d, rtol=1e-6, atol=1e-7)
    # appended nodes are valid edges: inside the workspace
    assert np.all(s[1:, :w] >= cfg.ws_min) and np.all(s[1:, :w] <= cfg.ws_max)
    if st["goal_index"] >= 0:
        g = st["goal_index"]
        assert c[g] < cfg.goal_thresh
        # the goal is the smallest index reaching it among the last iteration's nodes
        path = out["path"]
        assert len(path) == st["path_rows"]
        node, rows = g, []
        while node != -1:
            rows.append(s[node])
            node = p[node]
        assert np.array_equal(np.array(rows), path)
        assert abs(st["path_cost"] - s[[x for x in _chain(p, g) if p[x] != -1], D - 1].sum()) < 1e-5
    return n


def _chain(p, g):
    out = []
    while g != -1:
        out.append(g)
        g = p[g]
    return out


@pytest.mark.parametrize("name,seed,kw", [("di6d", 0, dict(capacity=1 << 15)),
                                          ("dubins5d", 1, dict(capacity=1 << 14, max_iter=40)),
                                          ("quad12d", 0, dict(capacity=1 << 12, max_iter=8)),
                                          ("unicycle2d", 2, dict())])
def test_tree_invariants(name, seed, kw, orc):
    q = W.make_query(name, seed, **kw)
    out = orc.plan(q.cfg, q.start, q.goal, q.obstacles, log_iters=True)
    _check_tree(q, out)
    log = out["iter_log"]
    st = out["stats"]
    # launch regime of KGMT.cu:183-201 replayed from the logged multiset sizes
    tree = 1
    B = q.cfg.children_per_entry
    for R, nc, nxt, nreact in log:
        avail = q.cfg.capacity - tree
        per = B if R * B <= avail else max(1, min(avail // R if R else 0, B))
        assert nc == min(R * per, q.cfg.capacity)
        assert nxt <= avail
        tree += nxt
    assert tree == st["tree_size"]
    assert log[0][0] == q.cfg.root_repeat
    assert st["expanded"] == int(log[:, 1].sum())


def test_determinism_and_seed_sensitivity(orc):
    q = W.make_query("di6d", 3, capacity=1 << 14)
    a = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
    b = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
    assert np.array_equal(a["samples"].view(np.uint32), b["samples"].view(np.uint32))
    c = orc.plan(q.cfg.replace(seed=q.cfg.seed + 7), q.start, q.goal, q.obstacles)
    assert not np.array_equal(a["samples"][:100], c["samples"][:100])
    d = orc.plan(q.cfg.replace(query_id=5), q.start, q.goal, q.obstacles)
    assert not np.array_equal(a["samples"][:100], d["samples"][:100])


def test_capacity_clamp_and_full_tree(orc):
    # tiny capacity: appends are clamped (Q9) and the plan ends, unsuccessful,
    # in the iteration that fills the tree
    q = W.make_query("di6d", 6, capacity=300, max_iter=6)
    box = np.array([[0.85, 0.85, 0.85, 1.0, 1.0, 1.0]], np.float32)  # goal unreachable
    out = orc.plan(q.cfg, q.start, q.goal, box, log_iters=True)
    st = out["stats"]
    assert st["tree_size"] == 300 and st["goal_index"] == -1 and st["iterations"] < 6
    full = np.cumsum(out["iter_log"][:, 2]) + 1
    assert np.all(full <= 300) and full[-1] == 300 and np.all(full[:-1] < 300)
    assert out["iter_log"][-1][2] == 300 - full[-2] if len(full) > 1 else True
    _check_tree(q, out)


def test_first_iteration_replay(orc):
    """Iteration 1 rebuilt from the oracle's own stage functions and numpy:
    root expanded root_repeat x 32 times, every valid child in a fresh
    subregion joins the frontier (scores start at 0), first-come order."""
    q = W.make_query("di6d", 2, capacity=1 << 14, max_iter=1)
    cfg = q.cfg
    S, Cd, D = dims(cfg.model)
    out = orc.plan(cfg, q.start, q.goal, q.obstacles, log_iters=True)
    R = cfg.root_repeat
    n = R * cfg.children_per_entry
    ctrl, steps, uf = orc.draw_children(cfg, cfg.query_id, 1, 0, n)
    X = np.tile(q.start[:S, None], (1, 1)).astype(np.float32)
    child, valid, r1, r2 = orc.propagate_children(cfg, np.zeros(n, np.int32), X, ctrl, steps, q.obstacles)
    R1, R2 = cfg.r1_len ** 3, cfg.r1_len ** 3 * cfg.r2_len ** 3
    fnext = valid.astype(bool)  # score 0 < u never holds; only fresh subregions count
    flagged = np.flatnonzero(fnext)
    assert out["stats"]["tree_size"] == 1 + len(flagged)
    assert np.array_equal(out["samples"][1:].view(np.uint32), child[:, flagged].T.copy().view(np.uint32))
    tot = np.bincount(r1, minlength=R1)
    val = np.bincount(r1[valid == 1], minlength=R1)
    assert np.array_equal(out["total"], tot) and np.array_equal(out["valid"], val)
    act = np.zeros(R2, np.uint8)
    act[r2[valid == 1]] = 1
    assert np.array_equal(out["score"].view(np.uint32), orc.score(cfg, tot, val, act).view(np.uint32))


def test_plan_many_matches_plan(orc):
    qs = [W.make_query("di6d", s, capacity=1 << 13) for s in range(4)]
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    stats, ms = orc.plan_many(cfgs, starts, goals, obs, offs, cnts, threads=2)
    for q, st in zip(qs, stats):
        ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles)["stats"]
        assert st == ref
    assert np.all(ms > 0)
