"""Whole plans on the GPU against the CPU oracle: same tree, node by node.

Small capacities keep the oracle to seconds; the GPU runs exactly the same
iterations, so every sample, parent, cost, region counter and score must be
identical, and so must the path."""
import numpy as np
import pytest

from kgmt_amd import workloads as W
from kgmt_amd._lib import KgmtError
from kgmt_amd.abi import KGMT_DI_EXACT
from kgmt_amd.planner import KGMT, BatchKGMT

pytestmark = pytest.mark.gpu


def _same_plan(gpu: KGMT, res, ref):
    st = ref["stats"]
    assert res.iterations == st["iterations"]
    assert res.tree_size == st["tree_size"]
    assert res.expanded == st["expanded"]
    assert (res.goal_index if res.success else -1) == st["goal_index"]
    assert res.path_rows == st["path_rows"]
    s, p, c = gpu.tree()
    assert np.array_equal(p, ref["parents"])
    assert np.array_equal(s.view(np.uint32), ref["samples"].view(np.uint32))
    assert np.array_equal(c.view(np.uint32), ref["costs"].view(np.uint32))
    t, v, sc = gpu.regions()
    assert np.array_equal(t, ref["total"]) and np.array_equal(v, ref["valid"])
    assert np.array_equal(sc.view(np.uint32), ref["score"].view(np.uint32))
    if res.success:
        assert np.array_equal(gpu.path().view(np.uint32), ref["path"].view(np.uint32))
        assert np.float32(res.path_cost) == np.float32(st["path_cost"])
        assert np.float32(res.path_length) == np.float32(st["path_length"])


PLAN_CASES = [
    ("di6d", 0, dict(capacity=1 << 14)),
    ("di6d", 1, dict(capacity=1 << 15, integrator=KGMT_DI_EXACT)),
    ("dubins5d", 0, dict(capacity=1 << 15, max_iter=40)),
    ("dubins5d", 2, dict(capacity=5000, max_iter=30)),
    ("quad12d", 0, dict(capacity=1 << 13, max_iter=12)),
    ("unicycle2d", 0, dict()),
    ("unicycle2d", 4, dict(capacity=20000)),
    ("di6d", 2, dict(capacity=1 << 15, r1_len=20, r2_len=4)),  # 8000 regions: many distinct keys per tile
]


@pytest.mark.parametrize("name,seed,kw", PLAN_CASES)
def test_plan_matches_oracle(name, seed, kw, orc):
    q = W.make_query(name, seed, **kw)
    ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
    g = KGMT(q.cfg)
    res = g.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
    _same_plan(g, res, ref)
    # a second plan on the same context starts from scratch; then graph replay
    res2 = g.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
    _same_plan(g, res2, ref)
    g.set_graph(True)
    for _ in range(2):
        res3 = g.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
        _same_plan(g, res3, ref)


def test_full_tree_ends_the_plan(orc):
    q = W.make_query("dubins5d", 1, capacity=3000, max_iter=60)
    box = np.array([[0.8, 0.8, 0.8, 1.0, 1.0, 1.0]], np.float32)
    ref = orc.plan(q.cfg, q.start, q.goal, box)
    assert ref["stats"]["tree_size"] == 3000 and ref["stats"]["iterations"] < 60
    g = KGMT(q.cfg)
    res = g.plan(q.start, q.goal, box, 1)
    _same_plan(g, res, ref)


def test_plans_back_to_back_and_synchronize(orc):
    """A plan returns once its result is published, leaving launches queued
    past its last iteration to drain: the next plan on the same context, the
    getters and synchronize() still see the right state."""
    qa = W.make_query("unicycle2d", 0)
    qb = W.make_query("unicycle2d", 4)  # same configuration apart from the seed
    ra = orc.plan(qa.cfg, qa.start, qa.goal, qa.obstacles)
    rb = orc.plan(qb.cfg, qb.start, qb.goal, qb.obstacles)
    g = KGMT(qa.cfg)
    for _ in range(2):
        g.set_seed(qa.cfg.seed)
        res = g.plan(qa.start, qa.goal, qa.obstacles, len(qa.obstacles))
        assert res.tree_size == ra["stats"]["tree_size"] and res.iterations == ra["stats"]["iterations"]
        g.set_seed(qb.cfg.seed)
        res = g.plan(qb.start, qb.goal, qb.obstacles, len(qb.obstacles))
        _same_plan(g, res, rb)
    g.synchronize()


def test_long_max_iter_plans_synchronously(orc):
    """A configuration whose longest possible path does not fit the host path
    buffer (max_iter 200000) publishes no early result: the plan waits for its
    launches and copies the path from the device, with the same answers."""
    q = W.make_query("unicycle2d", 0, max_iter=200000)
    ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
    g = KGMT(q.cfg)
    res = g.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
    _same_plan(g, res, ref)


@pytest.mark.parametrize("block", [512, 256])
def test_max_obstacles_plan_matches_oracle(block, orc):
    """1500 boxes, the most a query may carry: the scene and the per-wave
    candidate lists fill the most LDS a propagate workgroup uses."""
    q = W.make_query("dubins5d", 3, capacity=4000, max_iter=10)
    obs = W.random_boxes(np.random.default_rng(5), 1500, 0.005, 0.03)
    ref = orc.plan(q.cfg, q.start, q.goal, obs)
    g = KGMT(q.cfg)
    g.set_launch(propagate_block=block, propagate_grid=256, chunk_grid=128)
    res = g.plan(q.start, q.goal, obs, len(obs))
    _same_plan(g, res, ref)


def test_profile_counts_children_and_flags(orc):
    """Profiling leaves the plan unchanged and counts the children and the
    flagged children (every flagged child is appended while the tree has
    room, so they are the tree's nodes after the root; a full tree appends
    fewer than it flags)."""
    for name, seed, kw, full in [("unicycle2d", 0, dict(), False),
                                 ("dubins5d", 1, dict(capacity=3000, max_iter=60), True)]:
        q = W.make_query(name, seed, **kw)
        obs = np.array([[0.8, 0.8, 0.8, 1.0, 1.0, 1.0]], np.float32) if full else q.obstacles
        ref = orc.plan(q.cfg, q.start, q.goal, obs)
        g = KGMT(q.cfg)
        for level in (1, 2):
            g.set_profiling(level)
            res = g.plan(q.start, q.goal, obs, len(obs))
            p = g.profile()
            g.set_profiling(0)
            _same_plan(g, res, ref)
            assert p["children"] == res.expanded
            if full:
                assert res.tree_size == q.cfg.capacity and p["flagged"] >= res.tree_size - 1
            else:
                assert res.tree_size < q.cfg.capacity and p["flagged"] == res.tree_size - 1
            assert p["kernels"]["propagate"]["launches"] >= res.iterations
            if level == 2:  # two launches per single-query iteration
                k = p["kernels"]
                assert k["update"]["launches"] == k["propagate"]["launches"]
                assert k["finish"]["launches"] == 0


def test_plan_device_obstacles(orc):
    torch = pytest.importorskip("torch")
    q = W.make_query("di6d", 4, capacity=1 << 14)
    ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
    g = KGMT(q.cfg)
    d_obs = torch.from_numpy(q.obstacles.reshape(-1)).cuda()
    res = g.plan(q.start, q.goal, int(d_obs.data_ptr()), len(q.obstacles))
    _same_plan(g, res, ref)


def test_plan_no_obstacles_and_unreachable(orc):
    q = W.make_query("di6d", 5, capacity=1 << 13, max_iter=15)
    empty = np.zeros((0, 6), np.float32)
    ref = orc.plan(q.cfg, q.start, q.goal, empty)
    g = KGMT(q.cfg)
    res = g.plan(q.start, q.goal, empty, 0)
    _same_plan(g, res, ref)
    # goal enclosed by a box: never reached; the plan ends when the tree is full
    box = np.array([[0.85, 0.85, 0.85, 1.0, 1.0, 1.0]], np.float32)
    ref = orc.plan(q.cfg, q.start, q.goal, box)
    res = g.plan(q.start, q.goal, box, 1)
    assert not res.success and res.tree_size == q.cfg.capacity
    _same_plan(g, res, ref)
    # with a large capacity it runs all max_iter iterations instead
    big = q.cfg.replace(capacity=1 << 17, max_iter=6)
    ref = orc.plan(big, q.start, q.goal, box)
    g2 = KGMT(big)
    res = g2.plan(q.start, q.goal, box, 1)
    assert not res.success and res.iterations == 6
    _same_plan(g2, res, ref)


def test_degenerate_starts_match_oracle(orc):
    q = W.make_query("di6d", 5, capacity=1 << 13, max_iter=15)
    g = KGMT(q.cfg)
    # start inside a box: every edge from the root collides, the tree stays at
    # the root, which is re-activated every iteration until max_iter
    s = np.asarray(q.start, np.float32)
    box = np.concatenate([s[:3] - 0.05, s[:3] + 0.05])[None].astype(np.float32)
    ref = orc.plan(q.cfg, q.start, q.goal, box)
    res = g.plan(q.start, q.goal, box, 1)
    assert not res.success and res.tree_size == 1 and res.iterations == q.cfg.max_iter
    _same_plan(g, res, ref)
    # start at the goal: only nodes appended by an iteration are goal-tested (KGMT.cu:342), so
    # the plan ends at the first child within the threshold, not at the root
    empty = np.zeros((0, 6), np.float32)
    ref = orc.plan(q.cfg, q.start, q.start, empty)
    res = g.plan(q.start, q.start, empty, 0)
    assert res.success and res.goal_index > 0 and res.iterations == 1
    _same_plan(g, res, ref)
    # a single iteration
    one = q.cfg.replace(max_iter=1)
    ref = orc.plan(one, q.start, q.goal, q.obstacles)
    g1 = KGMT(one)
    res = g1.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
    assert res.iterations == 1
    _same_plan(g1, res, ref)
    # the same queries in one batch, beside an ordinary one
    qs = [W.Query(q.cfg, q.start, q.goal, box), W.Query(q.cfg, q.start, q.start, empty),
          W.Query(one, q.start, q.goal, q.obstacles), q]
    b = BatchKGMT([x.cfg for x in qs])
    results, _ = b.plan(*W.pack_batch(qs)[1:])
    for i, x in enumerate(qs):
        ref = orc.plan(x.cfg, x.start, x.goal, x.obstacles)
        st, r = ref["stats"], results[i]
        assert (r.iterations, r.tree_size, r.goal_index if r.success else -1) == \
            (st["iterations"], st["tree_size"], st["goal_index"]), i
        s, p, c = b.tree(i)
        assert np.array_equal(p, ref["parents"]), i
        assert np.array_equal(s.view(np.uint32), ref["samples"].view(np.uint32)), i
        if r.success:
            assert np.array_equal(b.path(i).view(np.uint32), ref["path"].view(np.uint32)), i


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_oracle(use_graph, orc):
    qs = W.make_batch(12, seed=3, max_iter=25)
    for q in qs:  # keep the oracle fast
        q.cfg = q.cfg.replace(capacity=min(q.cfg.capacity, 1 << 13))
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    b = BatchKGMT(cfgs)
    results, fin = b.plan(starts, goals, obs, offs, cnts, use_graph=use_graph)
    for i, q in enumerate(qs):
        ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
        st = ref["stats"]
        r = results[i]
        assert r.iterations == st["iterations"], i
        assert r.tree_size == st["tree_size"], i
        assert (r.goal_index if r.success else -1) == st["goal_index"], i
        s, p, c = b.tree(i)
        assert np.array_equal(p, ref["parents"]), i
        assert np.array_equal(s.view(np.uint32), ref["samples"].view(np.uint32)), i
        if r.success:
            assert np.array_equal(b.path(i).view(np.uint32), ref["path"].view(np.uint32)), i
    assert np.all(fin > 0)
    # replaying the same batch gives the same answers
    results2, _ = b.plan(starts, goals, obs, offs, cnts, use_graph=use_graph)
    assert [r.tree_size for r in results2] == [r.tree_size for r in results]


@pytest.mark.parametrize("shape", [dict(propagate_block=128, propagate_grid=64, chunk_grid=32),
                                   dict(propagate_block=512, propagate_grid=2048, chunk_grid=1024),
                                   dict(propagate_block=256, propagate_grid=1, chunk_grid=1),
                                   dict(propagate_block=256, lanes_per_child=2, propagate_grid=512, chunk_grid=256),
                                   dict(propagate_block=256, lanes_per_child=4, propagate_grid=2048, chunk_grid=1024),
                                   dict(propagate_block=256, lanes_per_child=4, propagate_grid=3, chunk_grid=5)])
def test_launch_shape_does_not_change_results(shape, orc):
    """Workgroups of 128/256/512 threads, 1, 2 or 4 lanes per child (tiles
    of 64 to 512 children) and any persistent grid give the oracle's trees,
    plain and batched, with and without the graph."""
    q = W.make_query("dubins5d", 4, capacity=1 << 14, max_iter=30)
    ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
    g = KGMT(q.cfg)
    g.set_launch(**shape)
    for graph in (False, True):
        g.set_graph(graph)
        res = g.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
        _same_plan(g, res, ref)
    qs = W.make_batch(8, seed=11, max_iter=20)
    for qq in qs:
        qq.cfg = qq.cfg.replace(capacity=min(qq.cfg.capacity, 1 << 13))
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    b = BatchKGMT(cfgs)
    b.set_launch(**shape)
    results, _ = b.plan(starts, goals, obs, offs, cnts, use_graph=True)
    for i, qq in enumerate(qs):
        st = orc.plan(qq.cfg, qq.start, qq.goal, qq.obstacles)["stats"]
        assert (results[i].iterations, results[i].tree_size) == (st["iterations"], st["tree_size"]), i


def test_set_launch_rejects_bad_shapes():
    q = W.make_query("di6d", 0, capacity=1 << 12)
    g = KGMT(q.cfg)
    for bad in (dict(propagate_block=64), dict(propagate_block=1024), dict(propagate_grid=1 << 21),
                dict(propagate_block=128, lanes_per_child=2), dict(lanes_per_child=3),
                dict(propagate_block=512, lanes_per_child=4)):
        with pytest.raises(KgmtError):
            g.set_launch(**bad)
    g.set_launch(propagate_block=512)
    res = g.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
    assert res.tree_size >= 1


def test_dubins_airplane_plan_matches_oracle(orc):
    """The reference's 6D Dubins airplane (speed as a state, acceleration as
    a control): whole plans node by node, plain and batched."""
    import math
    from kgmt_amd.abi import KGMT_MODEL_DUBINS_AIRPLANE
    base = W.make_query("dubins5d", 5, capacity=1 << 14, max_iter=25)
    pi = math.pi
    cfg = base.cfg.replace(model=KGMT_MODEL_DUBINS_AIRPLANE, dt=0.05, steps_min=2, steps_max=12, v_min=0.05,
                           v_max=0.4, ctrl_min=[-1.0, -0.5, -0.3, 0], ctrl_max=[1.0, 0.5, 0.3, 0])
    start = np.zeros(10, np.float32)
    start[:5] = base.start[:5]
    start[5] = 0.2
    goal = np.zeros(10, np.float32)
    goal[:3] = base.goal[:3]
    ref = orc.plan(cfg, start, goal, base.obstacles)
    g = KGMT(cfg)
    res = g.plan(start, goal, base.obstacles, len(base.obstacles))
    _same_plan(g, res, ref)
    assert ref["stats"]["tree_size"] > 100


def test_unicycle_batch_and_device_obstacles(orc):
    """Planar unicycle queries batched (rows of 4 floats per box), with and
    without the graph, and a plain plan from device-resident planar rows."""
    import torch
    qs = [W.make_query("unicycle2d", s, capacity=1 << 14, query_id=s) for s in range(6)]
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    assert obs.shape[1] == 4
    b = BatchKGMT(cfgs)
    refs = [orc.plan(q.cfg, q.start, q.goal, q.obstacles) for q in qs]
    for use_graph in (False, True):
        results, _ = b.plan(starts, goals, obs, offs, cnts, use_graph=use_graph)
        for i, ref in enumerate(refs):
            assert (results[i].iterations, results[i].tree_size) == (ref["stats"]["iterations"],
                                                                     ref["stats"]["tree_size"]), i
            s, p, _ = b.tree(i)
            assert np.array_equal(s.view(np.uint32), ref["samples"].view(np.uint32)), i
    q = qs[2]
    g = KGMT(q.cfg)
    d = torch.from_numpy(q.obstacles.reshape(-1)).cuda()
    res = g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles))
    _same_plan(g, res, refs[2])


def test_batch_rejects_mixed_workspace_dims():
    cfgs = [W.config("unicycle2d", capacity=1 << 12), W.config("di6d", capacity=1 << 12)]
    with pytest.raises(KgmtError):
        BatchKGMT(cfgs)


def _odd_boxes(w, rng):
    """Boxes the separating-axis test must treat exactly as the reference's
    comparisons do: NaN faces, infinite faces (slabs), flat and inverted
    boxes, next to ordinary ones."""
    nan, inf = float("nan"), float("inf")
    base = W.random_boxes(rng, 40, 0.03, 0.12, keep_clear=[(np.full(w, 0.05), 0.08), (np.full(w, 0.95), 0.08)],
                          dim=w).astype(np.float64)
    odd = []
    for k in range(12):
        lo = rng.uniform(0.2, 0.8, w)
        hi = lo + rng.uniform(0.02, 0.1, w)
        b = np.concatenate([lo, hi])
        a = k % w
        if k % 6 == 0:
            b[a] = nan                      # NaN low face
        elif k % 6 == 1:
            b[w + a] = nan                  # NaN high face
        elif k % 6 == 2:
            b[a], b[w + a] = -inf, 0.1      # slab open below
        elif k % 6 == 3:
            b[w + a] = inf                  # slab open above
        elif k % 6 == 4:
            b[w + a] = b[a]                 # flat box
        else:
            b[a], b[w + a] = b[w + a], b[a]  # inverted on one axis
        odd.append(b)
    return np.concatenate([base, np.asarray(odd)]).astype(np.float32)


@pytest.mark.parametrize("name", ["dubins5d", "di6d", "unicycle2d"])
def test_unusual_obstacles_match_oracle(name, orc):
    """Whole plans (candidate prefilter, register-held candidates, face-gap
    test) against the oracle's literal comparisons, on scenes with NaN and
    infinite faces, flat and inverted boxes."""
    rng = np.random.default_rng(21)
    q = W.make_query(name, 7, capacity=1 << 14, max_iter=40)
    obs = _odd_boxes(q.cfg.wdim, rng)
    ref = orc.plan(q.cfg, q.start, q.goal, obs)
    g = KGMT(q.cfg)
    res = g.plan(q.start, q.goal, obs, len(obs))
    _same_plan(g, res, ref)
    assert ref["stats"]["tree_size"] > 50


def test_long_dubins_edges_match_oracle(orc):
    """Edges of 100-300 steps: the rotated sin/cos drift over many steps and
    the edge bound's per-step padding keep every candidate."""
    q = W.make_query("dubins5d", 8, capacity=1 << 13, max_iter=12, dt=0.005, steps_min=100, steps_max=300)
    ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles)
    g = KGMT(q.cfg)
    res = g.plan(q.start, q.goal, q.obstacles, len(q.obstacles))
    _same_plan(g, res, ref)
    assert ref["stats"]["tree_size"] > 20
