"""Pins of the oracle's planar unicycle (statePropagator.cu:13-71) and of the
planar workspace (W_DIM = 2) pieces it needs (no GPU).

Independent checks: closed-form sums of the Euler recursion (straight lines
and constant turns), FP32 against the FP64 twin, a lattice brute force of the
planar segment test and its equality with the 3D test on boxes spanning
z [-1, 1], exhaustive region maps, and the 2D maps against the already
pinned 3D maps restricted to z = 0.
"""
import math

import numpy as np
import pytest

from kgmt_amd import workloads as W
from kgmt_amd.abi import dims


def _free(cfg):
    return cfg.replace(ws_min=-1e6, ws_max=1e6)


def test_unicycle_dims_and_preset():
    cfg = W.config("unicycle2d")
    assert dims(cfg.model) == (4, 2, 7) and cfg.wdim == 2 and cfg.regions == cfg.r1_len ** 2


def _straight(x0, a, h, n):
    """Euler with zero steering: heading fixed, v_k = v0 + k a h."""
    x, y, th, v = x0
    dist = h * (n * v + a * h * n * (n - 1) / 2.0)
    return np.array([x + math.cos(th) * dist, y + math.sin(th) * dist, th, v + n * a * h])


def _turn(x0, delta, length, h, n):
    """Euler with a = 0: heading advances by phi = (v / L) tan(delta) h per
    step; sum_k cos(th0 + k phi) = sin(n phi / 2) / sin(phi / 2) cos(th0 + (n - 1) phi / 2)."""
    x, y, th, v = x0
    phi = v / length * math.tan(delta) * h
    if abs(phi) < 1e-300:
        cs, sn = n * math.cos(th), n * math.sin(th)
    else:
        k = math.sin(n * phi / 2) / math.sin(phi / 2)
        cs, sn = k * math.cos(th + (n - 1) * phi / 2), k * math.sin(th + (n - 1) * phi / 2)
    return np.array([x + v * h * cs, y + v * h * sn, th + n * phi, v])


def test_unicycle_closed_forms(orc):
    cfg = _free(W.config("unicycle2d"))
    h = float(np.float32(cfg.dt))
    rng = np.random.default_rng(11)
    n = 400
    x0 = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(-3, 3, n), rng.uniform(-0.5, 0.5, n)], 1)
    u = np.stack([rng.uniform(-0.2, 0.2, n), rng.uniform(-1.4, 1.4, n)], 1)
    straight = np.arange(n) % 2 == 0
    u[straight, 1] = 0.0          # zero steering
    u[~straight, 0] = 0.0         # zero acceleration
    steps = rng.integers(1, 11, n).astype(np.int32)
    x1, valid = orc.propagate_f64(cfg, x0, u, steps, np.zeros((0, 4)))
    assert np.all(valid == 1)
    for i in range(n):
        want = (_straight(x0[i], u[i, 0], h, steps[i]) if straight[i]
                else _turn(x0[i], u[i, 1], cfg.uni_length, h, steps[i]))
        assert np.allclose(x1[i, :4], want, rtol=0, atol=1e-12), i
        assert x1[i, 4] == u[i, 0] and x1[i, 5] == u[i, 1] and x1[i, 6] == h * steps[i]
    # the FP32 restatement within float resolution of the FP64 twin
    X = x0.T.astype(np.float32)
    U = u.T.astype(np.float32)
    child, v32, _, _ = orc.propagate_children(cfg, np.arange(n, dtype=np.int32), X, U, steps,
                                              np.zeros((0, 4), np.float32))
    ref, _ = orc.propagate_f64(cfg, X.T.astype(np.float64), U.T.astype(np.float64), steps, np.zeros((0, 4)))
    assert np.all(v32 == 1)
    assert np.max(np.abs(child[:4].T - ref[:, :4]) / np.maximum(np.abs(ref[:, :4]), 1.0)) < 2e-6


def test_unicycle_fp32_matches_fp64_with_obstacles(orc):
    q = W.make_query("unicycle2d", 3)
    cfg = q.cfg
    n = 4000
    rng = np.random.default_rng(12)
    X = np.zeros((4, n), np.float32)
    X[:2] = rng.uniform(0.05, 0.95, (2, n))
    X[2] = rng.uniform(-math.pi, math.pi, n)
    X[3] = rng.uniform(-0.6, 0.6, n)
    ctrl, steps, _ = orc.draw_children(cfg, 0, 1, 0, n)
    c32, v32, r1, r2 = orc.propagate_children(cfg, np.arange(n, dtype=np.int32), X, ctrl, steps, q.obstacles)
    c64, v64 = orc.propagate_f64(cfg, X.T.astype(np.float64), ctrl.T.astype(np.float64), steps,
                                 q.obstacles.astype(np.float64))
    assert 0.2 < v32.mean() < 0.98          # the boxes and walls matter
    assert (v32 == v64).mean() > 0.995
    both = (v32 == v64) & (v32 == 1)
    assert np.max(np.abs(c32[:4, both].T - c64[both, :4]) / np.maximum(np.abs(c64[both, :4]), 1.0)) < 1e-5
    # regions of the children are the 2D maps of their positions
    a, b = orc.regions(cfg, np.stack([c32[0], c32[1], np.zeros(n, np.float32)], 1))
    assert np.array_equal(a, r1) and np.array_equal(b, r2)


def test_unicycle_draws(orc):
    cfg = W.config("unicycle2d")
    ctrl, steps, uf = orc.draw_children(cfg, 0, 1, 0, 100000)
    assert ctrl.shape == (2, 100000)
    for i in range(2):
        assert ctrl[i].min() >= cfg.ctrl_min[i] and ctrl[i].max() <= cfg.ctrl_max[i]
        assert abs(ctrl[i].mean() - (cfg.ctrl_min[i] + cfg.ctrl_max[i]) / 2) < 0.01 * (cfg.ctrl_max[i] - cfg.ctrl_min[i])
    assert steps.min() == cfg.steps_min and steps.max() == cfg.steps_max


def _lattice_hit2(p0, p1, box):
    """Brute force as for the 3D test: coordinates are multiples of 0.25, so
    the segment box's points at step 0.125 decide whether it meets the open box."""
    lo, hi = np.minimum(p0, p1), np.maximum(p0, p1)
    axes = [np.arange(lo[d], hi[d] + 1e-9, 0.125) if hi[d] > lo[d] else np.array([lo[d]]) for d in range(2)]
    return any(box[0] < x < box[2] and box[1] < y < box[3] for x in axes[0] for y in axes[1])


def test_planar_segment_test_lattice_and_embedding(orc):
    rng = np.random.default_rng(13)
    grid = np.arange(0, 1.01, 0.25)
    for _ in range(80):
        a, b = rng.choice(grid, 2), rng.choice(grid, 2)
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        hi = np.where(hi == lo, np.minimum(lo + 0.25, 1.0), hi)
        box = np.concatenate([lo, hi]).astype(np.float32)
        box3 = np.array([[box[0], box[1], -1.0, box[2], box[3], 1.0]], np.float32)
        p0 = rng.choice(grid, (40, 2)).astype(np.float32)
        p1 = rng.choice(grid, (40, 2)).astype(np.float32)
        z = np.zeros((40, 1), np.float32)
        got3 = orc.segment_valid(np.hstack([p0, z]), np.hstack([p1, z]), box3)
        for i in range(40):
            want = not _lattice_hit2(p0[i], p1[i], box)
            assert orc.segment_valid2(p0[i], p1[i], box) == want
            assert bool(got3[i]) == want
    # touching an edge is free (strict inequalities); no boxes: free
    box = np.array([0.2, 0.2, 0.4, 0.4], np.float32)
    assert orc.segment_valid2(np.float32([0.4, 0.3]), np.float32([0.5, 0.3]), box)
    assert not orc.segment_valid2(np.float32([0.39, 0.3]), np.float32([0.5, 0.3]), box)
    assert orc.segment_valid2(np.float32([0.39, 0.3]), np.float32([0.5, 0.3]), np.zeros((0, 4), np.float32))


def test_planar_region_maps(orc):
    cfg = W.config("unicycle2d")
    L, L2 = cfg.r1_len, cfg.r2_len
    size = (cfg.ws_max - cfg.ws_min) / L
    idx = np.stack(np.meshgrid(np.arange(L), np.arange(L), indexing="ij"), -1).reshape(-1, 2)
    centres = np.zeros((len(idx), 3), np.float32)
    centres[:, :2] = cfg.ws_min + (idx + 0.5) * size
    centres[:, 2] = 123.0                      # z (theta for the unicycle) plays no part
    r1, r2 = orc.regions(cfg, centres)
    assert np.array_equal(r1, idx[:, 0] * L + idx[:, 1])           # y fastest (Graph.cu:105-114 with W_DIM 2)
    sub = np.stack(np.meshgrid(np.arange(L2), np.arange(L2), indexing="ij"), -1).reshape(-1, 2)
    for cell in idx[::7]:
        pts = np.zeros((len(sub), 3), np.float32)
        pts[:, :2] = cfg.ws_min + cell * size + (sub + 0.5) * size / L2
        a, b = orc.regions(cfg, pts)
        r = cell[0] * L + cell[1]
        assert np.all(a == r)
        assert np.array_equal(b, r * L2 ** 2 + sub[:, 0] * L2 + sub[:, 1])
    # against the 3D maps of the same grid at z = ws_min: r3 = r2d * L, sub3 = ... * L2
    cfg3 = W.config("di6d", r1_len=L, r2_len=L2)
    rng = np.random.default_rng(14)
    pts = rng.uniform(-0.2, 1.2, (5000, 3)).astype(np.float32)
    pts[:, 2] = 0.0
    a2, b2 = orc.regions(cfg, pts)
    a3, b3 = orc.regions(cfg3, pts)
    assert np.array_equal(a3, a2 * L)
    assert np.array_equal(b3 - a3 * L2 ** 3, (b2 - a2 * L2 ** 2) * L2)


@pytest.mark.parametrize("seed", [0, 1])
def test_unicycle_plan_replays(seed, orc):
    """Every tree node is its parent propagated with the node's own controls
    and step count (bit for bit, and valid); the goal node lies within the
    threshold of the goal in the plane."""
    q = W.make_query("unicycle2d", seed, capacity=1 << 15)
    cfg = q.cfg
    out = orc.plan(cfg, q.start, q.goal, q.obstacles)
    s, p = out["samples"], out["parents"]
    n = len(s)
    assert n > 1 and p[0] == -1
    k = np.rint(s[1:, 6] / np.float32(cfg.dt)).astype(np.int32)
    child, valid, _, _ = orc.propagate_children(cfg, p[1:].astype(np.int32), np.ascontiguousarray(s[:, :4].T),
                                                np.ascontiguousarray(s[1:, 4:6].T), k, q.obstacles)
    assert np.all(valid == 1)
    assert np.array_equal(child.T.view(np.uint32), s[1:].view(np.uint32))
    d = np.hypot(s[:, 0].astype(np.float64) - q.goal[0], s[:, 1].astype(np.float64) - q.goal[1])
    assert np.allclose(out["costs"][1:], d[1:], rtol=1e-6, atol=1e-7)
    st = out["stats"]
    if st["goal_index"] >= 0:
        assert out["costs"][st["goal_index"]] < cfg.goal_thresh
        path = out["path"]
        seg = np.hypot(np.diff(path[:, 0].astype(np.float64)), np.diff(path[:, 1].astype(np.float64)))
        assert abs(st["path_length"] - seg.sum()) < 1e-4
