"""Planner configurations and synthetic workspaces of the benchmark workloads.

Configurations mirror kgmt_config_preset() of the C-ABI (a CPU test checks
that both agree). Workspaces are generated with numpy's PCG64 seeded by
(seed, stream), so every run and every machine sees the same boxes.

Workloads:
  di6d      6D double integrator, unit cube, 20 boxes sides U[0.05,0.2], 2^17 nodes
  dubins5d  5D Dubins airplane, 100 boxes sides U[0.02,0.1], 2^18 nodes
  quad12d   12D quadcopter, three slotted walls + 50 boxes, 2^19 nodes
  unicycle2d planar unicycle, unit square, 30 boxes sides U[0.05,0.15], 2^17 nodes
  batch     512 queries, 50/50 double integrator / quadcopter, capacity
            log-uniform in [2^12, 2^17], 10..500 boxes, start/goal >= 0.6 apart
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from .abi import (KGMT_DI_REFERENCE, KGMT_MAX_SAMPLE, KGMT_MODEL_DOUBLE_INTEGRATOR, KGMT_MODEL_DUBINS_5D,
                  KGMT_MODEL_QUADCOPTER, KGMT_MODEL_UNICYCLE, KgmtConfig, dims)

PI = float(np.float32(math.pi))


def _base(**kw) -> KgmtConfig:
    c = KgmtConfig()
    c.epsilon = 1e-2
    c.children_per_entry = 32
    c.root_repeat = 5
    c.repeat_sparse = 15
    c.repeat_dense = 1
    c.sparse_valid_below = 10
    c.max_iter = 300
    c.quad_mass, c.quad_ix, c.quad_iy, c.quad_iz = 1.0, 1.0, 1.0, 2.0
    c.quad_nu, c.quad_mu, c.gravity = 10e-3, 2e-6, -9.81
    c.pitch_min, c.pitch_max = -PI / 3.0, PI / 3.0
    c.uni_length = 1.0
    c.seed = 1
    return c.replace(**kw)


def config(name: str, **overrides) -> KgmtConfig:
    """Planner configuration of a named workload (see module docstring)."""
    if name == "reference_quad":  # include/config/config.h:209-307
        c = _base(model=KGMT_MODEL_QUADCOPTER, capacity=400000, dt=0.1, steps_min=1, steps_max=10,
                  goal_thresh=5.0, ws_min=0.0, ws_max=100.0, v_min=-30.0, v_max=30.0,
                  ctrl_min=[0.0, -PI, -PI, -PI], ctrl_max=[30.0, PI, PI, PI], r1_len=8, r2_len=2)
    elif name == "di6d":
        c = _base(model=KGMT_MODEL_DOUBLE_INTEGRATOR, integrator=KGMT_DI_REFERENCE, capacity=1 << 17, dt=0.05,
                  steps_min=2, steps_max=10, goal_thresh=0.05, ws_min=0.0, ws_max=1.0, v_min=-1.0, v_max=1.0,
                  ctrl_min=[-1.0, -1.0, -1.0, 0.0], ctrl_max=[1.0, 1.0, 1.0, 0.0], r1_len=8, r2_len=4)
    elif name == "dubins5d":
        c = _base(model=KGMT_MODEL_DUBINS_5D, capacity=1 << 18, dt=0.02, steps_min=5, steps_max=25,
                  goal_thresh=0.05, ws_min=0.0, ws_max=1.0, v_min=0.0, v_max=1.0,
                  ctrl_min=[-1.0, -0.5, 0.2, 0.0], ctrl_max=[1.0, 0.5, 0.4, 0.0], r1_len=16, r2_len=2)
    elif name == "quad12d":
        c = _base(model=KGMT_MODEL_QUADCOPTER, capacity=1 << 19, dt=0.01, steps_min=10, steps_max=50,
                  goal_thresh=0.05, ws_min=0.0, ws_max=1.0, v_min=-5.0, v_max=5.0,
                  ctrl_min=[0.0, -0.05, -0.05, -0.05], ctrl_max=[19.62, 0.05, 0.05, 0.05],
                  quad_ix=0.01, quad_iy=0.01, quad_iz=0.02, r1_len=10, r2_len=4)
    elif name == "unicycle2d":  # steering / UNI_LENGTH of config.h:266-270, A of config.h:34-35
        c = _base(model=KGMT_MODEL_UNICYCLE, capacity=1 << 17, dt=0.1, steps_min=1, steps_max=10,
                  goal_thresh=0.05, ws_min=0.0, ws_max=1.0, v_min=0.0, v_max=0.0,
                  ctrl_min=[-0.2, -PI / 2.0, 0.0, 0.0], ctrl_max=[0.2, PI / 2.0, 0.0, 0.0], r1_len=16, r2_len=4)
    else:
        raise KeyError(name)
    return c.replace(**overrides) if overrides else c


PRESET_OF = {"reference_quad": 0, "di6d": 1, "dubins5d": 2, "quad12d": 3, "unicycle2d": 4}


# ---------------------------------------------------------------------------
# workspaces
# ---------------------------------------------------------------------------
@dataclass
class Query:
    cfg: KgmtConfig
    start: np.ndarray          # D floats
    goal: np.ndarray           # D floats
    obstacles: np.ndarray      # (n, 2W) float32 rows: min corner, max corner (W = 2 planar, else 3)
    name: str = ""
    meta: dict = field(default_factory=dict)


def _rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.default_rng([int(seed), int(stream)])


def box_ball_distance(boxes: np.ndarray, c: np.ndarray) -> np.ndarray:
    w = boxes.shape[1] // 2
    lo, hi = boxes[:, :w], boxes[:, w:]
    d = np.maximum(np.maximum(lo - c, 0.0), c - hi)
    return np.sqrt((d * d).sum(axis=1))


def random_boxes(rng, n, side_lo, side_hi, keep_clear=(), ws=(0.0, 1.0), dim: int = 3) -> np.ndarray:
    """n boxes with centre U(ws)^dim and sides U(side_lo, side_hi), clipped to
    the workspace; a box that comes closer than r to any (centre, r) in
    keep_clear is drawn again."""
    out = np.zeros((n, 2 * dim), np.float64)
    for i in range(n):
        for _ in range(10000):
            c = rng.uniform(ws[0], ws[1], dim)
            s = rng.uniform(side_lo, side_hi, dim)
            b = np.concatenate([np.clip(c - s / 2, ws[0], ws[1]), np.clip(c + s / 2, ws[0], ws[1])])
            if all(box_ball_distance(b[None], np.asarray(p))[0] >= r for p, r in keep_clear):
                out[i] = b
                break
        else:
            raise RuntimeError("could not place a box away from start/goal")
    return out.astype(np.float32)


def _sample(cfg: KgmtConfig, pos) -> np.ndarray:
    """Sample row at rest at position `pos` (W floats)."""
    s = np.zeros(cfg.dims[2], np.float32)
    s[:cfg.wdim] = pos
    return s


def make_query(name: str, seed: int, **cfg_overrides) -> Query:
    """One query of workload `name` (di6d, dubins5d, quad12d, unicycle2d) for `seed`."""
    cfg = config(name, seed=seed + 1, **cfg_overrides)
    rng = _rng(seed, {"di6d": 1, "dubins5d": 2, "quad12d": 3, "unicycle2d": 5}[name])
    w = cfg.wdim
    start, goal = np.full(w, 0.05), np.full(w, 0.95)
    clear = [(start, 0.05 + 1e-3), (goal, cfg.goal_thresh + 1e-3)]
    if name == "unicycle2d":
        obs = random_boxes(rng, 30, 0.05, 0.15, clear, dim=2)
    elif name == "di6d":
        obs = random_boxes(rng, 20, 0.05, 0.2, clear)
    elif name == "dubins5d":
        obs = random_boxes(rng, 100, 0.02, 0.1, clear)
    else:
        walls = []
        gaps = []
        for wx in (0.25, 0.5, 0.75):
            gy = rng.uniform(0.1, 0.9 - 0.08)
            walls.append([wx - 0.01, 0.0, 0.0, wx + 0.01, gy, 1.0])
            walls.append([wx - 0.01, gy + 0.08, 0.0, wx + 0.01, 1.0, 1.0])
            gaps.append((np.array([wx, gy + 0.04, 0.5]), 0.06))
        # keep random boxes out of a tube around each gap's centre line
        rnd = []
        while len(rnd) < 50:
            b = random_boxes(rng, 1, 0.02, 0.1, clear)[0]
            if any(b[0] < c[0] + 0.06 and b[3] > c[0] - 0.06 and b[1] < c[1] + 0.06 and b[4] > c[1] - 0.06
                   for c, _ in gaps):
                continue
            rnd.append(b)
        obs = np.concatenate([np.asarray(walls, np.float32), np.asarray(rnd, np.float32)])
    return Query(cfg, _sample(cfg, start), _sample(cfg, goal), obs.astype(np.float32), f"{name}/{seed}")


def make_batch(n_queries: int = 512, seed: int = 0, max_iter: int = 300):
    """Mixed batch: 50/50 double integrator / quadcopter, capacity
    log-uniform in [2^12, 2^17], 10..500 boxes, start/goal >= 0.6 apart."""
    rng = _rng(seed, 4)
    out = []
    for q in range(n_queries):
        model = "di6d" if q % 2 == 0 else "quad12d"
        cap = int(round(2.0 ** rng.uniform(12, 17)))
        n_obs = int(rng.integers(10, 501))
        while True:
            s = rng.uniform(0.1, 0.9, 3)
            g = rng.uniform(0.1, 0.9, 3)
            if np.linalg.norm(s - g) >= 0.6:
                break
        cfg = config(model, capacity=cap, seed=seed * 100003 + q + 1, max_iter=max_iter, query_id=q)
        clear = [(s, 0.06), (g, cfg.goal_thresh + 0.01)]
        obs = random_boxes(rng, n_obs, 0.02, 0.06, clear)
        out.append(Query(cfg, _sample(cfg, s), _sample(cfg, g), obs, f"batch/{q}", {"model": model}))
    return out


@dataclass
class RefineBatch:
    cfg: KgmtConfig            # quadcopter parameters, dt, control bounds
    params: object             # KgmtAdjointParams
    steps: np.ndarray          # (n,) int32 trajectory lengths
    x0: np.ndarray             # (12, n) initial states
    goal: np.ndarray           # (3, n) goals
    u: np.ndarray              # (max_steps, 4, n) initial control sequences
    obstacles: np.ndarray      # (m, 6) boxes shared by all trajectories


# Cost weights of the refinement workload. "configured": the objective as
# the workload states it, J = sum_t |u_t|^2 + softplus obstacle penalty
# (margin 0.02) + |p_T - goal|^2. "hover": effort measured from hover thrust
# and scaled down, goal weighted up, so that the effort term does not reduce
# to pulling thrust towards zero (with |u|^2 about u = 0 the cost is almost
# all hover thrust, m g per step, which Adam at lr 1e-2 lowers by at most
# 0.5 in 50 steps).
OBJECTIVES = {
    "configured": dict(w_effort=1.0, w_obstacle=1.0, w_goal=1.0, hover_ref=False),
    "hover": dict(w_effort=1e-3, w_obstacle=1.0, w_goal=100.0, hover_ref=True),
}


def refine_params(cfg: KgmtConfig, objective: str = "configured", **kw):
    """Parameters of the refinement workload (softplus temperature 0.01,
    margin 0.02; Adam with lr 1e-2 for 50 steps, controls clamped to their
    bounds) for one of OBJECTIVES."""
    from .abi import KgmtAdjointParams
    o = OBJECTIVES[objective]
    p = KgmtAdjointParams()
    p.w_effort, p.w_obstacle, p.w_goal = o["w_effort"], o["w_obstacle"], o["w_goal"]
    p.margin, p.tau = 0.02, 0.01
    hover = -cfg.gravity * cfg.quad_mass if o["hover_ref"] else 0.0
    for i, v in enumerate([hover, 0.0, 0.0, 0.0]):
        p.u_ref[i] = v
    p.lr, p.beta1, p.beta2, p.adam_eps = 1e-2, 0.9, 0.999, 1e-8
    p.iterations, p.clamp = 50, 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def make_refine_batch(n: int = 4096, seed: int = 0, steps_lo: int = 20, steps_hi: int = 80,
                      objective: str = "configured") -> RefineBatch:
    """Trajectory refinement workload: n quadcopter control sequences of
    length U{steps_lo..steps_hi} at dt 0.01 from rest in the unit cube,
    controls hover thrust + noise, goals 0.05-0.15 away, 50 shared boxes;
    cost weights of `objective` (OBJECTIVES)."""
    cfg = config("quad12d")
    rng = _rng(seed, 5)
    steps = rng.integers(steps_lo, steps_hi + 1, n).astype(np.int32)
    T = int(steps.max())
    x0 = np.zeros((12, n), np.float32)
    x0[:3] = rng.uniform(0.25, 0.75, (3, n))
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0, keepdims=True)
    goal = (x0[:3] + d * rng.uniform(0.05, 0.15, n)).astype(np.float32)
    hover = -cfg.gravity * cfg.quad_mass
    u = np.zeros((T, 4, n), np.float32)
    u[:, 0] = np.clip(hover + rng.normal(0, 1.0, (T, n)), cfg.ctrl_min[0], cfg.ctrl_max[0])
    u[:, 1:] = np.clip(rng.normal(0, 0.005, (T, 3, n)), cfg.ctrl_min[1], cfg.ctrl_max[1])
    obs = random_boxes(rng, 50, 0.02, 0.08)
    return RefineBatch(cfg, refine_params(cfg, objective), steps, x0, goal, u, obs)


def pack_batch(queries):
    """Arrays for kgmt_batch_plan: configs, starts/goals (n, 17), obstacle
    rows (2W floats; every query of a batch has the same W), per-query
    offsets and counts."""
    n = len(queries)
    w = queries[0].cfg.wdim if n else 3
    if any(q.cfg.wdim != w for q in queries):
        raise ValueError("a batch cannot mix planar and 3D queries")
    starts = np.zeros((n, KGMT_MAX_SAMPLE), np.float32)
    goals = np.zeros((n, KGMT_MAX_SAMPLE), np.float32)
    offs = np.zeros(n, np.uint32)
    cnts = np.zeros(n, np.uint32)
    rows = []
    o = 0
    for i, q in enumerate(queries):
        starts[i, :len(q.start)] = q.start
        goals[i, :len(q.goal)] = q.goal
        offs[i], cnts[i] = o, len(q.obstacles)
        rows.append(q.obstacles.reshape(-1, 2 * w))
        o += len(q.obstacles)
    obs = np.concatenate(rows).astype(np.float32) if rows else np.zeros((0, 2 * w), np.float32)
    return [q.cfg for q in queries], starts, goals, np.ascontiguousarray(obs), offs, cnts


# ---------------------------------------------------------------------------
# obstacle CSV files (include/config/obstacles/*/obstacles.csv format)
# ---------------------------------------------------------------------------
_NUMBER = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")


def read_obstacles_csv(path: str, workspace_dim: int = 3) -> np.ndarray:
    """Boxes of an obstacle CSV, read as readObstaclesFromCSV
    (src/helper/helper.cu:13-39) reads them: each line is parsed number by
    number (leading whitespace skipped, one comma right after a number
    skipped) until a token is not a number; every value of the file in order,
    grouped 2*workspace_dim per box (a trailing partial group is dropped). A
    last line without a newline is read too."""
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    vals = []
    with open(path, "r") as f:
        for line in f:
            line = line.rstrip("\n")
            pos = 0
            while True:
                m = _NUMBER.match(line, pos)
                if not m:
                    break
                vals.append(float(np.float32(float(m.group(1)))))
                pos = m.end()
                if line.startswith(",", pos):
                    pos += 1
    k = 2 * workspace_dim
    n = len(vals) // k
    return np.asarray(vals[: n * k], np.float32).reshape(n, k)
