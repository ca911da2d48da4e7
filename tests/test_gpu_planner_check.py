"""The collision check the planner's propagate kernel actually runs, on
identical pre-generated buffers, against the CPU oracle's plain loop over
every box at every step (isMotionValid, collisionCheck.cu:49-62, inside the
models' step loops, statePropagator.cu).

kgmt_propagate_collide_planner runs children in waves of 64 with the
planner's per-wave step sort and calls the planner's own propagate_two_phase
(edge box from a dry run, or for the Dubins models from an interval bound;
wave candidate lists; per-lane candidates; whole-scene fallback) or, for the
quadcopter, the one-pass check. Bar: states, verdicts and regions bit for bit,
on the 5D Dubins workload's own trees and boxes, a 1500-box scene, NaN /
infinite / inverted / flat boxes and non-finite parent states, every model.
The Dubins interval bound is checked on its own: every position the edge
reaches lies inside the edge box it reports, for long edges, fast and slow
speeds, steep pitches and large angles."""
import ctypes as C
import math

import numpy as np
import pytest

from kgmt_amd import workloads as W
from kgmt_amd.abi import (KGMT_CHECK_CANDIDATES, KGMT_CHECK_DRY_RUN, KGMT_CHECK_ONE_PASS, KGMT_CHECK_SCENE,
                          KGMT_MODEL_DUBINS_5D, KGMT_MODEL_DUBINS_AIRPLANE, KGMT_MODEL_QUADCOPTER, dims)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tests.gpu_util import bits, call, dev, empty, host, ptr, stream  # noqa: E402
from tests.test_gpu_parity import CASES, dubins6d_config, obstacles_for, parent_states  # noqa: E402


def run_planner_check(cfg, pidx, X, ctrl, steps, obs, want_box=False):
    S, Cd, D = dims(cfg.model)
    n, npar = len(pidx), X.shape[1]
    child = empty((D, n), torch.float32)
    valid = empty(n, torch.uint8)
    r1 = empty(n, torch.int32)
    r2 = empty(n, torch.int32)
    chk = empty(n, torch.uint8)
    box = empty((6, n), torch.float32) if want_box else None
    obs = np.ascontiguousarray(obs, np.float32)
    d_obs = dev(obs.reshape(-1)) if len(obs) else None
    d_pidx, d_X, d_ctrl, d_steps = dev(pidx), dev(np.ascontiguousarray(X)), dev(ctrl), dev(steps)
    call("kgmt_propagate_collide_planner", C.byref(cfg), ptr(d_pidx), ptr(d_X), npar, ptr(d_ctrl), ptr(d_steps), n,
         ptr(d_obs), len(obs), ptr(child), ptr(valid), ptr(r1), ptr(r2), ptr(chk), ptr(box), stream())
    return host(child), host(valid), host(r1), host(r2), host(chk), host(box) if want_box else None


def assert_bit_exact(cfg, pidx, X, ctrl, steps, obs, orc):
    want = orc.propagate_children(cfg, pidx, X, ctrl, steps, obs)
    child, valid, r1, r2, chk, _ = run_planner_check(cfg, pidx, X, ctrl, steps, obs)
    assert np.array_equal(valid, want[1]), f"{np.count_nonzero(valid != want[1])} verdicts differ"
    assert np.array_equal(r1, want[2]) and np.array_equal(r2, want[3])
    # bit for bit, except that any NaN equals any NaN (a NaN made inside the
    # step, e.g. sin of an infinite angle, has the sign and payload of the
    # processor that made it)
    same = (bits(child) == bits(want[0])) | (np.isnan(child) & np.isnan(want[0]))
    mism = np.flatnonzero(~same.all(axis=0))
    assert mism.size == 0, f"{mism.size} children differ, first {mism[:5]}"
    return want, chk


def clustered_parents(rng, npar, n, per=32):
    """32 children per expansion entry, as the planner's first regime makes them."""
    return np.repeat(rng.integers(0, npar, n // per + 1), per)[:n].astype(np.int32)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("layout", ["clustered", "spread"])
def test_planner_check_bit_exact(case, layout, orc):
    cfg = CASES[case]()
    n, npar = 40000, 1500
    X = parent_states(cfg, npar)
    rng = np.random.default_rng(13)
    pidx = clustered_parents(rng, npar, n) if layout == "clustered" else rng.integers(0, npar, n).astype(np.int32)
    ctrl, steps, _ = orc.draw_children(cfg, 0, 1, 0, n)
    obs = obstacles_for(cfg)
    want, chk = assert_bit_exact(cfg, pidx, X, ctrl, steps, obs, orc)
    assert 0 < want[1].sum() < n
    paths = np.bincount(chk, minlength=4)
    print(f"{case}/{layout}: checks scene/candidates/dry/one-pass = {paths.tolist()}", flush=True)
    if cfg.model == KGMT_MODEL_QUADCOPTER:
        assert paths[KGMT_CHECK_ONE_PASS] == n
    elif layout == "clustered":
        assert paths[KGMT_CHECK_CANDIDATES] > 0
        assert paths[KGMT_CHECK_CANDIDATES] + paths[KGMT_CHECK_DRY_RUN] > n // 2


def _tree_parents(orc, q, iters):
    """Nodes of the oracle's own tree after `iters` iterations: the parent
    states the planner really expands."""
    cfg = q.cfg.replace(max_iter=iters)
    out = orc.plan(cfg, q.start, q.goal, q.obstacles)
    S = dims(cfg.model)[0]
    return np.ascontiguousarray(out["samples"][:, :S].T), out["stats"]["tree_size"]


@pytest.mark.parametrize("seed", [0, 3, 7])
def test_planner_check_on_dubins5d_workload(seed, orc):
    """The 5D Dubins workload's own scene (100 boxes) and its own tree's
    nodes as parents, children drawn as the planner draws them."""
    q = W.make_query("dubins5d", seed)
    X, ntree = _tree_parents(orc, q, 4)
    rng = np.random.default_rng(seed)
    n = 60000
    pidx = clustered_parents(rng, ntree, n)
    ctrl, steps, _ = orc.draw_children(q.cfg, 0, 5, 0, n)
    want, chk = assert_bit_exact(q.cfg, pidx, X, ctrl, steps, q.obstacles, orc)
    paths = np.bincount(chk, minlength=4)
    print(f"dubins5d seed {seed}: {ntree} parents, valid {want[1].mean():.3f}, checks {paths.tolist()}", flush=True)
    assert paths[KGMT_CHECK_CANDIDATES] > 0 and 0 < want[1].sum() < n


@pytest.mark.parametrize("case", ["di_ref", "dubins5d", "dubins6d", "unicycle"])
def test_planner_check_1500_boxes(case, orc):
    """The largest scene a query may carry, small boxes and large ones,
    children clustered by parent and spread: the wave candidate lists stay
    within 128 boxes in some waves and overflow in others (then the wave
    falls back to the whole scene); both checks must run, bit-exact."""
    cfg = CASES[case]()
    rng = np.random.default_rng(31)
    span = cfg.ws_max - cfg.ws_min
    n, npar = 30000, 900
    X = parent_states(cfg, npar, seed=4)
    ctrl, steps, _ = orc.draw_children(cfg, 0, 3, 0, n)
    seen = np.zeros(4, np.int64)
    for sides in ((0.005, 0.03), (0.05, 0.15)):
        obs = W.random_boxes(rng, 1500, sides[0] * span, sides[1] * span, ws=(cfg.ws_min, cfg.ws_max),
                             dim=cfg.wdim)
        for layout in ("spread", "clustered"):
            pidx = rng.integers(0, npar, n).astype(np.int32) if layout == "spread" else clustered_parents(rng, npar, n)
            want, chk = assert_bit_exact(cfg, pidx, X, ctrl, steps, obs, orc)
            paths = np.bincount(chk, minlength=4)
            seen += paths
            print(f"{case} 1500 boxes {sides} {layout}: checks {paths.tolist()}, valid {want[1].mean():.3f}",
                  flush=True)
    assert seen[KGMT_CHECK_SCENE] > 0 and seen[KGMT_CHECK_CANDIDATES] > 0, seen


@pytest.mark.parametrize("case", ["di_ref", "dubins5d", "dubins6d", "unicycle", "quad12d"])
def test_planner_check_odd_boxes_and_states(case, orc):
    """NaN, infinite, inverted and flat boxes, boxes outside the workspace,
    and parents with NaN or infinite coordinates (their edges cannot be
    bounded, so the wave falls back to the whole scene)."""
    cfg = CASES[case]()
    rng = np.random.default_rng(8)
    span = cfg.ws_max - cfg.ws_min
    lo = cfg.ws_min
    obs = W.random_boxes(rng, 60, 0.03 * span, 0.2 * span, ws=(cfg.ws_min, cfg.ws_max), dim=cfg.wdim)
    w = cfg.wdim
    obs[3, 0] = np.nan                                    # NaN low face: that face never separates
    obs[5, w + 1] = np.nan                                # NaN high face
    obs[9, :] = np.concatenate([np.full(w, lo + 0.6 * span), np.full(w, lo + 0.4 * span)])  # inverted
    obs[11, :] = np.concatenate([np.full(w, lo - 5 * span), np.full(w, lo - 4 * span)])     # outside
    obs[13, w] = obs[13, 0]                               # zero thickness on x
    obs[17, :w] = -np.inf                                 # infinite low corner
    obs[19, w:] = np.inf                                  # infinite high corner
    obs = obs.astype(np.float32)
    n, npar = 20000, 700
    X = parent_states(cfg, npar, seed=12)
    X[0, 5] = np.nan
    X[1, 9] = np.inf
    X[:3, 30] = np.nan
    if X.shape[0] > 4:
        X[3, 40] = np.nan    # non-finite angles (Dubins yaw / pitch): no edge bound either
        X[4, 41] = np.inf
    ctrl, steps, _ = orc.draw_children(cfg, 0, 4, 0, n)
    pidx = clustered_parents(rng, npar, n)
    pidx[:64] = 5            # a whole wave of NaN parents
    pidx[100] = 9            # one lane of a wave
    pidx[2000:2064] = 30
    pidx[3008:3072] = 40     # a whole wave of parents with a NaN state component
    pidx[4000] = 41
    want, chk = assert_bit_exact(cfg, pidx, X, ctrl, steps, obs, orc)
    paths = np.bincount(chk, minlength=4)
    print(f"{case} odd boxes/states: checks {paths.tolist()}", flush=True)
    if cfg.model != KGMT_MODEL_QUADCOPTER:
        assert np.all(chk[:64] == KGMT_CHECK_SCENE)  # the NaN wave cannot be bounded
        if X.shape[0] > 4:
            assert np.all(chk[3008:3072] == KGMT_CHECK_SCENE)


def test_out_of_range_parents_and_ragged_sizes(orc):
    cfg = CASES["dubins5d"]()
    npar = 50
    X = parent_states(cfg, npar)
    obs = obstacles_for(cfg)
    for n in (1, 63, 64, 65, 257):
        rng = np.random.default_rng(n)
        pidx = rng.integers(0, npar, n).astype(np.int32)
        pidx[::7] = -1
        pidx[3::11] = npar
        ctrl, steps, _ = orc.draw_children(cfg, 0, 1, 0, n)
        child, valid, r1, r2, chk, _ = run_planner_check(cfg, pidx, X, ctrl, steps, obs)
        bad = (pidx < 0) | (pidx >= npar)
        assert np.all(r1[bad] == -1) and np.all(r2[bad] == -1) and np.all(valid[bad] == 0)
        good = ~bad
        if good.any():
            want = orc.propagate_children(cfg, pidx[good], X, ctrl[:, good], steps[good], obs)
            assert np.array_equal(valid[good], want[1])
            assert np.array_equal(bits(child[:, good]), bits(want[0]))


def _loose(cfg):
    """Same dynamics with every bound check off: the edge runs all its steps."""
    return cfg.replace(ws_min=-1e30, ws_max=1e30, v_min=-1e30, v_max=1e30, pitch_min=-1e30, pitch_max=1e30)


def _dubins_stress(model):
    pi = math.pi
    if model == KGMT_MODEL_DUBINS_5D:
        return [W.config("dubins5d"),
                W.config("dubins5d", dt=0.1, steps_min=10, steps_max=50, ctrl_min=[-3.0, -1.5, 0.01, 0.0],
                         ctrl_max=[3.0, 1.5, 2.0, 0.0])]
    return [dubins6d_config(),
            dubins6d_config().replace(steps_min=10, steps_max=50, ctrl_min=[-3.0, -1.5, -1.0, 0.0],
                                      ctrl_max=[3.0, 1.5, 1.0, 0.0], v_max=3.0, pitch_min=-pi, pitch_max=pi)]


@pytest.mark.parametrize("model", [KGMT_MODEL_DUBINS_5D, KGMT_MODEL_DUBINS_AIRPLANE])
def test_dubins_edge_box_holds_every_position(model, orc):
    """Every position a Dubins edge reaches, step by step to its last step
    (bound checks off, so no edge stops early), lies inside the edge box the
    planner's check derives from the interval bound."""
    for k, cfg in enumerate(_dubins_stress(model)):
        S, Cd, D = dims(cfg.model)
        n = 20000
        rng = np.random.default_rng(40 + k)
        X = np.zeros((S, n), np.float32)
        X[:3] = rng.uniform(0.0, 1.0, (3, n))
        X[3] = np.where(rng.random(n) < 0.2, rng.uniform(-1000, 1000, n), rng.uniform(-math.pi, math.pi, n))
        X[4] = np.where(rng.random(n) < 0.3, rng.choice([-1, 1], n) * rng.uniform(1.5, 1.5708, n),
                        rng.uniform(-1.0, 1.0, n))
        if S == 6:
            X[5] = rng.uniform(0.0, 2.0, n)
        pidx = np.arange(n, dtype=np.int32)
        ctrl, steps, _ = orc.draw_children(cfg, 0, 2 + k, 0, n)
        _, _, _, _, chk, box = run_planner_check(cfg, pidx, X, ctrl, steps, np.zeros((0, 6), np.float32),
                                                 want_box=True)
        lo, hi = box[:3], box[3:]
        assert np.isfinite(box).all()
        loose = _loose(cfg)
        worst = np.inf
        for s in range(1, int(steps.max()) + 1):
            st = np.minimum(steps, s).astype(np.int32)
            pos = orc.propagate_children(loose, pidx, X, ctrl, st, np.zeros((0, 6), np.float32))[0][:3]
            inside = (pos >= lo) & (pos <= hi)
            assert inside.all(), (f"config {k}, step {s}: {np.count_nonzero(~inside.all(0))} positions outside, "
                                  f"first child {np.flatnonzero(~inside.all(0))[0]}")
            worst = min(worst, float(np.min(np.minimum(pos - lo, hi - pos))))
        print(f"model {model} config {k}: {n} edges up to {int(steps.max())} steps inside their boxes, "
              f"smallest margin {worst:.3g}", flush=True)
