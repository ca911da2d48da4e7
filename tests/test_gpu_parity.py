"""Parity of the HIP path with the CPU oracle on identical inputs.

Bar: bit-exact for everything (integer/index work trivially; float work
because kernels and oracle evaluate the same IEEE operations in the same
order, with contraction off). Each stage is checked in isolation on the same
pre-generated buffers, then whole plans are compared node by node.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from kgmt_amd import _lib

from kgmt_amd.abi import (KGMT_DI_EXACT, KGMT_MODEL_DOUBLE_INTEGRATOR, KGMT_MODEL_DUBINS_5D,
                          KGMT_MODEL_DUBINS_AIRPLANE, KGMT_MODEL_QUADCOPTER, KGMT_MODEL_UNICYCLE, dims)
from kgmt_amd import workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tests.gpu_util import bits, call, dev, empty, host, ptr, stream  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dubins6d_config(**kw):
    """The reference's Dubins-airplane block (config.h:108-203) on a unit cube."""
    pi = math.pi
    return W.config("dubins5d").replace(model=KGMT_MODEL_DUBINS_AIRPLANE, dt=0.1, steps_min=1, steps_max=10,
                                        v_min=0.0, v_max=0.3, ctrl_min=[-pi / 4, -pi / 4, -0.3, 0],
                                        ctrl_max=[pi / 4, pi / 4, 0.3, 0], r1_len=8, r2_len=2, **kw)


CASES = {
    "di_ref": lambda: W.make_query("di6d", 3).cfg,
    "di_exact": lambda: W.make_query("di6d", 3).cfg.replace(integrator=KGMT_DI_EXACT),
    "dubins6d": lambda: dubins6d_config(),
    "dubins5d": lambda: W.make_query("dubins5d", 3).cfg,
    "quad12d": lambda: W.make_query("quad12d", 3).cfg,
    "quad_reference": lambda: W.config("reference_quad"),
    "unicycle": lambda: W.make_query("unicycle2d", 3).cfg,
}


def obstacles_for(cfg, seed=5):
    rng = np.random.default_rng(seed)
    span = cfg.ws_max - cfg.ws_min
    return W.random_boxes(rng, 60, 0.03 * span, 0.15 * span, ws=(cfg.ws_min, cfg.ws_max), dim=cfg.wdim)


def parent_states(cfg, n, seed=7):
    rng = np.random.default_rng(seed)
    S, _, _ = dims(cfg.model)
    lo, hi = cfg.ws_min, cfg.ws_max
    X = np.zeros((S, n), np.float32)
    X[:3] = rng.uniform(lo + 0.05 * (hi - lo), lo + 0.95 * (hi - lo), (3, n))
    if cfg.model == KGMT_MODEL_UNICYCLE:
        X[2] = rng.uniform(-math.pi, math.pi, n)
        X[3] = rng.uniform(-0.6, 0.6, n)
    elif cfg.model == KGMT_MODEL_DOUBLE_INTEGRATOR:
        X[3:6] = rng.uniform(-0.5, 0.5, (3, n))
    elif cfg.model == KGMT_MODEL_DUBINS_AIRPLANE:
        X[3] = rng.uniform(-math.pi, math.pi, n)
        X[4] = rng.uniform(-0.5, 0.5, n)
        X[5] = rng.uniform(0.05, 0.3, n)
    elif cfg.model == KGMT_MODEL_DUBINS_5D:
        X[3] = rng.uniform(-math.pi, math.pi, n)
        X[4] = rng.uniform(-0.5, 0.5, n)
    else:
        X[3:6] = rng.uniform(-0.3, 0.3, (3, n))
        X[6:9] = rng.uniform(-1.0, 1.0, (3, n))
        X[9:12] = rng.uniform(-1.0, 1.0, (3, n))
    return X


@pytest.mark.parametrize("case", list(CASES))
def test_draws_bit_exact(case, orc):
    cfg = CASES[case]()
    S, Cd, D = dims(cfg.model)
    n, u0 = 5000, 1234
    ctrl = empty((Cd, n), torch.float32)
    steps = empty(n, torch.int32)
    uf = empty(n, torch.float32)
    call("kgmt_draw_controls", C.byref(cfg), 3, 17, u0, n, ptr(ctrl), ptr(steps), ptr(uf), stream())
    oc, os_, ou = orc.draw_children(cfg, 3, 17, u0, n)
    assert np.array_equal(bits(host(ctrl)), bits(oc))
    assert np.array_equal(host(steps), os_)
    assert np.array_equal(bits(host(uf)), bits(ou))
    assert host(steps).min() >= cfg.steps_min and host(steps).max() <= cfg.steps_max


@pytest.mark.parametrize("case", list(CASES))
def test_propagate_collide_bit_exact(case, orc):
    cfg = CASES[case]()
    S, Cd, D = dims(cfg.model)
    n, npar = 40000, 1500
    X = parent_states(cfg, npar)
    rng = np.random.default_rng(11)
    pidx = rng.integers(0, npar, n).astype(np.int32)
    ctrl, steps, _ = orc.draw_children(cfg, 0, 1, 0, n)
    obs = obstacles_for(cfg)
    want_child, want_valid, want_r1, want_r2 = orc.propagate_children(cfg, pidx, X, ctrl, steps, obs)
    child = empty((D, n), torch.float32)
    valid = empty(n, torch.uint8)
    r1 = empty(n, torch.int32)
    r2 = empty(n, torch.int32)
    # every device buffer stays referenced until the call has finished
    d_obs, d_pidx, d_X, d_ctrl, d_steps = dev(obs.reshape(-1)), dev(pidx), dev(X), dev(ctrl), dev(steps)
    call("kgmt_propagate_collide", C.byref(cfg), ptr(d_pidx), ptr(d_X), npar, ptr(d_ctrl), ptr(d_steps),
         n, ptr(d_obs), len(obs), ptr(child), ptr(valid), ptr(r1), ptr(r2), stream())
    got = host(child)
    # verdicts, regions and states all bit-exact
    assert np.array_equal(host(valid), want_valid), "collision/validity verdicts differ"
    assert np.array_equal(host(r1), want_r1)
    assert np.array_equal(host(r2), want_r2)
    mism = np.flatnonzero((bits(got) != bits(want_child)).any(axis=0))
    assert mism.size == 0, f"{mism.size} children differ, first {mism[:5]}"
    # the workload is non-trivial: both outcomes occur
    assert 0 < want_valid.sum() < n


def test_region_update_bit_exact(orc):
    cfg = W.make_query("di6d", 0).cfg
    R1 = cfg.r1_len ** 3
    R2 = R1 * cfg.r2_len ** 3
    rng = np.random.default_rng(3)
    n = 50000
    # spatially clustered regions (as children of one parent are)
    r1 = np.repeat(rng.integers(0, R1, n // 32 + 1), 32)[:n].astype(np.int32)
    r2 = (r1 * cfg.r2_len ** 3 + rng.integers(0, cfg.r2_len ** 3, n)).astype(np.int32)
    valid = (rng.random(n) < 0.7).astype(np.uint8)
    uf = rng.random(n).astype(np.float32)
    score = rng.random(R1).astype(np.float32)
    ap = (rng.random(R2) < 0.3).astype(np.uint8)
    an = ap.copy()
    total = rng.integers(0, 50, R1).astype(np.int32)
    vc = (total // 2).astype(np.int32)
    t0, v0 = total.copy(), vc.copy()
    want_f = orc.region_update(cfg, r1, r2, valid, uf, score, ap, an, total, vc)

    def words(b):
        w = np.packbits(b.astype(np.uint8), bitorder="little")
        w = np.concatenate([w, np.zeros((-len(w)) % 4, np.uint8)])
        return w.view(np.uint32)

    d_ap = dev(words(ap))
    d_an = dev(words(ap))
    d_t = dev(t0)
    d_v = dev(v0)
    fn = empty(n, torch.uint8)
    d_r1, d_r2, d_valid, d_uf, d_score = dev(r1), dev(r2), dev(valid), dev(uf), dev(score)
    call("kgmt_region_update", C.byref(cfg), ptr(d_r1), ptr(d_r2), ptr(d_valid), ptr(d_uf), n,
         ptr(d_score), ptr(d_ap), ptr(d_an), ptr(d_t), ptr(d_v), ptr(fn), stream())
    assert np.array_equal(host(fn), want_f)
    assert np.array_equal(host(d_t), total)
    assert np.array_equal(host(d_v), vc)
    got_an = np.unpackbits(host(d_an).view(np.uint8), bitorder="little")[:R2]
    assert np.array_equal(got_an, an)


@pytest.mark.parametrize("name", ["di6d", "dubins5d", "quad12d"])
def test_score_bit_exact(name, orc):
    cfg = W.make_query(name, 0).cfg
    R1 = cfg.r1_len ** 3
    P = cfg.r2_len ** 3
    rng = np.random.default_rng(5)
    total = rng.integers(0, 400000, R1).astype(np.int32)
    total[rng.random(R1) < 0.3] = 0
    vc = (total * rng.random(R1)).astype(np.int32)
    act = (rng.random(R1 * P) < 0.4).astype(np.uint8)
    want = orc.score(cfg, total, vc, act)
    w = np.packbits(act, bitorder="little")
    w = np.concatenate([w, np.zeros((-len(w)) % 4, np.uint8)]).view(np.uint32)
    out = empty(R1, torch.float32)
    d_total, d_vc, d_w = dev(total), dev(vc), dev(w)
    call("kgmt_score", C.byref(cfg), ptr(d_total), ptr(d_vc), ptr(d_w), ptr(out), stream())
    assert np.array_equal(bits(host(out)), bits(want))


@pytest.mark.parametrize("n", [0, 1, 1023, 1024, 1025, 300001])
def test_compact_and_repeat(n):
    rng = np.random.default_rng(n)
    flags = (rng.random(n) < 0.37).astype(np.uint8)
    out = empty(max(n, 1), torch.int32)
    cnt = empty(1, torch.int32)
    d_flags = dev(flags) if n else None
    call("kgmt_compact", ptr(d_flags), n, ptr(out), ptr(cnt), stream())
    want = np.flatnonzero(flags)
    assert host(cnt)[0] == len(want)
    assert np.array_equal(host(out)[: len(want)], want)
    ids = rng.integers(0, 1 << 20, n).astype(np.uint32)
    counts = rng.integers(0, 16, n).astype(np.uint32)
    rep = np.repeat(ids, counts)
    cap = max(len(rep) - 7, 0)
    o2 = empty(max(cap, 1), torch.int32)
    tot = empty(1, torch.int32)
    d_ids = dev(ids.view(np.int32)) if n else None
    d_counts = dev(counts.view(np.int32)) if n else None
    call("kgmt_repeat", ptr(d_ids), ptr(d_counts), n, ptr(o2), cap, ptr(tot), stream())
    assert host(tot)[0] == len(rep)
    assert np.array_equal(host(o2)[:cap].view(np.uint32), rep[:cap])


@pytest.mark.parametrize("kind", ["edge_boxes", "max_boxes", "reference_house"])
def test_scene_grouping_keeps_verdicts(kind, orc):
    """Boxes are grouped and reordered on the device side; verdicts must stay
    identical to the oracle's plain loop over all boxes, including inverted
    boxes, NaN coordinates, boxes beyond the workspace and the maximum count."""
    import os
    cfg = CASES["di_ref"]()
    rng = np.random.default_rng(21)
    if kind == "edge_boxes":
        obs = W.random_boxes(rng, 40, 0.05, 0.3)
        obs[3] = [0.6, 0.2, 0.2, 0.4, 0.5, 0.5]          # inverted on x
        obs[7] = [np.nan, 0.1, 0.1, 0.3, 0.3, 0.3]       # NaN face: that face never separates
        obs[11] = [-5, -5, -5, -4, -4, -4]               # far outside
        obs[19] = [0.2, 0.2, 0.2, 0.2, 0.9, 0.9]         # zero thickness
        obs = obs.astype(np.float32)
    elif kind == "max_boxes":
        obs = W.random_boxes(rng, 1500, 0.005, 0.03)
    else:
        here = os.path.dirname(os.path.abspath(__file__))
        obs = W.read_obstacles_csv(os.path.join(here, "golden", "obstacles", "house.csv"))
    S, Cd, D = dims(cfg.model)
    n, npar = 30000, 900
    X = parent_states(cfg, npar, seed=9)
    pidx = rng.integers(0, npar, n).astype(np.int32)
    ctrl, steps, _ = orc.draw_children(cfg, 0, 2, 0, n)
    want = orc.propagate_children(cfg, pidx, X, ctrl, steps, obs)
    child = empty((D, n), torch.float32)
    valid = empty(n, torch.uint8)
    r1 = empty(n, torch.int32)
    r2 = empty(n, torch.int32)
    d_obs, d_pidx, d_X, d_ctrl, d_steps = dev(obs.reshape(-1)), dev(pidx), dev(X), dev(ctrl), dev(steps)
    call("kgmt_propagate_collide", C.byref(cfg), ptr(d_pidx), ptr(d_X), npar, ptr(d_ctrl), ptr(d_steps),
         n, ptr(d_obs), len(obs), ptr(child), ptr(valid), ptr(r1), ptr(r2), stream())
    assert np.array_equal(host(valid), want[1])
    assert np.array_equal(bits(host(child)), bits(want[0]))
    if kind == "edge_boxes":
        # a NaN face never separates, so the box acts as if that face were at -inf
        alt = obs.copy()
        alt[7, 0] = -np.inf
        assert np.array_equal(orc.propagate_children(cfg, pidx, X, ctrl, steps, alt)[1], want[1])


def test_too_many_obstacles_is_an_error():
    from kgmt_amd import _lib
    cfg = CASES["di_ref"]()
    obs = np.zeros((1501, 6), np.float32)
    d_obs = dev(obs.reshape(-1))
    buf = empty(16, torch.float32)
    rc = _lib.lib().kgmt_propagate_collide(C.byref(cfg), ptr(buf), ptr(buf), 1, ptr(buf), ptr(buf), 1,
                                           ptr(d_obs), 1501, ptr(buf), ptr(buf), ptr(buf), ptr(buf), stream())
    assert rc == -4


def _scene_numpy(rows):
    """Scene layout restated in numpy (float32 arithmetic): Morton order of
    the box centres over the bounds of the finite centres (10 bits per axis,
    ties by input order), groups of 8, group bounds min/max per axis or NaN
    where a member has a NaN face, dummies lo = +inf / hi = -inf."""
    rows = np.asarray(rows, np.float32).reshape(-1, 6)
    n = len(rows)
    per = 8 * 4
    ngrp = (n + per - 1) // per * 4
    c = np.float32(0.5) * (rows[:, :3] + rows[:, 3:])
    with np.errstate(invalid="ignore", over="ignore"):
        fin = np.isfinite(c)
        lo = np.array([c[fin[:, d], d].min() if fin[:, d].any() else np.inf for d in range(3)], np.float32)
        hi = np.array([c[fin[:, d], d].max() if fin[:, d].any() else -np.inf for d in range(3)], np.float32)
        t = np.zeros_like(c)
        for d in range(3):
            if hi[d] > lo[d]:
                t[:, d] = (c[:, d] - lo[d]) / (hi[d] - lo[d])
        t = np.where(t >= 0, t, np.float32(0))
        t = np.minimum(t, np.float32(1))
        q = (t * np.float32(1023)).astype(np.uint32)
    key = np.zeros(n, np.uint64)
    for b in range(9, -1, -1):
        for d in range(3):
            key = (key << np.uint64(1)) | ((q[:, d] >> b) & 1).astype(np.uint64)
    order = np.lexsort((np.arange(n), key))
    boxes = np.empty((ngrp * 8, 2, 4), np.float32)
    boxes[:, 0, :3], boxes[:, 1, :3] = np.inf, -np.inf
    boxes[:, :, 3] = 0
    boxes[:n, 0, :3] = rows[order, :3]
    boxes[:n, 1, :3] = rows[order, 3:]
    grp = np.zeros((ngrp, 2, 4), np.float32)
    for g in range(ngrp):
        m = boxes[g * 8: min(n, (g + 1) * 8)]
        for d in range(3):
            l, h = m[:, 0, d], m[:, 1, d]
            if len(m) == 0:
                grp[g, 0, d], grp[g, 1, d] = np.inf, -np.inf
            elif np.isnan(l).any() or np.isnan(h).any():
                grp[g, 0, d] = grp[g, 1, d] = np.nan
            else:
                grp[g, 0, d], grp[g, 1, d] = l.min(), h.max()
    return ngrp, np.concatenate([grp.reshape(-1), boxes.reshape(-1)])


@pytest.mark.parametrize("kind", ["random_100", "random_1500", "nan_and_inverted", "ties", "one", "house"])
def test_device_scene_layout(kind):
    """kgmt_build_scene against the numpy restatement, bit for bit."""
    rng = np.random.default_rng(5)
    if kind.startswith("random"):
        rows = W.random_boxes(rng, int(kind.split("_")[1]), 0.02, 0.1)
    elif kind == "nan_and_inverted":
        rows = W.random_boxes(rng, 40, 0.02, 0.1)
        rows[3, 1] = np.nan
        rows[17, 5] = np.nan
        rows[9] = [0.6, 0.6, 0.6, 0.4, 0.4, 0.4]
        rows[21, :3] = -np.inf
    elif kind == "ties":
        rows = np.tile(np.array([[0.1, 0.2, 0.3, 0.2, 0.3, 0.4]], np.float32), (37, 1))
        rows[::5] += 0.25
    elif kind == "one":
        rows = np.array([[0.1, 0.2, 0.3, 0.2, 0.3, 0.4]], np.float32)
    else:
        rows = W.read_obstacles_csv(os.path.join(GOLDEN, "obstacles", "house.csv"))
    rows = np.ascontiguousarray(rows, np.float32)
    n = len(rows)
    want_g, want = _scene_numpy(rows)
    d_rows = dev(rows.reshape(-1))
    out = empty(want.size, torch.float32)
    ng = C.c_uint32()
    call("kgmt_build_scene", ptr(d_rows), n, ptr(out), want.size, C.byref(ng), stream())
    assert ng.value == want_g
    got = host(out)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), np.flatnonzero(~same)[:10]
    # a buffer one scene entry short is refused
    fn = getattr(_lib.lib(), "kgmt_build_scene")
    assert fn(ptr(d_rows), n, ptr(out), want.size - 8, C.byref(ng), stream()) == -4
