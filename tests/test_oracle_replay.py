"""Whole plans of the CPU oracle against the numpy restatement of the
reference's per-iteration rules (tests/kgmt_rules.py), every iteration from
the first to the last: the repeat counts 15/1 chosen by a region's valid
count (KGMT.cu:336-339), the re-activation of old nodes when U <= score +
fAccept (KGMT.cu:364-376), the order of the expansion multiset (ascending
node index, KGMT.cu:176-181), the launch regime when capacity runs short
(KGMT.cu:183-201), the capacity clamp and the smallest-index goal hit.

The cases are chosen so that every rule acts (the log of the replay is
checked for it), and the mutation test builds altered copies of the oracle
(one rule changed each) and requires the replay to tell each apart from the
real one."""
import os
import subprocess

import numpy as np
import pytest

from kgmt_amd import workloads as W
from kgmt_rules import (capture_updates, frontier_update, reference_faccept, react_uniforms, replay, run_update,
                        update_variants)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "kgmt_oracle.c")

CASES = {
    "di6d": ("di6d", 0, dict(capacity=1 << 13, max_iter=60)),
    "di6d_goal": ("di6d", 0, dict(capacity=1 << 14)),
    "dubins5d": ("dubins5d", 1, dict(capacity=1 << 13, max_iter=60)),
    "quad12d": ("quad12d", 0, dict(capacity=1 << 12, max_iter=12)),
    "unicycle2d": ("unicycle2d", 2, dict(capacity=1 << 13, max_iter=60)),
}


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _same(ref, got):
    """Field-by-field comparison; returns the first differing field or None."""
    for k in ("iterations", "tree_size", "goal_index", "path_rows", "expanded"):
        if ref["stats"][k] != got["stats"][k]:
            return k
    for k in ("path_cost", "path_length"):
        if np.float32(ref["stats"][k]).view(np.uint32) != np.float32(got["stats"][k]).view(np.uint32):
            return k
    for k in ("samples", "parents", "costs", "path", "total", "valid", "score"):
        if not np.array_equal(_bits(ref[k]), _bits(got[k])):
            return k
    return None


@pytest.fixture(scope="module")
def replays(orc):
    out = {}
    for name, (wl, seed, kw) in CASES.items():
        q = W.make_query(wl, seed, **kw)
        out[name] = (q, *replay(orc, q))
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_whole_plan_replay(name, replays, orc):
    q, rep, log = replays[name]
    ref = orc.plan(q.cfg, q.start, q.goal, q.obstacles, log_iters=True)
    assert _same(ref, rep) is None, _same(ref, rep)
    il = ref["iter_log"]
    assert [tuple(r) for r in il] == [(g["R"], g["children"], g["next"], g["react"]) for g in log]
    # the rules the replay pins act in this plan, from iteration 2 on
    assert len(log) >= 3
    later = log[1:]
    assert sum(g["react"] for g in later) > 0
    assert sum(g["react_by_faccept"] for g in later) > 0  # decided by the + fAccept term
    assert sum(g["sparse"] for g in later) > 0 and sum(g["dense"] for g in later) > 0
    assert any(g["react"] > 0 and g["next"] > 0 for g in later)  # both parts of the next multiset


def test_rules_reach_the_capacity_regimes(replays):
    """The small-capacity plans run the reduced-children regime and the
    clamp of accepted children; the goal case stops at a goal hit."""
    logs = [log for _, _, log in replays.values()]
    assert any(g["per"] < 32 for log in logs for g in log)
    assert any(g["clamped"] for log in logs for g in log)
    q, rep, log = replays["di6d_goal"]
    assert rep["stats"]["goal_index"] >= 0 and log[-1]["hits"] >= 1


def test_react_uniforms_are_the_oracle_draws(orc):
    cfg = W.config("dubins5d", seed=0x1234567890AB, query_id=3)
    U = react_uniforms(orc, cfg, 7, 300)
    want = np.array([orc.draw_react(cfg, 3, 7, i) for i in range(300)], np.float32)
    assert np.array_equal(U.view(np.uint32), want.view(np.uint32))


def test_faccept_rule_matches_oracle(orc):
    cfg = W.config("di6d", capacity=1 << 13)
    for itr in (1, 2, 17, 300):
        for tree, nxt in ((1, 160), (4000, 100), (8000, 192), (8191, 1)):
            got = np.float32(orc.faccept(cfg, itr, tree, nxt))
            assert got.view(np.uint32) == reference_faccept(cfg, itr, tree, nxt).view(np.uint32)


# altered copies of the oracle: (name, original text, replacement)
MUTATIONS = [
    ("repeat counts swapped", "cfg->repeat_sparse : cfg->repeat_dense", "cfg->repeat_dense : cfg->repeat_sparse"),
    ("sparse test <=", "validc[r1] < cfg->sparse_valid_below", "validc[r1] <= cfg->sparse_valid_below"),
    ("fAccept dropped", "U <= score[orc_region(cfg, samples + (size_t)i * D)] + fA",
     "U <= score[orc_region(cfg, samples + (size_t)i * D)]"),
    ("re-activated order reversed", "for(uint32_t k = 0; k < nReact && nM < N; ++k) M[nM++] = react[k];",
     "for(uint32_t k = nReact; k-- > 0 && nM < N;) M[nM++] = react[k];"),
    ("re-activation ignores the frontier flag", "if(frontier[i] == 0)", "if(1)"),
    ("largest goal index", "if(dg < cfg->goal_thresh && best < 0) best", "if(dg < cfg->goal_thresh) best"),
    ("subregion flag read after this iteration's writes", "fnext[u]    = (uf < score[r1]) || !actPrev[r2];",
     "fnext[u]    = (uf < score[r1]) || !actNext[r2];"),
]


def _build_variant(tmp_path, src, orc):
    path = tmp_path / "kgmt_oracle_variant.c"
    path.write_text(src)
    so = tmp_path / "libkgmt_oracle_variant.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-mfma", "-fopenmp",
                    f"-I{os.path.join(ROOT, 'oracle')}", str(path), "-o", str(so), "-lm"], check=True)
    return orc.load(str(so))


def _differences(lib, replays, orc):
    out = []
    for name, (q, rep, _) in replays.items():
        field = _same(orc.plan(q.cfg, q.start, q.goal, q.obstacles, lib_=lib), rep)
        if field:
            out.append((name, field))
    return out


def test_unaltered_copy_agrees(replays, tmp_path, orc):
    """Control for the mutation test: the same recipe on the unaltered source
    agrees with the replay everywhere."""
    lib = _build_variant(tmp_path, open(ORACLE_SRC).read(), orc)
    assert _differences(lib, replays, orc) == []


@pytest.mark.parametrize("label,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_replay_detects_altered_oracle(label, old, new, replays, tmp_path, orc):
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"mutation anchor not found once: {old}"
    lib = _build_variant(tmp_path, src.replace(old, new), orc)
    assert _differences(lib, replays, orc), f"the replay did not detect: {label}"


UPDATE_CASES = {
    "di6d": ("di6d", 0, dict(capacity=1 << 13, max_iter=60), (1, 2, 5)),
    "dubins5d": ("dubins5d", 1, dict(capacity=1 << 13, max_iter=60), (2, 4)),
    "quad12d": ("quad12d", 0, dict(capacity=1 << 12, max_iter=12), (2, 6)),
    "unicycle2d": ("unicycle2d", 2, dict(capacity=1 << 13, max_iter=60), (3,)),
}


def update_states(orc):
    """Captured frontier-update inputs of mid-plan iterations, plus the edge
    cases derived from each (capacity clamp, goal hits, no children, no
    re-activation candidates)."""
    out = []
    for name, (wl, seed, kw, its) in UPDATE_CASES.items():
        for st in capture_updates(orc, W.make_query(wl, seed, **kw), its):
            out.append((f"{name}/itr{st['itr']}", st))
            out += [(f"{name}/itr{st['itr']}/{k}", v) for k, v in update_variants(st).items()]
    return out


def compare_update(want, got):
    """First differing output of two frontier updates, or None."""
    for k in ("next", "R", "goal", "react"):
        if int(want[k]) != int(got[k]):
            return k
    for k in ("samples", "parents", "costs", "frontier", "M"):
        if not np.array_equal(_bits(np.asarray(want[k])), _bits(np.asarray(got[k]))):
            return k
    return None


def oracle_update(orc, cfg, itr, tree, samples, parents, costs, frontier, child, parent_u, fnext, valid, score, goal):
    return orc.frontier_update(cfg, itr, tree, samples, parents, costs, frontier, child, parent_u, fnext, valid,
                               score, goal)


def test_frontier_update_stage_against_rules(orc):
    """orc_frontier_update (the oracle's stage function, the one orc_plan
    calls) against the numpy rule on mid-plan states and edge cases."""
    states = update_states(orc)
    seen = {"clamped": 0, "goal": 0, "react": 0}
    for name, st in states:
        want = run_update(frontier_update, orc, st)
        got = run_update(oracle_update, orc, st)
        assert compare_update(want, got) is None, (name, compare_update(want, got))
        seen["clamped"] += want["clamped"]
        seen["goal"] += want["goal"] >= 0
        seen["react"] += want["react"] > 0
    assert all(v > 0 for v in seen.values()), seen
