"""The reference planner's per-iteration rules restated in numpy, written
from the reference source independently of the C oracle, to pin the oracle's
whole-plan function (orc_plan) iteration by iteration.

Only stage functions that are pinned on their own are borrowed from the
oracle: the Philox counter RNG (known-answer vectors), the child draws, the
propagate + collide + region step of every child (closed forms, FP64 twins,
brute-force segment tests, exhaustive region maps) and the region score
(numpy FP64 closed form). Everything that decides the tree is restated here:

  expansion multiset   {i repeated repeat[i] times} in ascending i: the
                       exclusive scan + repeatInd of KGMT.cu:176-181,
                       Planner.cu:68-80
  launch regime        32 children per entry, or min(avail // R, 32) >= 1
                       when capacity is short, KGMT.cu:183-201
  children             child u of entry u // per, parent's frontier flag
                       cleared, KGMT.cu:226-258
  counters, flags      total/valid per region, next-frontier flag = valid and
                       (u < score[r1] or subregion not yet active), subregion
                       set active; the active set is read as of the start of
                       the iteration, KGMT.cu:249-256
  compaction           accepted children in ascending child order, clamped to
                       the free capacity, KGMT.cu:383-385
  fAccept              KGMT.cu:387-388, evaluated as the reference's host does
  re-activation        every old node with frontier 0 and U <= score[region]
                       + fAccept rejoins with repeat 1, KGMT.cu:364-376
  append               new node t = treeSize + j copies accepted child j,
                       cost = distance to the goal, repeat 15 if the region's
                       valid count < 10 else 1, KGMT.cu:318-340
  goal                 the smallest new index within the goal threshold,
                       path walked goal -> root, KGMT.cu:342-361
"""
from __future__ import annotations

import numpy as np

from kgmt_amd.abi import dims

F32 = np.float32


def reference_faccept(cfg, itr, tree, nxt):
    """KGMT.cu:387-388 as the reference's host evaluates it: treeAddSize and
    h_itr_ * EPSILON in float, pow(treeAddSize, 5) in double (std::pow of a
    float promotes), the product in double, stored into a float."""
    a = F32(1) - F32(tree + nxt) / F32(cfg.capacity)
    first = F32(itr) * F32(cfg.epsilon)
    return F32(np.float64(first) * np.power(np.float64(a), 5.0))


def u01(words):
    """curand_uniform's range (0, 1]: the top 24 bits plus one, times 2^-24."""
    return ((words >> np.uint32(8)) + np.uint32(1)).astype(F32) * F32(2.0 ** -24)


def react_uniforms(orc, cfg, itr, n):
    """Re-activation uniforms of nodes 0..n-1 in iteration itr: word 0 of
    Philox4x32-10 with counter (node, itr, 2, query_id) and the 64-bit seed
    as key."""
    idx = np.arange(n, dtype=np.uint32)
    ctr = np.stack([idx, np.full(n, itr, np.uint32), np.full(n, 2, np.uint32),
                    np.full(n, cfg.query_id, np.uint32)], 1)
    key = np.tile(np.array([cfg.seed & 0xFFFFFFFF, cfg.seed >> 32], np.uint32), (n, 1))
    return u01(orc.philox(ctr, key)[:, 0])


def goal_distance(pos, goal, w):
    """helper.cuh:138-152 over the W position coordinates, in float32 in
    source order."""
    d = pos[:, 0] - F32(goal[0])
    acc = d * d
    d = pos[:, 1] - F32(goal[1])
    acc = acc + d * d
    if w == 3:
        d = pos[:, 2] - F32(goal[2])
        acc = acc + d * d
    return np.sqrt(acc)


def frontier_update(orc, cfg, itr, tree, samples, parents, costs, frontier, child, parent_u, fnext, valid, score,
                    goal):
    """The frontier update of iteration itr on copies of the arrays: samples
    (capacity, D), parents, costs, frontier (capacity), child (D, n),
    parent_u, fnext (n), region valid counts and scores. Returns a dict with
    the updated arrays, the repeat count of every node, the next expansion
    list (its first capacity entries), its size R, the appended count, the
    goal node (-1) and counts of the rules that acted."""
    N, D, w = cfg.capacity, cfg.dims[2], cfg.wdim
    samples, parents, costs = samples.copy(), parents.copy(), costs.copy()
    frontier = frontier.astype(bool).copy()
    acc = np.flatnonzero(np.asarray(fnext) != 0)
    avail = N - tree
    nxt = min(len(acc), avail)
    fa = reference_faccept(cfg, itr, tree, nxt)
    rep = np.zeros(N, np.int64)
    # re-activation of old nodes (repeat counts were zeroed first); the
    # region of an old node is recomputed from its sample, as the reference does
    U = react_uniforms(orc, cfg, itr, tree)
    pos = np.zeros((tree, 3), F32)
    pos[:, :w] = samples[:tree, :w]
    sc = score[orc.regions(cfg, pos)[0]]
    react = ~frontier[:tree] & (U <= sc + fa)
    by_fa = react & (U > sc)
    frontier[:tree][react] = True
    rep[:tree][react] = 1
    # append the accepted children
    u = acc[:nxt]
    new = tree + np.arange(nxt)
    samples[new] = child[:, u].T
    parents[new] = parent_u[u]
    d = goal_distance(samples[new, :w], goal, w)
    costs[new] = d
    frontier[new] = True
    cpos = np.zeros((nxt, 3), F32)
    cpos[:, :w] = samples[new, :w]
    r1 = orc.regions(cfg, cpos)[0]
    sparse = valid[r1] < cfg.sparse_valid_below
    rep[new] = np.where(sparse, cfg.repeat_sparse, cfg.repeat_dense)
    hits = np.flatnonzero(d < F32(cfg.goal_thresh))
    M = np.repeat(np.arange(tree + nxt), rep[:tree + nxt])
    return {"samples": samples, "parents": parents, "costs": costs, "frontier": frontier.astype(np.uint8),
            "rep": rep, "M": M[:N], "R": len(M), "next": nxt, "goal": tree + int(hits[0]) if len(hits) else -1,
            "react": int(react.sum()), "react_by_faccept": int(by_fa.sum()), "sparse": int(sparse.sum()),
            "dense": int((~sparse).sum()), "clamped": len(acc) > avail, "hits": len(hits)}


def replay(orc, q, hook=None):
    """Plan query q with the rules above. Returns (out, log): out has the
    fields of oracle.plan's result; log has one dict per iteration counting
    how often each rule acted. hook(itr, inputs) (optional) receives the
    inputs of every iteration's frontier update."""
    cfg = q.cfg
    S, Cd, D = dims(cfg.model)
    w = cfg.wdim
    N, B = cfg.capacity, cfg.children_per_entry
    R1 = cfg.regions
    R2 = R1 * cfg.r2_len ** w
    goal = np.asarray(q.goal, F32)
    samples = np.zeros((N, D), F32)
    parents = np.zeros(N, np.int32)
    costs = np.zeros(N, F32)
    frontier = np.zeros(N, bool)
    rep = np.zeros(N, np.int64)
    samples[0] = q.start[:D]
    parents[0] = -1
    frontier[0] = True
    rep[0] = cfg.root_repeat
    total = np.zeros(R1, np.int32)
    valid = np.zeros(R1, np.int32)
    score = np.zeros(R1, F32)
    active = np.zeros(R2, np.uint8)
    tree, itr, goal_idx, expanded = 1, 0, -1, 0
    log = []
    while itr < cfg.max_iter:
        itr += 1
        M = np.repeat(np.arange(tree), rep[:tree])
        R = len(M)
        avail = N - tree
        per = B if R * B <= avail else max(1, min(avail // R if R else 0, B))
        nC = min(R * per, N)
        parent = M[np.arange(nC) // per].astype(np.int32)
        ctrl, steps, uf = orc.draw_children(cfg, cfg.query_id, itr, 0, nC)
        child, ok, r1, r2 = orc.propagate_children(cfg, parent, samples[:tree, :S].T, ctrl, steps, q.obstacles)
        expanded += nC
        frontier[parent] = False
        okb = ok.astype(bool)
        total += np.bincount(r1, minlength=R1).astype(np.int32)
        valid += np.bincount(r1[okb], minlength=R1).astype(np.int32)
        fnext = okb & ((uf < score[r1]) | (active[r2] == 0))
        active = active.copy()
        active[r2[okb]] = 1
        score = orc.score(cfg, total, valid, active)
        inputs = dict(itr=itr, tree=tree, samples=samples, parents=parents, costs=costs,
                      frontier=frontier.astype(np.uint8), child=child, parent_u=parent, fnext=fnext.astype(np.uint8),
                      valid=valid.copy(), score=score.copy(), total=total.copy(), children=nC)
        if hook:
            hook(itr, inputs)
        f = frontier_update(orc, cfg, itr, tree, samples, parents, costs, frontier, child, parent, fnext, valid, score,
                            goal)
        samples, parents, costs = f["samples"], f["parents"], f["costs"]
        frontier, rep = f["frontier"].astype(bool), f["rep"]
        nxt = f["next"]
        log.append({"R": R, "children": nC, "per": per, "next": nxt, "react": f["react"],
                    "react_by_faccept": f["react_by_faccept"], "sparse": f["sparse"], "dense": f["dense"],
                    "clamped": f["clamped"], "hits": f["hits"]})
        tree += nxt
        if f["goal"] >= 0:
            goal_idx = f["goal"]
            break
        if tree == N:
            break
    stats = {"iterations": itr, "tree_size": tree, "goal_index": goal_idx, "path_rows": 0,
             "path_cost": F32(0), "path_length": F32(0), "expanded": expanded}
    path = np.zeros((0, D), F32)
    if goal_idx >= 0:
        rows, n, prev = [], goal_idx, None
        cost, length = F32(0), F32(0)
        while n != -1:
            row = samples[n]
            rows.append(row)
            if parents[n] != -1:
                cost = F32(cost + row[D - 1])
            if prev is not None:
                length = F32(length + goal_distance(prev[None, :w], row[:w], w)[0])
            prev = row
            n = parents[n]
        path = np.array(rows, F32)
        stats.update(path_rows=len(rows), path_cost=cost, path_length=length)
    out = {"stats": stats, "samples": samples[:tree], "parents": parents[:tree], "costs": costs[:tree],
           "path": path, "total": total, "valid": valid, "score": score}
    return out, log


def capture_updates(orc, q, iterations):
    """Inputs of the frontier update at the given iterations of query q's
    plan (a list of dicts, with the query's configuration and goal)."""
    want = set(iterations)
    out = []

    def hook(itr, inputs):
        if itr in want:
            out.append(dict(inputs, cfg=q.cfg, goal=np.asarray(q.goal, F32),
                            samples=inputs["samples"].copy(), parents=inputs["parents"].copy(),
                            costs=inputs["costs"].copy(), child=inputs["child"].copy(),
                            parent_u=inputs["parent_u"].copy()))

    replay(orc, q, hook)
    return out


def update_variants(state):
    """Edge cases built from one captured state: the capacity cut so the
    accepted children overflow it, the goal moved onto accepted children
    (several hits), no children at all, every old node already in the
    frontier."""
    cfg, tree = state["cfg"], state["tree"]
    flagged = np.flatnonzero(state["fnext"])
    D = cfg.dims[2]
    out = {}
    # capacity cut to the tree plus half the accepted children, children
    # truncated to fit it (the planner never has more children than capacity)
    k = len(state["fnext"])
    while True:
        f = int(np.count_nonzero(state["fnext"][:k]))
        cap = tree + max(1, f // 2)
        if k <= cap:
            break
        k = cap
    if f >= 4 and cap < cfg.capacity:
        c = dict(state, cfg=cfg.replace(capacity=cap))
        c["samples"] = state["samples"][:cap].copy()
        c["parents"] = state["parents"][:cap].copy()
        c["costs"] = state["costs"][:cap].copy()
        c["frontier"] = state["frontier"][:cap].copy()
        c["child"] = state["child"][:, :k].copy()
        c["parent_u"] = state["parent_u"][:k].copy()
        c["fnext"] = state["fnext"][:k].copy()
        out["clamped"] = c
    if len(flagged) >= 4:
        g = dict(state)
        goal = np.zeros(D, F32)
        goal[:cfg.wdim] = state["child"][:cfg.wdim, flagged[len(flagged) // 3]]
        g["goal"] = goal
        out["goal_hits"] = g
    e = dict(state)
    e["child"] = np.zeros((D, 0), F32)
    e["parent_u"] = np.zeros(0, np.int32)
    e["fnext"] = np.zeros(0, np.uint8)
    out["no_children"] = e
    f = dict(state)
    f["frontier"] = state["frontier"].copy()
    f["frontier"][:tree] = 1
    out["no_reactivation"] = f
    return out


def run_update(fn, orc, st):
    """Call a frontier-update implementation (kgmt_rules.frontier_update style
    signature) on a captured state."""
    return fn(orc, st["cfg"], st["itr"], st["tree"], st["samples"], st["parents"], st["costs"], st["frontier"],
              st["child"], st["parent_u"], st["fnext"], st["valid"], st["score"], st["goal"])
