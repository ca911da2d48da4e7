"""Benchmark of the KGMT expansion path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dubins5d]
                    [--seeds 200] [--batch 512] [--no-cpu] [--no-plain]

A step is one pass of the plain single-query planner over the workload's
fixed seed set: by default the 200 queries (seeds 0..199) of the 5D Dubins
airplane workload (100 boxes, 2^18 tree nodes, 16^3/2^3 region grid, dt 0.02,
5..25 steps per edge), planned one after the other. W untimed passes, then K
timed ones, so the figure does not depend on which seeds K happens to cover.
Obstacles are device-resident before the timed region (passed as device
pointers, as the reference's plan() takes them); start/goal are host arrays
like the reference's h_initial/h_goal.

One JSON line on stdout: queries/s (value), median time-to-solution over the
solved queries, the propagate kernel's roofline figures measured with HIP
events on the planner's stream (around every propagate launch while the seed
set is planned once more: the event pairs add dispatch gaps, so the timed
passes have none), the CPU oracle's throughput on a bounded sample of the same
seeds, and (--batch N) the batched planner on N mixed queries against the same
queries run one by one through the plain planner, with the batched propagate
kernel's roofline.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import struct
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 157.3
# FP32 vector issue without FMA or packed math: 256 CUs x 4 SIMDs x 16 lanes
# x 2.4 GHz (the 157.3 TFLOP/s peak counts FMA as 2 and packed FP32 as 2)
VALU_PEAK_TLANEOPS = VALU_PEAK_TFLOPS / 4

WORKLOAD_DESC = {
    "dubins5d": "5D Dubins airplane, plain single-query planner, 100 boxes sides U[0.02,0.1], 2^18 nodes, grid 16^3/2^3, dt 0.02, 5-25 steps/edge",
    "di6d": "6D double integrator, plain single-query planner, 20 boxes sides U[0.05,0.2], 2^17 nodes, grid 8^3/4^3, dt 0.05, 2-10 steps/edge",
    "quad12d": "12D quadcopter, plain single-query planner, 3 slotted walls + 50 boxes, 2^19 nodes, grid 10^3/4^3, dt 0.01, 10-50 steps/edge",
}
BOXES = {"dubins5d": 100, "di6d": 20, "quad12d": 56}
DEFAULT_SEEDS = {"dubins5d": 200, "di6d": 50, "quad12d": 200}


def algorithmic_bytes_per_child(model: int) -> int:
    """Algorithmic HBM bytes of one expanded child, SoA FP32 with
    pre-generated controls: parent index 4, parent state 4S, controls 4C and
    u16 steps, child state+controls+duration 4(S+C+1), child parent index 4,
    valid/frontier flag 1, regions r1 and r2 8. 95 B for the double
    integrator, 87 B for the 5D Dubins airplane, 151 B for the quadcopter."""
    from kgmt_amd.abi import dims
    S, C, D = dims(model)
    return 4 + 4 * S + 4 * C + 2 + 4 * D + 4 + 1 + 8


def moved_bytes(model: int, children: int, flagged: int) -> int:
    """Bytes this kernel moves by design: per child the expansion-list entry
    4 B, parent state 4S, region 4 B and next-frontier flag 1 B; per flagged
    child (the only ones the tree append reads) the row of 4D and the parent
    index 4 B. Controls come from the counter-based RNG in registers."""
    from kgmt_amd.abi import dims
    S, C, D = dims(model)
    return children * (4 + 4 * S + 4 + 1) + flagged * (4 * D + 4)


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def fixture_agreement(name, res):
    """Seeds whose outcome (iterations, tree size, goal node, path rows, path
    cost bits) equals the committed oracle record of the same seed
    (tests/golden/plan_stats_<workload>.json, written by the CPU oracle);
    None when there is no record file."""
    path = os.path.join(ROOT, "tests", "golden", f"plan_stats_{name}.json")
    try:
        with open(path) as f:
            recs = {r[0]: r for r in json.load(f)["records"]}
    except (OSError, ValueError, KeyError):
        return None
    same = checked = 0
    for seed, r in enumerate(res):
        rec = recs.get(seed)
        if rec is None:
            continue
        checked += 1
        same += (rec[1], rec[2], rec[3], rec[4], rec[5]) == (
            r["iterations"], r["tree_size"], r["goal_index"], r["path_rows"],
            f32_bits(r["path_cost"]) if r["success"] else rec[5])
    return {"seeds_checked": checked, "seeds_identical": same}


def run_plain(name, n_seeds, steps, warmup, launch=None, profile=True):
    import torch
    from kgmt_amd import workloads as W
    from kgmt_amd.planner import KGMT

    qs = [W.make_query(name, s) for s in range(n_seeds)]
    d_obs = [torch.from_numpy(q.obstacles.reshape(-1)).cuda() for q in qs]
    g = KGMT(qs[0].cfg)
    if launch:
        g.set_launch(**launch)

    # the queries' arguments as the planner takes them (device obstacle
    # addresses, seeds), ready before the timed region
    args = [(q.cfg.seed, q.start, q.goal, int(d.data_ptr()), len(q.obstacles)) for q, d in zip(qs, d_obs)]

    def one_pass():
        out = []
        for seed, start, goal, ptr, n in args:
            g.set_seed(seed)
            out.append(g.plan(start, goal, ptr, n))
        return out

    for _ in range(warmup):
        one_pass()
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    passes = [one_pass() for _ in range(steps)]
    g.synchronize()  # the planner's own stream (torch's synchronize does not cover it)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    passes = [[r.as_dict() for r in p] for p in passes]
    res = passes[0]
    for p in passes[1:]:
        if [(r["tree_size"], r["iterations"]) for r in p] != [(r["tree_size"], r["iterations"]) for r in res]:
            raise RuntimeError("plans differ between passes over the same seeds")
    prof = breakdown = None
    if profile:
        # propagate launch durations: the seed set planned once more with HIP
        # events on the planner's stream around every propagate launch (an
        # event pair adds ~15 us of dispatch gaps per iteration, so the timed
        # passes have none); the plans are identical, so are the launches
        g.set_profiling(1)
        for i, (q, d) in enumerate(zip(qs, d_obs)):
            g.set_seed(q.cfg.seed)
            r = g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles))
            if (r.tree_size, r.iterations) != (res[i]["tree_size"], res[i]["iterations"]):
                raise RuntimeError(f"seed {i}: the profiled plan differs from the timed one")
        prof = g.profile()
        # per-kernel breakdown from events around every kernel, first 10 seeds
        g.set_profiling(2)
        for q, d in zip(qs[:10], d_obs[:10]):
            g.set_seed(q.cfg.seed)
            g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles))
        p2 = g.profile()
        breakdown = {k: round(v["ms"] / max(v["launches"], 1) * 1e3, 2) for k, v in p2["kernels"].items()}
        g.set_profiling(0)
    g.synchronize()
    return qs[0].cfg, passes, dt, prof, breakdown


def measured_traffic(key):
    """HBM bytes per child of a propagate kernel measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (scaled by a 256 MiB copy) and committed
    with the commands that produced them; None when not measured."""
    path = os.path.join(ROOT, "profiles", "k_propagate_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def run_cpu(name, n_queries, threads):
    from kgmt_amd import workloads as W
    from oracle import oracle as O

    qs = [W.make_query(name, s) for s in range(n_queries)]
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    t0 = time.perf_counter()
    stats, ms = O.plan_many(cfgs, starts, goals, obs, offs, cnts, threads=threads)
    dt = time.perf_counter() - t0
    return stats, ms, dt


def run_batch(n, seed=0, tune=True, launch=None, sequential=True, profile=True):
    import numpy as np
    import torch
    from kgmt_amd import autotune
    from kgmt_amd import workloads as W
    from kgmt_amd.abi import KGMT_MODEL_QUADCOPTER
    from kgmt_amd.planner import KGMT, BatchKGMT

    qs = W.make_batch(n, seed=seed)
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    tuned = None
    if launch is None and tune:
        # launch shape chosen on a calibration batch of the same mix (other seed)
        cq = W.make_batch(n, seed=seed + 1000)
        tuned = autotune.tune_batch(*W.pack_batch(cq), use_graph=True, reps=2, cache=None)
        launch = tuned["launch"]
    b = BatchKGMT(cfgs)
    if launch:
        b.set_launch(**launch)
    b.plan(starts, goals, obs, offs, cnts, use_graph=True)  # warm-up: graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_b, fin = b.plan(starts, goals, obs, offs, cnts, use_graph=True)
    torch.cuda.synchronize()
    dt_b = time.perf_counter() - t0
    out = {
        "queries": n,
        "mix": "50% 6D double integrator / 50% 12D quadcopter, capacity log-U[2^12,2^17], 10-500 boxes",
        "qps_batched": n / dt_b,
        "median_tts_ms_batched": statistics.median([float(f) for f, r in zip(fin, res_b) if r.success] or [0.0]),
        "success_rate_batched": float(np.mean([r.success for r in res_b])),
        "hip_graph": True,
        "launch": launch or "default",
        "autotune": {"candidates": len(tuned["trials"]), "calibration": "same mix, seed + 1000",
                     "best_s": round(tuned["seconds"], 5),
                     "default_s": round(tuned["trials"][0]["seconds"], 5)} if tuned else None,
    }
    if profile:
        # propagate launches of the same batch, HIP events around each (no graph)
        b.set_profiling(1)
        res_e, _ = b.plan(starts, goals, obs, offs, cnts, use_graph=False)
        prof = b.profile()
        b.set_profiling(0)
        if [(r.tree_size, r.iterations) for r in res_e] != [(r.tree_size, r.iterations) for r in res_b]:
            raise RuntimeError("the profiled batch differs from the timed one")
        pk = prof["kernels"]["propagate"]
        bytes_alg = sum(algorithmic_bytes_per_child(c.model) * r.expanded for c, r in zip(cfgs, res_e))
        children = sum(r.expanded for r in res_e)
        quad_children = sum(r.expanded for c, r in zip(cfgs, res_e) if c.model == KGMT_MODEL_QUADCOPTER)
        avg_ms = pk["ms"] / max(pk["launches"], 1)
        ach = bytes_alg / (pk["ms"] * 1e-3) / 1e9
        tr = measured_traffic("batch512")
        per_launch = children / max(pk["launches"], 1)
        out["roofline"] = {
            "kernel": "k_propagate<mixed>",
            "bound": "hbm",
            "achieved": round(ach, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5),
            "traffic": round(tr["hbm_bytes_per_child"] * per_launch) if tr else None,
            "bytes_per_child": round(bytes_alg / max(children, 1), 2),
            "children_per_launch": round(per_launch, 1),
            "quad_fraction": round(quad_children / max(children, 1), 4),
            "avg_launch_us": round(avg_ms * 1e3, 2),
            "launches": pk["launches"],
            "note": "achieved = algorithmic bytes per child (95 B double integrator, 151 B quadcopter, per child "
                    "counted by model) x children / summed event-timed propagate launch durations",
        }
        if tr and "valu_insts_per_child" in tr:
            v = tr["valu_insts_per_child"] * 64 * per_launch / (avg_ms * 1e-3) / 1e12
            out["roofline"]["valu"] = {"achieved": round(v, 2), "peak": round(VALU_PEAK_TLANEOPS, 2),
                                       "unit": "T lane-ops/s", "frac": round(v / VALU_PEAK_TLANEOPS, 4)}
    if sequential:
        # the same queries, one by one through the plain planner (contexts
        # made beforehand: allocation is not part of a plan)
        ctxs = [KGMT(q.cfg) for q in qs]
        d_obs = [torch.from_numpy(q.obstacles.reshape(-1)).cuda() for q in qs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res_p = []
        for g, q, d in zip(ctxs, qs, d_obs):
            res_p.append(g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles)))
        for g in ctxs:
            g.synchronize()
        torch.cuda.synchronize()
        dt_p = time.perf_counter() - t0
        same = 0
        for i, (a, c) in enumerate(zip(res_b, res_p)):
            ok = (a.tree_size, a.success, a.iterations, a.goal_index, a.path_rows, f32_bits(a.path_cost),
                  f32_bits(a.path_length)) == (c.tree_size, c.success, c.iterations, c.goal_index, c.path_rows,
                                               f32_bits(c.path_cost), f32_bits(c.path_length))
            if ok and a.success:
                ok = np.array_equal(b.path(i).view(np.uint32), ctxs[i].path().view(np.uint32))
            same += ok
        out.update({
            "qps_plain_sequential": n / dt_p,
            "speedup": dt_p / dt_b,
            "median_tts_ms_plain": statistics.median([float(r.time_ms) for r in res_p if r.success] or [0.0]),
            "success_rate_plain": float(np.mean([r.success for r in res_p])),
            "identical_results": same == n,
            "identical_queries": same,
            "identical_fields": "tree size, success, iterations, goal node, path rows, path cost and length bits, "
                                "every path row's bits",
        })
        del ctxs
    agree = fixture_agreement("batch512", [r.as_dict() for r in res_b])
    if agree:
        out["oracle_fixture"] = agree
    return out


def run_refine(n=4096, seed=0, cpu_sample=0, objective="configured"):
    """Adam refinement of n quadcopter control sequences (T ~ U{20..80},
    dt 0.01, 50 steps, lr 1e-2) through kgmt_refine, cost weights of
    `objective` (workloads.OBJECTIVES); the CPU figure is the oracle's FP64
    cost+adjoint on cpu_sample trajectories, scaled to the 50 steps, one
    thread."""
    import numpy as np
    import torch
    from kgmt_amd import adjoint as A
    from kgmt_amd import workloads as W

    rb = W.make_refine_batch(n, seed=seed, objective=objective)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    u0, x0, goal, steps, obs = d(rb.u), d(rb.x0), d(rb.goal), d(rb.steps), d(rb.obstacles)
    u = u0.clone()
    A.refine(u, x0, goal, steps, obs, rb.cfg, rb.params)  # warm-up
    u = u0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = A.refine(u, x0, goal, steps, obs, rb.cfg, rb.params)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    h = hist.cpu().numpy()
    p = rb.params
    out = {
        "trajectories": n,
        "objective": f"{objective}: {p.w_effort:g}*|u - u_ref|^2 (u_ref thrust {p.u_ref[0]:g}) + "
                     f"{p.w_obstacle:g}*softplus obstacle penalty (margin {p.margin:g}, tau {p.tau:g}) + "
                     f"{p.w_goal:g}*|p_T - goal|^2",
        "adam_steps": int(rb.params.iterations),
        "ms": round(dt * 1e3, 3),
        "trajectories_per_s": round(n / dt, 1),
        "mean_cost_before": round(float(h[0].mean()), 5),
        "mean_cost_after": round(float(h[-1].mean()), 5),
        "mean_cost_reduction": round(1.0 - float(h[-1].mean()) / float(h[0].mean()), 4),
        "dtype": "f32",
    }
    if cpu_sample:
        from oracle import oracle as O
        t0 = time.perf_counter()
        for i in range(cpu_sample):
            T = int(rb.steps[i])
            O.traj_grad_f64(rb.cfg, rb.params, rb.x0[:, i], rb.u[:T, :, i], rb.goal[:, i], rb.obstacles)
        per = (time.perf_counter() - t0) / cpu_sample * rb.params.iterations
        out["cpu_baseline"] = {"value": round(1.0 / per, 2), "unit": "trajectories/s", "cores": 1, "kind": "port",
                               "sample": f"FP64 oracle cost+adjoint of {cpu_sample} trajectories, x50 steps"}
    return out


def parse_launch(s):
    if not s:
        return None
    blk, lanes, grd, chk = (int(x) for x in s.split(","))
    return {"propagate_block": blk, "lanes_per_child": lanes, "propagate_grid": grd, "chunk_grid": chk}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed passes over the seed set")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes over the seed set")
    ap.add_argument("--workload", default="dubins5d", choices=sorted(WORKLOAD_DESC))
    ap.add_argument("--seeds", type=int, default=0, help="queries per pass, seeds 0..n-1 (default: 200; 50 for di6d)")
    ap.add_argument("--batch", type=int, default=512, help="mixed batch size (0: skip)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain planner (batch or refine only)")
    ap.add_argument("--no-batch-seq", action="store_true",
                    help="skip the batch's one-by-one comparison through the plain planner")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (no autotuning)")
    ap.add_argument("--launch", default="",
                    help="plain planner launch shape BLOCK,LANES,GRID,CHUNKGRID (skips tuning)")
    ap.add_argument("--batch-launch", default="", help="batched planner launch shape (skips tuning)")
    ap.add_argument("--refine", type=int, default=4096, help="trajectories of the adjoint refinement (0: skip)")
    ap.add_argument("--cpu-queries", type=int, default=32, help="CPU sample: seeds 0..n-1 of the same set")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 256 MiB three times (scales the FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args(argv)

    if a.copy_calibration:
        import torch
        n = 64 * 1024 * 1024  # floats: 256 MiB
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 4, "launches": 3}), flush=True)
        return 0

    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1",
                          "n_gpus_requested": a.gpus}), flush=True)
        return 2

    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        return 1

    line = {
        "metric": "planning queries/sec + median time-to-solution, batched vs plain, 1xMI355X",
        "value": None,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
    }
    n_seeds = a.seeds or DEFAULT_SEEDS[a.workload]
    if not a.no_plain:
        plain_launch = parse_launch(a.launch)
        if plain_launch is None and not a.no_tune:
            from kgmt_amd import autotune
            from kgmt_amd import workloads as W
            tq = [W.make_query(a.workload, s) for s in range(1000, 1008)]
            plain_launch = autotune.tune_plain(tq, reps=2, cache=None)["launch"]
        print(f"bench: plain {a.workload}, {n_seeds} seeds x {a.warmup}+{a.steps} passes", file=sys.stderr,
              flush=True)
        cfg, passes, dt, prof, breakdown = run_plain(a.workload, n_seeds, a.steps, a.warmup, launch=plain_launch)
        res = passes[0]
        n_plans = sum(len(p) for p in passes)
        qps = n_plans / dt
        tts_solved = [r["time_ms"] for p in passes for r in p if r["success"]]
        tts_all = [r["time_ms"] for p in passes for r in p]
        succ = [r for r in res if r["success"]]

        # roofline of the dominant kernel (propagate)
        pk = prof["kernels"]["propagate"]
        avg_ms = pk["ms"] / max(pk["launches"], 1)
        per_launch = prof["children"] / max(pk["launches"], 1)
        b_child = algorithmic_bytes_per_child(cfg.model)
        achieved = b_child * prof["children"] / (pk["ms"] * 1e-3) / 1e9
        moved = moved_bytes(cfg.model, prof["children"], prof["flagged"])
        achieved_moved = moved / (pk["ms"] * 1e-3) / 1e9
        tr = measured_traffic(a.workload)
        valu = None
        if tr and "valu_insts_per_child" in tr:
            # issue rate of the same launches: VALU wave instructions per child
            # (rocprofv3 SQ_INSTS_VALU) x 64 lanes x children per launch
            v_ach = tr["valu_insts_per_child"] * 64 * per_launch / (avg_ms * 1e-3) / 1e12
            valu = {"achieved": round(v_ach, 2), "peak": round(VALU_PEAK_TLANEOPS, 2), "unit": "T lane-ops/s",
                    "frac": round(v_ach / VALU_PEAK_TLANEOPS, 4),
                    "source": "profiles/k_propagate_traffic.json valu_insts_per_child (SQ_INSTS_VALU)"}
        roofline = {
            "kernel": "k_propagate",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": round(tr["hbm_bytes_per_child"] * per_launch) if tr else None,
            "traffic_source": "profiles/k_propagate_traffic.json (rocprofv3 FETCH_SIZE, WRITE_SIZE per child, "
                              "scaled by a 256 MiB copy)" if tr else None,
            "bytes_per_child": b_child,
            "children_per_launch": round(per_launch, 1),
            "avg_launch_us": round(avg_ms * 1e3, 2),
            "moved": {"bytes_per_child": round(moved / max(prof["children"], 1), 2),
                      "achieved": round(achieved_moved, 2), "frac": round(achieved_moved / HBM_PEAK_GBS, 5),
                      "flagged_fraction": round(prof["flagged"] / max(prof["children"], 1), 4),
                      "note": "bytes the kernel moves by design: rows are written for flagged children only"},
            "kernel_avg_us": breakdown,
            "valu": valu,
            "note": "achieved = algorithmic bytes per expanded child (parent idx, parent state, controls+steps, "
                    "child sample, child parent idx, flag, r1, r2) x children per launch / average launch; launch "
                    "durations from HIP events around every propagate launch while the seed set is planned once "
                    "more (identical plans); the kernel is bound by the VALU issue and latency of its RK4 steps "
                    "and box tests, not by HBM (see valu)",
        }
        line.update({
            "value": round(qps, 3),
            "ms_per_step": round(dt / a.steps * 1e3, 4),
            "data": f"synthetic: seeded random boxes; one step = one pass over seeds 0..{n_seeds - 1} "
                    f"({n_seeds} queries); {a.warmup} untimed + {a.steps} timed passes",
            "config": {
                "workload": WORKLOAD_DESC[a.workload],
                "seeds_per_step": n_seeds,
                "capacity": cfg.capacity,
                "boxes": BOXES[a.workload],
                "max_iter": cfg.max_iter,
                "parallelism": "single GPU",
                "launch": plain_launch or "default",
            },
            "median_tts_ms": round(statistics.median(tts_solved), 4) if tts_solved else None,
            "median_tts_note": "wall time of kgmt_plan, solved queries only, over every timed pass",
            "median_tts_ms_all": round(statistics.median(tts_all), 4),
            "success_rate": len(succ) / len(res),
            "mean_iterations": statistics.mean(r["iterations"] for r in res),
            "mean_path_cost": statistics.mean(r["path_cost"] for r in succ) if succ else None,
            "oracle_fixture": fixture_agreement(a.workload, res),
            "roofline": roofline,
        })

        if not a.no_cpu:
            threads = 16  # one GPU's share of the host: 16 cores
            nq = min(a.cpu_queries, n_seeds)
            stats, ms, cdt = run_cpu(a.workload, nq, threads)
            same = sum(1 for s, r in zip(stats, res) if (s["tree_size"], s["iterations"]) ==
                       (r["tree_size"], r["iterations"]))
            gpu_ms = sum(r["time_ms"] for r in res[:nq])
            line["cpu_baseline"] = {
                "value": round(nq / cdt, 4),
                "unit": "queries/s",
                "cores": threads,
                "cpu": cpu_model(),
                "kind": "port",
                "sample": f"seeds 0..{nq - 1} of the timed seed set, one query per OpenMP thread on {threads} "
                          f"threads, oracle built gcc -O3 -march=x86-64-v3; median "
                          f"{statistics.median(ms):.0f} ms/query",
                "gpu_same_sample_qps": round(nq / (gpu_ms * 1e-3), 2),
                "identical_outcomes": same,
            }

    if a.batch:
        print(f"bench: batch {a.batch}", file=sys.stderr, flush=True)
        line["batched"] = run_batch(a.batch, tune=not a.no_tune, launch=parse_launch(a.batch_launch),
                                    sequential=not a.no_batch_seq)
        if a.no_plain:
            line["value"] = round(line["batched"]["qps_batched"], 3)
            line["config"] = {"workload": "batched planner, " + line["batched"]["mix"], "parallelism": "single GPU"}
    if a.refine:
        print(f"bench: refine {a.refine}", file=sys.stderr, flush=True)
        line["refine"] = run_refine(a.refine, cpu_sample=0 if a.no_cpu else 64, objective="configured")
        line["refine_hover_referenced"] = run_refine(a.refine, cpu_sample=0, objective="hover")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
