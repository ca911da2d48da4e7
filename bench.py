"""Benchmark of the KGMT expansion path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dubins5d]
                    [--batch 512] [--no-cpu]

A step is one planning query of the workload (5D Dubins airplane by default:
100 boxes, 2^18 tree nodes, 16^3/2^3 region grid, dt 0.02, 5..25 steps per
edge), seed = step index. Obstacles are device-resident before the timed
region (passed as device pointers, as the reference's plan() takes them);
start/goal are host arrays like the reference's h_initial/h_goal.

One JSON line on stdout: queries/s (value), median time-to-solution, the
propagate kernel's roofline figures measured with HIP events on the planner's
stream (around every propagate launch while the timed queries are planned a
second time: the event pairs add dispatch gaps, so the timed run has none),
the CPU oracle's throughput on a bounded sample of the same workload, and
(--batch N) the batched planner on N mixed queries against the same queries
run one by one through the plain planner.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
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


def propagate_bytes(model: int, children: int, flagged: int) -> int:
    """Algorithmic HBM bytes of the propagate kernel: per child the
    expansion-list entry 4 B, parent state 4S, region 4 B and next-frontier
    flag 1 B; per flagged child (the only ones the tree append reads) the
    child sample 4D and parent index 4 B. Controls come from the counter-based
    RNG in registers, so no control bytes are read."""
    from kgmt_amd.abi import dims
    S, C, D = dims(model)
    return children * (4 + 4 * S + 4 + 1) + flagged * (4 * D + 4)


def run_plain(name, steps, warmup, profile=1, launch=None):
    import torch
    from kgmt_amd import workloads as W
    from kgmt_amd.planner import KGMT

    qs = [W.make_query(name, s) for s in range(warmup + steps)]
    d_obs = [torch.from_numpy(q.obstacles.reshape(-1)).cuda() for q in qs]
    g = KGMT(qs[0].cfg)
    if launch:
        g.set_launch(**launch)
    for q, d in zip(qs[:warmup], d_obs[:warmup]):
        g.set_seed(q.cfg.seed)
        g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = []
    for q, d in zip(qs[warmup:], d_obs[warmup:]):
        g.set_seed(q.cfg.seed)
        r = g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles))
        res.append(r.as_dict())
    g.synchronize()  # the planner's own stream (torch's synchronize does not cover it)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # propagate launch durations: the same queries planned again with HIP
    # events on the planner's stream around every propagate launch (an event
    # pair adds ~15 us of dispatch gaps per iteration, so the timed run above
    # has none); the plans are identical, so are the launches
    prof = None
    if profile:
        g.set_profiling(1)
        for i, (q, d) in enumerate(zip(qs[warmup:], d_obs[warmup:])):
            g.set_seed(q.cfg.seed)
            r = g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles))
            if (r.tree_size, r.iterations) != (res[i]["tree_size"], res[i]["iterations"]):
                raise RuntimeError(f"query {i}: the profiled plan differs from the timed one")
        prof = g.profile()
    # per-kernel breakdown from an untimed pass over the first queries with
    # events around every kernel
    breakdown = None
    if profile:
        g.set_profiling(2)
        for q, d in zip(qs[warmup:warmup + min(steps, 10)], d_obs[warmup:warmup + min(steps, 10)]):
            g.set_seed(q.cfg.seed)
            g.plan(q.start, q.goal, int(d.data_ptr()), len(q.obstacles))
        p2 = g.profile()
        breakdown = {k: round(v["ms"] / max(v["launches"], 1) * 1e3, 2) for k, v in p2["kernels"].items()}
    return qs[0].cfg, res, dt, prof, breakdown


def measured_traffic(name):
    """HBM bytes per child of the propagate kernel measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (scaled by a 256 MiB copy) and committed
    with the command that produced them; None when not measured."""
    path = os.path.join(ROOT, "profiles", "k_propagate_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def run_cpu(name, n_queries, threads):
    from kgmt_amd import workloads as W
    from kgmt_amd.abi import KGMT_MAX_SAMPLE
    import numpy as np
    from oracle import oracle as O

    qs = [W.make_query(name, s) for s in range(n_queries)]
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    t0 = time.perf_counter()
    stats, ms = O.plan_many(cfgs, starts, goals, obs, offs, cnts, threads=threads)
    dt = time.perf_counter() - t0
    return stats, ms, dt


def run_batch(n, seed=0, tune=True):
    import numpy as np
    import torch
    from kgmt_amd import autotune
    from kgmt_amd import workloads as W
    from kgmt_amd.planner import KGMT, BatchKGMT

    qs = W.make_batch(n, seed=seed)
    cfgs, starts, goals, obs, offs, cnts = W.pack_batch(qs)
    tuned = None
    if tune:
        # launch shape chosen on a calibration batch of the same mix (other seed)
        cq = W.make_batch(n, seed=seed + 1000)
        tuned = autotune.tune_batch(*W.pack_batch(cq), use_graph=True, reps=2, cache=None)
    b = BatchKGMT(cfgs)
    if tuned:
        b.set_launch(**tuned["launch"])
    b.plan(starts, goals, obs, offs, cnts, use_graph=True)  # warm-up: graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_b, fin = b.plan(starts, goals, obs, offs, cnts, use_graph=True)
    torch.cuda.synchronize()
    dt_b = time.perf_counter() - t0
    # the same queries, one by one through the plain planner (contexts made
    # beforehand: allocation is not part of a plan)
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
    same = all(a.tree_size == c.tree_size and a.success == c.success and a.iterations == c.iterations
               for a, c in zip(res_b, res_p))
    succ_b = float(np.mean([r.success for r in res_b]))
    succ_p = float(np.mean([r.success for r in res_p]))
    tts_b = [float(f) for f, r in zip(fin, res_b) if r.success]
    tts_p = [float(r.time_ms) for r in res_p if r.success]
    del ctxs
    return {
        "queries": n,
        "mix": "50% 6D double integrator / 50% 12D quadcopter, capacity log-U[2^12,2^17], 10-500 boxes",
        "qps_batched": n / dt_b,
        "qps_plain_sequential": n / dt_p,
        "speedup": dt_p / dt_b,
        "median_tts_ms_batched": statistics.median(tts_b) if tts_b else None,
        "median_tts_ms_plain": statistics.median(tts_p) if tts_p else None,
        "success_rate_batched": succ_b,
        "success_rate_plain": succ_p,
        "identical_results": same,
        "hip_graph": True,
        "launch": tuned["launch"] if tuned else "default",
        "autotune": {"candidates": len(tuned["trials"]), "calibration": "same mix, seed + 1000",
                     "best_s": round(tuned["seconds"], 5),
                     "default_s": round(tuned["trials"][0]["seconds"], 5)} if tuned else None,
    }


def run_refine(n=4096, seed=0, cpu_sample=0):
    """Adam refinement of n quadcopter control sequences (T ~ U{20..80},
    dt 0.01, 50 steps, lr 1e-2) through kgmt_refine; the CPU figure is the
    oracle's FP64 cost+adjoint on cpu_sample trajectories, scaled to the
    50 steps, one thread."""
    import numpy as np
    import torch
    from kgmt_amd import adjoint as A
    from kgmt_amd import workloads as W

    rb = W.make_refine_batch(n, seed=seed)
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
    out = {
        "trajectories": n,
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dubins5d", choices=sorted(WORKLOAD_DESC))
    ap.add_argument("--batch", type=int, default=512, help="mixed batch size (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (no autotuning)")
    ap.add_argument("--launch", default="",
                    help="plain planner launch shape BLOCK,LANES,GRID,CHUNKGRID (skips tuning)")
    ap.add_argument("--refine", type=int, default=4096, help="trajectories of the adjoint refinement (0: skip)")
    ap.add_argument("--cpu-queries", type=int, default=0, help="CPU sample size (default: one per thread)")
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

    plain_launch = None
    if a.launch:
        blk, lanes, grd, chk = (int(x) for x in a.launch.split(","))
        plain_launch = {"propagate_block": blk, "lanes_per_child": lanes, "propagate_grid": grd, "chunk_grid": chk}
    elif not a.no_tune:
        from kgmt_amd import autotune
        from kgmt_amd import workloads as W
        tq = [W.make_query(a.workload, s) for s in range(1000, 1008)]
        plain_launch = autotune.tune_plain(tq, reps=2, cache=None)["launch"]
    cfg, res, dt, prof, breakdown = run_plain(a.workload, a.steps, a.warmup, launch=plain_launch)
    tts = [r["time_ms"] for r in res]
    succ = [r for r in res if r["success"]]
    qps = len(res) / dt

    # roofline of the dominant kernel (propagate)
    pk = prof["kernels"]["propagate"]
    bytes_total = propagate_bytes(cfg.model, prof["children"], prof["flagged"])
    B = bytes_total / max(prof["children"], 1)
    avg_ms = pk["ms"] / max(pk["launches"], 1)
    achieved = bytes_total / (pk["ms"] * 1e-3) / 1e9
    per_launch = prof["children"] / max(pk["launches"], 1)
    tr = measured_traffic(a.workload)
    traffic = round(tr["hbm_bytes_per_child"] * per_launch) if tr else None
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
        "traffic": traffic,
        "traffic_source": "profiles/k_propagate_traffic.json (rocprofv3 FETCH_SIZE x2.0, WRITE_SIZE x1.0, per child)"
        if tr else None,
        "bytes_per_child": round(B, 2),
        "flagged_fraction": round(prof["flagged"] / max(prof["children"], 1), 4),
        "children_per_launch": round(per_launch, 1),
        "avg_launch_us": round(avg_ms * 1e3, 2),
        "kernel_avg_us": breakdown,
        "valu": valu,
        "note": "avg_launch_us from HIP events around every propagate launch in a second pass over the timed queries (identical plans); achieved = algorithmic bytes (per child: list entry, parent state, region, flag; per flagged child: row and parent index) / avg launch; propagate+collide is bound by the VALU issue of its RK4 steps and box tests (~2 working waves per SIMD; see valu), not by HBM; HBM fraction reported as required",
    }

    cpu = None
    if not a.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        nq = a.cpu_queries or threads
        stats, ms, cdt = run_cpu(a.workload, nq, threads)
        cpu = {
            "value": round(nq / cdt, 4),
            "unit": "queries/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{nq} queries (seeds 0..{nq - 1}) of the same workload, one query per OpenMP thread, "
                      f"median {statistics.median(ms):.0f} ms/query single-thread",
        }

    batch = run_batch(a.batch, tune=not a.no_tune) if a.batch else None
    refine = run_refine(a.refine, cpu_sample=0 if a.no_cpu else 64) if a.refine else None

    line = {
        "metric": "planning queries/sec + median time-to-solution, batched vs plain, 1xMI355X",
        "value": round(qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / len(res) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random boxes; seeds 0..steps-1)",
        "config": {
            "workload": WORKLOAD_DESC[a.workload],
            "capacity": cfg.capacity,
            "boxes": 100 if a.workload == "dubins5d" else None,
            "max_iter": cfg.max_iter,
            "parallelism": "single GPU",
            "launch": plain_launch or "default",
        },
        "median_tts_ms": round(statistics.median(tts), 4),
        "success_rate": len(succ) / len(res),
        "mean_iterations": statistics.mean(r["iterations"] for r in res),
        "mean_path_cost": statistics.mean(r["path_cost"] for r in succ) if succ else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batch,
        "refine": refine,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
