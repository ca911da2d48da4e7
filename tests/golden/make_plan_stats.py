"""Writes tests/golden/plan_stats_<workload>.json: the CPU oracle's planning
outcome for every seed of the statistical parity sets (50 seeds of the 6D
double integrator, 200 of the 5D Dubins airplane, 200 of the 12D
quadcopter, 200 of the planar unicycle, and the 512 queries of the mixed
double-integrator / quadcopter batch workload, workloads.make_batch(512,
seed=0): capacities log-uniform in [2^12, 2^17], 10..500 boxes), one record
per seed (per query for the batch):

  [seed, iterations, tree_size, goal_index, path_rows, path_cost bits,
   path_length bits, expanded]

(float32 values as their uint32 bit patterns). The GPU planner is compared
against these records seed by seed and as distributions.

    python tests/golden/make_plan_stats.py [workload ...] [--threads N] [--chunk K]

Seeds are planned K at a time and the records so far are kept in
plan_stats_<workload>.partial.json, so an interrupted run resumes where it
stopped; the final file is written when every seed is done.
"""
import argparse
import json
import os
import struct
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEEDS = {"di6d": 50, "dubins5d": 200, "quad12d": 200, "unicycle2d": 200, "batch512": 512}


def queries(name, n):
    from kgmt_amd import workloads as W
    if name == "batch512":
        return W.make_batch(n, seed=0)
    return [W.make_query(name, s) for s in range(n)]


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=sorted(SEEDS))
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--chunk", type=int, default=0, help="seeds per oracle call (0: all at once)")
    a = ap.parse_args(argv)
    from kgmt_amd import workloads as W
    from oracle import oracle as O

    for name in a.workloads:
        n = SEEDS[name]
        qs = queries(name, n)
        partial = os.path.join(HERE, f"plan_stats_{name}.partial.json")
        rows = []
        if os.path.exists(partial):
            with open(partial) as f:
                rows = json.load(f)
        chunk = a.chunk or n
        t0 = time.perf_counter()
        while len(rows) < n:
            part = qs[len(rows):len(rows) + chunk]
            cfgs, starts, goals, obs, offs, cnts = W.pack_batch(part)
            stats, ms = O.plan_many(cfgs, starts, goals, obs, offs, cnts, threads=a.threads)
            base = len(rows)
            rows += [[base + s, st["iterations"], st["tree_size"], st["goal_index"], st["path_rows"],
                      f32_bits(st["path_cost"]), f32_bits(st["path_length"]), st["expanded"]]
                     for s, st in enumerate(stats)]
            with open(partial, "w") as f:
                json.dump(rows, f)
            print(f"{name}: {len(rows)}/{n} seeds, {time.perf_counter() - t0:.0f} s", flush=True)
        dt = time.perf_counter() - t0
        out = {"workload": name, "seeds": n, "max_iter": qs[0].cfg.max_iter,
               "capacity": qs[0].cfg.capacity if name != "batch512" else [q.cfg.capacity for q in qs],
               "fields": ["seed", "iterations", "tree_size", "goal_index", "path_rows", "path_cost_bits",
                          "path_length_bits", "expanded"],
               "records": rows}
        with open(os.path.join(HERE, f"plan_stats_{name}.json"), "w") as f:
            json.dump(out, f, separators=(",", ":"))
            f.write("\n")
        if os.path.exists(partial):
            os.remove(partial)
        ok = sum(1 for r in rows if r[3] >= 0)
        print(f"{name}: {n} seeds, {ok} solved, {dt:.1f} s on {a.threads} threads", flush=True)


if __name__ == "__main__":
    main()
