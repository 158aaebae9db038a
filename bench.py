#!/usr/bin/env python3
"""Element-steps/s of BookLeaf's FP64 Lagrangian step on one MI355X.

Workload: Sedov blast, 1024x1024 quads on [0,1.2]^2 with interior nodes
jittered by U(-0.1,0.1) cell widths (seed 12345), gamma=1.4, rho=1,
E=0.25 in the origin corner element, reflective walls (a 500-step run;
--steps K times K of them after --warmup W untimed ones).  Inputs are
resident in HBM before the timed region; dt never leaves the device inside
it (bl_run_graph: one captured hipGraph per step).

Prints one JSON line (see the keys below).  One GPU only: --gpus N>1 exits.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--runner", choices=["run", "run_graph"], default="run_graph")
    ap.add_argument("--n", type=int, default=1024, help="mesh is n x n")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=50)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the path is single-partition "
              "(see DESIGN.md, 'How the path would shard'); --gpus %d refused" % args.gpus)
        sys.exit(2)
    from bookleaf_f90_amd import problem as P
    from bookleaf_f90_amd.device import BookLeaf

    prob = P.sedov(n=args.n, nsteps=args.warmup + args.steps)
    g = BookLeaf(prob)
    nel = prob.nel
    print("# workload sedov %dx%d nel=%d nnd=%d" % (args.n, args.n, nel, prob.nnd),
          file=sys.stderr, flush=True)
    if args.warmup:
        g.run(args.warmup, runner=args.runner)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(args.steps, runner=args.runner)
    g.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    value = nel * args.steps / sec

    # per-kernel durations: HIP events on the library's stream around every
    # launch of a further profile region of the same run
    g.timing(True)
    g.timing_reset()
    g.run(args.profile_steps, runner="run")
    kern = {}
    for name in g.timing_names():
        ms, n = g.timing_query(name)
        if n:
            kern[name] = dict(avg_ms=ms, launches=n, total_ms=ms * n)
    g.timing(False)
    dom = max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None

    out = {
        "metric": "element-steps/s FP64 at 1M cells, %HBM peak; learned vs plain ALE at equal L1(rho)",
        "value": value,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "sedov_lagrangian_1024x1024_jitter0.1_seed12345_500steps",
                   "nel": nel, "nnd": prob.nnd, "runner": args.runner},
        "kernels": kern,
        "dominant_kernel": dom,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
