#!/usr/bin/env python3
"""Element-steps/s of BookLeaf's FP64 Lagrangian step on one MI355X.

Workload: Sedov blast, 1024x1024 quads on [0,1.2]^2, interior nodes jittered
by U(-0.1,0.1) cell widths (seed 12345), gamma=1.4, rho=1, E=0.25 in the
origin corner element, reflective walls, anti-hourglass coefficient 0.01
(kappaall of the reference's own Sedov deck, input/sedov_lag_small; without
it the deposit element is crushed by hourglass modes and dt falls below
dt_min before step 500), a 500-step run (W untimed warm-up steps, then K
timed ones).  Inputs are resident in HBM before the timed
region and dt never leaves the device inside it (bl_run_graph: one captured
hipGraph per step, dt selected on the device).

Reported next to the throughput:
  roofline      the dominant kernel's algorithmic HBM bytes per launch (the
                distinct dataAPI array elements it must read or write, see
                KERNEL_BYTES) over its average duration, timed with HIP events
                on the library's stream in a profile pass of the same run;
                traffic is the rocprofv3 FETCH_SIZE+WRITE_SIZE of that kernel
                from profiles/traffic_<workload>.json when present
  cpu_baseline  the reference's own Fortran (oracle/_ref, flang -O2, serial)
                timed on this host, one core, on the first --cpu-steps steps of
                the same workload ("port": the C oracle when _ref is absent)
  cpu_baseline_throughput  the same on every host core at once: --cpu-procs
                independent copies of that sample, their element-steps/s summed
  profile_pass.at_rest_fraction  the share of elements whose four nodes share
                one velocity (they skip the limiter pass) over the profile pass
  driver_boundary_GBps  the throughput expressed in the per-driver
                algorithmic bytes of the reference's call graph
                (1664 B/element + 272 B/node per Lagrangian step, +84 B/element
                with the anti-hourglass force): an equivalent rate, not a
                bandwidth, since the fused step does not move those bytes
  fused_floor   a Lagrangian step against the state-once floor (144 B/element
                + 68 B/node): the floor's time at the HBM peak over the step's
                time (step_frac), the step's algorithmic bytes and its measured
                counter traffic (profiles/traffic_<workload>.json) over it;
                null for a step with a remap, whose state the floor omits

--workload tgv runs a Taylor-Green vortex on the same 1024x1024 size with
every element in motion from the first step (no element skips the limiter),
beside the Sedov headline where the blast has reached a small share of the
mesh by step 500.

One GPU only: --gpus N with N > 1 exits.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
KAPPA = 0.01       # anti-hourglass coefficient (input/sedov_lag_small kappaall)
# algorithmic bytes per launch of each fused kernel: (bytes/element, bytes/node)
KERNEL_BYTES = {
    # every element: ielnd 16, cs2 8, the at-rest mask 0.25 (read and
    # written, one bit) | node x,y,u,v 32; plus ZEROED_BYTES per element
    # whose zero viscosity it writes.  The limiter's stencil (ielel, outer
    # nodes, neighbour rho and cs2) is read by f1_limiter for the elements on
    # F1's worklist only, so f1_limiter has no entry here
    "f1_getdt": (24.25, 32),
    # ielnd 16, ielmat 4, ielreg 4 (hourglass kappa), p, m, e 24, cnmass 32,
    # the at-rest mask 0.125; writes corner force x/y, weight, mass term 128
    # | node x,y,u,v 32; plus MOVING_BYTES (cnviscx/y) per element in motion
    "f2_force": (208.125, 32),
    # CSR corner ids 16, corner force x/y, weight, mass term 128 | ptr 4,
    # indtype 4, x,y,u,v 32, writes x,y,u,v 32, ubar,vbar 16
    "f3_accel": (144, 88),
    # ielnd 16, ielmat 4, m, e 16, corner force x/y 64; writes e, rho, p,
    # cs2 32 | x,y 16, ubar,vbar 16
    "f4_state": (132, 32),
    # ensemble F1: F1's bytes and the element instance id 2
    "e1_getdt": (26.25, 32),
    # remap (ALE step), per element | per node (the neighbours and their
    # faces come from one packed int32 per face, 16 B an element):
    # ielnd, packed ielel/ielfc 32, cnwt 32, rho 8; writes dv 8, face mass
    # flux 32 | x, y, mesh velocity 32
    "r1_volflux": (112, 32),
    # packed ielel/ielfc 16, face mass flux 32, V, m 16, cnmass 32, e 8;
    # writes v0, m0, m 24, face energy flux 32
    "r2_massflux": (160, 0),
    # ielnd, packed ielel/ielfc 32, energy flux 32, m, v0, m0 24, e r/w 16,
    # cnmass r/w 64, face mass flux 32, ielmat 4; writes corner m0, dm, mass
    # flux, m0 (corner-major) 128, rho, p, cs2 24 | x, y, mesh velocity 32
    "r3_energy_update": (356, 32),
    # packed ielel/ielfc 16, corner m0, dm 64, ielnd 16; writes corner fluxes
    # 64 | u, v 16 (a corner with zero mass flux reads only its flux: at rest
    # fewer bytes)
    "r4_momflux": (160, 16),
    "r5_nodes": (136, 124),
    # node relaxation sweeps: stencil table 32, x, y 16, writes 16 | the last
    # sweep also reads the Lagrangian x, y 16
    "relax_sweep": (0, 64),
    "relax_velocity": (0, 80),
    # learned relaxation: per relaxed node the relaxed-node list 4, stencil
    # table 32, x, y 16; writes mesh velocity 16, one-sweep displacement 8,
    # guard mark 4 (the neighbours' x, y come through the cache)
    "mlp_nodes": (0, 80),
    # the check: ielnd 16 | x, y, mesh velocity 32
    "learned_check": (16, 32),
    # the limiter pass: only the elements on F1's worklist (LIMITER_BYTES);
    # the ensemble's pass reads its instances' worklist segments the same way
    "f1_limiter": (0, 0),
    "e1_limiter": (0, 0),
}
# the limiter pass, per listed element: worklist entry 4, ielnd 16, ielel
# 16, outer-node table 32, own rho, cs2 16; writes cnviscx/y 64, elvisc 8
# (the neighbours' rho, cs2 and the far nodes come through the cache); and
# the x, y, u, v of the nodes it touches, 32 B/node in the proportion of
# elements listed
LIMITER_BYTES = (156, 32)
# data-dependent parts, from the run's own counters (bl_run_stats): F1 and
# E1 write cnviscx/y and elvisc (72 B) for an element at rest only when it
# moved in the step before (or on a call's first step); F2 reads cnviscx/y
# (64 B) only for an element in motion
ZEROED_BYTES = {"f1_getdt": 72, "e1_getdt": 72}
MOVING_BYTES = {"f2_force": 64}
# a Lagrangian step of pure cells (and every ensemble step) folds F1 (E1)
# into F4 (F1 then runs once per run call): F4 also reads u, v at x^{n+1}
# (16 B/node) and writes the next step's at-rest mask (1/8 B/element), and
# F2 writes the zeros instead of F1; E1's instance id read moves into F4
FOLD_BYTES = {"f4_state": (0.125, 16)}


# The tile step (a pure-cell Lagrangian step: F2, F3 and F4 per tile of
# elements and the ring of elements around their nodes, bl_tiles.hip),
# priced from bl_get_tiling's counts: nent entries (each element once in its
# own tile and once in the ring of every other tile holding one of its
# nodes).
#  lag_tile: per entry its element id and node slots 8, node ids 16; per
#    element once its at-rest byte 1, p, m, e 24, ielmat, ielreg 8, cnmass
#    32, the next at-rest byte written 1, and rho, e, p, cs2 written 32 |
#    per node once the tile's node list 4, x, y, u, v 32, indtype 4; x, y,
#    u, v, ubar, vbar written 48; per element in motion cnviscx/y 64; per
#    element whose zero viscosity it writes 72.  Ring entries and the nodes
#    a tile shares with others read their values a second time through the
#    cache: not counted.
#  f1_seg (F1 of a fresh run call's first step): per entry 8 and per element
#    ielnd 16, cs2 8, the at-rest byte written 1 | x, y, u, v 32.
TILE_KERNELS = ("lag_tile", "f1_seg")


def tile_bytes(k, nel, nnd, moving, zeroed, tiles):
    ntile, _, nent = tiles
    if k == "lag_tile":
        # (per node: its entry's corner-mask byte; per tile: its bounds)
        return 24 * nent + 98 * nel + 89 * nnd + 16 * ntile + 64 * moving + 72 * zeroed
    return 8 * nent + 25 * nel + 32 * nnd


# before a remap F4 also writes the corner weights and the volume at x^{n+1}
# (cnwt 32, elvolume 8), which the remap reads
REMAP_F4_BYTES = 40


def kernel_bytes(k, nel, nnd, moving, zeroed, fold=False, tiles=None, remap=False):
    """algorithmic bytes of one launch of kernel k; moving and zeroed are the
    per-step averages of bl_run_stats; tiles = (tiles, boundary elements,
    entries) of the tile step; remap: the step ends with the remap"""
    if k in TILE_KERNELS:
        return tile_bytes(k, nel, nnd, moving, zeroed, tiles)
    if k in ("f1_limiter", "e1_limiter"):
        return LIMITER_BYTES[0] * moving + LIMITER_BYTES[1] * nnd * min(1.0, moving / nel)
    be, bn = KERNEL_BYTES[k]
    if remap and k == "f4_state":
        be += REMAP_F4_BYTES
    if fold:
        fe, fn = FOLD_BYTES.get(k, (0, 0))
        be, bn = be + fe, bn + fn
        zb = 72 if k == "f2_force" else 0
    else:
        zb = ZEROED_BYTES.get(k, 0)
    return be * nel + bn * nnd + zb * zeroed + MOVING_BYTES.get(k, 0) * moving


# per Lagrangian step at the driver boundary: 1664 B/element + 272 B/node,
# +84 B/element with the anti-hourglass force on
DRIVER_BOUNDARY = (1664 + 84, 272)
# Eulerian: the Lagrangian step without the hourglass force plus the remap's
# 504 B/element + 108 B/node
DRIVER_BOUNDARY_EUL = (1664 + 504, 272 + 108)
# the fused floor: a Lagrangian step's element and node state read once and
# written once, 144 B/element + 68 B/node, with no intermediate in HBM
FUSED_FLOOR = (144, 68)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 490 for sedov, 190 for ensemble, 290 else)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--runner", choices=["run", "run_graph"], default="run_graph")
    ap.add_argument("--n", type=int, default=1024, help="mesh is n x n (sedov)")
    ap.add_argument("--workload", choices=["sedov", "tgv", "sod2mat", "relax", "ensemble",
                                           "learned"],
                    default="sedov",
                    help="sedov: Lagrangian 1024^2 Sedov (the headline); tgv: Lagrangian "
                         "1024^2 Taylor-Green vortex, every element in motion from the first "
                         "step; sod2mat: Eulerian "
                         "(ALE remap every step) two-material Sod on 2048x256 with mixed "
                         "interface cells; relax: two-material Sod 1024x512 with ALE node "
                         "relaxation; ensemble: 512 Sod/Sedov instances of 64^2..512^2 "
                         "stepped together; learned: the relax problem with the network's "
                         "node motion (trained here first) against 8 plain sweeps")
    ap.add_argument("--instances", type=int, default=512, help="ensemble size")
    ap.add_argument("--diagnostics", action="store_true",
                    help="record the region table and Sod L1 on the device every step")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="learned: skip the equal-time comparison with the plain relaxation")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape of the limiter pass instead of "
                         "timing the candidates (bl_autotune) before the last warm-up step")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="processes of the throughput CPU baseline (one per host core, at "
                         "most the host's core count); 0 skips it")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="launch the plain nel-double copy kernel (hydro_setpredictor) this "
                         "many times, to scale profiler byte counters")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"sedov": 490, "tgv": 490, "ensemble": 190}.get(args.workload, 290)
    return args


def cpu_baseline(prob, steps):
    from oracle import refio
    # the reference has no node-relaxation mesh motion: the C oracle runs it
    if refio.available() and not prob.relax:
        with tempfile.TemporaryDirectory() as d:
            fin, fout = os.path.join(d, "p.bin"), os.path.join(d, "o.bin")
            refio.write_problem(prob, fin, nsteps=steps)
            r = subprocess.run(["bash", "-c", 'ulimit -s unlimited; exec "$0" "$@"',
                                refio.REF_EXE, fin, fout, "2"], capture_output=True,
                               text=True, timeout=900)
            m = re.search(r"steps=\s*(\d+) seconds=\s*(\S+) elsteps_per_s=\s*(\S+)", r.stdout)
            if r.returncode == 0 and m:
                return {"value": float(m.group(3)), "unit": "element-steps/s", "cores": 1,
                        "kind": "reference",
                        "sample": "oracle/_ref (the reference's Fortran, flang -O2 "
                                  "-ffp-contract=off, serial) on steps 1-%d of the same "
                                  "%s run, wall clock of the step loop "
                                  "(%.2f s)" % (int(m.group(1)), prob.name, float(m.group(2)))}
    from oracle.pyoracle import Oracle
    o = Oracle(prob)
    t0 = time.perf_counter()
    o.run(steps)
    sec = time.perf_counter() - t0
    return {"value": prob.nel * steps / sec, "unit": "element-steps/s", "cores": 1,
            "kind": "port", "sample": "C oracle, steps 1-%d of the same run (%.2f s)"
            % (steps, sec)}


def cpu_baseline_throughput(prob, steps, nproc):
    """The reference's Fortran (oracle/_ref, serial) as nproc independent
    processes at once on nproc host cores, each running the same bounded
    sample of the workload: the sum of their own element-steps/s (the host's
    throughput on independent copies of the problem)."""
    from oracle import refio
    nproc = min(nproc, os.cpu_count() or 1)
    if nproc < 1 or not refio.available() or prob.relax:
        return None
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "p.bin")
        refio.write_problem(prob, fin, nsteps=steps)
        procs = [subprocess.Popen(["bash", "-c", 'ulimit -s unlimited; exec "$0" "$@"',
                                   refio.REF_EXE, fin, os.path.join(d, "o%d.bin" % k), "2"],
                                  stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
                 for k in range(nproc)]
        rates = []
        for pr in procs:
            out, _ = pr.communicate(timeout=900)
            m = re.search(r"elsteps_per_s=\s*(\S+)", out)
            if pr.returncode != 0 or not m:
                return None
            rates.append(float(m.group(1)))
    return {"value": sum(rates), "unit": "element-steps/s", "cores": nproc, "kind": "reference",
            "sample": "oracle/_ref (the reference's Fortran, flang -O2 -ffp-contract=off, "
                      "serial), %d processes at once on %d host cores, each on steps 1-%d of "
                      "the same %s run; the sum of their own element-steps/s (slowest %.3g, "
                      "fastest %.3g)" % (nproc, nproc, steps, prob.name, min(rates),
                                          max(rates))}


def cpu_baseline_ensemble(params, steps):
    """The reference's Fortran (oracle/_ref, serial) on the ensemble's first
    P instances, one process per instance, all P at once on P host cores;
    the throughput is the sum of the processes' own element-steps/s."""
    from bookleaf_f90_amd import problem as P
    from oracle import refio
    if not refio.available():
        return None
    nproc = min(16, os.cpu_count() or 1, len(params))
    with tempfile.TemporaryDirectory() as d:
        procs = []
        for k in range(nproc):
            fin, fout = os.path.join(d, "p%d.bin" % k), os.path.join(d, "o%d.bin" % k)
            refio.write_problem(P.ensemble_instance(params[k]), fin, nsteps=steps)
            procs.append(subprocess.Popen(["bash", "-c", 'ulimit -s unlimited; exec "$0" "$@"',
                                           refio.REF_EXE, fin, fout, "2"],
                                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                          text=True))
        rates, nels = [], 0
        for k, pr in enumerate(procs):
            out, _ = pr.communicate(timeout=900)
            m = re.search(r"elsteps_per_s=\s*(\S+)", out)
            if pr.returncode != 0 or not m:
                return None
            rates.append(float(m.group(1)))
            nels += params[k]["nx"] * params[k]["ny"]
    return {"value": sum(rates), "unit": "element-steps/s", "cores": nproc, "kind": "reference",
            "sample": "oracle/_ref (the reference's Fortran, flang -O2 -ffp-contract=off, "
                      "serial), %d processes at once, one per instance 1-%d of the ensemble "
                      "(%d elements together), steps 1-%d each; the sum of their own "
                      "element-steps/s" % (nproc, nproc, nels, steps)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library steps one mesh partition per "
              "GPU and has no exchange between partitions; --gpus %d refused" % args.gpus)
        sys.exit(2)
    from bookleaf_f90_amd import problem as P
    from bookleaf_f90_amd.device import BookLeaf, Ensemble

    prob = None
    if args.workload == "ensemble":
        params = P.ensemble_params(args.instances, seed=2025)
        wname = ("ensemble_%d_sod_sedov_instances_64^2..512^2_seed2025_%dsteps"
                 % (args.instances, args.warmup + args.steps))
    elif args.workload == "sedov":
        prob = P.sedov(n=args.n, nsteps=args.warmup + args.steps, kappa=KAPPA)
        wname = "sedov_lagrangian_%dx%d_jitter0.1_seed12345_kappa0.01_%dsteps" % (
            args.n, args.n, args.warmup + args.steps)
    elif args.workload == "tgv":
        prob = P.taylor_green(n=args.n, nsteps=args.warmup + args.steps)
        wname = "taylor_green_lagrangian_%dx%d_u1_p10_%dsteps" % (
            args.n, args.n, args.warmup + args.steps)
    elif args.workload == "sod2mat":
        prob = P.sod2mat_mixed(nx=2048, ny=256, nsteps=args.warmup + args.steps)
        wname = "sod2mat_eulerian_2048x256_mixed_interface_seed2024_%dsteps" % (
            args.warmup + args.steps)
    elif args.workload == "learned":
        from bookleaf_f90_amd import learned as L
        net, train_info = L.train()
        prob = L.relax_problem(args.warmup + args.steps)
        wname = ("sod2mat_ale_learned_equi_relaxation_1024x512_mixed_interface_jitter0.1_"
                 "seed2024_nsweep%d_om%g_%dsteps" % (L.RELAX[1], L.RELAX[2],
                                                     args.warmup + args.steps))
    else:
        from bookleaf_f90_amd import abi
        prob = P.with_relaxation(
            P.sod2mat_mixed(nx=1024, ny=512, nsteps=args.warmup + args.steps, eulerian=False,
                            jitter=0.1), abi.RELAX_EQUI, nsweep=2, om=0.5)
        wname = ("sod2mat_ale_equi_relaxation_1024x512_mixed_interface_jitter0.1_seed2024_"
                 "nsweep2_om0.5_%dsteps" % (args.warmup + args.steps))
    if prob is None:
        g = Ensemble(params)
        nel, nnd = g.nel, g.nnd
        run = g.run
    else:
        nel, nnd = prob.nel, prob.nnd
        g = BookLeaf(prob)
        if args.workload == "learned":
            g.relax_learned(net)
        run = lambda n: g.run(n, runner=args.runner)  # noqa: E731
    print("# workload %s nel=%d nnd=%d" % (wname, nel, nnd), file=sys.stderr, flush=True)
    if args.diagnostics:
        g.set_diagnostics(True, maxrec=args.warmup + args.steps + args.profile_steps)
    # the launch shape is tuned on the state the timed steps start from; the
    # last warm-up step then captures the step graphs with it
    tune = None
    counted = [0, 0, 0]  # steps, moving, zeroed over every run call

    def count(n):
        mv, zr = g.run_stats()
        counted[0] += n
        counted[1] += mv
        counted[2] += zr

    if args.warmup:
        run(args.warmup - 1)
        count(args.warmup - 1)
        if not args.no_autotune:
            tune = g.autotune(reps=5)
            print("# autotune %s" % tune, file=sys.stderr, flush=True)
        run(1)
        count(1)
    g.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    g.synchronize()
    sec = time.perf_counter() - t0
    count(args.steps)
    value = nel * args.steps / sec
    print("# timed %d steps in %.4f s" % (args.steps, sec), file=sys.stderr, flush=True)

    # per-kernel durations: HIP events around every launch on the library's
    # stream, over a further profile pass of the same run and kernels
    g.timing(True)
    g.timing_reset()
    run(args.profile_steps)
    kern = {}
    for name in g.timing_names():
        ms, n = g.timing_query(name)
        if n:
            kern[name] = {"avg_ms": ms, "launches": n, "total_ms": ms * n}
    g.timing(False)
    mv, zr = g.run_stats()
    count(args.profile_steps)
    moving, zeroed = mv / max(args.profile_steps, 1), zr / max(args.profile_steps, 1)
    # F1 folded into F4: F1 launched once per run call, not once per step
    f1name = "e1_getdt" if "e1_getdt" in kern else "f1_getdt"
    tiles = None
    if "lag_tile" in kern:
        on, ntile, nbnd, nent = g.get_tiling()
        tiles = (ntile, nbnd, nent)
    # (the tile step always folds F1 into it: f1_seg runs once per call)
    fold = tiles is not None or (f1name in kern and "f4_state" in kern and
                                 kern[f1name]["launches"] < kern["f4_state"]["launches"])
    priced = set(KERNEL_BYTES) | (set(TILE_KERNELS) if tiles else set())
    remap = "r1_volflux" in kern
    dom = max((k for k in kern if k in priced), key=lambda k: kern[k]["total_ms"],
              default=None)
    roof = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    traffic_file = tf if os.path.exists(tf) and args.n == 1024 else None
    if dom:
        alg = kernel_bytes(dom, nel, nnd, moving, zeroed, fold, tiles, remap)
        avg_s = kern[dom]["avg_ms"] * 1e-3
        traffic = None
        if traffic_file is not None:
            traffic = json.load(open(traffic_file)).get(dom)
        roof = {"kernel": dom, "bound": "hbm", "achieved": alg / avg_s / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": alg / avg_s / HBM_PEAK,
                "traffic": traffic, "algorithmic_bytes": alg, "avg_ms": kern[dom]["avg_ms"]}
        for k in kern:
            if k in priced:
                kern[k]["alg_GBps"] = (kernel_bytes(k, nel, nnd, moving, zeroed, fold, tiles, remap) /
                                       (kern[k]["avg_ms"] * 1e-3) / 1e9)
    for _ in range(args.calibrate_copy):
        g.hydro_setpredictor()  # k_copy_r: reads and writes 8*nel bytes
    g.synchronize()
    cpu_tp = None
    if args.no_cpu_baseline:
        cpu = None
    elif prob is None:
        cpu = cpu_baseline_ensemble(params, min(args.cpu_steps, 10))
    else:
        cpu = cpu_baseline(prob, args.cpu_steps)
        if args.cpu_procs > 0:
            cpu_tp = cpu_baseline_throughput(prob, args.cpu_steps, args.cpu_procs)
    dbe, dbn = (DRIVER_BOUNDARY if args.workload in ("sedov", "tgv", "ensemble")
                else DRIVER_BOUNDARY_EUL)
    step_bytes = dbe * nel + dbn * nnd
    # algorithmic bytes of the kernels one step launches (from the profile pass)
    fused_bytes = sum(kernel_bytes(k, nel, nnd, moving, zeroed, fold, tiles, remap) *
                      kern[k]["launches"] / max(args.profile_steps, 1)
                      for k in kern if k in priced)
    floor_bytes = FUSED_FLOOR[0] * nel + FUSED_FLOOR[1] * nnd
    # measured HBM bytes of one step (rocprofv3 counters, per launch) when the
    # traffic file matches this size
    step_traffic = None
    if traffic_file is not None:
        tdict = json.load(open(traffic_file))
        step_traffic = sum(tdict[k] * kern[k]["launches"] / max(args.profile_steps, 1)
                           for k in kern if k in tdict)
    out = {
        "metric": "element-steps/s FP64 at 1M cells, %HBM peak; learned vs plain ALE at "
                  "equal L1(rho)",
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
        "config": {"workload": wname, "nel": nel, "nnd": nnd, "runner": args.runner,
                   "diagnostics": bool(args.diagnostics),
                   "tile_step": (None if tiles is None else
                                 {"tiles": tiles[0], "entries": tiles[2]}),
                   "launch": tune if tune is not None else g.get_tuning()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        # the same reference on every host core at once (independent copies)
        "cpu_baseline_throughput": cpu_tp,
        "speedup_vs_cpu_throughput": (value / cpu_tp["value"]) if cpu_tp else None,
        "step_hbm_frac": fused_bytes / (sec / args.steps) / HBM_PEAK,
        # an equivalent rate, not a bandwidth: fusion does not move these bytes
        "driver_boundary_GBps": step_bytes / (sec / args.steps) / 1e9,
        # the Lagrangian step against the state-once floor: step_frac is the
        # floor's time at the HBM peak over the measured step time (a remap
        # step's own state is not in the floor: null there)
        "fused_floor": ({"bytes_per_step": floor_bytes,
                         "ms_at_peak": 1e3 * floor_bytes / HBM_PEAK,
                         "step_frac": floor_bytes / (sec / args.steps) / HBM_PEAK,
                         "algorithmic_over_floor": fused_bytes / floor_bytes,
                         "step_traffic_bytes": step_traffic,
                         "traffic_over_floor": (step_traffic / floor_bytes
                                                if step_traffic else None)}
                        if not remap else None),
        "kernels": kern,
        # elements in motion and zero viscosity writes per step (bl_run_stats),
        # over the profile pass and over every run call of this command
        "profile_pass": {"steps": args.profile_steps, "moving_per_step": moving,
                         "at_rest_fraction": 1.0 - moving / nel,
                         "zeroed_per_step": zeroed, "f1_folded_into_f4": fold},
        "all_calls": {"steps": counted[0], "moving_per_step": counted[1] / max(counted[0], 1),
                      "at_rest_fraction": 1.0 - counted[1] / max(counted[0], 1) / nel,
                      "zeroed_per_step": counted[2] / max(counted[0], 1)},
    }
    if args.workload == "learned":
        s_, gd, fb, rs = g.learned_stats(full=True)
        out["learned"] = {"training": train_info, "timed_run": {"guarded_steps": gd,
                                                              "fallback_steps": fb},
                          "equal_time": None if args.no_compare else L.compare(net),
                          # plain EQUI (om 0.25) at 1..8 sweeps and the network
                          # against the 8-sweep run: time and L1 of each
                          "pareto": None if args.no_compare else L.pareto(net)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
