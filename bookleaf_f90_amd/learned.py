"""The learned node relaxation's standard workload and training recipe.

Problem: the two-material interface problem (problem.sod2mat_mixed) on
1024x512 quads, jittered by 0.1 cell widths, ALE on every step with EQUI
node relaxation of 8 sweeps of om 0.25 as the plain method the network
imitates.  Training: (stencil, displacement) pairs of that plain relaxation
from three 512x256 runs of the same problem with other interface seeds
(every 40th relaxed node per step, 120 steps each, up to 1e6 pairs), then
Adam on the device.  Comparison: the plain relaxation, the network, and one
plain sweep (the method at the network's cost, without the network), each
run from t = 0 to the same time; the error measure is the relative L1 of
density (volume-weighted) and energy (mass-weighted) against the plain run.
"""
import time

import numpy as np

from . import abi
from . import problem as P

RELAX = (abi.RELAX_EQUI, 8, 0.25)  # type, sweeps, om of the plain method
TRAIN_SEEDS = (1, 2, 3)
T_END = 0.015


def relax_problem(nsteps, seed=2024, nx=1024, ny=512, nsweep=None, tend=1.0e30):
    rtype, k, om = RELAX
    p = P.with_relaxation(P.sod2mat_mixed(nx=nx, ny=ny, nsteps=nsteps, eulerian=False,
                                          jitter=0.1, seed=seed), rtype,
                          nsweep=k if nsweep is None else nsweep, om=om)
    p.config["time_end"] = tend
    return p


def train(pairs=1_000_000, iters=1500, nx=512, ny=256, steps=120, stride=40, seed=1):
    """A network trained on the plain relaxation's pairs; returns (net,
    info)."""
    from .device import BookLeaf, RelaxNet
    t0 = time.perf_counter()
    net = RelaxNet(seed=seed, capacity=pairs)
    for s in TRAIN_SEEDS:
        g = BookLeaf(relax_problem(steps, seed=s, nx=nx, ny=ny))
        g.relax_collect(net, stride=stride)
        g.run(steps, runner="run_graph")
        g.relax_collect(None)
        g.close()
    t1 = time.perf_counter()
    h1 = net.train(iters, batch=8192, lr=1e-3, seed=7)
    h2 = net.train(iters, batch=8192, lr=2e-4, seed=8)
    t2 = time.perf_counter()
    return net, {"pairs": net.npairs(), "collect_s": t1 - t0, "train_s": t2 - t1,
                 "adam_steps": 2 * iters, "batch": 8192, "loss_first": float(h1[0]),
                 "loss_last": float(h2[-1])}


def l1(st, ref):
    """(density, energy) relative L1 against ref, volume- and
    mass-weighted"""
    out = []
    for k, wk in (("eldensity", "elvolume"), ("elenergy", "elmass")):
        w = ref[wk]
        out.append(float(np.sum(np.abs(st[k] - ref[k]) * w) / np.sum(np.abs(ref[k]) * w)))
    return out


def run_to(mode, net=None, tend=T_END, nx=1024, ny=512, nsweep=None, reps=1):
    """One run from t = 0 to tend: mode "plain" (the plain relaxation, with
    nsweep sweeps in place of the method's 8 when given), "learned" (net) or
    "one_sweep".  With reps > 1 the run is made reps times from the same
    start and "seconds" is the median wall time (every repeat gives the same
    bits).  Returns (final fields, info); info has "error" when the run
    stopped on one."""
    from .device import BookLeaf, BookLeafError
    p = relax_problem(1_000_000, nx=nx, ny=ny, tend=tend,
                      nsweep=1 if mode == "one_sweep" else nsweep)
    times = []
    for rep in range(reps):
        g = BookLeaf(p)
        if mode == "learned":
            g.relax_learned(net)
        info = {}
        try:
            g.run(1, stop_at_time_end=True, runner="run_graph")
            g.synchronize()
            t0 = time.perf_counter()
            while g.ts.time < tend:
                g.run(20, stop_at_time_end=True, runner="run_graph")
            g.synchronize()
            times.append(time.perf_counter() - t0)
        except BookLeafError as e:
            info["error"] = str(e)
        if "error" in info or rep == reps - 1:
            break
        g.close()
    if times:
        info["seconds"] = float(np.median(times))
        info["seconds_all"] = times
    info.update(steps=int(g.ts.nstep), time=g.ts.time)
    if mode == "learned":
        s_, gd, fb, rs = g.learned_stats(full=True)
        info.update(guarded_steps=gd, fallback_steps=fb, reasons=rs)
    st = None
    if "error" not in info:
        st = {k: g.get(k)[:p.nel] for k in ("eldensity", "elenergy", "elvolume", "elmass")}
    g.close()
    return st, info


def compare(net, tend=T_END, nx=1024, ny=512, untrained=None):
    """plain, learned, one sweep (and an untrained network when given) run to
    tend; each entry with its L1 against plain and its speed-up"""
    ref, plain = run_to("plain", tend=tend, nx=nx, ny=ny)
    out = {"t_end": tend, "plain": plain}
    runs = [("learned", net), ("one_sweep", None)]
    if untrained is not None:
        runs.append(("untrained", untrained))
    for name, m in runs:
        st, info = run_to("learned" if m is not None else name, m, tend=tend, nx=nx, ny=ny)
        if st is not None and ref is not None:
            info["L1_density"], info["L1_energy"] = l1(st, ref)
            info["speedup_vs_plain"] = plain["seconds"] / info["seconds"]
        out[name] = info
    return out


# the plain settings of the cost/accuracy curve: EQUI at the method's om with
# these sweep counts, and the accuracy bar for "as good as the method"
PARETO_SWEEPS = (1, 2, 3, 4, 6, 8)
PARETO_TOL = 1.0e-3


def pareto(net, sweeps=PARETO_SWEEPS, tend=T_END, nx=1024, ny=512, tol=PARETO_TOL, reps=3):
    """The plain relaxation at each sweep count of sweeps (om 0.25) and the
    network, each run from t = 0 to tend, with wall time (the median of reps
    runs) and L1 (density, energy) against the 8-sweep run, the method the
    network imitates.  Returns {"plain": [rows by sweep count], "learned":
    row, "cheapest_plain": the sweep count of the fastest plain setting
    within tol of the 8-sweep run in both L1s, "learned_within_tol",
    "learned_speedup_vs_cheapest"}."""
    ref, ref_info = run_to("plain", tend=tend, nx=nx, ny=ny, reps=reps)
    rows = []
    for k in sweeps:
        if k == RELAX[1]:
            st, info = ref, dict(ref_info)
        else:
            st, info = run_to("plain", tend=tend, nx=nx, ny=ny, nsweep=k, reps=reps)
        info["nsweep"] = k
        if st is not None and ref is not None:
            info["L1_density"], info["L1_energy"] = l1(st, ref)
        rows.append(info)
    st, lr = run_to("learned", net, tend=tend, nx=nx, ny=ny, reps=reps)
    if st is not None and ref is not None:
        lr["L1_density"], lr["L1_energy"] = l1(st, ref)
    ok = [r for r in rows if "error" not in r and r["L1_density"] <= tol and
          r["L1_energy"] <= tol]
    cheapest = min(ok, key=lambda r: r["seconds"]) if ok else None
    out = {"t_end": tend, "tol": tol, "plain": rows, "learned": lr,
           "cheapest_plain": cheapest["nsweep"] if cheapest else None,
           "learned_within_tol": ("error" not in lr and lr.get("L1_density", 1.0) <= tol and
                                  lr.get("L1_energy", 1.0) <= tol)}
    if cheapest and "error" not in lr:
        out["learned_speedup_vs_cheapest"] = cheapest["seconds"] / lr["seconds"]
    return out
