#!/usr/bin/env python3
"""Write the full-size fixtures tests/golden/full_*.npz from the reference's own build.

The benchmark configurations at their full sizes and lengths, run by
oracle/_ref/bookleaf_ref (the reference's Fortran compiled from its sources by
oracle/build_ref.py, driven by its own inf_solver_getdt / lagstep / alestep):

  c0_sod_200x20            Sod 200x20 on [0,1]x[0,0.1], Lagrangian, to t = 0.2
                           (stop at time_end), CFL 0.5, default viscosity
  c1_sedov_1024_kappa0.01  Sedov 1024^2 jittered (seed 12345), anti-hourglass
                           0.01, Lagrangian, 500 steps
  c2_sod2mat_mixed_2048x256  two-material Sod with mixed interface cells,
                           Eulerian ALE, 300 steps

Each fixture holds only what stays small at a million elements: the
reference's dt / idtel / dt-control / time sequence over the whole run,
SHA-256 digests of every dataAPI field at a few checkpoint steps (a dump
streams through the digest without being kept), hydro_dr_shortprint's tables
at the end, digests of the problem inputs (so a test can tell that the
problem generator still gives the reference's inputs), and, for c0 (4000
elements), the final density, energy, pressure and node fields in full.
Digests are over the canonical form of make_golden.py (-0.0 -> +0.0, one NaN).

usage: python tests/golden/make_fullsize.py [name ...]   (needs oracle/_ref)
"""
import json
import mmap
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bookleaf_f90_amd import abi  # noqa: E402
from bookleaf_f90_amd import problem as P  # noqa: E402
from tests.golden.make_golden import UNSET, digest  # noqa: E402

KAPPA = 0.01
INPUTS = ("ielnd", "ndx", "ndy", "indtype", "ielreg", "ielmat", "density", "energy", "ndu",
          "ndv")
FULL = ("eldensity", "elenergy", "elpressure", "elvolume", "ndx", "ndy", "ndu", "ndv")


def problems():
    """name -> (problem, [(run length, dump_every)], full final fields)"""
    return {
        "c0_sod_200x20": (P.sod(), [(None, 0)], True),
        "c1_sedov_1024_kappa0.01": (P.sedov(n=1024, nsteps=500, kappa=KAPPA),
                                    [(25, 0), (500, 250)], False),
        "c2_sod2mat_mixed_2048x256": (P.sod2mat_mixed(nx=2048, ny=256, nsteps=300),
                                      [(20, 0), (300, 150)], False),
    }


def input_digests(p):
    out = {k: digest(np.asarray(getattr(p, k))) for k in INPUTS}
    if p.ncp:
        for k in ("imxel", "imxfcp", "imxncp", "icpmat", "frvolume", "cpdensity", "cpenergy"):
            out[k] = digest(np.asarray(getattr(p, k)))
    return out


def scan_dumps(path, keep=()):
    """Stream ref_driver's dump file: returns (dumps, steps, shortprint),
    each dump a dict of label/nstep/time and {field: [count, digest]}, plus
    the arrays named in keep of the last dump."""
    dumps, steps, sp, kept = [], [], {}, {}
    with open(path, "rb") as f, mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ) as m:
        pos = 0

        def take(dtype, n=1):
            nonlocal pos
            a = np.frombuffer(m, dtype, n, pos)
            pos += a.nbytes
            return a

        while pos < len(m):
            marker = int(take(np.int32)[0])
            if marker == -9:
                break
            if marker == -1:
                nstep, idtel = (int(v) for v in take(np.int32, 2))
                dt, time = (float(v) for v in take(np.float64, 2))
                sdt = bytes(take(np.uint8, 8)).decode().strip()
                steps.append((nstep, idtel, dt, time, sdt))
                continue
            if marker == 2:
                nr, nm = (int(v) for v in take(np.int32, 2))
                sp["region"] = take(np.float64, 9 * nr).reshape(9, nr).T.copy()
                sp["material"] = take(np.float64, 9 * nm).reshape(9, nm).T.copy()
                continue
            assert marker == 1, marker
            label, nstep, idtel = (int(v) for v in take(np.int32, 3))
            dt, time, dts = (float(v) for v in take(np.float64, 3))
            d = dict(label=label, nstep=nstep, time=time, fields={})
            kept = {}
            while True:
                fid, kind = (int(v) for v in take(np.int32, 2))
                count = int(take(np.int64)[0])
                if fid == 0:
                    break
                a = take({1: np.float64, 2: np.int32, 3: np.uint8}[kind], count)
                name = abi.NAME[fid]
                if name not in UNSET:
                    d["fields"][name] = [count, digest(a)]
                if name in keep:
                    kept[name] = a.copy()
                del a
            dumps.append(d)
    return dumps, steps, sp, kept


def run_ref(p, nsteps, dump_every, workdir, keep=()):
    from oracle import refio
    fin, fout = os.path.join(workdir, "p.bin"), os.path.join(workdir, "dump.bin")
    refio.write_problem(p, fin, nsteps=nsteps, dump_every=dump_every)
    r = subprocess.run(["bash", "-c", 'ulimit -s unlimited; exec "$0" "$@"', refio.REF_EXE, fin,
                        fout, "0"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference run failed:\n" + r.stdout + r.stderr)
    out = scan_dumps(fout, keep)
    os.remove(fout)
    return out


def make(name, p, runs, full):
    meta = dict(name=name, nel=p.nel, nnd=p.nnd, ncp=p.ncp, config=p.config,
                stop_at_time_end=int(p.stop_at_time_end), inputs=input_digests(p),
                checkpoints={})
    arrays = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as wd:
        for nsteps, every in runs:
            dumps, steps, sp, kept = run_ref(p, nsteps, every, wd, FULL if full else ())
            for d in dumps:
                if d["label"] > 0:
                    meta["checkpoints"][str(d["nstep"])] = dict(time=d["time"],
                                                                fields=d["fields"])
            # the longest run's step records and final tables
            if "ref_dt" not in arrays or len(steps) > len(arrays["ref_dt"]):
                arrays["ref_dt"] = np.array([s[2] for s in steps])
                arrays["ref_idt"] = np.array([s[1] for s in steps], np.int32)
                arrays["ref_time"] = np.array([s[3] for s in steps])
                meta["dt_names"] = [s[4] for s in steps]
                arrays["final_shortprint_region"] = sp["region"]
                arrays["final_shortprint_material"] = sp["material"]
                arrays.update({"final_" + k: v for k, v in kept.items()})
    meta["nsteps"] = len(arrays["ref_dt"])
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    out = os.path.join(HERE, "full_" + name + ".npz")
    np.savez_compressed(out, **arrays)
    print("%-28s %4d steps, checkpoints %s, %d bytes" % (
        name, meta["nsteps"], sorted(int(k) for k in meta["checkpoints"]), os.path.getsize(out)),
        flush=True)


def main():
    from oracle import refio
    if not refio.available():
        sys.exit("oracle/_ref/bookleaf_ref is not built (python oracle/build_ref.py)")
    want = sys.argv[1:]
    for name, (p, runs, full) in problems().items():
        if want and name not in want:
            continue
        make(name, p, runs, full)


if __name__ == "__main__":
    main()
