"""Problem files and dump reader for the reference build oracle/_ref/bookleaf_ref.

TEST INFRASTRUCTURE ONLY.  ``write_problem`` serialises a Problem in the
stream layout read by oracle/ref/ref_driver.f90; ``read_dumps`` parses the
dataAPI dumps it writes; ``run_reference`` runs the binary.
"""
import os
import subprocess
import tempfile

import numpy as np

from bookleaf_f90_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
REF_EXE = os.path.join(HERE, "_ref", "bookleaf_ref")
# the same program with the hot-path driver modules replaced by the drop-in
# modules of bookleaf_f90_amd/fortran/dropin (runs the step on the GPU)
DROPIN_EXE = os.path.join(HERE, "_ref", "bookleaf_ref_dropin")


def available(exe=REF_EXE):
    return os.path.exists(exe)


def build():
    subprocess.run(["python3", os.path.join(HERE, "build_ref.py")], check=True,
                   capture_output=True)
    return available()


def write_problem(p, path, nsteps=None, dump_every=0):
    c = p.config
    i4 = lambda *v: np.array(v, np.int32).tobytes()
    f8 = lambda *v: np.array(v, np.float64).tobytes()
    with open(path, "wb") as f:
        f.write(i4(p.nel, p.nnd, p.nmat, p.nreg, p.nsteps if nsteps is None else nsteps,
                   dump_every, p.stop_at_time_end, c["zeul"], c["adv_type"], c["zhg"],
                   c["zsp"]))
        f.write(f8(c["zcut"], c["zerocut"], c["dencut"], c["accut"], c["cvisc1"], c["cvisc2"],
                   c["cfl_sf"], c["div_sf"], c["ccut"], c["pcut"], c["time_start"],
                   c["time_end"], c["dt_g"], c["dt_min"], c["dt_max"], c["dt_initial"]))
        f.write(np.asarray(p.kappareg, np.float64).tobytes())
        f.write(np.asarray(p.pmeritreg, np.float64).tobytes())
        f.write(np.asarray(p.zdtnotreg, np.int32).tobytes())
        f.write(np.asarray(p.zmidlength, np.int32).tobytes())
        f.write(np.asarray(p.eos_itype[1:], np.int32).tobytes())
        f.write(np.asarray(p.eos_param[abi.NEOSPARAM:], np.float64).tobytes())
        for a, t in ((p.ielnd, np.int32), (p.ielreg, np.int32), (p.ielmat, np.int32),
                     (p.indtype, np.int32), (p.ndx, np.float64), (p.ndy, np.float64),
                     (p.ndu, np.float64), (p.ndv, np.float64), (p.density, np.float64),
                     (p.energy, np.float64)):
            f.write(np.asarray(a, t).tobytes())
        # mixed cells (flattened component lists)
        sz = p.sizes()
        f.write(i4(p.ncp, p.nmx, sz.mcp, sz.mmx))
        if p.ncp:
            for a, t in ((p.imxel, np.int32), (p.imxfcp, np.int32), (p.imxncp, np.int32),
                         (p.icpmat, np.int32), (p.icpnext, np.int32), (p.icpprev, np.int32),
                         (p.frvolume, np.float64), (p.cpdensity, np.float64),
                         (p.cpenergy, np.float64)):
                f.write(np.asarray(a, t).tobytes())


def read_dumps(path):
    """Returns (dumps, steps): dumps is a list of dicts with 'label', 'nstep',
    'idtel', 'dt', 'time', 'dts' and one array per dataAPI field name (the
    last dump of a full run also holds 'shortprint': region and material
    tables of shape (nflag, 9), columns in abi.FLAG_TOTALS order); steps is a list of (nstep, idtel, dt, time, sdt)."""
    data = open(path, "rb").read()
    pos = 0

    def take(dtype, n=1):
        nonlocal pos
        a = np.frombuffer(data, dtype, n, pos)
        pos += a.nbytes
        return a

    dumps, steps, shortprint = [], [], {}
    while pos < len(data):
        marker = int(take(np.int32)[0])
        if marker == -9:
            break
        if marker == -1:
            nstep, idtel = take(np.int32, 2)
            dt, time = take(np.float64, 2)
            sdt = bytes(take(np.uint8, 8)).decode()
            steps.append((int(nstep), int(idtel), float(dt), float(time), sdt))
            continue
        if marker == 2:
            # hydro_dr_shortprint's tables: 9 columns over the regions, then
            # over the materials, stored column by column
            nr, nm = (int(v) for v in take(np.int32, 2))
            shortprint["region"] = take(np.float64, 9 * nr).reshape(9, nr).T.copy()
            shortprint["material"] = take(np.float64, 9 * nm).reshape(9, nm).T.copy()
            continue
        assert marker == 1, marker
        label, nstep, idtel = (int(v) for v in take(np.int32, 3))
        dt, time, dts = (float(v) for v in take(np.float64, 3))
        d = dict(label=label, nstep=nstep, idtel=idtel, dt=dt, time=time, dts=dts)
        while True:
            fid, kind = (int(v) for v in take(np.int32, 2))
            count = int(take(np.int64)[0])
            if fid == 0:
                break
            dt_ = {1: np.float64, 2: np.int32, 3: np.uint8}[kind]
            d[abi.NAME[fid]] = take(dt_, count).copy()
        dumps.append(d)
    if shortprint and dumps:
        dumps[-1]["shortprint"] = shortprint
    return dumps, steps


def run_reference(p, mode=0, nsteps=None, dump_every=0, timeout=3600, exe=REF_EXE):
    """Run oracle/_ref/bookleaf_ref (or exe) on problem p; returns (dumps,
    steps, stdout)."""
    if not available(exe):
        raise RuntimeError(exe + " not built")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "p.bin"), os.path.join(d, "dump.bin")
        write_problem(p, fin, nsteps=nsteps, dump_every=dump_every)
        env = dict(os.environ)
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec \"$0\" \"$@\"", exe,
                            fin, fout, str(mode)], capture_output=True, text=True,
                           timeout=timeout, env=env)
        if r.returncode != 0:
            raise RuntimeError("reference run failed (%d):\n%s%s" % (r.returncode, r.stdout,
                                                                     r.stderr))
        dumps, steps = read_dumps(fout) if mode != 2 else ([], [])
        return dumps, steps, r.stdout
