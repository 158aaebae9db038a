"""A Fortran host (tests/fortran/fortran_host.f90) drives the library through
the iso_c_binding module bookleaf_f90_amd/fortran/bookleaf_hip_mod.f90: the
reference's solver sequence over the per-driver entry points, then the graph
runner.  Its dt sequence and final state must equal the oracle's bit for bit,
and the conservation tables it reads through bl_shortprint the oracle's to a
summation-order tolerance."""
import os
import subprocess

import numpy as np
import pytest

from bookleaf_f90_amd import problem as P
from bookleaf_f90_amd import build
from tests import golden_util as G

pytestmark = pytest.mark.gpu
EXE = os.path.join(build.FORTRAN_BUILD, "fortran_host")


def _read(path, nel, nnd, nreg, nmat):
    data = open(path, "rb").read()
    n = int(np.frombuffer(data, np.int32, 1)[0])
    pos = 4
    out = {}
    for name, dt, cnt in (("dt", np.float64, n), ("idt", np.int32, n), ("ctl", np.int32, n),
                          ("eldensity", np.float64, nel), ("elenergy", np.float64, nel),
                          ("elpressure", np.float64, nel), ("ndx", np.float64, nnd),
                          ("ndy", np.float64, nnd), ("ndu", np.float64, nnd),
                          ("ndv", np.float64, nnd)):
        out[name] = np.frombuffer(data, dt, cnt, pos)
        pos += out[name].nbytes
    for name, nf in (("region", nreg), ("material", nmat)):
        out[name] = np.frombuffer(data, np.float64, 9 * nf, pos).reshape(nf, 9)
        pos += out[name].nbytes
    assert pos == len(data)
    return out


@pytest.mark.parametrize("case", ["sod_lag", "sod2mat_eul"])
def test_fortran_host_matches_oracle(case, tmp_path, oracle_lib):
    from oracle import refio
    from oracle.pyoracle import Oracle
    if not os.path.exists(EXE):
        pytest.fail("fortran_host not built: run __graft_entry__.build()")
    if case == "sod_lag":
        p, nsteps, nhost = P.sod(nx=24, ny=4, nsteps=30), 30, 12
    else:
        p, nsteps, nhost = P.sod2mat(nx=32, ny=4, nsteps=20), 20, 7
    print("fortran host case", case, flush=True)
    fin, fout = str(tmp_path / "p.bin"), str(tmp_path / "out.bin")
    refio.write_problem(p, fin, nsteps=nsteps)
    r = subprocess.run([EXE, fin, fout, str(nhost)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f = _read(fout, p.nel, p.nnd, p.nreg, p.nmat)
    o = Oracle(p)
    dt, idt, ctl = o.run(nsteps)
    assert np.array_equal(f["dt"], dt)
    assert np.array_equal(f["idt"], idt)
    assert np.array_equal(f["ctl"], ctl)
    for k in ("eldensity", "elenergy", "elpressure", "ndx", "ndy", "ndu", "ndv"):
        assert np.array_equal(f[k], o.get(k)[:len(f[k])]), k
    reg, mat = o.shortprint()
    G.assert_table_matches(f["region"], reg, "region")
    G.assert_table_matches(f["material"], mat, "material")
