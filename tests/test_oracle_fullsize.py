"""The CPU oracle against the reference's full-size fixtures (written by
tests/golden/make_fullsize.py from oracle/_ref, the reference's own Fortran
driven by its inf_solver_* loop).

The problem generators must still give the reference's inputs (digests), and
the oracle must reproduce the reference bit for bit: the whole of c0 (Sod
200x20 to t = 0.2, 395 steps), and the first checkpoint of c1 (Sedov 1024^2,
kappa 0.01, 25 steps) and c2 (two-material Eulerian Sod 2048x256 with mixed
cells, 20 steps).  The c0 density is also held to the analytic Sod solution
of the reference's check (check_kn_tests.f90:64-111) through the scaled L1.
"""
import numpy as np
import pytest

from tests import fullsize as F

CASES = ("c0_sod_200x20", "c1_sedov_1024_kappa0.01", "c2_sod2mat_mixed_2048x256")


@pytest.mark.parametrize("name", CASES)
def test_generators_give_the_reference_inputs(name):
    meta, _ = F.load(name)
    p = F.problem(name)
    assert (p.nel, p.nnd, p.ncp) == (meta["nel"], meta["nnd"], meta["ncp"])
    assert F.input_mismatches(meta, p) == []


def test_fixture_lengths():
    """c0 runs to t = 0.2 (time_end reached, stop_at_time_end), c1 500 steps,
    c2 300 steps; the fixtures hold every dt of the run."""
    m0, a0 = F.load("c0_sod_200x20")
    assert a0["ref_time"][-1] >= 0.2 > a0["ref_time"][-2]
    m1, a1 = F.load("c1_sedov_1024_kappa0.01")
    m2, a2 = F.load("c2_sod2mat_mixed_2048x256")
    assert len(a1["ref_dt"]) == 500 and len(a2["ref_dt"]) == 300
    assert sorted(int(k) for k in m1["checkpoints"]) == [25, 250, 500]
    assert sorted(int(k) for k in m2["checkpoints"]) == [20, 150, 300]
    # time is the running sum of dt, as time_dr_end forms it
    for a in (a0, a1, a2):
        assert np.array_equal(np.cumsum(a["ref_dt"]), a["ref_time"])


def test_c0_whole_run_bitexact(oracle_lib):
    from oracle.pyoracle import Oracle
    meta, a = F.load("c0_sod_200x20")
    o = Oracle(F.problem("c0_sod_200x20"))
    dt, idt, ctl = o.run(100000, stop_at_time_end=True)
    assert np.array_equal(dt, a["ref_dt"]) and np.array_equal(idt, a["ref_idt"])
    assert [c.strip() for c in F.dt_names(ctl)] == meta["dt_names"]
    n = len(dt)
    assert F.checkpoint_mismatches(meta, n, o.get) == []
    for k in ("eldensity", "elenergy", "elpressure", "ndx", "ndu"):
        assert np.array_equal(o.get(k), a["final_" + k]), k
    reg, mat = o.shortprint()
    assert np.array_equal(reg, a["final_shortprint_region"])
    assert np.array_equal(mat, a["final_shortprint_material"])


def test_c0_density_against_the_analytic_solution(oracle_lib):
    """check_kn_sod's volume-weighted L1 of the reference's final density
    against sod(100 (x - 0.5), 100 t), the analytic solution of the
    reference's check scaled to the unit tube.  The oracle's L1 of the
    reference's own final state (the fixture's fields put into the oracle)
    is small, and a wrong time or shift makes it several times larger."""
    from oracle.pyoracle import Oracle
    meta, a = F.load("c0_sod_200x20")
    o = Oracle(F.problem("c0_sod_200x20"))
    for k in ("eldensity", "elenergy", "ndx", "ndy"):
        o.set(k, a["final_" + k])
    o.geometry_getgeometry()  # corner coordinates and volumes of the final mesh
    assert np.array_equal(o.get("elvolume"), a["final_elvolume"])
    t = float(a["ref_time"][-1])
    l1, _ = o.sod_l1(time=t, x0=0.5, scale=100.0)
    assert 0.0 < l1 < 5.0e-3
    wrong_time, _ = o.sod_l1(time=0.5 * t, x0=0.5, scale=100.0)
    wrong_shift, _ = o.sod_l1(time=t, x0=0.45, scale=100.0)
    assert wrong_time > 3.0 * l1 and wrong_shift > 3.0 * l1


@pytest.mark.parametrize("name,nstep", [("c1_sedov_1024_kappa0.01", 25),
                                        ("c2_sod2mat_mixed_2048x256", 20)])
def test_first_checkpoint_bitexact(name, nstep, oracle_lib):
    from oracle.pyoracle import Oracle
    meta, a = F.load(name)
    o = Oracle(F.problem(name))
    dt, idt, ctl = o.run(nstep)
    assert np.array_equal(dt, a["ref_dt"][:nstep])
    assert np.array_equal(idt, a["ref_idt"][:nstep])
    assert F.checkpoint_mismatches(meta, nstep, o.get) == []
