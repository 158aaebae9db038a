"""Device diagnostics against the oracle: hydro_dr_shortprint's region and
material tables (bl_shortprint), check_kn_sod's L1 (bl_sod_l1), and the
per-step records the step writes on the device (bl_set_diagnostics /
bl_diagnostics).

Bar: minima and maxima exact; sums (a fixed two-level tree on the device, a
serial loop in the oracle and the reference) and L1 norms to rtol 1e-12.  The
L1 values the reference recorded for its sod_sm_{lag,eul}_small decks print
the same to the 8 digits they were recorded with.  At full size the records
are checked against what the mathematics fixes: a Lagrangian region's mass
identical in every record (element masses do not change and the tree is
fixed), total energy of Sod between reflecting walls conserved to rounding,
the region volumes summing to the domain's area, record times equal to the
running sum of the dt history, and the last record equal to what
bl_shortprint / bl_sod_l1 give for the final state.
"""
import numpy as np
import pytest

from bookleaf_f90_amd import problem as P
from tests import golden_util as G
from tests.helpers import small_problems

pytestmark = pytest.mark.gpu
PROBLEMS = small_problems()
L1_RTOL = 1e-12


def _pair(p):
    from bookleaf_f90_amd.device import BookLeaf
    from oracle.pyoracle import Oracle
    print("diag case", p.name, "nel", p.nel, flush=True)
    return BookLeaf(p), Oracle(p)


def _sod_frame(p):
    """(x0, scale) placing the Sod interface of a [0, L] tube at 50 on the
    reference's [0, 100] tube (sod() is self-similar in x/t)."""
    length = float(np.max(p.ndx) - np.min(p.ndx))
    return 0.5 * length, 100.0 / length


@pytest.mark.parametrize("runner", ["run_graph", "run"])
@pytest.mark.parametrize("name", ["sod_lag", "sod2mat_eul", "sedov_hg", "sod2mat_mixed_eul",
                                  "deck_sod_mm", "noh_div"])
def test_shortprint_matches_oracle(name, runner, oracle_lib):
    p = PROBLEMS[name]
    g, o = _pair(p)
    n = min(p.nsteps, 30)
    g.run(n, runner=runner)
    o.run(n)
    reg_g, mat_g = g.shortprint()
    reg_o, mat_o = o.shortprint()
    G.assert_table_matches(reg_g, reg_o, "region")
    G.assert_table_matches(mat_g, mat_o, "material")


@pytest.mark.parametrize("eulerian,expected", [(False, "5.19756949E-03"),
                                               (True, "6.74455780E-03")])
def test_sod_deck_l1(eulerian, expected, oracle_lib):
    """The reference's recorded L1(rho) of sod_sm_{lag,eul}_small at t=20."""
    p = P.sod_deck_small(eulerian=eulerian)
    g, o = _pair(p)
    dt, _, _ = g.run(5000, stop_at_time_end=True, runner="run_graph")
    assert len(dt) == 384
    o.run(5000, stop_at_time_end=True)
    lg, lo = g.sod_l1(), o.sod_l1()
    assert "%.8E" % lg[0] == expected
    np.testing.assert_allclose(lg, lo, rtol=L1_RTOL, atol=0.0)


def _oracle_records(p, nsteps, x0, scale, stop):
    from oracle.pyoracle import Oracle
    o = Oracle(p)
    recs = []
    for _ in range(nsteps):
        if len(o.run(1, stop_at_time_end=stop)[0]) == 0:
            break
        reg, _ = o.shortprint()
        recs.append((o.ts.nstep, o.ts.time, reg, o.sod_l1(x0=x0, scale=scale)))
    return recs


@pytest.mark.parametrize("name,runner", [("deck_sod_small", "run_graph"),
                                         ("sod_eul", "run_graph"),
                                         ("sod2mat_mixed_eul", "run_graph"),
                                         ("deck_sod_mm", "run_graph"),
                                         ("sod_lag", "run")])
def test_step_records_match_oracle(name, runner, oracle_lib):
    from bookleaf_f90_amd.device import BookLeaf
    p = PROBLEMS[name]
    print("diag records", name, runner, "nel", p.nel, flush=True)
    x0, scale = _sod_frame(p)
    stop = bool(p.stop_at_time_end)
    n = min(p.nsteps, 400)
    g = BookLeaf(p)
    g.set_diagnostics(True, x0=x0, scale=scale, maxrec=n + 8)
    # two batches: the records continue across calls
    g.run(n // 3, stop_at_time_end=stop, runner=runner)
    g.run(n - n // 3, stop_at_time_end=stop, runner=runner)
    d = g.diagnostics()
    ref = _oracle_records(p, n, x0, scale, stop)
    assert len(d["nstep"]) == len(ref) == g.ts.nstep
    for k, (nstep, time, reg, l1) in enumerate(ref):
        assert d["nstep"][k] == nstep
        assert d["time"][k] == time
        G.assert_table_matches(d["region"][k], reg, "record %d" % k)
        np.testing.assert_allclose(d["l1"][k], l1, rtol=L1_RTOL, atol=0.0)


def test_records_stop_at_capacity_and_reset(oracle_lib):
    from bookleaf_f90_amd.device import BookLeaf
    p = PROBLEMS["sod_lag"]
    g = BookLeaf(p)
    g.set_diagnostics(True, maxrec=10)
    g.run(25, runner="run_graph")
    d = g.diagnostics()
    assert list(d["nstep"]) == list(range(1, 11))
    g.set_diagnostics(True, maxrec=10)  # restart the record
    g.run(3, runner="run_graph")
    assert list(g.diagnostics()["nstep"]) == [26, 27, 28]
    g.set_diagnostics(False)
    g.run(2, runner="run_graph")
    assert g.ts.nstep == 30


def test_records_full_size_lagrangian():
    """Sod 2048x256 between reflecting walls, 60 steps."""
    from bookleaf_f90_amd.device import BookLeaf
    p = P.sod(nx=2048, ny=256, nsteps=60)
    print("diag full size", p.name, flush=True)
    g = BookLeaf(p)
    reg0, _ = g.shortprint()
    x0, scale = _sod_frame(p)
    g.set_diagnostics(True, x0=x0, scale=scale, maxrec=64)
    dt, _, _ = g.run(60, runner="run_graph")
    d = g.diagnostics()
    assert list(d["nstep"]) == list(range(1, 61))
    assert np.array_equal(d["time"], np.cumsum(dt))
    mass = d["region"][:, :, 1]
    assert np.array_equal(mass, np.broadcast_to(reg0[:, 1], mass.shape))
    energy = d["region"][:, :, 2].sum(axis=1) + d["region"][:, :, 3].sum(axis=1)
    e0 = reg0[:, 2].sum() + reg0[:, 3].sum()
    np.testing.assert_allclose(energy, e0, rtol=1e-12, atol=0.0)
    np.testing.assert_allclose(d["region"][:, :, 0].sum(axis=1), 1.0 * 0.1, rtol=1e-12)
    # the last record is the final state's table and L1
    reg, _ = g.shortprint()
    assert np.array_equal(d["region"][-1], reg)
    assert np.array_equal(d["l1"][-1], np.array(g.sod_l1(x0=x0, scale=scale)))
    assert np.all(np.isfinite(d["l1"])) and np.all(d["l1"] > 0.0)


def test_records_full_size_eulerian_mixed():
    """The two-material Eulerian Sod with mixed cells at 2048x256: the remap
    conserves mass and the mesh (hence every region's volume) is fixed."""
    from bookleaf_f90_amd.device import BookLeaf
    p = P.sod2mat_mixed(nsteps=20)
    print("diag full size", p.name, flush=True)
    g = BookLeaf(p)
    reg0, _ = g.shortprint()
    g.set_diagnostics(True, maxrec=32)
    g.run(20, runner="run_graph")
    d = g.diagnostics()
    assert len(d["nstep"]) == 20
    np.testing.assert_allclose(d["region"][:, :, 1].sum(axis=1), reg0[:, 1].sum(), rtol=1e-12)
    np.testing.assert_allclose(d["region"][:, :, 0], np.broadcast_to(reg0[:, 0], (20, p.nreg)),
                               rtol=1e-13)
