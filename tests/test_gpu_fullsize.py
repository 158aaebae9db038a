"""The benchmark configurations at full size and full length on the GPU,
bit for bit against the reference's own build (the digests, dt sequences and
fields of tests/golden/make_fullsize.py, taken from oracle/_ref driven by the
reference's inf_solver_* loop).

The graph runner is driven through the call sequences bench.py makes: run
calls of different lengths (both node-set parities and both at-rest masks),
the autotuner between them (its trials write the viscosity fields and the
worklist, and must put the state back), and a single-step call before the
timed run.  The fused step does not write the phase scratch the reference
uses only inside a step, so those ids are left out of its comparisons; the
host-driven per-driver runner is compared on every field.
"""
import numpy as np
import pytest

from tests import fullsize as F
from tests.test_gpu_graph import ALE_SCRATCH, LAG_SCRATCH

pytestmark = pytest.mark.gpu

C0, C1, C2 = "c0_sod_200x20", "c1_sedov_1024_kappa0.01", "c2_sod2mat_mixed_2048x256"


def _device(name):
    from bookleaf_f90_amd.device import BookLeaf
    meta, a = F.load(name)
    p = F.problem(name)
    print("case", name, "nel", p.nel, flush=True)
    return meta, a, BookLeaf(p)


class _Runs:
    """Run calls on one handle, collecting the dt history."""

    def __init__(self, g, runner="run_graph", stop=False):
        self.g, self.runner, self.stop = g, runner, stop
        self.dt, self.idt, self.ctl = [], [], []

    def __call__(self, n):
        dt, idt, ctl = self.g.run(n, stop_at_time_end=self.stop, runner=self.runner)
        self.dt += list(dt)
        self.idt += list(idt)
        self.ctl += list(ctl)
        return len(dt)

    def check_history(self, meta, a):
        n = len(self.dt)
        assert np.array_equal(self.dt, a["ref_dt"][:n])
        assert np.array_equal(self.idt, a["ref_idt"][:n])
        assert [c.strip() for c in F.dt_names(self.ctl)] == meta["dt_names"][:n]
        assert self.g.ts.nstep == n and self.g.ts.time == a["ref_time"][n - 1]


def test_c0_sod_to_time_end_graph(oracle_lib):
    """c0 to t = 0.2 through bl_run_graph with stop_at_time_end, in three run
    calls and a tuning between them: the reference's 395 dts, every field, the
    final fields in full, the shortprint tables and the analytic L1."""
    from oracle.pyoracle import Oracle
    meta, a, g = _device(C0)
    r = _Runs(g, stop=True)
    assert r(100) == 100
    g.autotune()
    r(1)
    r(1000)
    r.check_history(meta, a)
    n = len(r.dt)
    assert n == meta["nsteps"]
    assert F.checkpoint_mismatches(meta, n, g.get, skip=LAG_SCRATCH) == []
    for k in ("eldensity", "elenergy", "elpressure", "elvolume", "ndx", "ndy", "ndu", "ndv"):
        assert np.array_equal(g.get(k), a["final_" + k]), k
    reg, mat = g.shortprint()
    for mine, ref in ((reg, a["final_shortprint_region"]), (mat, a["final_shortprint_material"])):
        assert np.allclose(mine, ref, rtol=1e-12, atol=0.0)
        assert np.array_equal(mine[:, 5:], ref[:, 5:])  # minima and maxima exact
    o = Oracle(F.problem(C0))
    for k in ("eldensity", "elenergy", "ndx", "ndy"):
        o.set(k, a["final_" + k])
    o.geometry_getgeometry()
    t = float(a["ref_time"][-1])
    l1g = g.sod_l1(time=t, x0=0.5, scale=100.0)
    l1o = o.sod_l1(time=t, x0=0.5, scale=100.0)
    assert l1g == pytest.approx(l1o, rel=1e-12)
    assert l1g[0] < 5.0e-3


def test_c0_sod_to_time_end_per_driver():
    """c0 through the host-driven per-driver runner (bl_run): every field,
    the phase scratch included."""
    meta, a, g = _device(C0)
    r = _Runs(g, runner="run", stop=True)
    r(1000)
    r.check_history(meta, a)
    assert F.checkpoint_mismatches(meta, len(r.dt), g.get) == []


def test_c1_sedov_bench_sequence(oracle_lib):
    """c1 (1024^2 Sedov, kappa 0.01) for all 500 steps in bench.py's call
    sequences: --warmup 5 --steps 20 (run 4, tune, run 1, run 20), then
    --warmup 10 from there (run 9, tune, run 1, run 215), then the rest in
    calls of 100, 1 and 149.  The reference's dt sequence and every field at
    steps 25, 250 and 500."""
    meta, a, g = _device(C1)
    r = _Runs(g)
    r(4)
    g.autotune()
    r(1)
    r(20)
    r.check_history(meta, a)
    assert F.checkpoint_mismatches(meta, 25, g.get, skip=LAG_SCRATCH) == []
    r(9)
    g.autotune()
    r(1)
    r(215)
    r.check_history(meta, a)
    assert F.checkpoint_mismatches(meta, 250, g.get, skip=LAG_SCRATCH) == []
    r(100)
    r(1)
    r(149)
    r.check_history(meta, a)
    assert len(r.dt) == 500
    assert F.checkpoint_mismatches(meta, 500, g.get, skip=LAG_SCRATCH) == []
    reg, _ = g.shortprint()
    ref = a["final_shortprint_region"]
    assert np.allclose(reg, ref, rtol=1e-12, atol=0.0)
    assert np.array_equal(reg[:, 5:], ref[:, 5:])


def test_c2_eulerian_mixed_graph():
    """c2 (2048x256 two-material Eulerian Sod, mixed cells) for all 300
    steps through bl_run_graph: calls of 20, 130, 1, a tuning, then 149."""
    meta, a, g = _device(C2)
    r = _Runs(g)
    r(20)
    r.check_history(meta, a)
    assert F.checkpoint_mismatches(meta, 20, g.get, skip=ALE_SCRATCH) == []
    r(130)
    r.check_history(meta, a)
    assert F.checkpoint_mismatches(meta, 150, g.get, skip=ALE_SCRATCH) == []
    r(1)
    g.autotune()
    r(149)
    r.check_history(meta, a)
    assert len(r.dt) == 300
    assert F.checkpoint_mismatches(meta, 300, g.get, skip=ALE_SCRATCH) == []


def test_c2_eulerian_mixed_per_driver():
    """The first 20 steps of c2 through the per-driver kernels (bl_run):
    every field at full size, the alestep scratch included."""
    meta, a, g = _device(C2)
    r = _Runs(g, runner="run")
    r(20)
    r.check_history(meta, a)
    assert F.checkpoint_mismatches(meta, 20, g.get) == []
