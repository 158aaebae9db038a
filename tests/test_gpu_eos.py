"""The Tait and JWL equations of state (eos_kn_forms.f90:28-106) on the GPU.

These two forms call pow (Tait) and exp (JWL).  The oracle and the reference
build both call glibc's, and agree bit for bit (test_oracle_ref.py); the
device's pow and exp are other correctly-rounded-to-within-an-ulp
implementations, so the GPU is held to stated tolerances here rather than to
bit equality:

  EOS_RTOL, EOS_ATOL   one geteos over a sweep of states against the oracle's:
                       a few ulps of the largest term of each form
  RUN_DT_RTOL          each dt of 100 steps (Lagrangian and Eulerian, both
                       runners) against the oracle's
  RUN_NORM_RTOL        every field after those 100 steps, normwise:
                       max |gpu - oracle| <= RUN_NORM_RTOL * max |oracle|
                       (the two components of the velocity, and of the
                       viscous force, against their larger one: the flow is
                       along x, and the y components are roundoff about 0).
                       The ulp-level differences of every step's pressure
                       are carried through the flow; the viscosity, a
                       limited function of velocity differences, has the
                       largest relative error where it is itself ~1e-17, so
                       a normwise bound is the meaningful one
Every other part of the step stays bit-exact: the ideal-gas problems of the
other GPU tests cover it.
"""
import numpy as np
import pytest

from bookleaf_f90_amd import problem as P
from tests.test_gpu_graph import ALE_SCRATCH, LAG_SCRATCH

pytestmark = pytest.mark.gpu

EOS_RTOL, EOS_ATOL = 4.0e-15, 4.0e-15
RUN_DT_RTOL = 1.0e-11
RUN_NORM_RTOL = 1.0e-10
COMPARED = ("eldensity", "elenergy", "elpressure", "elcs2", "elvolume", "ndx", "ndy", "ndu",
            "ndv", "cnviscx", "cnviscy", "elvisc")


def _pair(p):
    from bookleaf_f90_amd.device import BookLeaf
    from oracle.pyoracle import Oracle
    print("case", p.name, "nel", p.nel, flush=True)
    return BookLeaf(p), Oracle(p)


def test_geteos_sweep_within_ulps(oracle_lib):
    """hydro_dr_geteos on both materials over densities 0.3..3 x rho0 and
    energies 0..4: the device's pressure and sound speed within a few ulps
    of the oracle's (glibc pow and exp)."""
    p = P.tait_jwl(nx=100, ny=40, nsteps=1)
    g, o = _pair(p)
    rng = np.random.default_rng(5)
    mat = p.ielmat
    rho0 = np.where(mat == 1, P.JWL_PARAM[5], P.TAIT_PARAM[2])
    rho = rho0 * rng.uniform(0.3, 3.0, p.nel)
    en = rng.uniform(0.0, 4.0, p.nel)
    for s in (g, o):
        s.set("eldensity", rho)
        s.set("elenergy", en)
        s.hydro_geteos()
    for f in ("elpressure", "elcs2"):
        a, b = g.get(f), o.get(f)
        np.testing.assert_allclose(a, b, rtol=EOS_RTOL, atol=EOS_ATOL, err_msg=f)
        print(f, "max |diff|", np.max(np.abs(a - b)), "bit-equal", np.mean(a == b), flush=True)


@pytest.mark.parametrize("eulerian", [False, True])
@pytest.mark.parametrize("runner", ["run_graph", "run"])
def test_tait_jwl_run_within_tolerance(eulerian, runner, oracle_lib):
    p = P.tait_jwl(eulerian=eulerian, nsteps=100)
    g, o = _pair(p)
    hg = g.run(100, runner=runner)
    ho = o.run(100)
    np.testing.assert_allclose(hg[0], ho[0], rtol=RUN_DT_RTOL, atol=0.0)
    assert np.array_equal(hg[2], ho[2])  # the same dt control every step
    skip = ALE_SCRATCH if eulerian else LAG_SCRATCH
    for f in COMPARED:
        if f in skip:
            continue
        a, b = g.get(f), o.get(f)
        scale = np.max(np.abs(b))
        for pair in (("ndu", "ndv"), ("cnviscx", "cnviscy")):
            if f in pair:
                scale = max(np.max(np.abs(o.get(pair[0]))), np.max(np.abs(o.get(pair[1]))))
        err = float(np.max(np.abs(a - b)) / scale)
        print("%-10s normwise relative difference %.3e" % (f, err), flush=True)
        assert err <= RUN_NORM_RTOL, f
    if not eulerian:
        assert np.array_equal(g.get("elmass"), o.get("elmass"))
