"""Host side of ensembles: the instance parameters of the c3-style ensemble
and the standalone problem of one instance, which must be the reference
decks' own builders (problem.sod, problem.sedov) for their states."""
import numpy as np

from bookleaf_f90_amd import abi
from bookleaf_f90_amd import problem as P


def test_ensemble_params_ranges_and_determinism():
    a = P.ensemble_params(64, seed=7)
    assert a == P.ensemble_params(64, seed=7)
    assert a != P.ensemble_params(64, seed=8)
    for q in a:
        assert 64 <= q["nx"] == q["ny"] <= 512
        assert 1.3 <= q["gamma"] <= 1.67
        if q["kind"] == abi.INSTANCE_SOD:
            assert 2.0 <= q["rho_l"] / q["rho_r"] <= 20.0
            pl = (q["gamma"] - 1.0) * q["rho_l"] * q["e_l"]
            pr = (q["gamma"] - 1.0) * q["rho_r"] * q["e_r"]
            assert 2.0 * (1 - 1e-12) <= pl / pr <= 100.0 * (1 + 1e-12)
    kinds = [q["kind"] for q in a]
    assert 0 < sum(kinds) < len(kinds)


def test_sod_instance_is_the_sod_builder():
    q = dict(nx=40, ny=40, kind=abi.INSTANCE_SOD, length=1.0, height=1.0, gamma=1.4,
             kappa=0.0, rho_l=1.0, e_l=2.5, rho_r=0.125, e_r=2.0)
    a = P.ensemble_instance(q)
    b = P.sod(nx=40, ny=40, length=1.0, height=1.0)
    for k in ("ielnd", "ndx", "ndy", "indtype", "ielreg", "density", "energy"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_sedov_instance_is_the_sedov_builder():
    q = dict(nx=32, ny=32, kind=abi.INSTANCE_SEDOV, length=1.2, height=1.2, gamma=1.4,
             kappa=0.01, rho_l=1.0, e_l=0.25, rho_r=1.0, e_r=0.0)
    a = P.ensemble_instance(q)
    b = P.sedov(n=32, jitter=0.0, kappa=0.01)
    for k in ("ielnd", "ndx", "ndy", "indtype", "density", "energy"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert a.energy[0] > 0.0 and np.count_nonzero(a.energy) == 1
