"""Shared problems and comparison helpers for the parity tests."""
import numpy as np

from bookleaf_f90_amd import abi
from bookleaf_f90_amd import problem as P


def small_problems():
    """Small cases covering every branch on the hot path."""
    out = {}
    out["sod_lag"] = P.sod(nx=20, ny=4, nsteps=40)
    out["sod_eul"] = P.sod(nx=12, ny=5, eulerian=True, nsteps=40)
    out["sedov_hg"] = P.sedov(n=10, nsteps=40, kappa=0.01)
    p = P.sod(nx=16, ny=4, nsteps=30)
    p.pmeritreg[:] = 0.1
    p.__post_init__()
    out["sod_sp"] = p
    p = P.sod(nx=16, ny=4, nsteps=30)
    p.zmidlength[:] = 1
    out["sod_midlength"] = p
    out["sod2mat_eul"] = P.sod2mat(nx=32, ny=4, nsteps=30)
    out["sedov_jitter_eul"] = _sedov_eul(12)
    out["deck_sod_small"] = P.sod_deck_small()
    # dt controls: CFL-limited Sod, DIV- then CFL-limited Noh
    out["sod_cfl"] = P.sod(nx=50, ny=5, dt_initial=1.0e-3, nsteps=150)
    out["noh_div"] = P.noh(n=20, dt_initial=1.0e-2, nsteps=80)
    # mixed cells: the sod_mm_lag_small deck (Lagrangian, a column of 50/50
    # mixed cells) and a two-material Eulerian Sod with a mixed interface
    out["deck_sod_mm"] = P.sod_mm_deck_small()
    out["sod2mat_mixed_eul"] = P.sod2mat_mixed(nx=32, ny=6, nsteps=30)
    return out


def _sedov_eul(n):
    p = P.sedov(n=n, nsteps=30)
    p.config["zeul"] = 1
    p.__post_init__()
    p.name = "sedov_eul_%d" % n
    return p


# driver call sequence of one step (inf_solver_getdt / lagstep / alestep) with
# the labels of oracle/ref/ref_driver.f90's granular dumps
def granular_sequence(s):
    def pred():
        s.ts.zcorrector = 0
        s.ts.dts = 0.5 * s.ts.dt
        s.hydro_setpredictor()

    def corr():
        s.ts.zcorrector = 1
        s.ts.dts = s.ts.dt
        s.hydro_setcorrector()

    def scat():
        s.hydro_initacceleration()
        s.hydro_scatteracceleration()

    seq = [(100, s.getdt), (200, pred), (201, s.hydro_getforce), (202, s.hydro_getstate),
           (203, s.hydro_getenergy), (204, s.hydro_geteos), (300, corr),
           (301, s.hydro_getforce), (302, scat), (303, s.hydro_getacceleration),
           (304, s.hydro_applyacceleration), (305, s.hydro_getstate),
           (306, s.hydro_getenergy), (307, s.hydro_geteos)]
    if s.problem.config.get("zexist"):
        seq += [(400, s.ale_getmeshstatus), (401, s.ale_getfluxvolume), (402, s.ale_advect),
                (403, s.ale_hydroupdate)]
    return seq


# component scratch the reference allocates without initial values
# (data_reset_cpquant, data.f90:369-377): only the entries a driver has
# written are defined
CP_SCRATCH_UNSET = {"icpscratch11", "icpscratch12", "rcpscratch23", "rcpscratch24"}
CP_SCRATCH_COUNT = {"rcpscratch11": 1, "rcpscratch21": 4, "rcpscratch22": 4}


def diff_fields(a, b, names=None, ncp=None, skip=()):
    """Names of fields whose contents are not bit-identical (NaN == NaN, and
    -0.0 == +0.0 as IEEE equality).  With ncp given (mixed cells), the
    component scratch compares only the entries of the ncp components."""
    bad = []
    for n in names or abi.FIELDS:
        if n not in a or n not in b or n in skip:
            continue
        if ncp is not None and n in CP_SCRATCH_UNSET:
            continue
        x, y = np.asarray(a[n]), np.asarray(b[n])
        m = min(len(x), len(y))
        if ncp is not None and n in CP_SCRATCH_COUNT:
            m = min(m, CP_SCRATCH_COUNT[n] * ncp)
        if not np.array_equal(x[:m], y[:m], equal_nan=(x.dtype.kind == "f")):
            bad.append((n, int(np.sum(x[:m] != y[:m]))))
    return bad
