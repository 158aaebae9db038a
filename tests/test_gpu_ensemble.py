"""Ensembles on the device (bl_ensemble_create / bl_ensemble_run) against each
instance run alone: the generated mesh, tables and initial state against the
host set-up of the same instance (bl_init_mesh, bl_build_maps, inf_init's
drivers), and every instance's dt sequence and final state against the
oracle running that instance by itself, bit for bit."""
import numpy as np
import pytest

from bookleaf_f90_amd import abi
from bookleaf_f90_amd import problem as P
from tests.helpers import diff_fields
from tests.test_gpu_graph import LAG_SCRATCH

pytestmark = pytest.mark.gpu

INIT_FIELDS = ("ielnd", "ielel", "ielfc", "indtype", "ielreg", "ielmat", "ielsort1", "ielsort2",
               "ndx", "ndy", "ndu", "ndv", "eldensity", "elenergy", "elpressure", "elcs2",
               "elvolume", "elmass", "cnwt", "cnmass", "cnx", "cny", "a1", "a2", "a3", "b1",
               "b2", "b3")
STATE_FIELDS = ("eldensity", "elenergy", "elpressure", "elcs2", "elvolume", "elmass", "elvisc",
                "cnwt", "cnmass", "cnviscx", "cnviscy", "cnx", "cny", "a1", "a3", "b1", "b3",
                "ndx", "ndy", "ndu", "ndv", "rscratch23", "rscratch24")


# a larger first dt than the ensemble default, so that the small instances'
# own CFL limits take over from the first steps
CFG = dict(P.ENSEMBLE_CONFIG, dt_initial=1.0e-2)


def _small(seed=11, n=5):
    return P.ensemble_params(n, seed=seed, nmin=16, nmax=26)


def test_generated_ensemble_matches_host_setup():
    from bookleaf_f90_amd.device import BookLeaf, Ensemble
    params = _small()
    print("ensemble generation", [(q["kind"], q["nx"]) for q in params], flush=True)
    ens = Ensemble(params, config=CFG)
    full = {f: ens.get(f) for f in INIT_FIELDS}
    for k, q in enumerate(params):
        b = BookLeaf(P.ensemble_instance(q, config=CFG))
        for f in INIT_FIELDS:
            mine = ens.instance_field(f, k, full[f])
            ref = b.get(f)[:len(mine)]
            assert np.array_equal(mine, ref), (k, f)


@pytest.mark.parametrize("seed,tiles", [(11, True), (12, True), (11, False)])
def test_ensemble_run_matches_each_instance_alone(seed, tiles, oracle_lib):
    """The tile step (tiles within instances, each tile's dt its
    instance's) and the three-kernel step (bl_set_tiling off)."""
    from bookleaf_f90_amd.device import Ensemble
    from oracle.pyoracle import Oracle
    params = _small(seed=seed, n=6)
    print("ensemble run", tiles, [(q["kind"], q["nx"]) for q in params], flush=True)
    ens = Ensemble(params, config=CFG)
    ens.set_tiling(tiles)
    h1 = ens.run(15)
    h2 = ens.run(25)
    on, ntile, _, nent = ens.get_tiling()
    assert on == tiles and (ntile >= len(params) if tiles else True)
    dt, idt, ctl = (np.concatenate([a, b]) for a, b in zip(h1, h2))
    times, last_dt, nstep = ens.times()
    state = {f: ens.get(f) for f in STATE_FIELDS}
    for k, q in enumerate(params):
        o = Oracle(P.ensemble_instance(q, config=CFG))
        odt, oidt, octl = o.run(40)
        assert np.array_equal(dt[:, k], odt), k
        assert np.array_equal(idt[:, k], oidt), k
        assert np.array_equal(ctl[:, k], octl), k
        assert times[k] == o.ts.time and last_dt[k] == o.ts.dt and nstep[k] == 40
        mine = {f: ens.instance_field(f, k, state[f]) for f in STATE_FIELDS}
        assert diff_fields(mine, o.state(), [f for f in STATE_FIELDS
                                             if f not in LAG_SCRATCH]) == [], k


def test_ensemble_medium_spot_checks(oracle_lib):
    """64 instances of 64^2..128^2 elements; three instances checked against
    the oracle after 20 steps."""
    from bookleaf_f90_amd.device import Ensemble
    from oracle.pyoracle import Oracle
    params = P.ensemble_params(64, seed=3, nmin=64, nmax=128)
    print("ensemble medium nel", sum(q["nx"] * q["ny"] for q in params), flush=True)
    ens = Ensemble(params)
    dt, idt, ctl = ens.run(20)
    state = {f: ens.get(f) for f in ("eldensity", "elenergy", "ndx", "ndu")}
    for k in (0, 31, 63):
        o = Oracle(P.ensemble_instance(params[k]))
        odt, oidt, octl = o.run(20)
        assert np.array_equal(dt[:, k], odt) and np.array_equal(idt[:, k], oidt)
        for f in state:
            assert np.array_equal(ens.instance_field(f, k, state[f]),
                                  o.get(f)[:len(ens.instance_field(f, k, state[f]))]), (k, f)


def test_ensemble_argument_errors():
    from bookleaf_f90_amd.device import BookLeafError, Ensemble
    bad = [dict(_small()[0], nx=8, ny=8)]  # 64 elements < 256
    with pytest.raises(BookLeafError):
        Ensemble(bad)
    cfg = dict(P.ENSEMBLE_CONFIG, zexist=1)
    with pytest.raises(BookLeafError):
        Ensemble(_small(), config=cfg)


def _dt_sequences(params, nsteps):
    from oracle.pyoracle import Oracle
    return np.array([Oracle(P.ensemble_instance(q, config=CFG)).run(nsteps)[0] for q in params])


@pytest.mark.parametrize("fail_step", [3, 8])
def test_ensemble_dt_min_failure_is_deterministic(fail_step, oracle_lib):
    """An instance whose dt falls below dt_min (time_dr_advance.f90:122-127)
    stops the whole ensemble at that step, whatever order the per-instance
    selections run in: every instance's dt of the failing step is selected
    (its nstep and time advance, as time_dr_end advances them before its
    check), no instance's state advances, and the error names the smallest
    failing instance.  dt_min is placed, from the oracle's dt sequences,
    below every dt of the steps before fail_step and above at least one of
    that step's (two instances at step 3, one at step 8).  Two runs give the
    same bits, and each instance holds the oracle's state after
    fail_step - 1 steps of it alone (the viscosity fields aside: the failing
    step's limiter pass has written those of its moving elements; and the
    corrector corner-force scratch rscratch23/24, which on the tile step only
    a batch's planned last step writes)."""
    from bookleaf_f90_amd.device import BookLeafError, Ensemble
    from oracle.pyoracle import Oracle
    params = _small(seed=11, n=6)
    D = _dt_sequences(params, fail_step)
    s = fail_step - 1
    dt_min = D[:, :s].min()
    failing = [k for k in range(len(params)) if D[k, s] < dt_min]
    assert failing, "no instance falls below dt_min at this step"
    cfg = dict(CFG, dt_min=dt_min)
    fields = [f for f in STATE_FIELDS
              if f not in ("cnviscx", "cnviscy", "elvisc", "rscratch23", "rscratch24")]
    runs = []
    for rep in range(2):
        print("ensemble dt_min failure at step", fail_step, "rep", rep, flush=True)
        ens = Ensemble(params, config=cfg)
        with pytest.raises(BookLeafError) as ei:
            ens.run(20)
        assert "dt < dt_min (instance %d)" % failing[0] in str(ei.value)
        dt, idt, ctl = ens.last_history
        times, last_dt, nstep = ens.times()
        runs.append((dt, idt, ctl, times, nstep, {f: ens.get(f) for f in fields}))
    for a, b in zip(runs[0][:5], runs[1][:5]):
        assert np.array_equal(a, b)
    assert diff_fields(runs[0][5], runs[1][5]) == []
    dt, idt, ctl, times, nstep, state = runs[0]
    assert dt.shape[0] == fail_step
    assert np.all(nstep == fail_step)
    assert np.array_equal(dt[s], D[:, s])
    for k, q in enumerate(params):
        o = Oracle(P.ensemble_instance(q, config=CFG))
        odt, _, _ = o.run(s)
        assert np.array_equal(dt[:s, k], odt), k
        assert times[k] == o.ts.time + D[k, s], k
        mine = {f: ens.instance_field(f, k, state[f]) for f in fields}
        assert diff_fields(mine, o.state(), [f for f in fields if f not in LAG_SCRATCH]) == [], k
