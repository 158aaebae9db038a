"""The drop-in under the reference's own solver, on the GPU.

oracle/_ref/bookleaf_ref_dropin is the reference's Fortran (its
inf_solver_getdt / lagstep / alestep and everything they call, unchanged)
with the hot-path driver modules replaced by bookleaf_f90_amd/fortran/dropin
(same module and procedure names; their bodies call libbookleaf_hip.so),
linked with the same host program as oracle/_ref/bookleaf_ref, the
reference's unswapped build (oracle/build_ref.py --dropin).  Both run the same
problem file; every dataAPI field after every driver call of step 1
(mode 1), and the dt sequence, periodic dumps of every field and the
shortprint tables over a run (mode 0) must be identical bit for bit.  The
swapped build keeps the fields on the GPU between driver calls and brings
them back only for its dumps.
"""
import os

import numpy as np
import pytest

from tests.helpers import small_problems, diff_fields

pytestmark = pytest.mark.gpu
PROBLEMS = small_problems()
# the component forces, unset in the reference until the first getforce
# (driver 201) writes them: both builds hold whatever their memory held
UNSET_BEFORE_FORCE = ("rcpscratch21", "rcpscratch22")


@pytest.fixture(scope="module")
def refio():
    from oracle import refio as r
    for exe in (r.REF_EXE, r.DROPIN_EXE):
        if not r.available(exe):
            pytest.skip(os.path.basename(exe) + " not built (needs the reference sources at "
                        "build time)")
    return r


def _runs(refio, p, **kw):
    print("case", p.name, kw, flush=True)
    ref = refio.run_reference(p, **kw)
    gpu = refio.run_reference(p, exe=refio.DROPIN_EXE, **kw)
    return ref, gpu


@pytest.mark.parametrize("name", sorted(PROBLEMS))
def test_dropin_every_driver_call(name, refio):
    p = PROBLEMS[name]
    (ref, _, _), (gpu, _, _) = _runs(refio, p, mode=1)
    assert [d["label"] for d in gpu] == [d["label"] for d in ref]
    for a, b in zip(gpu, ref):
        assert (a["nstep"], a["dt"], a["time"], a["idtel"]) == (b["nstep"], b["dt"], b["time"],
                                                              b["idtel"]), a["label"]
        skip = UNSET_BEFORE_FORCE if a["label"] < 201 else ()
        assert diff_fields(a, b, ncp=p.ncp, skip=skip) == [], "after driver %d" % a["label"]


@pytest.mark.parametrize("name", sorted(PROBLEMS))
def test_dropin_run(name, refio):
    p = PROBLEMS[name]
    n = min(p.nsteps, 60) if name != "deck_sod_small" else 384
    (ref, rs, _), (gpu, gs, _) = _runs(refio, p, mode=0, nsteps=n, dump_every=20)
    assert gs == rs  # (nstep, idtel, dt, time, dt control) of every step
    assert [d["label"] for d in gpu] == [d["label"] for d in ref]
    for a, b in zip(gpu, ref):
        skip = UNSET_BEFORE_FORCE if a["label"] == 0 else ()
        assert diff_fields(a, b, ncp=p.ncp, skip=skip) == [], "dump %d" % a["label"]
    for t in ("region", "material"):
        assert np.array_equal(gpu[-1]["shortprint"][t], ref[-1]["shortprint"][t])
