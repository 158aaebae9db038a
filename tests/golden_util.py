"""Load the golden fixtures of tests/golden/ (written from the reference's own
build by tests/golden/make_golden.py)."""
import glob
import hashlib
import json
import os

import numpy as np

from bookleaf_f90_amd import abi
from bookleaf_f90_amd.problem import Problem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INPUTS = ("ielnd", "ndx", "ndy", "indtype", "ielreg", "ielmat", "density", "energy", "ndu",
          "ndv", "eos_itype", "eos_param", "kappareg", "pmeritreg", "zdtnotreg", "zmidlength")
MIXED = ("imxel", "imxfcp", "imxncp", "icpmat", "icpnext", "icpprev", "frvolume", "cpdensity",
         "cpenergy")


def names():
    """The step records (make_golden.py); meshgen_*.npz are the mesh
    generator's (make_meshgen.py), full_*.npz the full-size runs
    (make_fullsize.py)."""
    return sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(f).startswith(("meshgen_", "full_")))


def load_meshgen(name):
    z = np.load(os.path.join(GOLDEN, "meshgen_%s.npz" % name))  # plain arrays, no pickle
    return {k: z[k] for k in z.files}


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))  # plain arrays, no pickle
    meta = json.loads(bytes(z["meta"]).decode())
    kw = {k: z["in_" + k].copy() for k in INPUTS}
    kw.update({k: z["in_" + k].copy() for k in MIXED if "in_" + k in z})
    p = Problem(name=name, nmat=meta["nmat"], nreg=meta["nreg"], config=dict(meta["config"]),
                nsteps=meta["nsteps"], stop_at_time_end=meta["stop_at_time_end"], **kw)
    return p, meta, z


def canonical(a):
    a = np.asarray(a)
    if a.dtype.kind == "f":
        a = np.where(a == 0.0, 0.0, a)
        a = np.where(np.isnan(a), np.nan, a)
    return np.ascontiguousarray(a)


def digest(a):
    return hashlib.sha256(canonical(a).tobytes()).hexdigest()[:32]


def digest_mismatches(state, record, skip=()):
    """Fields of state whose first len entries do not match the recorded digest."""
    bad = []
    for n, (count, d) in record.items():
        if n in skip or n not in state:
            continue
        if digest(np.asarray(state[n])[:count]) != d:
            bad.append(n)
    return bad


# hydro_dr_shortprint table columns: 0-4 sums, 5-8 minima/maxima
TABLE_SUM_RTOL = 1e-12


def assert_table_matches(got, ref, what=""):
    """A device shortprint table against a serial one: the minima and maxima
    exactly, the sums to TABLE_SUM_RTOL (a tree of positive terms in another
    order, |error| <= n eps sum)."""
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape, what
    assert np.array_equal(got[:, 5:], ref[:, 5:]), what
    np.testing.assert_allclose(got[:, :5], ref[:, :5], rtol=TABLE_SUM_RTOL, atol=0.0,
                               err_msg=what)
