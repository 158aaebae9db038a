"""The full-size fixtures written by tests/golden/make_fullsize.py from the
reference's own build: loading, and comparing a state against a checkpoint's
digests."""
import json
import os

import numpy as np

from bookleaf_f90_amd import problem as P
from tests.golden.make_golden import digest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KAPPA = 0.01


def load(name):
    z = np.load(os.path.join(GOLDEN, "full_" + name + ".npz"))  # plain arrays, no pickle
    meta = json.loads(bytes(z["meta"]).decode())
    return meta, {k: z[k] for k in z.files if k != "meta"}


def problem(name):
    """The benchmark configuration a fixture was made from (the same builders
    make_fullsize.py calls)."""
    if name == "c0_sod_200x20":
        return P.sod()
    if name == "c1_sedov_1024_kappa0.01":
        return P.sedov(n=1024, nsteps=500, kappa=KAPPA)
    if name == "c2_sod2mat_mixed_2048x256":
        return P.sod2mat_mixed(nx=2048, ny=256, nsteps=300)
    raise KeyError(name)


def input_mismatches(meta, p):
    out = []
    for k, d in meta["inputs"].items():
        if digest(np.asarray(getattr(p, k))) != d:
            out.append(k)
    return out


# registered with one entry per component by the reference (data.f90:360-361),
# which zeroes and reads four per component; this registry holds (4, mcp):
# the reference's entries are the first ones
LONGER_HERE = {"cpviscx", "cpviscy"}


def checkpoint_mismatches(meta, nstep, get, skip=()):
    """Fields whose digest (or length) differs from the reference's at step
    nstep; get(name) returns the field's full array."""
    ck = meta["checkpoints"][str(nstep)]
    bad = []
    for name, (count, d) in sorted(ck["fields"].items()):
        if name in skip:
            continue
        a = np.asarray(get(name))
        if name in LONGER_HERE and len(a) > count:
            a = a[:count]
        if len(a) != count or digest(a) != d:
            bad.append(name)
    return bad


def dt_names(ctl):
    from bookleaf_f90_amd import abi
    return [abi.DT_NAMES[int(c)] for c in ctl]
