"""Node relaxation on the device (bl_relax_mesh, and bl_set_relaxation as the
ALE step's mesh motion) against the oracle, bit for bit: the relaxed node
positions for every stencil type, the dt sequence and every persistent field
of ALE runs in both runners, at small and at full size (1024x512)."""
import numpy as np
import pytest

from bookleaf_f90_amd import abi
from bookleaf_f90_amd import problem as P
from tests import golden_util as G
from tests.helpers import diff_fields
from tests.test_gpu_graph import compared

pytestmark = pytest.mark.gpu


def _pair(p):
    from bookleaf_f90_amd.device import BookLeaf
    from oracle.pyoracle import Oracle
    print("relax case", p.name, "nel", p.nel, flush=True)
    return BookLeaf(p), Oracle(p)


def _jittered(nx, ny, jitter=0.3):
    ielnd, ndx, ndy, indtype = P.structured_mesh(nx, ny, 0.0, 1.0, 0.0, ny / nx, jitter=jitter)
    return P.mesh_problem("jit%dx%d" % (nx, ny), ielnd, ndx, ndy, indtype)


def _same_nodes(g, o, nnd):
    for k in ("ndx", "ndy"):
        assert np.array_equal(g.get(k)[:nnd], o.get(k)[:nnd]), k


@pytest.mark.parametrize("rtype", [abi.RELAX_LIN1, abi.RELAX_LIN2, abi.RELAX_EQUI])
@pytest.mark.parametrize("nsweep,om,tol", [(1, 1.0, 0.0), (3, 0.7, 0.0), (4, 1.0, 1e-3)])
def test_relax_mesh_bitexact(rtype, nsweep, om, tol, oracle_lib):
    p = _jittered(40, 24)
    g, o = _pair(p)
    g.relax_mesh(rtype, nsweep, om, tol)
    o.relax_mesh(rtype, nsweep, om, tol)
    _same_nodes(g, o, p.nnd)
    assert not np.array_equal(g.get("ndx")[:p.nnd], p.ndx)


@pytest.mark.parametrize("name", ["lin1_annulus", "lin2_annulus"])
def test_relax_mesh_on_reference_generator_meshes(name, oracle_lib):
    m = G.load_meshgen(name)
    p = P.mesh_problem(name, m["ielnd"], m["ndx"], m["ndy"], m["indtype"])
    g, o = _pair(p)
    g.relax_mesh(abi.RELAX_EQUI, 5, 1.0, 0.0)
    o.relax_mesh(abi.RELAX_EQUI, 5, 1.0, 0.0)
    _same_nodes(g, o, p.nnd)


def test_relax_mesh_full_size(oracle_lib):
    p = _jittered(1024, 512)
    g, o = _pair(p)
    g.relax_mesh(abi.RELAX_EQUI, 5, 1.0, 0.0)
    o.relax_mesh(abi.RELAX_EQUI, 5, 1.0, 0.0)
    _same_nodes(g, o, p.nnd)


def _ale_cases():
    return {
        "sod2mat_jit": P.with_relaxation(
            P.sod2mat(nx=32, ny=8, nsteps=25, eulerian=False, jitter=0.2),
            abi.RELAX_EQUI, nsweep=1, om=0.5),
        "sod2mat_mixed_jit": P.with_relaxation(
            P.sod2mat_mixed(nx=32, ny=8, nsteps=25, eulerian=False, jitter=0.2),
            abi.RELAX_EQUI, nsweep=2, om=0.6),
        "sedov_lin1": P.with_relaxation(P.sedov(n=14, nsteps=25), abi.RELAX_LIN1, nsweep=3,
                                        om=0.5, tol=1e-9),
    }


@pytest.mark.parametrize("runner", ["run_graph", "run"])
@pytest.mark.parametrize("name", sorted(_ale_cases()))
def test_ale_relaxation_bitexact(name, runner, oracle_lib):
    p = _ale_cases()[name]
    g, o = _pair(p)
    hg = g.run(p.nsteps, runner=runner)
    ho = o.run(p.nsteps)
    for a, b in zip(hg, ho):
        assert np.array_equal(a, b)
    assert (g.ts.time, g.ts.dt, g.ts.nstep) == (o.ts.time, o.ts.dt, o.ts.nstep)
    names = compared(p) if runner == "run_graph" else None
    assert diff_fields(g.state(), o.state(), names, ncp=p.ncp or None) == []
    # the mesh moved away from the Lagrangian one
    assert not np.array_equal(g.get("ndx")[:p.nnd], p.ndx)


def test_ale_relaxation_full_size(oracle_lib):
    """The two-material interface problem at 1024x512 (mixed cells, jittered
    mesh) with EQUI relaxation, 3 steps of the graph runner."""
    p = P.with_relaxation(P.sod2mat_mixed(nx=1024, ny=512, nsteps=3, eulerian=False,
                                          jitter=0.1), abi.RELAX_EQUI, nsweep=2, om=0.5)
    g, o = _pair(p)
    hg = g.run(3, runner="run_graph")
    ho = o.run(3)
    for a, b in zip(hg, ho):
        assert np.array_equal(a, b)
    assert diff_fields(g.state(), o.state(), compared(p), ncp=p.ncp) == []


def test_relaxation_argument_errors(oracle_lib):
    from bookleaf_f90_amd.device import BookLeaf, BookLeafError
    g = BookLeaf(_jittered(8, 8))
    for args in ((4, 1, 1.0, 0.0), (3, 0, 1.0, 0.0), (3, 1, 0.0, 0.0), (3, 1, 1.0, -1.0)):
        with pytest.raises(BookLeafError):
            g.set_relaxation(*args)
    with pytest.raises(BookLeafError):
        g.relax_mesh(abi.RELAX_OFF, 1, 1.0, 0.0)
