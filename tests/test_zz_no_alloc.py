"""bl_hydro_getdt_noalloc against bl_hydro_getdt, through the C-ABI on an MI355X.

The allocation-free entry point runs the same kernels as bl_hydro_getdt with
its partial minima/maxima in a library-wide device array instead of the
handle's buffers.  Bar: bit-identical -- the CFL and DIV controls (value,
element, kind) and every dataAPI field the call writes -- and the same
controls as the oracle's hydro_dr_getdt on the same state.  Two calls in a
row, the second on a smaller mesh than the first, so that the second reads
its partials from an array that still holds the first call's.
"""
import pytest

from bookleaf_f90_amd import problem as P
from tests.helpers import diff_fields

pytestmark = pytest.mark.gpu


def _cases():
    # (problem, steps before the call): a 128^2 Sedov mesh with hourglass
    # control (64 workgroups of partials), then a 32x6 two-material mesh with
    # mixed cells (one workgroup)
    return [(P.sedov(n=128, nsteps=30, kappa=0.01), 12),
            (P.sod2mat_mixed(nx=32, ny=6, nsteps=30), 8)]


def _controls(cfl, div):
    return (cfl.rdt, cfl.idt, cfl.sdt), (div.rdt, div.idt, div.sdt)


def test_noalloc_getdt_matches_getdt_twice_in_a_row(oracle_lib):
    from bookleaf_f90_amd.device import BookLeaf
    from oracle.pyoracle import Oracle
    sizes = []
    for p, steps in _cases():
        print("case", p.name, "nel", p.nel, "steps", steps, flush=True)
        a, b, o = BookLeaf(p), BookLeaf(p), Oracle(p)
        for s in (a, b, o):
            s.run(steps)
        assert diff_fields(a.state(), b.state()) == []
        ref = _controls(*a.hydro_getdt())
        got = _controls(*b.hydro_getdt_noalloc())
        assert got == ref
        assert got == _controls(*o.hydro_getdt())
        # the fields the call writes (corner velocities, edge differences,
        # viscosity, cs2 + viscosity, the CFL terms) are the same bits
        assert diff_fields(a.state(), b.state()) == []
        assert diff_fields(b.state(), o.state()) == []
        sizes.append(p.nel)
    assert sizes[1] < sizes[0]


def test_noalloc_getdt_repeated_on_one_handle(oracle_lib):
    """Consecutive steps driven with the allocation-free timestep give the
    oracle's dt sequence and state."""
    from bookleaf_f90_amd.device import BookLeaf
    from oracle.pyoracle import Oracle
    p = P.sedov(n=24, nsteps=20, kappa=0.01)
    g, o = BookLeaf(p), Oracle(p)
    for k in range(6):
        cg = g.hydro_getdt_noalloc()
        co = o.hydro_getdt()
        assert _controls(*cg) == _controls(*co), "step %d" % k
        g.time_select(*cg)
        o.time_select(*co)
        g.lagstep()
        o.lagstep()
        assert (g.ts.dt, g.ts.time, g.ts.idtel) == (o.ts.dt, o.ts.time, o.ts.idtel)
    assert diff_fields(g.state(), o.state()) == []



def test_noalloc_getdt_refuses_a_mesh_beyond_its_workspace():
    """A handle of BL_NOALLOC_MAX_EL + 256 elements (about 15 GB of fields,
    created through the C-ABI with nothing uploaded): the allocation-free
    timestep refuses it with the reason before any launch, and the handle
    stays usable (bl_destroy frees it)."""
    import ctypes as C
    from bookleaf_f90_amd import abi
    from bookleaf_f90_amd.device import load_library
    lib = load_library()
    n = abi.NOALLOC_MAX_EL + 256
    print("case noalloc limit, nel", n, flush=True)
    sz = abi.Sizes(nel=n, nel1=n, nel2=n, nnd=n, nnd1=n, nnd2=n, nsz=n, nmat=1, nreg=1)
    h = C.c_void_p()
    assert lib.bl_create(C.byref(sz), C.byref(h)) == abi.SUCCESS, lib.bl_last_error().decode()
    try:
        zd, zi = (C.c_double * 1)(0.0), (C.c_int32 * 1)(0)
        assert lib.bl_set_regions(h, 1, zd, zd, zi, zi) == abi.SUCCESS
        cfl, div, err = abi.Dt(), abi.Dt(), abi.Error()
        rc = lib.bl_hydro_getdt_noalloc(h, C.byref(cfl), C.byref(div), C.byref(err))
        assert rc == abi.FAILURE
        msg = lib.bl_last_error().decode()
        assert "exceed the workspace maximum" in msg and str(abi.NOALLOC_MAX_EL) in msg, msg
    finally:
        assert lib.bl_destroy(h) == abi.SUCCESS
