"""The oracle's node relaxation (the stencil of the mesh generator's SOR,
setup_mesh.F90:795-877) pinned three ways:

  - LIN1 and LIN2 against meshes the reference's own generator converged
    (tests/golden/meshgen_*.npz, written from oracle/_ref by
    make_meshgen.py): every interior node must be a fixed point of one
    sweep to the generator's tolerance, which also pins the logical frame
    (a transposed frame swaps LIN1 and LIN2 and leaves residuals of order
    the cell size);
  - EQUI, whose weights the reference cannot run (its weight check,
    setup_mesh.F90:648-657, halts on type 3), against what the mathematics
    fixes: the equipotential (Winslow) equations are solved by the inverse of
    a harmonic map, so Jacobi sweeps to convergence from the logical grid
    must approach the exact positions at second order.  A stencil with
    alpha and gamma swapped or beta of the wrong sign, planted in the
    oracle, converges elsewhere and fails the same test;
  - the ALE mesh-motion option on a state at rest: a conservative remap of a
    uniform state is the identity whatever the mesh motion.
"""
import numpy as np
import pytest

from bookleaf_f90_amd import abi
from bookleaf_f90_amd import problem as P
from tests import golden_util as G


def _oracle(p):
    from oracle.pyoracle import Oracle
    return Oracle(p)


def test_ring_frame_on_structured_mesh(oracle_lib):
    """k along element corner 1 -> 2 (x), l along corner 1 -> 4 (y); the
    table lists (lp,km) (lp,kk) (lp,kp) (ll,km) (ll,kp) (lm,km) (lm,kk)
    (lm,kp)."""
    nx, ny = 5, 4
    ielnd, ndx, ndy, indtype = P.structured_mesh(nx, ny, 0.0, 1.0, 0.0, 1.0)
    o = _oracle(P.mesh_problem("ring", ielnd, ndx, ndy, indtype))
    ring = o.relax_ring()
    nid = lambda i, j: i + j * (nx + 1)
    for j in range(ny + 1):
        for i in range(nx + 1):
            r = ring[nid(i, j)]
            if i in (0, nx) or j in (0, ny):
                assert np.all(r == -1)
                continue
            exp = [nid(i - 1, j + 1), nid(i, j + 1), nid(i + 1, j + 1), nid(i - 1, j),
                   nid(i + 1, j), nid(i - 1, j - 1), nid(i, j - 1), nid(i + 1, j - 1)]
            assert list(r) == exp


def _meshgen_problem(name):
    m = G.load_meshgen(name)
    return P.mesh_problem(name, m["ielnd"], m["ndx"], m["ndy"], m["indtype"]), float(m["tol"])


@pytest.mark.parametrize("name,rtype,other", [("lin1_annulus", abi.RELAX_LIN1, abi.RELAX_LIN2),
                                              ("lin2_annulus", abi.RELAX_LIN2, abi.RELAX_LIN1)])
def test_lin_stencil_fixed_point_of_reference_generator(name, rtype, other, oracle_lib):
    p, tol = _meshgen_problem(name)
    o = _oracle(p)
    x0, y0 = o.get("ndx")[:p.nnd], o.get("ndy")[:p.nnd]
    interior = p.indtype > 0
    assert interior.sum() == 7 * 11
    assert np.all(o.relax_ring()[interior, 0] >= 0)
    o.relax_mesh(rtype, 1, 1.0, 0.0)
    res = max(np.abs(o.get("ndx")[:p.nnd] - x0).max(), np.abs(o.get("ndy")[:p.nnd] - y0).max())
    assert res <= tol
    # the other direction's stencil is not satisfied by this mesh
    o.set("ndx", x0)
    o.set("ndy", y0)
    o.relax_mesh(other, 1, 1.0, 0.0)
    res = max(np.abs(o.get("ndx")[:p.nnd] - x0).max(), np.abs(o.get("ndy")[:p.nnd] - y0).max())
    assert res > 1.0e-3


def _harmonic_error(n, plant=0, sweeps=None):
    ielnd, ndx, ndy, indtype, ex, ey = P.harmonic_map(n, eps=0.25)
    o = _oracle(P.mesh_problem("harm%d" % n, ielnd, ndx, ndy, indtype))
    o.relax_plant(plant)
    o.relax_mesh(abi.RELAX_EQUI, sweeps or 12 * n * n, 1.0, 0.0)
    x, y = o.get("ndx")[:len(ndx)], o.get("ndy")[:len(ndx)]
    return max(np.abs(x - ex).max(), np.abs(y - ey).max())


def test_equi_converges_to_harmonic_map_at_second_order(oracle_lib):
    e = [_harmonic_error(n) for n in (8, 16, 32)]
    assert e[0] < 1e-4
    for a, b in zip(e, e[1:]):
        assert 3.0 < a / b < 5.0, e


@pytest.mark.parametrize("plant", [1, 2])
def test_harmonic_map_test_detects_planted_stencil_mistakes(plant, oracle_lib):
    e = [_harmonic_error(n, plant) for n in (8, 16)]
    assert not (e[0] < 1e-4 and 3.0 < e[0] / e[1] < 5.0), e


def test_equi_fixed_point_of_affine_grid(oracle_lib):
    """An affine image of a uniform grid solves every stencil exactly: one
    sweep moves no interior node beyond rounding."""
    ielnd, xi, eta, indtype = P.structured_mesh(9, 7, 0.0, 1.0, 0.0, 1.0)
    x, y = 2.0 * xi + 0.7 * eta + 1.0, -0.3 * xi + 1.5 * eta
    for rtype in (abi.RELAX_LIN1, abi.RELAX_LIN2, abi.RELAX_EQUI):
        o = _oracle(P.mesh_problem("affine", ielnd, x, y, indtype))
        o.relax_mesh(rtype, 1, 1.0, 0.0)
        assert np.abs(o.get("ndx")[:len(x)] - x).max() < 1e-14
        assert np.abs(o.get("ndy")[:len(x)] - y).max() < 1e-14


def test_tolerance_and_relaxation_factor(oracle_lib):
    """A node moves by om times its residual only when a residual exceeds
    tol (setup_mesh.F90:858-862)."""
    ielnd, ndx, ndy, indtype = P.structured_mesh(6, 6, 0.0, 1.0, 0.0, 1.0, jitter=0.2)
    p = P.mesh_problem("tol", ielnd, ndx, ndy, indtype)
    o = _oracle(p)
    o.relax_mesh(abi.RELAX_EQUI, 1, 1.0, 0.0)
    full = o.get("ndx")[:p.nnd] - ndx
    o = _oracle(p)
    o.relax_mesh(abi.RELAX_EQUI, 1, 0.5, 0.0)
    half = o.get("ndx")[:p.nnd] - ndx
    np.testing.assert_allclose(half, 0.5 * full, rtol=1e-12, atol=1e-17)
    o = _oracle(p)
    o.relax_mesh(abi.RELAX_EQUI, 1, 1.0, 1.0)
    assert np.array_equal(o.get("ndx")[:p.nnd], ndx)
    assert np.array_equal(o.get("ndy")[:p.nnd], ndy)


def test_relaxation_reduces_jitter(oracle_lib):
    ielnd, ndx, ndy, indtype = P.structured_mesh(16, 16, 0.0, 1.0, 0.0, 1.0, jitter=0.3)
    ux, uy = P.structured_mesh(16, 16, 0.0, 1.0, 0.0, 1.0)[1:3]
    o = _oracle(P.mesh_problem("jit", ielnd, ndx, ndy, indtype))
    dev = [np.abs(ndx - ux).max()]
    for _ in range(3):
        o.relax_mesh(abi.RELAX_EQUI, 40, 1.0, 0.0)
        dev.append(np.abs(o.get("ndx")[:len(ndx)] - ux).max())
    assert all(b < a for a, b in zip(dev, dev[1:]))


def test_ale_relaxation_keeps_uniform_state_at_rest(oracle_lib):
    """ALE with node relaxation on a jittered mesh: the mesh moves every step
    but the remap of a uniform state at rest leaves it uniform and at rest."""
    ielnd, ndx, ndy, indtype = P.structured_mesh(12, 10, 0.0, 1.2, 0.0, 1.0, jitter=0.25)
    p = P.mesh_problem("rest", ielnd, ndx, ndy, indtype, density=1.3, energy=2.0)
    P.with_relaxation(p, abi.RELAX_EQUI, nsweep=2, om=0.8)
    o = _oracle(p)
    o.run(5)
    nel = p.nel
    assert np.abs(o.get("ndx")[:p.nnd] - ndx).max() > 1e-3  # the mesh did move
    np.testing.assert_allclose(o.get("eldensity")[:nel], 1.3, rtol=1e-13)
    np.testing.assert_allclose(o.get("elenergy")[:nel], 2.0, rtol=1e-13)
    assert np.abs(o.get("ndu")[:p.nnd]).max() < 1e-13
    assert np.abs(o.get("ndv")[:p.nnd]).max() < 1e-13


def test_ale_relaxation_conserves_mass(oracle_lib):
    p = P.with_relaxation(P.sod2mat(nx=24, ny=6, nsteps=20, eulerian=False, jitter=0.2),
                          abi.RELAX_EQUI, nsweep=1, om=0.5)
    o = _oracle(p)
    reg0, _ = o.shortprint()
    o.run(20)
    reg, _ = o.shortprint()
    np.testing.assert_allclose(reg[:, 1].sum(), reg0[:, 1].sum(), rtol=1e-13)
    np.testing.assert_allclose(reg[:, 0].sum(), reg0[:, 0].sum(), rtol=1e-13)
