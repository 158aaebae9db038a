"""The tile step (bl_set_tiling): the predictor, node update and corrector of
a pure-cell Lagrangian step in one kernel per tile of nodes, the corner
terms kept in LDS, and a second kernel for the elements with nodes in two or
more tiles.  It must give the three-kernel step's bits (and so the
oracle's): every dt and every field, over run calls of several lengths
(both node and element sets, both at-rest byte masks), with the per-driver
entry points continuing from its state, and across switches between the two
step kinds (the element sets resynchronised)."""
import numpy as np
import pytest

from bookleaf_f90_amd import problem as P
from tests.helpers import small_problems, diff_fields
from tests.test_gpu_graph import compared

pytestmark = pytest.mark.gpu
SMALL = small_problems()


def _cases():
    out = {k: SMALL[k] for k in ("sod_lag", "sedov_hg", "sod_sp", "sod_midlength", "sod_cfl",
                                 "noh_div", "deck_sod_small")}
    out["sedov_96_jitter_hg"] = P.sedov(n=96, nsteps=40, kappa=0.01)
    out["taylor_green_80"] = P.taylor_green(n=80, nsteps=40)
    return out


CASES = _cases()


def _device(p, tiles):
    from bookleaf_f90_amd.device import BookLeaf
    g = BookLeaf(p)
    g.set_tiling(tiles)
    return g


@pytest.mark.parametrize("name", sorted(CASES))
def test_tiles_match_three_kernel_step_and_oracle(name, oracle_lib):
    from oracle.pyoracle import Oracle
    p = CASES[name]
    print("case", name, "nel", p.nel, flush=True)
    stop = bool(p.stop_at_time_end)
    runs = (7, 1, 12)
    hist = {}
    state = {}
    for tiles in (True, False):
        g = _device(p, tiles)
        h = [g.run(n, stop_at_time_end=stop, runner="run_graph") for n in runs]
        hist[tiles] = [np.concatenate([x[k] for x in h]) for k in range(3)]
        state[tiles] = g.state()
        on, ntile, nbnd, nent = g.get_tiling()
        if tiles:
            assert on and ntile >= 1 and nent >= p.nel and nbnd == 0
            print("tiles", ntile, "entries", nent, flush=True)
        else:
            assert not on
    o = Oracle(p)
    ho = o.run(sum(runs), stop_at_time_end=stop)
    for k in range(3):
        assert np.array_equal(hist[True][k], hist[False][k])
        assert np.array_equal(hist[True][k], ho[k])
    assert diff_fields(state[True], state[False], compared(p)) == []
    assert diff_fields(state[True], o.state(), compared(p)) == []


def test_per_driver_steps_continue_from_tile_state(oracle_lib):
    """After an odd number of tiled steps the registry's element fields are
    the second element set: the per-driver entry points must see them."""
    from oracle.pyoracle import Oracle
    p = P.sedov(n=40, nsteps=20, kappa=0.01)
    g = _device(p, True)
    o = Oracle(p)
    g.run(5, runner="run_graph")
    o.run(5)
    for _ in range(2):
        g.step()
        o.step()
    assert diff_fields(g.state(), o.state(), compared(p)) == []
    g.run(3, runner="run_graph")
    o.run(3)
    assert diff_fields(g.state(), o.state(), compared(p)) == []


def test_switching_step_kinds_between_runs(oracle_lib):
    """Tiled, three-kernel and tiled run calls in turn: a three-kernel call
    moves the node sets and keeps the elements in place, so the next tiled
    call first puts the element sets back in step with the node sets."""
    from oracle.pyoracle import Oracle
    p = P.sedov(n=48, nsteps=30, kappa=0.01)
    g = _device(p, True)
    o = Oracle(p)
    for tiles, n in ((True, 5), (False, 4), (True, 3), (False, 1), (True, 6)):
        g.set_tiling(tiles)
        g.run(n, runner="run_graph")
        o.run(n)
        assert diff_fields(g.state(), o.state(), compared(p)) == [], (tiles, n)


def test_mesh_without_a_tiling_keeps_three_kernel_step(oracle_lib):
    """A node that is corner 4 of two elements (some elements numbered from
    the opposite corner) has no tiling: the run uses the other step and
    still matches the oracle."""
    from oracle.pyoracle import Oracle
    p = P.sod(nx=10, ny=4, nsteps=20)
    q = p.ielnd.reshape(-1, 4)
    nx = 10
    for e in range(q.shape[0]):
        if (e % nx) % 2 == 1:
            q[e] = np.roll(q[e], 2)  # corners 3, 4, 1, 2: same orientation
    p.ielnd = q.ravel().copy()
    g = _device(p, True)
    o = Oracle(p)
    hg = g.run(20, runner="run_graph")
    ho = o.run(20)
    assert not g.get_tiling()[0]
    for a, b in zip(hg, ho):
        assert np.array_equal(a, b)
    assert diff_fields(g.state(), o.state(), compared(p)) == []
