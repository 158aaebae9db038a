"""The device-resident runner (bl_run_graph) against the reference's own
recorded results (tests/golden/, written from oracle/_ref): dt sequence and
fields after many steps, bit for bit, and the conservation tables of
hydro_dr_shortprint to a summation-order tolerance, with no oracle in
between."""
import numpy as np
import pytest

from bookleaf_f90_amd import abi
from tests import golden_util as G
from tests.test_gpu_graph import excluded

pytestmark = pytest.mark.gpu


# the Tait and JWL records call pow and exp, which the device computes
# within an ulp of glibc but not bit for bit: test_gpu_eos.py holds them to
# stated tolerances
BITEXACT = [n for n in G.names() if not n.startswith("tait_jwl")]


@pytest.mark.parametrize("name", BITEXACT)
def test_graph_runner_matches_reference_record(name):
    from bookleaf_f90_amd.device import BookLeaf
    p, meta, z = G.load(name)
    print("golden case", name, "nel", p.nel, flush=True)
    g = BookLeaf(p)
    dt, idt, ctl = g.run(meta["nsteps"], stop_at_time_end=bool(p.stop_at_time_end),
                         runner="run_graph")
    assert np.array_equal(dt, z["ref_dt"])
    assert np.array_equal(idt, z["ref_idt"])
    assert [abi.DT_NAMES[c].strip() for c in ctl] == meta["dt_names"]
    s = g.state()
    for k in ("eldensity", "elenergy", "elpressure", "ndx", "ndy", "ndu", "ndv"):
        assert np.array_equal(s[k][:len(z["final_" + k])], z["final_" + k]), k
    assert G.digest_mismatches(s, meta["final"], skip=excluded(p)) == []
    reg, mat = g.shortprint()
    G.assert_table_matches(reg, z["final_shortprint_region"], "region")
    G.assert_table_matches(mat, z["final_shortprint_material"], "material")
