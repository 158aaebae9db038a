#!/usr/bin/env python3
"""Write the mesh-generator fixtures of tests/golden/ from the reference's own
build.

Runs oracle/_ref/bookleaf_meshgen (the reference's setup_mesh_gen and
setup_mesh_transfer, compiled from its sources by oracle/build_ref.py) on
decks written here: a quarter annulus (inner radius 1, outer 2) with the
LIN1 and the LIN2 region type, so that the generator's SOR converges to a
mesh whose l-lines (LIN1) or k-lines (LIN2) are straight, radial and evenly
spaced while the other family follows the arcs.  Each fixture meshgen_<name>.npz
holds the deck text, the generator's tolerance and the transferred mesh
(ielnd, indtype, ndx, ndy).  (The generator's EQUI type cannot be run: its
weight check, setup_mesh.F90:648-657, halts on EQUI's unset weights.)

usage: python tests/golden/make_meshgen.py   (needs oracle/_ref built)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "bookleaf_meshgen")
TOL = 1.0e-12  # region_tol default (setup_mesh.F90:127)

DECK = """&MESH
 region_typ(1)='{typ}'
 region_dim(1,:)={nl},{nk}
 region_side(1,1,:)=1,1
 region_side(1,2,:)=1,2
 region_side(1,3,:)=1,3
 region_side(1,4,:)=1,4
{sides}/
"""

# the four sides starting from the bottom radial line (LIN2: the k-lines
# averaged are radial) or from the outer arc (LIN1: the l-lines are radial);
# straight lines along the other family would cut inside the inner arc
SIDES_K = """ segment_typ(:)='LINE','ARC_A','LINE','ARC_C'
 segment_bc(:)='SLIPY','FREE','SLIPX','FREE'
 segment_pos(1,:)=1.0,0.0,2.0,0.0
 segment_pos(2,:)=2.0,0.0,0.0,2.0,0.0,0.0
 segment_pos(3,:)=0.0,2.0,0.0,1.0
 segment_pos(4,:)=0.0,1.0,1.0,0.0,0.0,0.0
"""
SIDES_L = """ segment_typ(:)='ARC_A','LINE','ARC_C','LINE'
 segment_bc(:)='FREE','SLIPX','FREE','SLIPY'
 segment_pos(1,:)=2.0,0.0,0.0,2.0,0.0,0.0
 segment_pos(2,:)=0.0,2.0,0.0,1.0
 segment_pos(3,:)=0.0,1.0,1.0,0.0,0.0,0.0
 segment_pos(4,:)=1.0,0.0,2.0,0.0
"""

CASES = {"lin1_annulus": ("LIN1", 8, 12, SIDES_L), "lin2_annulus": ("LIN2", 8, 12, SIDES_K)}


def read_mesh(path):
    d = open(path, "rb").read()
    nel, nnd = (int(v) for v in np.frombuffer(d, np.int32, 2))
    o = 8
    out = {}
    for name, dt, n in (("ielnd", np.int32, 4 * nel), ("indtype", np.int32, nnd),
                        ("ielreg", np.int32, nel), ("ndx", np.float64, nnd),
                        ("ndy", np.float64, nnd)):
        out[name] = np.frombuffer(d, dt, n, o).copy()
        o += out[name].nbytes
    assert o == len(d)
    return out


def main():
    if not os.path.exists(EXE):
        sys.exit("oracle/_ref/bookleaf_meshgen is not built (python oracle/build_ref.py)")
    for name, (typ, nl, nk, sides) in CASES.items():
        deck = DECK.format(typ=typ, nl=nl, nk=nk, sides=sides)
        with tempfile.TemporaryDirectory() as d:
            fd, fm = os.path.join(d, "deck"), os.path.join(d, "mesh.bin")
            open(fd, "w").write(deck)
            r = subprocess.run([EXE, fd, fm], capture_output=True, text=True, timeout=60)
            if r.returncode != 0:
                sys.exit("meshgen failed: " + r.stdout + r.stderr)
            m = read_mesh(fm)
        out = os.path.join(HERE, "meshgen_%s.npz" % name)
        np.savez_compressed(out, deck=np.frombuffer(deck.encode(), np.uint8),
                            tol=np.array(TOL), **m)
        print("%-14s nel=%d nnd=%d  %d bytes" % (name, len(m["ielreg"]), len(m["ndx"]),
                                                 os.path.getsize(out)))


if __name__ == "__main__":
    main()
