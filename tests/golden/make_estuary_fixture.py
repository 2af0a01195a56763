"""Write tests/golden/estuary_mesh.npz from the reference's Estuary example
inputs (Examples/Estuary/tstinp: tst_grd.dat, tst_dep.dat, tst_obc.dat).

These are input data files of the reference's example deck (an 845-node,
1570-cell ideal estuary).  Only the mesh, depth and open-boundary node list
are kept; the triangles are stored clockwise, as the reference does after
reading the counter-clockwise grid file (mod_input.F:3739-3742).
Run from the repository root in the container that holds the reference.
"""
import os
import sys

import numpy as np

SRC = "/root/reference/Examples/Estuary/tstinp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "estuary_mesh.npz")


def main():
    lines = open(os.path.join(SRC, "tst_grd.dat")).read().split("\n")
    nn = int(lines[0].split("=")[1])
    nc = int(lines[1].split("=")[1])
    tri = np.array([[int(x) for x in l.split()[1:4]] for l in lines[2:2 + nc]], np.int32)
    xy = np.array([[float(x) for x in l.split()[1:3]] for l in lines[2 + nc:2 + nc + nn]])
    dep = open(os.path.join(SRC, "tst_dep.dat")).read().split("\n")
    h = np.array([float(l.split()[2]) for l in dep[1:1 + nn]])
    obc = open(os.path.join(SRC, "tst_obc.dat")).read().split("\n")
    nobc = int(obc[0].split("=")[1])
    obc_nodes = np.array([int(l.split()[1]) for l in obc[1:1 + nobc]], np.int32)
    nv_cw = tri[:, [0, 2, 1]].copy()
    np.savez_compressed(OUT, x=xy[:, 0], y=xy[:, 1], nv=nv_cw, h=h, obc=obc_nodes)
    print("wrote", OUT, nn, nc, nobc)


if __name__ == "__main__":
    sys.exit(main())
