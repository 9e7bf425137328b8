"""Models of the geometrically nonlinear static solves (Newton with the
Green-Lagrange brick8 force), shared by the golden-vector script and the
tests."""

import numpy as np

from misc_amd.meshes import cantilever_block

# (name, total lateral tip load in N): a 0.2 x 0.2 x 2 steel column clamped at
# z = 0, the load in +x spread evenly over the top nodes
CASES = [("column_2x2x8", 1.0e5)]


def column_model(P):
    m = cantilever_block(2, 2, 8, lx=0.2, ly=0.2, lz=2.0, pressure=0.0)
    top = (np.nonzero(np.isclose(m.nodes[:, 2], 2.0))[0] + 1).astype(np.int32)
    m.load_nodes = top
    m.load_vals = np.zeros((top.size, 7))
    m.load_vals[:, 0] = P / top.size
    m.node_sets["top"] = top
    return m
