"""Models of the thermo-elastic tests: the fixtures' models (rebuilt from
the parameters the fixture script records) and a block on statically
determinate supports, where a uniform temperature rise is a free expansion
u = T alpha x with zero stress (trilinear elements hold linear fields)."""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from make_thermoelastic import CASES, case_model  # noqa: E402,F401
from misc_amd.meshes import BRICK8, MPCTerm, Model, brick_block, iso_props  # noqa: E402


def free_block(nx=3, ny=2, nz=2, lx=0.3, ly=0.2, lz=0.2, alpha=1.5e-5, E=70e9, nu=0.3):
    """Supports: node (0,0,0) fixed, node (lx,0,0) held in y and z, node
    (0,ly,0) held in z (six single-node constraint equations)."""
    nodes, conn, colors = brick_block(nx, ny, nz, lx, ly, lz)
    Ne = conn.shape[0]
    m = Model(name=f"free{nx}x{ny}x{nz}", nodes=nodes, elements=conn, etype=np.full(Ne, BRICK8, np.int32),
              eprops=np.tile(iso_props(E, nu), (Ne, 1)), density=np.zeros(Ne), cond=np.zeros((Ne, 6)),
              colors=colors)
    NX, NY = nx + 1, ny + 1
    a, b, c = 1, NX, NX * (NY - 1) + 1
    fix = [(a, 1), (a, 2), (a, 3), (b, 2), (b, 3), (c, 3)]
    m.mpc = [MPCTerm(eqn=q + 1, nodes=np.array([nd], np.int32), dof=d, coef=1.0, is_set=False)
             for q, (nd, d) in enumerate(fix)]
    m.mpc_rhs = [0.0] * len(fix)
    m.therm_exp = np.tile([alpha, alpha, alpha, 0.0, 0.0, 0.0], (Ne, 1))
    return m
