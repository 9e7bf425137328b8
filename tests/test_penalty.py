"""The constraint penalty C^T C in the profile with multi-term equations
that share DOFs (several constraint rows reach one profile entry): one
thread per entry adds the products in convertToLTri's order
(Working/AStrO_bookKeeping.f90:776-790), so repeated builds are bit-identical
and the profile matches the oracle's sequential sum."""

import numpy as np
import pytest

from helpers import entry_rel
from misc_amd import astro
from misc_amd.meshes import MPCTerm, cantilever_block
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def tied_block():
    m = cantilever_block(4, 3, 2)
    top = np.arange(5 * 4 * 2 + 1, 5 * 4 * 3 + 1, dtype=np.int32)
    eqn = 4
    # chains of ties u_z(a) - u_z(b) = 0 and u_x(a) + 0.5 u_x(b) - 0.25 u_x(c) = 0 sharing nodes
    for a, b in zip(top[:-1], top[1:]):
        m.mpc.append(MPCTerm(eqn=eqn, nodes=np.array([a], np.int32), dof=3, coef=1.0, is_set=False))
        m.mpc.append(MPCTerm(eqn=eqn, nodes=np.array([b], np.int32), dof=3, coef=-1.0, is_set=False))
        m.mpc_rhs.append(0.0)
        eqn += 1
    for a, b, c in zip(top[:-2], top[1:-1], top[2:]):
        for nd, cf in ((a, 1.0), (b, 0.5), (c, -0.25)):
            m.mpc.append(MPCTerm(eqn=eqn, nodes=np.array([nd], np.int32), dof=1, coef=cf, is_set=False))
        m.mpc_rhs.append(0.0)
        eqn += 1
    return m


def test_profile_with_shared_constraint_dofs_is_deterministic_and_matches_oracle():
    m = tied_block()
    print("penalty: tied 4x3x2 block", flush=True)
    outs = []
    for _ in range(2):
        ctx = astro.Context(m)
        for _ in range(2):
            ctx.assemble()
            ctx.scale_mpc()
            ctx.build_profile()
            outs.append(ctx.getd("elMatLT"))
        ctx.close()
    for o in outs[1:]:
        assert np.array_equal(o.view(np.int64), outs[0].view(np.int64))
    om = O.OracleModel(m).prep().static_elastic(stop_after="profile")
    ref = om.getd("elMatLT")
    assert entry_rel(outs[0], ref, floor=1e-3 * np.abs(ref).max()) < 1e-12
