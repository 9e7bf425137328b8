"""Co-rotational shell4 internal force on the device
(astro_ctx_shell_internal_force) against the CPU oracle
(oracle/shell_corot.py) element by element, and its two limits on whole
plates: a finite rigid rotation gives no force, small displacements give the
assembled tangent times u."""

import numpy as np
import pytest

from misc_amd import astro
from misc_amd.meshes import shell_plate
from oracle import oracle as O
from oracle import shell_corot as S

from test_shell import rot_matrix, tilted_plate

pytestmark = pytest.mark.gpu


def _oracle_force(m, u, iters):
    om = O.OracleModel(m).prep()
    nd = om.geti("nDofIndex").reshape(-1, 6)
    F = np.zeros(om.scalar("elMatDim"))
    En = np.zeros(m.num_els)
    for e in range(m.num_els):
        nodes = m.elements[e, :4] - 1
        xyz = m.nodes[nodes]
        alpha = O.shell_frame(m.orient[e], xyz.T)
        abd = O.shell_abd(m.layers[e], m.z_offset[e])
        K32, _ = O.shell4_stiffness(xyz.T, alpha, abd)
        dofs = np.array([nd[nodes[a], v] for v in range(6) for a in range(4)]) - 1
        Fc, En[e] = S.condensed_force(xyz, alpha, abd, K32, u[dofs], iters=iters)
        np.add.at(F, dofs, Fc)
    return F, En, nd


def _plate_u(m, nd, n, rot_scale, disp_scale, seed):
    rng = np.random.default_rng(seed)
    u = np.zeros(n)
    for i in range(m.num_nodes):
        for v in range(6):
            if nd[i, v]:
                u[nd[i, v] - 1] = rng.normal() * (disp_scale if v < 3 else rot_scale)
    return u


@pytest.mark.parametrize("iters", [0, 4])
def test_corot_force_matches_oracle(iters):
    for m in (shell_plate(4, 2, nu=0.3), tilted_plate()):
        om = O.OracleModel(m).prep()
        nd = om.geti("nDofIndex").reshape(-1, 6)
        n = om.scalar("elMatDim")
        u = _plate_u(m, nd, n, 0.2, 2e-3, 3)
        print(f"corot: {m.name} iters={iters}", flush=True)
        Fr, Er, _ = _oracle_force(m, u, iters)
        ctx = astro.Context(m)
        F, E = ctx.shell_internal_force(u, iters=iters)
        ctx.close()
        assert np.abs(F - Fr).max() <= 1e-9 * np.abs(Fr).max()
        np.testing.assert_allclose(E, Er, rtol=1e-9, atol=1e-12 * np.abs(Er).max())


def test_corot_force_rigid_rotation_and_small_limit():
    m = tilted_plate()
    om = O.OracleModel(m).prep()
    om.static_elastic(stop_after="assemble")
    nd = om.geti("nDofIndex").reshape(-1, 6)
    n = om.scalar("elMatDim")
    ctx = astro.Context(m)
    ctx.assemble()
    ctx.sync()
    K = ctx.getd("elasticMat")
    rng, cols = om.geti("elMatRange"), om.geti("elMatCols")

    def Ku(u):
        out = np.zeros(n)
        for i in range(n):
            for s in range(rng[i], rng[i + 1]):
                if cols[s]:
                    out[i] += K[s] * u[cols[s] - 1]
        return out

    # a finite rigid rotation of the whole plate about a point off the plate
    th = np.array([0.5, -0.8, 0.3])
    R = rot_matrix(th, np.linalg.norm(th))
    c = np.array([0.1, -0.05, 0.2])
    u = np.zeros(n)
    for i in range(m.num_nodes):
        d = (R - np.eye(3)) @ (m.nodes[i] - c)
        for v in range(6):
            if nd[i, v]:
                u[nd[i, v] - 1] = d[v] if v < 3 else th[v - 3]
    print("corot: rigid rotation", flush=True)
    F, E = ctx.shell_internal_force(u, iters=4)
    scale = np.abs(Ku(1e-3 * u / np.abs(u).max())).max()
    assert np.abs(F).max() <= 1e-9 * scale
    assert np.abs(E).max() <= 1e-18
    # near rest the force is K u to first order
    u = _plate_u(m, nd, n, 1e-2, 1e-4, 4)
    errs = []
    for eps in (1e-2, 1e-3):
        F, _ = ctx.shell_internal_force(eps * u, iters=16)
        Kv = Ku(eps * u)
        errs.append(np.linalg.norm(F - Kv) / np.linalg.norm(Kv))
    ctx.close()
    assert errs[0] < 1e-3 and 7 < errs[0] / errs[1] < 13


def test_corot_force_at_the_solution_matches_the_linear_residual():
    """Without an argument the force is taken at the current solution.  For
    a small tip load it equals the applied load at the free DOFs (K u = F)
    up to the membrane force of the deflection, EA w'^2 / 2, which grows as
    the load squared (3.6e-7 N for 1e-6 N on this plate, whose membrane
    stiffness is (L / t)^2 ~ 2.5e5 times its bending stiffness)."""
    errs = []
    for load_n in (1e-9, 1e-8):
        m = shell_plate(8, 2, nu=0.0, tip_load=load_n)
        ctx = astro.Context(m, astro.Options(rtol=1e-9, maxCycles=100000))
        ctx.static_solve()
        F, _ = ctx.shell_internal_force(iters=16)
        F2, _ = ctx.shell_internal_force(ctx.getd("nodeDisp"), iters=16)
        ctx.close()
        assert np.array_equal(F, F2)
        om = O.OracleModel(m).prep()
        nd = om.geti("nDofIndex").reshape(-1, 6)
        load = np.zeros_like(F)
        for i in m.node_sets["tip"] - 1:
            load[nd[i, 2] - 1] = load_n
        free = np.ones(F.size, bool)
        for i in m.node_sets["root"] - 1:
            free[nd[i] - 1] = False
        errs.append(np.abs(F[free] - load[free]).max() / load_n)
    assert errs[0] < 2e-3 and 7 < errs[1] / errs[0] < 13
