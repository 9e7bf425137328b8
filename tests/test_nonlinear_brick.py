"""Geometrically nonlinear brick8 solves: what the literal reference's Newton
loop does (golden fixture), and the device's Newton solve on the
Green-Lagrange force against the oracle's force (GPU)."""

import os

import numpy as np
import pytest

from misc_amd import astro
from oracle import oracle as O

from nonlinear_cases import CASES, column_model

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reference_nonlinear_loop_repeats_the_linear_step():
    """D4 in the reference's own run: each element gathers the solution of
    nodes 1..8 (the clamped nodes here), so the residual of stepSolve's
    Newton loop never sees u and its 20 steps add the linear solution 20
    times.  The device path gathers each element's own nodes; its solve is
    pinned by the oracle's force below instead."""
    name, P = CASES[0]
    g = np.load(os.path.join(HERE, "golden", f"ref_nonlinear_{name}.npz"))
    assert int(g["nlSteps"][0]) == 20
    u, ul = g["nodeDisp"], g["linearDisp"]
    assert np.abs(u - 20.0 * ul).max() <= 1e-9 * np.abs(20.0 * ul).max()


def _oracle_force(m, nd, u):
    """Assembled Green-Lagrange internal force of the oracle."""
    R = np.zeros(u.size)
    for e in range(m.num_els):
        nodes = m.elements[e] - 1
        dofs = np.array([nd[nodes[a], c] - 1 for c in range(3) for a in range(8)])
        C = O.material_stiffness(m.eprops[e])
        R[dofs] += O.brick8_internal_force(m.nodes[nodes].T, C, u[dofs])
    return R


def _dense(ctx, n):
    K = ctx.getd("elasticMat")
    rng, cols = ctx.geti("elMatRange"), ctx.geti("elMatCols")
    A = np.zeros((n, n))
    for i in range(n):
        for s in range(rng[i], rng[i + 1]):
            if cols[s]:
                A[i, cols[s] - 1] += K[s]
    return A


@pytest.mark.gpu
def test_gl_tangent_is_the_derivative_of_the_oracle_force():
    """k_brick8_tangent_gl at u = 0 is k_brick8_tangent's matrix; at a
    displacement with 10 % strains it equals the central difference of the
    oracle's assembled Green-Lagrange force, column by column."""
    from misc_amd.meshes import cantilever_block

    m = cantilever_block(2, 2, 2, e_field=np.linspace(150e9, 250e9, 8), pressure=0.0)
    ctx = astro.Context(m)
    n = ctx.scalar("elMatDim")
    nd = ctx.geti("nDofIndex").reshape(-1, 3) if ctx.geti("nDofIndex").size == 3 * m.num_nodes else \
        ctx.geti("nDofIndex").reshape(-1, 6)
    ctx.assemble()
    ctx.sync()
    K0 = _dense(ctx, n)
    print("gl tangent: at rest", flush=True)
    ctx.nonlinear_tangent(np.zeros(n))
    scale = np.abs(K0).max()
    assert np.abs(_dense(ctx, n) - K0).max() <= 1e-12 * scale
    u = np.random.default_rng(5).standard_normal(n) * 0.05
    print("gl tangent: finite strains", flush=True)
    ctx.nonlinear_tangent(u)
    Kt = _dense(ctx, n)
    ctx.close()
    Kfd = np.zeros((n, n))
    for j in range(n):
        up, um = u.copy(), u.copy()
        up[j] += 1e-6
        um[j] -= 1e-6
        Kfd[:, j] = (_oracle_force(m, nd, up) - _oracle_force(m, nd, um)) / (up[j] - um[j])
    assert np.abs(Kt - K0).max() > 1e-2 * scale
    assert np.abs(Kt - Kt.T).max() <= 1e-12 * scale
    assert np.abs(Kt - Kfd).max() <= 1e-7 * scale


@pytest.mark.gpu
def test_nonlinear_column_reaches_the_oracle_equilibrium():
    """A 0.2 x 0.2 x 2 column under a lateral tip load of 2 MN (linear tip
    travel 0.12 m): the Newton solve on the Green-Lagrange force ends in a few
    quadratically shrinking steps at a displacement where the oracle's own
    assembled force balances the load at every free DOF to 1e-9, and whose
    tip drops as the column bends, which the linear solution does not."""
    m = column_model(2.0e6)
    ctx = astro.Context(m, astro.Options(rtol=1e-12))
    n = ctx.scalar("elMatDim")
    ndi = ctx.geti("nDofIndex")
    nd = ndi.reshape(-1, 6) if ndi.size == 6 * m.num_nodes else ndi.reshape(-1, 3)
    print("nonlinear column: linear solve", flush=True)
    ctx.static_solve()
    ul = ctx.getd("nodeDisp")
    ctx.setd("nodeDisp", np.zeros(n))
    steps = []
    print("nonlinear column: Newton steps", flush=True)
    for _ in range(12):
        u0 = ctx.getd("nodeDisp")
        assert ctx.nonlinear_solve(max_it=1, tol=1e-300) == 1
        steps.append(np.abs(ctx.getd("nodeDisp") - u0).max())
        if steps[-1] <= 1e-12 * np.abs(u0).max():
            break
    u = ctx.getd("nodeDisp")
    ctx.close()
    assert len(steps) <= 8, steps
    # the last full step shows the quadratic rate
    assert steps[-2] <= 1e2 * steps[-3] ** 2 / np.abs(u).max() + 1e-12 * np.abs(u).max(), steps
    F = np.zeros(n)
    for k, i in enumerate(m.load_nodes - 1):
        F[nd[i, 0] - 1] += m.load_vals[k, 0]
    R = _oracle_force(m, nd, u)
    free = np.ones(n, bool)
    for i in np.nonzero(np.isclose(m.nodes[:, 2], 0.0))[0]:
        free[nd[i, :3] - 1] = False
    assert np.abs((F - R)[free]).max() <= 1e-9 * np.abs(F).max()
    assert np.abs(u[~free]).max() <= 1e-12 * np.abs(u).max()
    # the geometric signature: the tip's lateral travel moves by 3e-4 of
    # itself, and the tip drops (the arc shortens the column by about
    # 0.6 dx^2 / L), which the linear solution does not show
    top = m.load_nodes[0] - 1
    dx, dxl = u[nd[top, 0] - 1], ul[nd[top, 0] - 1]
    dz, dzl = u[nd[top, 2] - 1], ul[nd[top, 2] - 1]
    print(f"tip dx {dx:.6e} (linear {dxl:.6e}), dz {dz:.6e} (linear {dzl:.6e})", flush=True)
    assert 1e-5 < abs(dx - dxl) / abs(dxl) < 1e-2
    assert dz - dzl < -0.2 * dxl ** 2 / 2.0


@pytest.mark.gpu
def test_nonlinear_free_expansion_is_the_exact_stretch():
    """A uniform temperature rise on statically determinate supports: with
    the Green-Lagrange strain, E = a + a^2 / 2 = T alpha on the diagonal for
    u = a x, so the nonlinear solve must stretch the block by
    a = sqrt(1 + 2 T alpha) - 1 exactly (zero stress), where the linear
    solve gives T alpha.  At T alpha = 1e-2 the two differ by 5e-5."""
    from thermoelastic_cases import free_block

    alpha, T0 = 1.0e-4, 100.0
    m = free_block(alpha=alpha)
    ctx = astro.Context(m, astro.Options(rtol=1e-12))
    ctx.setd("nodeTemp", np.full(m.num_nodes, T0))
    print("nonlinear free expansion", flush=True)
    it = ctx.nonlinear_solve(max_it=20, tol=1e-13)
    u = ctx.getd("nodeDisp")
    nd = ctx.geti("nDofIndex").reshape(-1, 6)
    ctx.close()
    assert it <= 8
    got = np.array([[u[nd[i, c] - 1] for c in range(3)] for i in range(m.num_nodes)])
    a = np.sqrt(1.0 + 2.0 * T0 * alpha) - 1.0
    assert abs(a - T0 * alpha) > 4e-5
    assert np.abs(got - a * m.nodes).max() <= 1e-11 * np.abs(a * m.nodes).max()
