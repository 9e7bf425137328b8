"""Stress and strain recovery of brick8 elements (r_getSolidStress,
Working/AStrO_r_elementEqns.f90:924-976).

The oracle is pinned on the CPU by closed forms: a linear displacement
field u = G x (reproduced exactly by the trilinear element on any affine
brick) has the Green-Lagrange strain (G + G^T + G^T G) / 2 at every point;
a finite rigid rotation has zero strain; the stress is C times the strain
for the isotropic C of the closed form.  On the GPU the device recovery of
every element is compared with the oracle's for a random displacement field
on a non-affine mesh, and the stresses of a solved block with the applied
pressure."""

import numpy as np
import pytest

from misc_amd.meshes import cantilever_block, iso_props
from oracle import oracle as O

RNG = np.random.default_rng(11)
POINTS = [(0.0, 0.0, 0.0), (0.5773502691896258, -0.5773502691896258, 0.5773502691896258), (1.0, -1.0, 1.0)]


def affine_brick():
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    A = np.array([[0.9, 0.2, 0.1], [0.05, 1.3, -0.2], [0.1, 0.0, 0.7]])
    return (base @ A.T + np.array([0.3, -0.1, 0.2])).T.copy()  # 3 x 8


def voigt_gl(G):
    E = 0.5 * (G + G.T + G.T @ G)
    return np.array([E[0, 0], E[1, 1], E[2, 2], 2 * E[0, 1], 2 * E[0, 2], 2 * E[1, 2]])


def ue_of(X, disp):
    """Component-major element DOFs of a nodal displacement function."""
    U = np.array([disp(X[:, n]) for n in range(8)])  # 8 x 3
    return U.T.ravel()


def test_linear_field_gives_the_green_lagrange_closed_form():
    X = affine_brick()
    C = O.material_stiffness(iso_props(200e9, 0.3))
    for scale in (1e-4, 0.3):
        G = scale * RNG.standard_normal((3, 3))
        ue = ue_of(X, lambda x: G @ x)
        for s in POINTS:
            st, sn = O.brick8_stress(X, C, ue, s)
            np.testing.assert_allclose(sn, voigt_gl(G), rtol=0, atol=1e-13 * max(1.0, np.abs(G).max() ** 2))
            np.testing.assert_allclose(st, C @ voigt_gl(G), rtol=1e-12, atol=1e-12 * np.abs(C @ voigt_gl(G)).max())


def test_rigid_rotation_has_no_strain():
    X = affine_brick()
    C = O.material_stiffness(iso_props(70e9, 0.33))
    a = np.array([0.3, -0.7, 0.5])
    th = 0.8
    k = a / np.linalg.norm(a)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    R = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    ue = ue_of(X, lambda x: (R - np.eye(3)) @ x + np.array([0.1, 0.2, -0.3]))
    for s in POINTS:
        _, sn = O.brick8_stress(X, C, ue, s)
        assert np.abs(sn).max() <= 1e-14


def test_isotropic_stress_closed_form():
    X = affine_brick()
    E, nu = 200e9, 0.3
    C = O.material_stiffness(iso_props(E, nu))
    G = 1e-4 * RNG.standard_normal((3, 3))
    st, sn = O.brick8_stress(X, C, ue_of(X, lambda x: G @ x), POINTS[0])
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    tr = sn[:3].sum()
    expect = np.concatenate([lam * tr + 2 * mu * sn[:3], mu * sn[3:]])
    np.testing.assert_allclose(st, expect, rtol=1e-12)


@pytest.mark.gpu
def test_gpu_stress_matches_oracle_on_random_field():
    from misc_amd import astro

    m = cantilever_block(6, 5, 4, lx=0.6, ly=0.5, lz=0.4)
    # a non-affine mesh: interior nodes moved
    X = m.nodes.copy()
    inner = np.all((X > 1e-9) & (X < np.array([0.6, 0.5, 0.4]) - 1e-9), axis=1)
    X[inner] += 0.02 * RNG.uniform(-1, 1, size=(inner.sum(), 3))
    m.nodes = X
    om = O.OracleModel(m).prep()
    om.static_elastic(stop_after="assemble")
    n = om.scalar("elMatDim")
    u = 1e-3 * RNG.standard_normal(n)
    om.setd("nodeDisp", u)
    ctx = astro.Context(m)
    ctx.setd("nodeDisp", u)
    print("stress recovery 6x5x4", flush=True)
    for s in POINTS:
        st, sn = ctx.element_stress(s)
        so, eo = om.element_stress(s)
        assert np.abs(sn - eo).max() <= 1e-12 * np.abs(eo).max()
        assert np.abs(st - so).max() <= 1e-12 * np.abs(so).max()
    ctx.close()


@pytest.mark.gpu
def test_gpu_stress_of_solved_block_carries_the_pressure():
    """The clamped block under 1 MPa on its top face: at the top surface of the
    top elements away from the edges sigma_33 is the applied -1 MPa, and the
    shear there vanishes."""
    from misc_amd import astro

    nx, ny, nz = 12, 12, 6
    m = cantilever_block(nx, ny, nz)
    ctx = astro.Context(m, astro.Options(rtol=1e-12))
    ctx.static_solve()
    st, _ = ctx.element_stress((0.0, 0.0, 1.0))
    e = np.arange(m.num_els)
    i, j, k = e % nx, (e // nx) % ny, e // (nx * ny)
    top = (k == nz - 1) & (i >= 3) & (i < nx - 3) & (j >= 3) & (j < ny - 3)
    s33 = st[top, 2]
    print(f"top sigma_33 {s33.min():.4e} .. {s33.max():.4e}", flush=True)
    assert np.all(np.abs(s33 + 1e6) <= 0.05e6)
    assert np.abs(st[top, 4:6]).max() <= 0.05e6
    ctx.close()
