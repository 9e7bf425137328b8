"""The co-rotational shell4 oracle (oracle/shell_corot.py) pinned by what
the mathematics fixes: a rotation matrix is orthogonal and its derivatives
are those of a complex step; a finite rigid motion strains nothing; near
rest the internal force tends to the linear tangent of the element (the
tangent is pinned against the reference's own element matrices elsewhere);
the force is the gradient of the energy (central differences)."""

import numpy as np

from oracle import oracle as O
from oracle import shell_corot as S

XYZ = np.array([[0, 0, 0], [0.1, 0.01, 0.0], [0.12, 0.1, 0.02], [0.0, 0.09, 0.01]], float)


def _element():
    alpha = O.shell_frame([1, 0, 0, 0, 1, 0], XYZ.T)
    abd = O.shell_abd([[70e9, 70e9, 0.33, 70e9 / 2.66, 0.002, 0.3], [120e9, 9e9, 0.3, 5e9, 0.001, 0.7]], 0.0004)
    K32, _ = O.shell4_stiffness(XYZ.T, alpha, abd)
    return alpha, abd, K32


def test_rotation_matrix_and_derivatives():
    for r in ([0.3, -0.2, 0.5], [1e-9, 2e-9, 0.0], [2.0, -1.5, 1.0]):
        r = np.array(r)
        R, dR, d2R = S.rotation(r)
        assert np.abs(R @ R.T - np.eye(3)).max() < 1e-15
        assert abs(np.linalg.det(R) - 1.0) < 1e-15
        # Rodrigues' formula with numpy's trigonometry
        t = np.linalg.norm(r)
        K = S._skew(r)
        Rr = np.eye(3) + np.sin(t) / t * K + (1 - np.cos(t)) / t ** 2 * K @ K
        assert np.abs(R - Rr).max() < 1e-14
        for i in range(3):
            rp = r.astype(complex)
            rp[i] += 1e-30j
            Rc, dRc, _ = S.rotation(rp)
            assert np.abs(Rc.imag / 1e-30 - dR[i]).max() < 1e-13
            for j in range(3):
                assert np.abs(dRc[j].imag / 1e-30 - d2R[j][i]).max() < 1e-13


def test_rigid_motion_is_strain_free():
    alpha, abd, K32 = _element()
    for th in ([0.4, -0.7, 0.25], [1.2, 0.3, -0.9]):
        th = np.array(th)
        Rr = S.rotation(th)[0]
        shift = np.array([0.01, -0.02, 0.03])
        u = np.array([(Rr - np.eye(3)) @ x + shift for x in XYZ])
        ue = np.concatenate([u.T.ravel(), np.tile(th, (4, 1)).T.ravel(), np.zeros(8)])
        assert np.abs(S.strains(XYZ, alpha, ue)).max() < 1e-14
        F = S.internal_force(XYZ, alpha, abd, ue)
        assert np.abs(F).max() < 1e-12 * np.abs(K32).max() * 0.01
        Fc, E = S.condensed_force(XYZ, alpha, abd, K32, ue[:24], iters=4)
        assert np.abs(Fc).max() < 1e-12 * np.abs(K32).max() * 0.01 and abs(E) < 1e-20


def test_small_displacements_follow_the_tangent():
    alpha, abd, K32 = _element()
    v = np.random.default_rng(0).normal(size=32) * np.r_[np.full(12, 1e-3), np.full(12, 1e-2), np.full(8, 1e-3)]
    errs = []
    for eps in (1e-2, 1e-3, 1e-4):
        F = S.internal_force(XYZ, alpha, abd, eps * v)
        errs.append(np.linalg.norm(F - K32 @ (eps * v)) / np.linalg.norm(K32 @ (eps * v)))
    # first order in eps: the quadratic terms of the kinematics
    assert errs[0] < 1e-3 and 8 < errs[0] / errs[1] < 12 and 8 < errs[1] / errs[2] < 12
    # the condensed force likewise tends to the condensed tangent
    _, Kc = O.shell4_stiffness(XYZ.T, alpha, abd)
    u24 = 1e-4 * v[:24]
    Fc, _ = S.condensed_force(XYZ, alpha, abd, K32, u24)
    assert np.linalg.norm(Fc - Kc @ u24) < 1e-3 * np.linalg.norm(Kc @ u24)


def test_force_is_the_gradient_of_the_energy():
    alpha, abd, _ = _element()
    ue = np.random.default_rng(1).normal(size=32) * np.r_[np.full(12, 5e-3), np.full(12, 0.2), np.full(8, 1e-3)]
    F = S.internal_force(XYZ, alpha, abd, ue)
    for d in (0, 7, 13, 20, 23, 27):
        h = 1e-6 * (1e-2 if d < 12 or d >= 24 else 1.0)
        up, um = ue.copy(), ue.copy()
        up[d] += h
        um[d] -= h
        fd = (S.energy(XYZ, alpha, abd, up) - S.energy(XYZ, alpha, abd, um)) / (2 * h)
        assert abs(fd - F[d]) <= 1e-6 * np.abs(F).max()


def test_chord_steps_bring_the_internal_dofs_to_equilibrium():
    alpha, abd, K32 = _element()
    u24 = np.random.default_rng(2).normal(size=24) * np.r_[np.full(12, 5e-3), np.full(12, 0.1)]
    Kii = K32[24:, 24:]
    ui = -np.linalg.solve(Kii, K32[24:, :24] @ u24)
    r0 = np.abs(S.internal_force(XYZ, alpha, abd, np.concatenate([u24, ui]))[24:]).max()
    for _ in range(16):
        ui = ui - np.linalg.solve(Kii, S.internal_force(XYZ, alpha, abd, np.concatenate([u24, ui]))[24:])
    r16 = np.abs(S.internal_force(XYZ, alpha, abd, np.concatenate([u24, ui]))[24:]).max()
    assert r16 < 1e-6 * r0
