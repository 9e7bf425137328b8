"""CPU oracle of the co-rotational shell4 kinematics and internal force.

Test infrastructure only (imported by the tests as the checker).

Restates Working/AStrO_r_elementEqns.f90:681-891 (r_getShellDef,
r_getInstOrient, r_getInstDof) and the shell branch of r_getElRuk
(:2245-2423) with the finite rotation of r_rotateAlpha
(Working/AStrO_r_designPropertyFunctions.f90:1075-1463):

* the element frame follows the mean nodal rotation tbar:
  instOri = alpha R(tbar)^T, R the rotation by the vector tbar
  (R = I + a(s) [r]x + b(s) [r]x^2, s = r.r, a = sin t / t, b = (1 - cos t) / t^2,
  t^2 = s; the reference's r_sin / r_cos are not reproduced, D1);
* local displacements instOri (u + X) - alpha X at the corner nodes, the
  internal (bubble) DOFs as they are;
* local rotations drIdrG (theta - tbar) + 1/2 d2rIdrG (theta - tbar)^2 with
  drIdrG(:, i) = axial(dInstOri/dtbar_i instOri^T) and the second
  derivatives likewise (axial(P) = (P23, P31, P12));
* the nine generalised strains of r_getShellDef at the four Gauss points,
  strain energy E = 1/2 sum w e^T ABD e, internal force R = dE/du.

a, b and their first two derivatives in s are power series (entire
functions), so every quantity is analytic in the DOFs and R is taken by a
complex step on each of the 32 element DOFs (24 corner DOFs, component-major
u v w tx ty tz over the 4 nodes, then 8 internal).  Pinned
(tests/test_shell_corot.py): a finite rigid rotation of an element gives zero
strains, energy and force; R(eps u) / eps tends to K32 u (the tangent at rest,
oracle.shell4_stiffness) with an O(eps) error; central differences of E;
R, dR and d2R against complex steps of R; orthogonality of R.
"""

from __future__ import annotations

import math

import numpy as np

_Q = 1.0 / math.sqrt(3.0)
_NT = 24  # series terms: |s| up to ~30 (rotations up to ~5 rad) at full precision


def _series(s):
    """a, b, a', b', a'', b'' at s (scalar, real or complex)."""
    a = b = da = db = d2a = d2b = 0.0
    for k in range(_NT - 1, -1, -1):
        ca = (-1.0) ** k / math.factorial(2 * k + 1)
        cb = (-1.0) ** k / math.factorial(2 * k + 2)
        a = a * s + ca
        b = b * s + cb
        if k >= 1:
            da = da * s + k * ca
            db = db * s + k * cb
        if k >= 2:
            d2a = d2a * s + k * (k - 1) * ca
            d2b = d2b * s + k * (k - 1) * cb
    return a, b, da, db, d2a, d2b


def _skew(r):
    return np.array([[0 * r[0], -r[2], r[1]], [r[2], 0 * r[0], -r[0]], [-r[1], r[0], 0 * r[0]]])


def rotation(r):
    """R(r), dR/dr_i (3), d2R/dr_i dr_j (3, 3) for a rotation vector r."""
    r = np.asarray(r)
    s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2]
    a, b, da, db, d2a, d2b = _series(s)
    K = _skew(r)
    K2 = K @ K
    E = [_skew(np.eye(3)[i]) for i in range(3)]
    R = np.eye(3) + a * K + b * K2
    dR = [da * 2 * r[i] * K + a * E[i] + db * 2 * r[i] * K2 + b * (E[i] @ K + K @ E[i]) for i in range(3)]
    d2R = [[None] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            d = (i == j) * 1.0
            t = (d2a * 4 * r[i] * r[j] + da * 2 * d) * K + da * 2 * (r[i] * E[j] + r[j] * E[i])
            t = t + (d2b * 4 * r[i] * r[j] + db * 2 * d) * K2
            t = t + db * 2 * (r[i] * (E[j] @ K + K @ E[j]) + r[j] * (E[i] @ K + K @ E[i]))
            t = t + b * (E[i] @ E[j] + E[j] @ E[i])
            d2R[i][j] = t
    return R, dR, d2R


def _axial(P):
    return np.array([P[1, 2], P[2, 0], P[0, 1]])


def inst_frame(alpha, tbar):
    """instOri and drIdrG (3 x 3) and d2rIdrG (3 x 3 x 3) of r_getInstOrient."""
    R, dR, d2R = rotation(tbar)
    I = alpha @ R.T
    d1 = np.stack([_axial((alpha @ dR[i].T) @ I.T) for i in range(3)], axis=1)
    d2 = np.zeros((3, 3, 3), dtype=np.result_type(I, d1))
    for i in range(3):
        for j in range(3):
            d2[:, i, j] = _axial((alpha @ d2R[i][j].T) @ I.T)
    return I, d1, d2


def geometry(xyz, alpha):
    """Per Gauss point (bit 0 -> s1): N (10), Nx (10, 2) in the element frame,
    and detJ, for the 4 bilinear and 6 bubble functions."""
    out = []
    loc = (alpha @ np.asarray(xyz).T).T  # (4, 3)
    for g in range(4):
        s1 = _Q if g & 1 else -_Q
        s2 = _Q if g & 2 else -_Q
        N = np.array([0.25 * (1 - s1) * (1 - s2), 0.25 * (1 + s1) * (1 - s2), 0.25 * (1 + s1) * (1 + s2),
                      0.25 * (1 - s1) * (1 + s2), 1 - s1 * s1, 1 - s2 * s2, (1 - s1 * s1) * (1 - s2),
                      (1 - s2 * s2) * (1 + s1), (1 - s1 * s1) * (1 + s2), (1 - s2 * s2) * (1 - s1)])
        Ns = np.array([[-0.25 * (1 - s2), -0.25 * (1 - s1)], [0.25 * (1 - s2), -0.25 * (1 + s1)],
                       [0.25 * (1 + s2), 0.25 * (1 + s1)], [-0.25 * (1 + s2), 0.25 * (1 - s1)],
                       [-2 * s1, 0.0], [0.0, -2 * s2], [-2 * s1 * (1 - s2), -(1 - s1 * s1)],
                       [1 - s2 * s2, -2 * s2 * (1 + s1)], [-2 * s1 * (1 + s2), 1 - s1 * s1],
                       [-(1 - s2 * s2), -2 * s2 * (1 - s1)]])
        J = np.zeros((3, 3))
        J[:, :2] = loc.T @ Ns[:4]
        J[2, 2] = 1.0
        Ji = np.linalg.inv(J)
        Nx = Ns @ Ji[:2, :2]
        out.append((N, Nx, np.linalg.det(J)))
    return out


# internal DOF k -> (local component, bubble function)
_INT = [(0, 4), (0, 5), (1, 4), (1, 5), (2, 6), (2, 7), (2, 8), (2, 9)]


def strains(xyz, alpha, ue):
    """The nine generalised strains at the 4 Gauss points (4, 9) for the 32
    element DOFs ue (real or complex)."""
    xyz = np.asarray(xyz)
    ue = np.asarray(ue)
    u = ue[:12].reshape(3, 4).T  # (node, comp)
    th = ue[12:24].reshape(3, 4).T
    tbar = th.mean(axis=0)
    I, d1, d2 = inst_frame(alpha, tbar)
    dt = (ue.dtype if np.iscomplexobj(ue) else np.result_type(I, ue))
    disp = np.zeros((3, 10), dtype=dt)
    rot = np.zeros((3, 4), dtype=dt)
    for a in range(4):
        disp[:, a] = I @ (u[a] + xyz[a]) - alpha @ xyz[a]
        d = th[a] - tbar
        rot[:, a] = d1 @ d + 0.5 * np.einsum("kij,i,j->k", d2, d, d)
    for k, (c, fn) in enumerate(_INT):
        disp[c, fn] = disp[c, fn] + ue[24 + k]
    out = []
    for N, Nx, _ in geometry(xyz, alpha):
        ux = disp @ Nx         # (3, 2)
        rx = rot @ Nx[:4]      # (3, 2)
        pr = rot @ N[:4]       # (3,)
        out.append(np.array([ux[0, 0], ux[1, 1], ux[0, 1] + ux[1, 0], rx[1, 0], -rx[0, 1], rx[1, 1] - rx[0, 0],
                             ux[2, 0] + pr[1], ux[2, 1] - pr[0], 2 * pr[2] - ux[1, 0] + ux[0, 1]]))
    return np.stack(out)


def energy(xyz, alpha, abd, ue):
    w = [g[2] for g in geometry(xyz, alpha)]
    e = strains(xyz, alpha, ue)
    return 0.5 * sum(w[g] * (e[g] @ (abd @ e[g])) for g in range(4))


def internal_force(xyz, alpha, abd, ue, h=1e-30):
    """R = dE/du over the 32 DOFs by complex step."""
    R = np.zeros(32)
    for d in range(32):
        up = np.asarray(ue, complex).copy()
        up[d] += 1j * h
        R[d] = energy(xyz, alpha, abd, up).imag / h
    return R


def condensed_force(xyz, alpha, abd, K32, u24, iters=16):
    """The internal DOFs from the linear recovery u_i = -K_ii^-1 K_ie u_e,
    then `iters` chord steps u_i -= K_ii^-1 R_i(u_e, u_i) with K_ii at rest,
    and R_c = R_e - K_ei K_ii^-1 R_i at the result; returns (R_c, E)."""
    Kie = K32[24:, :24]
    Kii = K32[24:, 24:]
    ui = -np.linalg.solve(Kii, Kie @ u24)
    for _ in range(iters):
        R = internal_force(xyz, alpha, abd, np.concatenate([u24, ui]))
        ui = ui - np.linalg.solve(Kii, R[24:])
    ue = np.concatenate([u24, ui])
    R = internal_force(xyz, alpha, abd, ue)
    Rc = R[:24] - K32[:24, 24:] @ np.linalg.solve(Kii, R[24:])
    return Rc, energy(xyz, alpha, abd, ue)
