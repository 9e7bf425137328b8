"""CPU oracle of the Newmark objective gradient.

Test infrastructure only: the tests import it as the checker of
astro_ctx_dynamic_objective_gradient; nothing in the product imports it.

Restates the dynamic branch of getTotaldLdD
(Working/AStrO_commandFunctions.f90:1416-1464) with augmentdLdu (:819-915)
and solveDynamicAdjoint (:917-956) for the linear Newmark step the device
takes (the tangent K + M / (beta dt^2) at rest, one solve per step,
r_getElRu's acceleration a = c1 (c2 (u - pu - dt pv) - (1 - 2 beta) pa),
c1 = 1 / (2 beta), c2 = 2 / dt^2, Working/AStrO_r_elementEqns.f90:2684-2690):

* J = sum over steps i = 1..N of the objective terms active at i dt;
* backward from step N: dL/du_i, dL/dv_i, dL/da_i gather dObj/du at u_i and,
  from step i + 1, the residual's dependence on the previous state (M lam
  times c1 c2, c1 c2 dt and c1 (1 - 2 beta), which augmentdLdu obtains by
  evaluating the residual with lam in the previous-state slot) with the
  Newmark update's own terms in accAdj and velAdj;
* velAdj = dL/dv, accAdj = -(dL/da + dt gamma velAdj) / (dt^2 beta),
  (K + M / (beta dt^2) + C^T C) lam = dL/du - accAdj;
* dJ/dd = sum_i dObj/dd(u_i) - lam_i . dR_i/dd, dR_i/dd holding the
  stiffness at u_i, the inertia at a_i (consistent mass, r_getElRum
  :2628-2653) and the nodal loads scaled by the step's factor.

The element mass is formed here in numpy (complex-capable for the complex
step).  Pinned (tests/test_newmark_adjoint.py) by central finite differences
of the whole forward-then-sum pipeline on the same linear Newmark steps.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from . import design_oracle as DO


def element_mass(X, rho):
    """Consistent mass (24 x 24, component-major) of one brick8: sum over
    the Gauss points of rho detJ N N^T on each component."""
    M8 = np.zeros((8, 8), dtype=np.result_type(X, rho))
    for s in DO.GAUSS:
        N, _, det = DO.point(X, s)
        M8 = M8 + rho * det * np.outer(N, N)
    M = np.zeros((24, 24), dtype=M8.dtype)
    for c in range(3):
        M[8 * c:8 * c + 8, 8 * c:8 * c + 8] = M8
    return M


class NewmarkAdjointOracle:
    def __init__(self, model, dv, beta=0.25, gamma=0.5, dt=1e-5, h=1e-30):
        self.model, self.dv = model, dv
        self.beta, self.gamma, self.dt = beta, gamma, dt
        self.do = DO.DesignOracle(model, dv, h)
        A0, b0 = self.do._system("elastic")
        n = self.do.n
        rows, cols, vals = [], [], []
        for e in range(model.num_els):
            X, dofs, _ = self.do._el(e, np.zeros(n))
            Me = element_mass(X, model.density[e])
            rows.append(np.repeat(dofs, 24))
            cols.append(np.tile(dofs, 24))
            vals.append(Me.ravel())
        self.M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
        self.Aeff = (A0 + self.M / (beta * dt * dt)).tocsc()
        self.F = b0
        self.n = n

    def forward(self, scales):
        """Linear Newmark from rest: (U, V, A), (N + 1) x n each."""
        b, dt, g = self.beta, self.dt, self.gamma
        c1, c2 = 0.5 / b, 2.0 / dt ** 2
        N = len(scales)
        U, V, A = (np.zeros((N + 1, self.n)) for _ in range(3))
        lu = spl.splu(self.Aeff)
        for k in range(N):
            u, v, a = U[k], V[k], A[k]
            rhs = scales[k] * self.F + self.M @ (c1 * (c2 * (u + dt * v) + (1 - 2 * b) * a))
            un = lu.solve(rhs)
            an = c1 * (c2 * (un - u - dt * v) - (1 - 2 * b) * a)
            U[k + 1], A[k + 1], V[k + 1] = un, an, v + dt * ((1 - g) * a + g * an)
        return U, V, A

    def _active(self, terms, windows, i):
        t = i * self.dt
        return [tm for q, tm in enumerate(terms) if windows is None or windows[q][0] <= t <= windows[q][1]]

    def objective(self, terms, U, windows=None):
        J = 0.0
        for i in range(1, U.shape[0]):
            act = self._active(terms, windows, i)
            if act:
                J += float(np.sum(self.do.objective(act, u=U[i])))
        return J

    def _dRdD(self, u, a, scale):
        do, m, h = self.do, self.model, self.do.h
        do.u = u
        D = do.dRdD()
        cat = [DO.DV_CATS[c] if isinstance(c, str) else int(c) for c in self.dv.category]
        for k, ck in enumerate(cat):
            if ck == 7:
                D[:, k] *= scale
        for e in range(m.num_els):
            X, dofs, _ = do._el(e, u)
            for k, (dep, dcond, drho, dX) in do._directions(e).items():
                if not (drho.any() or dX.any()):
                    continue
                f = element_mass(X + 1j * h * dX, m.density[e] + 1j * h * drho[0]) @ a[dofs]
                D[dofs, k] += f.imag / h
        return D

    def gradient(self, terms, U, A, scales, windows=None):
        b, dt, g = self.beta, self.dt, self.gamma
        c1, c2 = 0.5 / b, 2.0 / dt ** 2
        N = U.shape[0] - 1
        lu = spl.splu(self.Aeff)
        G = np.zeros(len(self.dv.category))
        lam_n = acc_n = vel_n = None
        for i in range(N, 0, -1):
            act = self._active(terms, windows, i)
            self.do.u = U[i]
            if act:
                dLdu, _, gi, _, _ = self.do.explicit_parts(act)
            else:
                dLdu, gi = np.zeros(self.n), np.zeros_like(G)
            dLdv, dLda = np.zeros(self.n), np.zeros(self.n)
            if lam_n is not None:
                Ml = self.M @ lam_n
                dLdu = dLdu + c1 * c2 * Ml + acc_n
                dLdv = dLdv + c1 * c2 * dt * Ml + dt * acc_n + vel_n
                dLda = dLda + c1 * (1 - 2 * b) * Ml + 0.5 * dt * dt * (1 - 2 * b) * acc_n + dt * (1 - g) * vel_n
            vel = dLdv
            acc = -(dLda + dt * g * vel) / (dt * dt * b)
            lam = lu.solve(dLdu - acc)
            G += gi - lam @ self._dRdD(U[i], A[i], scales[i - 1])
            lam_n, acc_n, vel_n = lam, acc, vel
        return G
