// astro_design.hip -- design-variable derivatives, objective terms and the
// adjoint right-hand side on gfx950 (brick8 elements, 2x2x2 Gauss).
//
// Derivatives with respect to design variables are complex steps, as the
// reference's c_ twins take them (Working/AStrO_c_elementEqns.f90,
// Working/AStrO_c_designPropertyFunctions.f90): every property a variable
// moves is perturbed by i h times its coefficient, the element quantity is
// evaluated in complex128 arithmetic and its imaginary part over h is the
// derivative, exact to O(h^2) with no cancellation.  The shape functions,
// the Jacobian inverse by adjugate and the LU inverse of the compliance
// matrix are the same algorithms as the real element kernels, written once
// for a scalar type T (double, or the complex type cx below).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "astro_cx.h"
#include "astro_design.h"

namespace astro {
namespace {

__device__ __forceinline__ double im(cx a) { return a.i; }
__device__ __forceinline__ double im(double) { return 0.0; }
__device__ __forceinline__ double re(cx a) { return a.r; }
__device__ __forceinline__ double re(double a) { return a; }
template <class T>
__device__ __forceinline__ T zero() {
    return T{};
}
template <>
__device__ __forceinline__ double zero<double>() {
    return 0.0;
}
template <>
__device__ __forceinline__ cx zero<cx>() {
    return cx{0.0, 0.0};
}

// Shape-function x-derivatives Nx and detJ of a brick8 with coordinates X at
// the natural point (s0, s1, s2): Working/AStrO_r_elementEqns.f90:27-36,
// 110-135 (Ns), 212-238 (adjugate inverse), 580-620 (Nx = Ns J^-1).
template <class T>
__device__ __forceinline__ void b8_point(double s0, double s1, double s2, const T (&X)[8][3], double (&N)[8],
                                         T (&Nx)[8][3], T& det) {
    double Ns[8][3];
    T J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[i][j] = zero<T>();
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double sg0 = ((n + 1) & 2) ? 1.0 : -1.0;
        const double sg1 = (n & 2) ? 1.0 : -1.0;
        const double sg2 = (n & 4) ? 1.0 : -1.0;
        const double f0 = 1.0 + sg0 * s0, f1 = 1.0 + sg1 * s1, f2 = 1.0 + sg2 * s2;
        N[n] = 0.125 * f0 * f1 * f2;
        Ns[n][0] = 0.125 * sg0 * f1 * f2;
        Ns[n][1] = 0.125 * sg1 * f0 * f2;
        Ns[n][2] = 0.125 * sg2 * f0 * f1;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) J[i][j] = J[i][j] + X[n][i] * Ns[n][j];
    }
    const T a11 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const T a12 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const T a13 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    det = J[0][0] * a11 + J[0][1] * a12 + J[0][2] * a13;
    const T di = 1.0 / det;
    T I[3][3];
    I[0][0] = di * a11;
    I[1][0] = di * a12;
    I[2][0] = di * a13;
    I[0][1] = di * (J[0][2] * J[2][1] - J[0][1] * J[2][2]);
    I[1][1] = di * (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
    I[2][1] = di * (J[0][1] * J[2][0] - J[0][0] * J[2][1]);
    I[0][2] = di * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
    I[1][2] = di * (J[0][2] * J[1][0] - J[0][0] * J[1][2]);
    I[2][2] = di * (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
        for (int j = 0; j < 3; ++j) Nx[n][j] = Ns[n][0] * I[0][j] + Ns[n][1] * I[1][j] + Ns[n][2] * I[2][j];
}

__device__ __forceinline__ double gauss_coord(int gp, int axis) {
    const double g = 0.57735026918962576451;  // 1/sqrt(3); the 2x2x2 weights are 1
    return ((gp >> axis) & 1) ? g : -g;
}

// C = S^-1 from the engineering constants (E1 E2 E3 nu12 nu13 nu23 G12 G13
// G23) by LU without pivoting, the procedure of r_getMaterialStiffness +
// r_GetInvLU (Working/AStrO_r_designPropertyFunctions.f90:58-78, 996-1073).
template <class T>
__device__ void b8_stiffness_from_constants(const T (&ep)[9], T (&C)[6][6]) {
    T S[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int k = 0; k < 6; ++k) S[r][k] = zero<T>();
    S[0][0] = 1.0 / ep[0];
    S[0][1] = -ep[3] / ep[0];
    S[0][2] = -ep[4] / ep[0];
    S[1][1] = 1.0 / ep[1];
    S[1][2] = -ep[5] / ep[1];
    S[2][2] = 1.0 / ep[2];
    S[3][3] = 1.0 / ep[6];
    S[4][4] = 1.0 / ep[7];
    S[5][5] = 1.0 / ep[8];
    S[1][0] = S[0][1];
    S[2][0] = S[0][2];
    S[2][1] = S[1][2];
    // Doolittle LU in place (unit lower L below the diagonal, U on and above)
    for (int i = 0; i < 6; ++i) {
        for (int j = i; j < 6; ++j)
            for (int k = 0; k < i; ++k) S[i][j] = S[i][j] - S[i][k] * S[k][j];
        const T f = 1.0 / S[i][i];
        for (int j = i + 1; j < 6; ++j) {
            for (int k = 0; k < i; ++k) S[j][i] = S[j][i] - S[j][k] * S[k][i];
            S[j][i] = S[j][i] * f;
        }
    }
    for (int c = 0; c < 6; ++c) {
        T z[6], x[6];
        for (int i = 0; i < 6; ++i) {
            z[i] = (i == c) ? zero<T>() + 1.0 : zero<T>();
            for (int k = 0; k < i; ++k) z[i] = z[i] - S[i][k] * z[k];
        }
        for (int i = 5; i >= 0; --i) {
            T dp = zero<T>();
            for (int k = i + 1; k < 6; ++k) dp = dp + S[i][k] * x[k];
            x[i] = (z[i] - dp) / S[i][i];
        }
        for (int i = 0; i < 6; ++i) C[i][c] = x[i];
    }
}

// Green-Lagrange strain (Voigt 11 22 33 12 13 23, engineering shear) of the
// displacement gradient ux[i][j] = du_i/dx_j (r_getNLStrain,
// Working/AStrO_r_elementEqns.f90:643-648).
template <class T>
__device__ __forceinline__ void gl_strain(const T (&ux)[3][3], T (&e)[6]) {
    e[0] = ux[0][0] + 0.5 * (ux[0][0] * ux[0][0] + ux[1][0] * ux[1][0] + ux[2][0] * ux[2][0]);
    e[1] = ux[1][1] + 0.5 * (ux[0][1] * ux[0][1] + ux[1][1] * ux[1][1] + ux[2][1] * ux[2][1]);
    e[2] = ux[2][2] + 0.5 * (ux[0][2] * ux[0][2] + ux[1][2] * ux[1][2] + ux[2][2] * ux[2][2]);
    e[3] = ux[0][1] + ux[1][0] + ux[0][0] * ux[0][1] + ux[1][0] * ux[1][1] + ux[2][0] * ux[2][1];
    e[4] = ux[0][2] + ux[2][0] + ux[0][0] * ux[0][2] + ux[1][0] * ux[1][2] + ux[2][0] * ux[2][2];
    e[5] = ux[1][2] + ux[2][1] + ux[0][2] * ux[0][1] + ux[1][2] * ux[1][1] + ux[2][2] * ux[2][1];
}

template <class T>
__device__ __forceinline__ void grad_u(const double (&ue)[3][8], const T (&Nx)[8][3], T (&ux)[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T a = zero<T>();
#pragma unroll
            for (int n = 0; n < 8; ++n) a = a + ue[i][n] * Nx[n][j];
            ux[i][j] = a;
        }
}

// Element quantity q and volume v of an objective term for one brick8
// (r_getSolidStress at the element centre, r_getElStrainEnergy,
// r_getElMassVol: Working/AStrO_r_elementEqns.f90:924-976, 1715-1744,
// 1377-1432).  comp is 1-based.
template <class T>
__device__ void element_quantity(int quantity, int comp, const T (&X)[8][3], const double (&ue)[3][8],
                                 const T (&C)[6][6], T rho, T& q, T& v) {
    double N[8];
    T Nx[8][3], det;
    v = zero<T>();
    T se = zero<T>();
    for (int gp = 0; gp < 8; ++gp) {
        b8_point(gauss_coord(gp, 0), gauss_coord(gp, 1), gauss_coord(gp, 2), X, N, Nx, det);
        v = v + det;
        if (quantity == OBJQ_STRAIN_ENERGY) {
            T ux[3][3], e[6];
            grad_u(ue, Nx, ux);
            gl_strain(ux, e);
            T w = zero<T>();
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                T s = zero<T>();
#pragma unroll
                for (int k = 0; k < 6; ++k) s = s + C[r][k] * e[k];
                w = w + s * e[r];
            }
            se = se + 0.5 * (w * det);
        }
    }
    if (quantity == OBJQ_VOLUME) {
        q = v;
    } else if (quantity == OBJQ_MASS) {
        q = rho * v;
    } else if (quantity == OBJQ_STRAIN_ENERGY) {
        q = se;
    } else {
        b8_point(0.0, 0.0, 0.0, X, N, Nx, det);
        T ux[3][3], e[6];
        grad_u(ue, Nx, ux);
        gl_strain(ux, e);
        if (quantity == OBJQ_STRAIN) {
            q = e[comp - 1];
        } else {
            T s = zero<T>();
#pragma unroll
            for (int k = 0; k < 6; ++k) s = s + C[comp - 1][k] * e[k];
            q = s;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// dR/dd of the linear static residual for one (element, variable) pair per
// thread: f = Im(K_e(d + i h) u_e) / h with K_e u_e = sum over the Gauss points
// of B^T C B u_e detJ, the coordinates and engineering constants perturbed
// along the pair's direction record (getdRelasticdD's element loop,
// Working/AStrO_commandFunctions.f90:1211-1250; solid branch of c_getElRuk,
// Working/AStrO_c_elementEqns.f90:2424-2529).  lamDot[p] = lambda_e . f
// (lambda may be null); with dRdD (n x nDV column-major) the pair adds f into
// its variable's column at the element's DOFs, so a launch must not hold two
// pairs of one variable whose elements share a node (one colour per launch).
// With acc (a Newmark step's acceleration) the element's inertia
// M_e(d) a_e = sum_g rho detJ N N^T a_e is part of the residual as well
// (r_getElRum, Working/AStrO_r_elementEqns.f90:2628-2653), its density and
// coordinates perturbed with the rest.
__global__ __launch_bounds__(128) void k_dv_pair_force(int nPairs, const int* __restrict__ pairEl,
                                                       const int* __restrict__ pairDV, const double* __restrict__ dir,
                                                       const int* __restrict__ conn, const double* __restrict__ X,
                                                       const double* __restrict__ eprops,
                                                       const int* __restrict__ rank, int nS,
                                                       const double* __restrict__ u, const double* __restrict__ acc,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ lam,
                                                       double h, double* __restrict__ lamDot,
                                                       double* __restrict__ dRdD, int64_t n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPairs) return;
    const int e = pairEl[p];
    const double* dv = dir + (size_t)p * DVD_STRIDE;
    cx xe[8][3];
    double ue[3][8];
    int dof[3][8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int nd = conn[8 * (size_t)e + a] - 1;
        const int r = rank[nd];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xe[a][c] = mkc(X[3 * (size_t)nd + c], h * dv[DVD_X + 3 * a + c]);
            dof[c][a] = r + c * nS - 1;
            ue[c][a] = u[dof[c][a]];
        }
    }
    cx ep[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) ep[k] = mkc(eprops[9 * (size_t)e + k], h * dv[DVD_EP + k]);
    cx C[6][6];
    b8_stiffness_from_constants(ep, C);
    cx f[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 8; ++a) f[c][a] = mkc(0.0, 0.0);
    double ae[3][8];
    if (acc) {
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) ae[c][a] = acc[dof[c][a]];
    }
    const cx rh = mkc(rho ? rho[e] : 0.0, h * dv[DVD_RHO]);
    for (int gp = 0; gp < 8; ++gp) {
        double N[8];
        cx Nx[8][3], det;
        b8_point(gauss_coord(gp, 0), gauss_coord(gp, 1), gauss_coord(gp, 2), xe, N, Nx, det);
        if (acc) {
            const cx rd = rh * det;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double na = 0.0;
#pragma unroll
                for (int b = 0; b < 8; ++b) na += N[b] * ae[c][b];
#pragma unroll
                for (int a = 0; a < 8; ++a) f[c][a] = f[c][a] + (N[a] * na) * rd;
            }
        }
        cx ux[3][3];
        grad_u(ue, Nx, ux);
        const cx eps[6] = {ux[0][0], ux[1][1], ux[2][2], ux[0][1] + ux[1][0], ux[0][2] + ux[2][0],
                           ux[1][2] + ux[2][1]};
        cx sg[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            cx s = mkc(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < 6; ++k) s = s + C[r][k] * eps[k];
            sg[r] = s * det;
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            f[0][a] = f[0][a] + Nx[a][0] * sg[0] + Nx[a][1] * sg[3] + Nx[a][2] * sg[4];
            f[1][a] = f[1][a] + Nx[a][1] * sg[1] + Nx[a][0] * sg[3] + Nx[a][2] * sg[5];
            f[2][a] = f[2][a] + Nx[a][2] * sg[2] + Nx[a][0] * sg[4] + Nx[a][1] * sg[5];
        }
    }
    const double invH = 1.0 / h;
    double ld = 0.0;
    double* col = dRdD ? dRdD + (int64_t)pairDV[p] * n : nullptr;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const double g = f[c][a].i * invH;
            if (lam) ld += lam[dof[c][a]] * g;
            if (col) col[dof[c][a]] += g;
        }
    if (lamDot) lamDot[p] = ld;
}

// dRt/dd of the thermal residual for one (element, variable) pair per thread:
// f = Im(K_t(d + i h) T_e) / h, K_t T_e = sum over the Gauss points of
// Nx k (Nx^T T_e) detJ with the conductivities and coordinates perturbed
// (getdRthermaldD, Working/AStrO_commandFunctions.f90:1016-1066; r_getElRtk,
// Working/AStrO_r_elementEqns.f90:2062-2124).  Thermal DOF = rank(node).
__global__ __launch_bounds__(128) void k_dv_pair_flux(int nPairs, const int* __restrict__ pairEl,
                                                      const int* __restrict__ pairDV, const double* __restrict__ dir,
                                                      const int* __restrict__ conn, const double* __restrict__ X,
                                                      const double* __restrict__ cond, const int* __restrict__ rank,
                                                      const double* __restrict__ T, const double* __restrict__ lam,
                                                      double h, double* __restrict__ lamDot,
                                                      double* __restrict__ dRtdD, int64_t n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPairs) return;
    const int e = pairEl[p];
    const double* dv = dir + (size_t)p * DVD_STRIDE;
    cx xe[8][3];
    double te[8];
    int dof[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int nd = conn[8 * (size_t)e + a] - 1;
        dof[a] = rank[nd] - 1;
        te[a] = T[dof[a]];
#pragma unroll
        for (int c = 0; c < 3; ++c) xe[a][c] = mkc(X[3 * (size_t)nd + c], h * dv[DVD_X + 3 * a + c]);
    }
    cx kv[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) kv[k] = mkc(cond[6 * (size_t)e + k], h * dv[DVD_COND + k]);
    const cx K[3][3] = {{kv[0], kv[3], kv[4]}, {kv[3], kv[1], kv[5]}, {kv[4], kv[5], kv[2]}};
    cx f[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) f[a] = mkc(0.0, 0.0);
    for (int gp = 0; gp < 8; ++gp) {
        double N[8];
        cx Nx[8][3], det;
        b8_point(gauss_coord(gp, 0), gauss_coord(gp, 1), gauss_coord(gp, 2), xe, N, Nx, det);
        cx g[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cx s = mkc(0.0, 0.0);
#pragma unroll
            for (int a = 0; a < 8; ++a) s = s + te[a] * Nx[a][j];
            g[j] = s;
        }
        cx q[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) q[i] = (K[i][0] * g[0] + K[i][1] * g[1] + K[i][2] * g[2]) * det;
#pragma unroll
        for (int a = 0; a < 8; ++a) f[a] = f[a] + Nx[a][0] * q[0] + Nx[a][1] * q[1] + Nx[a][2] * q[2];
    }
    const double invH = 1.0 / h;
    double acc = 0.0;
    double* col = dRtdD ? dRtdD + (int64_t)pairDV[p] * n : nullptr;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double g = f[a].i * invH;
        if (lam) acc += lam[dof[a]] * g;
        if (col) col[dof[a]] += g;
    }
    if (lamDot) lamDot[p] = acc;
}

// ---------------------------------------------------------------------------
// Objective terms.  k_obj_elements: per element of a term's set (0-based ids
// in items) its quantity q, volume v and, for the displacement-dependent
// quantities, dq/du_e (24 values per item, component-major like the element
// DOF table): the Green-Lagrange strain derivative of r_getNLStrain
// (Working/AStrO_r_elementEqns.f90:649-662), dStress = C dStrain, and for the
// strain energy the internal force sum stress . dStrain detJ (getdObjdU,
// Working/AStrO_objective.f90:661-894).  C is the assembly's material table
// times the element scale.
__global__ __launch_bounds__(128) void k_obj_elements(int quantity, int comp, int nItems,
                                                      const int* __restrict__ items, const int* __restrict__ conn,
                                                      const double* __restrict__ X, const int* __restrict__ elMat,
                                                      const double* __restrict__ elScale,
                                                      const double* __restrict__ matC, const double* __restrict__ rho,
                                                      const int* __restrict__ rank, int nS,
                                                      const double* __restrict__ u, double* __restrict__ q,
                                                      double* __restrict__ v, double* __restrict__ dqdu) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nItems) return;
    const int e = items[it];
    double xe[8][3], ue[3][8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int nd = conn[8 * (size_t)e + a] - 1;
        const int r = rank[nd];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xe[a][c] = X[3 * (size_t)nd + c];
            ue[c][a] = u[r + c * nS - 1];
        }
    }
    double C[6][6];
    const double sc = elScale[e];
    const double* Cm = matC + 36 * (size_t)elMat[e];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int k = 0; k < 6; ++k) C[r][k] = sc * Cm[r + 6 * k];
    double qe, ve;
    element_quantity<double>(quantity, comp, xe, ue, C, rho[e], qe, ve);
    q[it] = qe;
    v[it] = ve;
    if (!dqdu || quantity == OBJQ_MASS || quantity == OBJQ_VOLUME) return;
    double d[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 8; ++a) d[c][a] = 0.0;
    const int ng = quantity == OBJQ_STRAIN_ENERGY ? 8 : 1;
    for (int gp = 0; gp < ng; ++gp) {
        double N[8], Nx[8][3], det;
        if (ng == 8)
            b8_point(gauss_coord(gp, 0), gauss_coord(gp, 1), gauss_coord(gp, 2), xe, N, Nx, det);
        else
            b8_point(0.0, 0.0, 0.0, xe, N, Nx, det);
        double ux[3][3], eps[6];
        grad_u(ue, Nx, ux);
        gl_strain(ux, eps);
        // weights of the six strain derivatives: the stress times detJ (strain
        // energy), row comp of C (stress) or the unit vector (strain)
        double wv[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            if (quantity == OBJQ_STRAIN_ENERGY) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 6; ++k) s += C[r][k] * eps[k];
                wv[r] = s * det;
            } else if (quantity == OBJQ_STRESS) {
                wv[r] = C[comp - 1][r];
            } else {
                wv[r] = (r == comp - 1) ? 1.0 : 0.0;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double F0 = (c == 0 ? 1.0 : 0.0) + ux[c][0];
            const double F1 = (c == 1 ? 1.0 : 0.0) + ux[c][1];
            const double F2 = (c == 2 ? 1.0 : 0.0) + ux[c][2];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const double n0 = Nx[a][0], n1 = Nx[a][1], n2 = Nx[a][2];
                d[c][a] += wv[0] * (F0 * n0) + wv[1] * (F1 * n1) + wv[2] * (F2 * n2) + wv[3] * (F0 * n1 + F1 * n0) +
                           wv[4] * (F0 * n2 + F2 * n0) + wv[5] * (F1 * n2 + F2 * n1);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 8; ++a) dqdu[24 * (size_t)it + 8 * c + a] = d[c][a];
}

// Explicit design derivative of an element term's quantity and volume for
// one (element, variable) pair per thread (getdObjdD,
// Working/AStrO_objective.f90:1282-1560: c_getElStress, c_getElMassVol,
// c_getElStrainEnergy with c_dVec(k) + compStep): dq = Im q / h, dv = Im v / h
// at the current displacement, for the pairs pairIdx[0..nItems).
__global__ __launch_bounds__(128) void k_obj_pairs(int quantity, int comp, int nItems,
                                                   const int* __restrict__ pairIdx, const int* __restrict__ pairEl,
                                                   const double* __restrict__ dir, const int* __restrict__ conn,
                                                   const double* __restrict__ X, const double* __restrict__ eprops,
                                                   const double* __restrict__ rho, const int* __restrict__ rank,
                                                   int nS, const double* __restrict__ u, double h,
                                                   double* __restrict__ dq, double* __restrict__ dv) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nItems) return;
    const int p = pairIdx[it];
    const int e = pairEl[p];
    const double* dd = dir + (size_t)p * DVD_STRIDE;
    cx xe[8][3];
    double ue[3][8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int nd = conn[8 * (size_t)e + a] - 1;
        const int r = rank[nd];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xe[a][c] = mkc(X[3 * (size_t)nd + c], h * dd[DVD_X + 3 * a + c]);
            ue[c][a] = u[r + c * nS - 1];
        }
    }
    cx C[6][6];
    if (quantity == OBJQ_STRESS || quantity == OBJQ_STRAIN_ENERGY) {
        cx ep[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) ep[k] = mkc(eprops[9 * (size_t)e + k], h * dd[DVD_EP + k]);
        b8_stiffness_from_constants(ep, C);
    } else {
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int k = 0; k < 6; ++k) C[r][k] = mkc(0.0, 0.0);
    }
    const cx rh = mkc(rho[e], h * dd[DVD_RHO]);
    cx qe, ve;
    element_quantity<cx>(quantity, comp, xe, ue, C, rh, qe, ve);
    dq[it] = qe.i / h;
    dv[it] = ve.i / h;
}

// Node-set term (displacement or temperature, 'powerNorm'): val[i] =
// (x[dof_i] - t_i)^expo and dL/dx[dof_i] += coef expo (x - t)^(expo - 1)
// (Working/AStrO_objective.f90:125-158, 598-634).  tgtMode 0: t = 0, 1: t =
// tgt[0], 2: t = tgt[i].  The host rejects sets that list a node twice, so
// no two threads touch one entry.
__global__ __launch_bounds__(256) void k_obj_nodes(int nItems, const int* __restrict__ dofs,
                                                   const double* __restrict__ tgt, int tgtMode, double coef,
                                                   double expo, const double* __restrict__ x,
                                                   double* __restrict__ val, double* __restrict__ dLdx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nItems) return;
    const int d = dofs[i];
    const double t = tgtMode == 0 ? 0.0 : (tgtMode == 1 ? tgt[0] : tgt[i]);
    const double r = x[d] - t;
    val[i] = pow(r, expo);
    if (dLdx) dLdx[d] += coef * expo * pow(r, expo - 1.0);
}

// dL/du += w_i dq_i/du_e for the items slot[0..count) of one element colour
// (elements of one colour share no node, so the adds do not collide).
__global__ __launch_bounds__(256) void k_obj_scatter_du(int count, const int* __restrict__ slot,
                                                        const int* __restrict__ items, const double* __restrict__ w,
                                                        const double* __restrict__ dqdu,
                                                        const int* __restrict__ conn, const int* __restrict__ rank,
                                                        int nS, double* __restrict__ dLdu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 24) return;
    const int i = slot[t / 24], j = t % 24;
    const int c = j / 8, a = j % 8;
    const int e = items[i];
    const int nd = conn[8 * (size_t)e + a] - 1;
    dLdu[rank[nd] + c * nS - 1] += w[i] * dqdu[24 * (size_t)i + j];
}

// Heat flux and temperature gradient of every brick8 at the natural point
// (s0, s1, s2): Tx = Nx^T T_e, flux = -k Tx (r_getElFlux,
// Working/AStrO_r_elementEqns.f90:1746-1799).  Thermal DOF = rank(node).
__global__ __launch_bounds__(256) void k_brick8_flux(int E, const int* __restrict__ conn, const double* __restrict__ X,
                                                     const double* __restrict__ cond, const int* __restrict__ rank,
                                                     const double* __restrict__ T, double s0, double s1, double s2,
                                                     double* __restrict__ flux, double* __restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double xe[8][3], te[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int nd = conn[8 * (size_t)e + a] - 1;
        te[a] = T[rank[nd] - 1];
#pragma unroll
        for (int c = 0; c < 3; ++c) xe[a][c] = X[3 * (size_t)nd + c];
    }
    double N[8], Nx[8][3], det;
    b8_point(s0, s1, s2, xe, N, Nx, det);
    double g[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) acc += Nx[a][j] * te[a];
        g[j] = acc;
    }
    const double* k = cond + 6 * (size_t)e;
    const double K[3][3] = {{k[0], k[3], k[4]}, {k[3], k[1], k[5]}, {k[4], k[5], k[2]}};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        flux[3 * (size_t)e + i] = -(K[i][0] * g[0] + K[i][1] * g[1] + K[i][2] * g[2]);
        grad[3 * (size_t)e + i] = g[i];
    }
}

}  // namespace astro
