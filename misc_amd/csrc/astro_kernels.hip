// astro_kernels.hip -- device kernels of the AStrO hot path for gfx950 (CDNA4).
//
// All arithmetic is FP64.  Index conventions inside the kernels are 0-based;
// the reference's 1-based column numbers are converted where CSR columns are
// read.  A row-skyline profile row r covers columns s[r]..r and its entry
// (r, c) lives at lt[rowoff[r] + c].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "astro_cx.h"
#include "astro_dd.h"
#include "astro_kernels.h"

namespace astro {

// Values not yet produced in the current sweep hold this signalling-NaN bit
// pattern; arithmetic never yields it.
#define ASTRO_SENTINEL 0x7ff4d1c3a5b6c7e9LL

// A brick8 with nodal coordinates X at the natural point (s0, s1, s2): shape
// functions N, their x-derivatives Nx and detJ; brick8_gauss_point takes Gauss
// point gp (0..7, signs of the natural coordinates from its bits; the 2x2x2
// weights are 1).  Working/AStrO_r_elementEqns.f90:27-36,
// 110-135 (N, Ns), 212-238 (adjugate inverse), 580-620 (Nx = Ns J^-1).
__device__ __forceinline__ void brick8_point(double s0, double s1, double s2, const double (&sX)[8][3],
                                             double (&N)[8], double (&Nx)[8][3], double& w) {
    double Ns[8][3];
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double sg0 = ((n + 1) & 2) ? 1.0 : -1.0;  // nodes 2,3,6,7 at +s1
        const double sg1 = (n & 2) ? 1.0 : -1.0;        // nodes 3,4,7,8 at +s2
        const double sg2 = (n & 4) ? 1.0 : -1.0;        // nodes 5-8 at +s3
        const double f0 = 1.0 + sg0 * s0, f1 = 1.0 + sg1 * s1, f2 = 1.0 + sg2 * s2;
        N[n] = 0.125 * f0 * f1 * f2;
        Ns[n][0] = 0.125 * sg0 * f1 * f2;
        Ns[n][1] = 0.125 * sg1 * f0 * f2;
        Ns[n][2] = 0.125 * sg2 * f0 * f1;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) J[i][j] += sX[n][i] * Ns[n][j];
    }
    const double a11 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double a12 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double a13 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * a11 + J[0][1] * a12 + J[0][2] * a13;
    const double di = 1.0 / det;
    double I[3][3];
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
    w = det;
}
__device__ __forceinline__ void brick8_gauss_point(int gp, const double (&sX)[8][3], double (&N)[8],
                                                   double (&Nx)[8][3], double& w) {
    const double g = 0.57735026918962576451;  // 1/sqrt(3)
    brick8_point((gp & 1) ? g : -g, (gp & 2) ? g : -g, (gp & 4) ? g : -g, sX, N, Nx, w);  // weight 1
}

// ---------------------------------------------------------------------------
// Element tangents: brick8, 2x2x2 Gauss, B^T C B (+ massFactor * rho N^T N),
// one 64-lane wavefront per element, lane (a, b) owning the 3x3 block of node
// pair (a, b).  Working/AStrO_r_elementEqns.f90:2206-2533 (solid branch at
// u = 0) with shape functions :27-36, :110-135, Jacobian :580-620, B columns of
// r_getNLStrain :649-662 and the mass of r_getElRum :2628-2653.
//
// Elements of one launch share no node (colour classes), so the scatter into
// the CSR values is a plain read-modify-write.  The slot of column (node b,
// component d) in the rows of node a is elPos[e][a][b][d]; rows of node a start
// at csrRange[dof(a, c) - 1] with dof(a, c) = rank(a) + c * nS.
__global__ __launch_bounds__(64) void k_brick8_tangent(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const int* __restrict__ elMat, const double* __restrict__ elScale, const double* __restrict__ matC,
    const double* __restrict__ elRho, double massFactor, const int* __restrict__ rank, int nS,
    const int* __restrict__ csrRange, const uint8_t* __restrict__ elPos, double* __restrict__ val) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ double sX[8][3];
    __shared__ double sNx[8][8][3];
    __shared__ double sN[8][8];
    __shared__ double sW[8];
    __shared__ int sNode[8];
    if (lane < 8) {
        int nd = conn[8 * (size_t)e + lane] - 1;
        sNode[lane] = nd;
        sX[lane][0] = X[3 * (size_t)nd];
        sX[lane][1] = X[3 * (size_t)nd + 1];
        sX[lane][2] = X[3 * (size_t)nd + 2];
    }
    __syncthreads();
    if (lane < 8) brick8_gauss_point(lane, sX, sN[lane], sNx[lane], sW[lane]);
    __syncthreads();
    const int a = lane >> 3, b = lane & 7;
    const double sc = elScale[e];
    const double* Cm = matC + 36 * (size_t)elMat[e];
    double C[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) C[r][c] = sc * Cm[r + 6 * c];
    double K[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double mf = massFactor * elRho[e];
#pragma unroll 2
    for (int g = 0; g < 8; ++g) {
        const double w = sW[g];
        const double ax = sNx[g][a][0], ay = sNx[g][a][1], az = sNx[g][a][2];
        const double bx = sNx[g][b][0], by = sNx[g][b][1], bz = sNx[g][b][2];
        double D[6][3];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            D[r][0] = C[r][0] * bx + C[r][3] * by + C[r][4] * bz;
            D[r][1] = C[r][1] * by + C[r][3] * bx + C[r][5] * bz;
            D[r][2] = C[r][2] * bz + C[r][4] * bx + C[r][5] * by;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            K[0][d] += w * (ax * D[0][d] + ay * D[3][d] + az * D[4][d]);
            K[1][d] += w * (ay * D[1][d] + ax * D[3][d] + az * D[5][d]);
            K[2][d] += w * (az * D[2][d] + ax * D[4][d] + ay * D[5][d]);
        }
        if (mf != 0.0) {
            const double m = mf * w * sN[g][a] * sN[g][b];
            K[0][0] += m;
            K[1][1] += m;
            K[2][2] += m;
        }
    }
    const int ra = rank[sNode[a]];
    const uint8_t* pos = elPos + 192 * (size_t)e + 3 * lane;
    const int p0 = pos[0], p1 = pos[1], p2 = pos[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int row = ra + c * nS;  // 1-based DOF
        double* v = val + csrRange[row - 1];
        v[p0] += K[c][0];
        v[p1] += K[c][1];
        v[p2] += K[c][2];
    }
}

// E -= T alpha at a Gauss point with shape functions N: the thermal strain
// of the element's expansion tExp[6 e..] at the interpolated nodal
// temperature (T by rank; tExp null: none), stress = C (E - T alpha) as in
// Working/AStrO_r_elementEqns.f90:2433-2447.
__device__ __forceinline__ void brick8_subtract_thermal(int e, const double (&N)[8], const int (&sNode)[8],
                                                        const int* __restrict__ rank, const double* __restrict__ tExp,
                                                        const double* __restrict__ T, double (&E)[6]) {
    if (!tExp) return;
    double Tp = 0.0;
    for (int n = 0; n < 8; ++n) Tp += N[n] * T[rank[sNode[n]] - 1];
    for (int k = 0; k < 6; ++k) E[k] -= Tp * tExp[6 * (size_t)e + k];
}

// Green-Lagrange tangent of brick8 elements at the displacement u, the
// derivative of k_brick8_residual's internal force (the solid branch of
// r_getElRuk with the strain of r_getNLStrain, Working/AStrO_r_elementEqns.f90:
// 2206-2533, 622-680, which getElasticSolnLoad(1) assembles at every Newton
// step with nonLinearGeom): per Gauss point, with F = I + grad u and
// S = C E, the material part dE/du^T C dE/du, dE/du_(a,c) = sym(F^T e_c
// grad N_a) in engineering Voigt order (11, 22, 33, 12, 13, 23), and the
// geometric part grad N_a . S grad N_b on the diagonal of each 3x3 block.  At
// u = 0 it is k_brick8_tangent's B^T C B.  With tExp (may be null) S is
// C (E - T alpha), as in k_brick8_residual.  Lane (a, b) owns the 3x3 block of
// node pair (a, b), scattered as k_brick8_tangent's.
__global__ __launch_bounds__(64) void k_brick8_tangent_gl(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const int* __restrict__ elMat, const double* __restrict__ elScale, const double* __restrict__ matC,
    const int* __restrict__ rank, int nS, const double* __restrict__ u, const double* __restrict__ tExp,
    const double* __restrict__ T, const int* __restrict__ csrRange, const uint8_t* __restrict__ elPos,
    double* __restrict__ val) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ double sX[8][3];
    __shared__ double sNx[8][8][3];
    __shared__ double sN[8][8];
    __shared__ double sW[8];
    __shared__ double sU[3][8], sF[8][3][3], sS[8][6];
    __shared__ int sNode[8];
    if (lane < 8) {
        const int nd = conn[8 * (size_t)e + lane] - 1;
        sNode[lane] = nd;
        for (int c = 0; c < 3; ++c) sX[lane][c] = X[3 * (size_t)nd + c];
    }
    __syncthreads();
    if (lane < 24) {
        const int comp = lane >> 3, nd = lane & 7;
        sU[comp][nd] = u[rank[sNode[nd]] + comp * nS - 1];
    }
    if (lane < 8) brick8_gauss_point(lane, sX, sN[lane], sNx[lane], sW[lane]);
    __syncthreads();
    const double sc = elScale[e];
    const double* Cm = matC + 36 * (size_t)elMat[e];
    if (lane < 8) {
        const int g = lane;
        double ux[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
                for (int n = 0; n < 8; ++n) acc += sU[i][n] * sNx[g][n][j];
                ux[i][j] = acc;
                sF[g][i][j] = (i == j ? 1.0 : 0.0) + acc;
            }
        double E[6];
        E[0] = ux[0][0] + 0.5 * (ux[0][0] * ux[0][0] + ux[1][0] * ux[1][0] + ux[2][0] * ux[2][0]);
        E[1] = ux[1][1] + 0.5 * (ux[0][1] * ux[0][1] + ux[1][1] * ux[1][1] + ux[2][1] * ux[2][1]);
        E[2] = ux[2][2] + 0.5 * (ux[0][2] * ux[0][2] + ux[1][2] * ux[1][2] + ux[2][2] * ux[2][2]);
        E[3] = ux[0][1] + ux[1][0] + ux[0][0] * ux[0][1] + ux[1][0] * ux[1][1] + ux[2][0] * ux[2][1];
        E[4] = ux[0][2] + ux[2][0] + ux[0][0] * ux[0][2] + ux[1][0] * ux[1][2] + ux[2][0] * ux[2][2];
        E[5] = ux[1][2] + ux[2][1] + ux[0][2] * ux[0][1] + ux[1][2] * ux[1][1] + ux[2][2] * ux[2][1];
        brick8_subtract_thermal(e, sN[g], sNode, rank, tExp, T, E);
        for (int r = 0; r < 6; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 6; ++k) acc += Cm[r + 6 * k] * E[k];
            sS[g][r] = sc * acc;
        }
    }
    __syncthreads();
    const int a = lane >> 3, b = lane & 7;
    double C[6][6];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) C[r][c] = sc * Cm[r + 6 * c];
    double K[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int g = 0; g < 8; ++g) {
        const double w = sW[g];
        const double* xa = sNx[g][a];
        const double* xb = sNx[g][b];
        double Ba[3][6], Bb[3][6];
        for (int c = 0; c < 3; ++c) {
            const double f0 = sF[g][c][0], f1 = sF[g][c][1], f2 = sF[g][c][2];
            Ba[c][0] = xa[0] * f0;
            Ba[c][1] = xa[1] * f1;
            Ba[c][2] = xa[2] * f2;
            Ba[c][3] = xa[1] * f0 + xa[0] * f1;
            Ba[c][4] = xa[2] * f0 + xa[0] * f2;
            Ba[c][5] = xa[2] * f1 + xa[1] * f2;
            Bb[c][0] = xb[0] * f0;
            Bb[c][1] = xb[1] * f1;
            Bb[c][2] = xb[2] * f2;
            Bb[c][3] = xb[1] * f0 + xb[0] * f1;
            Bb[c][4] = xb[2] * f0 + xb[0] * f2;
            Bb[c][5] = xb[2] * f1 + xb[1] * f2;
        }
        for (int d = 0; d < 3; ++d) {
            double CB[6];
            for (int r = 0; r < 6; ++r) {
                double acc = 0.0;
                for (int k = 0; k < 6; ++k) acc += C[r][k] * Bb[d][k];
                CB[r] = acc;
            }
            for (int c = 0; c < 3; ++c) {
                double acc = 0.0;
                for (int r = 0; r < 6; ++r) acc += Ba[c][r] * CB[r];
                K[c][d] += w * acc;
            }
        }
        const double* S = sS[g];
        const double geo = S[0] * xa[0] * xb[0] + S[1] * xa[1] * xb[1] + S[2] * xa[2] * xb[2] +
                           S[3] * (xa[0] * xb[1] + xa[1] * xb[0]) + S[4] * (xa[0] * xb[2] + xa[2] * xb[0]) +
                           S[5] * (xa[1] * xb[2] + xa[2] * xb[1]);
        for (int c = 0; c < 3; ++c) K[c][c] += w * geo;
    }
    const int ra = rank[sNode[a]];
    const uint8_t* pos = elPos + 192 * (size_t)e + 3 * lane;
    const int p0 = pos[0], p1 = pos[1], p2 = pos[2];
    for (int c = 0; c < 3; ++c) {
        double* v = val + csrRange[ra + c * nS - 1];
        v[p0] += K[c][0];
        v[p1] += K[c][1];
        v[p2] += K[c][2];
    }
}

// ---------------------------------------------------------------------------
// Shell4 tangent at rest with static condensation, one 64-lane workgroup per
// element (colour classes: elements of one launch share no node).
// Working/AStrO_r_elementEqns.f90:2245-2423 and 2478-2494 (sum_gp detJ w
// dedU^T ABD dedU), generalised strains r_getShellDef :800-814, instantaneous
// DOFs r_getInstDof :703-891 linearised at rest with exact derivatives
// (u_L(a) = alpha (u_a - rbar x X_a), r_L(a) = alpha (r_a - rbar), rbar the
// mean nodal rotation), shape functions :44-55 / 143-162 (4 bilinear + 6
// bubbles carrying the 8 internal DOFs), condensation r_condenseMat
// :2722-2743.  DOF order of the element: (var, node) var-major, var 1..6 on
// nodes 1..4, then internal (1,5) (1,6) (2,5) (2,6) (3,7) (3,8) (3,9) (3,10).
// Rows of node a start at csrRange[dof(a, v) - 1], dof(a, v) = rank(a) +
// v * nSec + 2 * nSolid; the slot of (node b, var w) is elPos[e][(a*4+b)*6+w].
// Shared stage of the shell4 kernels: node ids and coordinates, frame, shape
// functions at the 4 Gauss points, and the 9 x 32 strain-displacement matrix B
// per Gauss point (lane = DOF, two Gauss points per lane half).  Leaves sB,
// sN, sNx, sW, sX, sAl, sNode in LDS; ends with a barrier.
__device__ __forceinline__ void shell4_stage_B(int e, int lane, const int* __restrict__ conn,
                                               const double* __restrict__ X, const double* __restrict__ elAlpha,
                                               double (&sX)[4][3], double (&sAl)[3][3], double (&sN)[4][10],
                                               double (&sNx)[4][10][2], double (&sW)[4], double (&sB)[4][9][32],
                                               int (&sNode)[4]) {
    if (lane < 4) {
        const int nd = conn[8 * (size_t)e + lane] - 1;
        sNode[lane] = nd;
        for (int c = 0; c < 3; ++c) sX[lane][c] = X[3 * (size_t)nd + c];
    }
    if (lane < 9) sAl[lane % 3][lane / 3] = elAlpha[9 * (size_t)e + lane];
    __syncthreads();
    if (lane < 4) {
        // Gauss point g: natural coordinates (+-1/sqrt3, +-1/sqrt3), bit 0 -> s1
        const int g = lane;
        const double q = 0.57735026918962576451;
        const double s1 = (g & 1) ? q : -q, s2 = (g & 2) ? q : -q;
        double N[10], Ns[10][2];
        N[0] = 0.25 * (1 - s1) * (1 - s2);
        N[1] = 0.25 * (1 + s1) * (1 - s2);
        N[2] = 0.25 * (1 + s1) * (1 + s2);
        N[3] = 0.25 * (1 - s1) * (1 + s2);
        N[4] = 1 - s1 * s1;
        N[5] = 1 - s2 * s2;
        N[6] = (1 - s1 * s1) * (1 - s2);
        N[7] = (1 - s2 * s2) * (1 + s1);
        N[8] = (1 - s1 * s1) * (1 + s2);
        N[9] = (1 - s2 * s2) * (1 - s1);
        Ns[0][0] = -0.25 * (1 - s2); Ns[0][1] = -0.25 * (1 - s1);
        Ns[1][0] = 0.25 * (1 - s2);  Ns[1][1] = -0.25 * (1 + s1);
        Ns[2][0] = 0.25 * (1 + s2);  Ns[2][1] = 0.25 * (1 + s1);
        Ns[3][0] = -0.25 * (1 + s2); Ns[3][1] = 0.25 * (1 - s1);
        Ns[4][0] = -2 * s1;          Ns[4][1] = 0.0;
        Ns[5][0] = 0.0;              Ns[5][1] = -2 * s2;
        Ns[6][0] = -2 * s1 * (1 - s2); Ns[6][1] = -(1 - s1 * s1);
        Ns[7][0] = 1 - s2 * s2;        Ns[7][1] = -2 * s2 * (1 + s1);
        Ns[8][0] = -2 * s1 * (1 + s2); Ns[8][1] = 1 - s1 * s1;
        Ns[9][0] = -(1 - s2 * s2);     Ns[9][1] = -2 * s2 * (1 - s1);
        // dx/ds of the local coordinates alpha X, third column e3 (flat element)
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 1}};
        for (int a = 0; a < 4; ++a) {
            double loc[3];
            for (int i = 0; i < 3; ++i) loc[i] = sAl[i][0] * sX[a][0] + sAl[i][1] * sX[a][1] + sAl[i][2] * sX[a][2];
            for (int i = 0; i < 3; ++i) {
                J[i][0] += loc[i] * Ns[a][0];
                J[i][1] += loc[i] * Ns[a][1];
            }
        }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double di = 1.0 / det;
        const double I00 = di * c00, I10 = di * c01;
        const double I01 = di * (J[0][2] * J[2][1] - J[0][1] * J[2][2]);
        const double I11 = di * (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
        for (int n = 0; n < 10; ++n) {
            sN[g][n] = N[n];
            sNx[g][n][0] = Ns[n][0] * I00 + Ns[n][1] * I10;
            sNx[g][n][1] = Ns[n][0] * I01 + Ns[n][1] * I11;
        }
        sW[g] = det;
    }
    __syncthreads();
    // B: lane = DOF d (0..31) for Gauss points 2 (lane >> 5) and 2 (lane >> 5) + 1
    {
        const int d = lane & 31;
        for (int g = 2 * (lane >> 5); g < 2 * (lane >> 5) + 2; ++g) {
            double ux[3][2] = {{0, 0}, {0, 0}, {0, 0}}, rx[3][2] = {{0, 0}, {0, 0}, {0, 0}}, rot[3] = {0, 0, 0};
            if (d >= 24) {
                const int k = d - 24;
                const int var = k < 2 ? 0 : (k < 4 ? 1 : 2);
                const int fn = k < 4 ? 4 + (k & 1) : 6 + (k - 4);
                ux[var][0] = sNx[g][fn][0];
                ux[var][1] = sNx[g][fn][1];
            } else if (d < 12) {
                const int var = d >> 2, nd = d & 3;
                for (int i = 0; i < 3; ++i) {
                    ux[i][0] = sAl[i][var] * sNx[g][nd][0];
                    ux[i][1] = sAl[i][var] * sNx[g][nd][1];
                }
            } else {
                const int q = (d >> 2) - 3, nd = d & 3;
                for (int a = 0; a < 4; ++a) {
                    // -(e_q x X_a) / 4 in local components
                    double ex[3] = {0, 0, 0};
                    if (q == 0) { ex[1] = -sX[a][2]; ex[2] = sX[a][1]; }
                    if (q == 1) { ex[0] = sX[a][2]; ex[2] = -sX[a][0]; }
                    if (q == 2) { ex[0] = -sX[a][1]; ex[1] = sX[a][0]; }
                    const double f = (a == nd ? 1.0 : 0.0) - 0.25;
                    for (int i = 0; i < 3; ++i) {
                        const double v = -0.25 * (sAl[i][0] * ex[0] + sAl[i][1] * ex[1] + sAl[i][2] * ex[2]);
                        ux[i][0] += v * sNx[g][a][0];
                        ux[i][1] += v * sNx[g][a][1];
                        rx[i][0] += sAl[i][q] * f * sNx[g][a][0];
                        rx[i][1] += sAl[i][q] * f * sNx[g][a][1];
                        rot[i] += sAl[i][q] * f * sN[g][a];
                    }
                }
            }
            sB[g][0][d] = ux[0][0];
            sB[g][1][d] = ux[1][1];
            sB[g][2][d] = ux[0][1] + ux[1][0];
            sB[g][3][d] = rx[1][0];
            sB[g][4][d] = -rx[0][1];
            sB[g][5][d] = rx[1][1] - rx[0][0];
            sB[g][6][d] = ux[2][0] + rot[1];
            sB[g][7][d] = ux[2][1] - rot[0];
            sB[g][8][d] = 2.0 * rot[2] - ux[1][0] + ux[0][1];
        }
    }
    __syncthreads();
}

// LU (no pivoting) in place of the internal block of the element matrix,
// rows and columns 24..31 (r_condenseMat, Working/AStrO_r_elementEqns.f90:
// 2722-2743).  One lane.
__device__ __forceinline__ void shell4_lu_internal(double (&sK)[32][33]) {
    for (int i = 24; i < 32; ++i) {
        for (int j = i; j < 32; ++j)
            for (int k = 24; k < i; ++k) sK[i][j] -= sK[i][k] * sK[k][j];
        const double f = 1.0 / sK[i][i];
        for (int j = i + 1; j < 32; ++j) {
            for (int k = 24; k < i; ++k) sK[j][i] -= sK[j][k] * sK[k][i];
            sK[j][i] *= f;
        }
    }
}

// x = K_ii^-1 rhs with the factors of shell4_lu_internal
__device__ __forceinline__ void shell4_solve_internal(const double (&sK)[32][33], const double* rhs, double* x) {
    double z[8];
    for (int i = 0; i < 8; ++i) {
        double v = rhs[i];
        for (int k = 0; k < i; ++k) v -= sK[24 + i][24 + k] * z[k];
        z[i] = v;
    }
    for (int i = 7; i >= 0; --i) {
        double v = z[i];
        for (int k = i + 1; k < 8; ++k) v -= sK[24 + i][24 + k] * x[k];
        x[i] = v / sK[24 + i][24 + i];
    }
}

// The 24 x 24 condensed matrix K_ee - K_ei K_ii^-1 K_ie of the 32 x 32
// element matrix sK (its internal block factored in place) added into the CSR
// values at the element's slots.  sXs: 24 x 8 scratch.  Starts with a barrier.
__device__ __forceinline__ void shell4_condense_scatter(int e, int lane, double (&sK)[32][33], double (&sXs)[24][8],
                                                        const int (&sNode)[4], const int* __restrict__ rank,
                                                        int nSolid, int nSec, const int* __restrict__ csrRange,
                                                        const uint8_t* __restrict__ elPos, double* __restrict__ val) {
    __syncthreads();
    if (lane == 0) shell4_lu_internal(sK);
    __syncthreads();
    if (lane < 24) {
        // x = K_ii^-1 K_i,c for nodal column c = lane
        double rhs[8], x[8];
        for (int i = 0; i < 8; ++i) rhs[i] = sK[24 + i][lane];
        shell4_solve_internal(sK, rhs, x);
        for (int i = 0; i < 8; ++i) sXs[lane][i] = x[i];
    }
    __syncthreads();
    for (int k = lane; k < 24 * 24; k += 64) {
        const int i = k / 24, j = k % 24;
        double v = sK[i][j];
#pragma unroll
        for (int m = 0; m < 8; ++m) v -= sK[i][24 + m] * sXs[j][m];
        const int vi = i >> 2, a = i & 3, vj = j >> 2, b = j & 3;
        const int row = rank[sNode[a]] + vi * nSec + 2 * nSolid;
        const int slot = elPos[192 * (size_t)e + (a * 4 + b) * 6 + vj];
        val[csrRange[row - 1] + slot] += v;
    }
}

__global__ __launch_bounds__(64) void k_shell4_tangent(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ elABD, const double* __restrict__ elAlpha, const int* __restrict__ rank, int nSolid,
    int nSec, const int* __restrict__ csrRange, const uint8_t* __restrict__ elPos, double* __restrict__ val) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ double sX[4][3], sAl[3][3], sD[9][9];
    __shared__ double sN[4][10], sNx[4][10][2], sW[4];
    __shared__ double sB[4][9][32], sDB[4][9][32];
    __shared__ double sK[32][33];
    __shared__ double sXs[24][8];
    __shared__ int sNode[4];
    for (int k = lane; k < 81; k += 64) sD[k % 9][k / 9] = elABD[81 * (size_t)e + k];
    shell4_stage_B(e, lane, conn, X, elAlpha, sX, sAl, sN, sNx, sW, sB, sNode);
    __syncthreads();
    for (int k = lane; k < 4 * 9 * 32; k += 64) {
        const int g = k / 288, r = (k / 32) % 9, d = k % 32;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 9; ++m) acc += sD[r][m] * sB[g][m][d];
        sDB[g][r][d] = sW[g] * acc;
    }
    __syncthreads();
    for (int k = lane; k < 32 * 32; k += 64) {
        const int i = k >> 5, j = k & 31;
        double acc = 0.0;
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int m = 0; m < 9; ++m) acc += sB[g][m][i] * sDB[g][m][j];
        sK[i][j] = acc;
    }
    shell4_condense_scatter(e, lane, sK, sXs, sNode, rank, nSolid, nSec, csrRange, elPos, val);
}

// ---------------------------------------------------------------------------
// Co-rotational shell4 strains at the four Gauss points for the 32 element
// DOFs of sU (24 corner DOFs component-major u v w tx ty tz over the 4 nodes,
// 8 internal) with DOF `pert` (< 0: none) carrying the imaginary part h:
// Working/AStrO_r_elementEqns.f90:681-891 (r_getShellDef, r_getInstOrient,
// r_getInstDof) with the finite rotation of r_rotateAlpha
// (Working/AStrO_r_designPropertyFunctions.f90:1075-1463).  The frame
// follows the mean nodal rotation t: instOri = alpha R(t)^T with
// R = I + a K + b K^2, K = [t]x, a = sin r / r, b = (1 - cos r) / r^2, r^2 = s =
// t.t, as power series in s (so R and its first and second derivatives are
// analytic in the DOFs; exact sin / cos, the reference's r_sin / r_cos are
// not reproduced, D1); local displacements instOri (u + X) - alpha X, local
// rotations drIdrG (theta - t) + 1/2 d2rIdrG (theta - t)^2 with
// drIdrG(:, i) = axial(d instOri / dt_i instOri^T).
__device__ void shell4_corot_strains(int pert, double h, const double (&sU)[32], const double (&sX)[4][3],
                                     const double (&sAl)[3][3], const double (&sN)[4][10],
                                     const double (&sNx)[4][10][2], cx (&ep)[4][9]) {
    cx ue[32];
    for (int k = 0; k < 32; ++k) ue[k] = mkc(sU[k], k == pert ? h : 0.0);
    cx tb[3];
    for (int c = 0; c < 3; ++c)
        tb[c] = 0.25 * (ue[12 + 4 * c] + ue[12 + 4 * c + 1] + ue[12 + 4 * c + 2] + ue[12 + 4 * c + 3]);
    const cx s = tb[0] * tb[0] + tb[1] * tb[1] + tb[2] * tb[2];
    // a, b, a', b', a'', b'' as power series in s (24 terms)
    cx a = mkc(0, 0), b = mkc(0, 0), da = mkc(0, 0), db = mkc(0, 0), d2a = mkc(0, 0), d2b = mkc(0, 0);
    cx pk = mkc(1, 0), pk1 = mkc(0, 0), pk2 = mkc(0, 0);
    double ca = 1.0, cb = 0.5;
    for (int k = 0; k < 24; ++k) {
        a = a + ca * pk;
        b = b + cb * pk;
        da = da + (k * ca) * pk1;
        db = db + (k * cb) * pk1;
        d2a = d2a + (k * (k - 1) * ca) * pk2;
        d2b = d2b + (k * (k - 1) * cb) * pk2;
        pk2 = pk1;
        pk1 = pk;
        pk = pk * s;
        ca *= -1.0 / ((2.0 * k + 2.0) * (2.0 * k + 3.0));
        cb *= -1.0 / ((2.0 * k + 3.0) * (2.0 * k + 4.0));
    }
    auto skew = [](const cx (&v)[3], cx (&M)[3][3]) {
        const cx z = mkc(0, 0);
        M[0][0] = z; M[0][1] = -v[2]; M[0][2] = v[1];
        M[1][0] = v[2]; M[1][1] = z; M[1][2] = -v[0];
        M[2][0] = -v[1]; M[2][1] = v[0]; M[2][2] = z;
    };
    auto mm = [](const cx (&A)[3][3], const cx (&B)[3][3], cx (&C)[3][3]) {
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) C[p][q] = A[p][0] * B[0][q] + A[p][1] * B[1][q] + A[p][2] * B[2][q];
    };
    cx K[3][3], K2[3][3], Em[3][3][3];
    skew(tb, K);
    mm(K, K, K2);
    for (int i = 0; i < 3; ++i) {
        cx ev[3] = {mkc(i == 0, 0), mkc(i == 1, 0), mkc(i == 2, 0)};
        skew(ev, Em[i]);
    }
    // instOri = alpha R^T
    cx I[3][3];
    {
        cx R[3][3];
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) R[p][q] = a * K[p][q] + b * K2[p][q] + (p == q ? 1.0 : 0.0);
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q)
                I[p][q] = sAl[p][0] * R[q][0] + sAl[p][1] * R[q][1] + sAl[p][2] * R[q][2];
    }
    // axial(alpha M^T I^T) for a derivative M of R
    auto axial_of = [&](const cx (&M)[3][3], cx (&w)[3]) {
        cx A[3][3];  // alpha M^T
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) A[p][q] = sAl[p][0] * M[q][0] + sAl[p][1] * M[q][1] + sAl[p][2] * M[q][2];
        auto P = [&](int p, int q) { return A[p][0] * I[q][0] + A[p][1] * I[q][1] + A[p][2] * I[q][2]; };
        w[0] = P(1, 2);
        w[1] = P(2, 0);
        w[2] = P(0, 1);
    };
    cx d1[3][3], d2[3][3][3];
    for (int i = 0; i < 3; ++i) {
        cx M[3][3], EK[3][3], KE[3][3];
        mm(Em[i], K, EK);
        mm(K, Em[i], KE);
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q)
                M[p][q] = (2.0 * tb[i]) * (da * K[p][q] + db * K2[p][q]) + a * Em[i][p][q] + b * (EK[p][q] + KE[p][q]);
        cx w[3];
        axial_of(M, w);
        for (int c = 0; c < 3; ++c) d1[c][i] = w[c];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
            cx M[3][3], EiK[3][3], KEi[3][3], EjK[3][3], KEj[3][3], EiEj[3][3], EjEi[3][3];
            mm(Em[i], K, EiK);
            mm(K, Em[i], KEi);
            mm(Em[j], K, EjK);
            mm(K, Em[j], KEj);
            mm(Em[i], Em[j], EiEj);
            mm(Em[j], Em[i], EjEi);
            const double dij = i == j ? 1.0 : 0.0;
            const cx fK = 4.0 * (d2a * tb[i] * tb[j]) + (2.0 * dij) * da;
            const cx fK2 = 4.0 * (d2b * tb[i] * tb[j]) + (2.0 * dij) * db;
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q)
                    M[p][q] = fK * K[p][q] + (2.0 * da) * (tb[i] * Em[j][p][q] + tb[j] * Em[i][p][q]) + fK2 * K2[p][q] +
                              (2.0 * db) * (tb[i] * (EjK[p][q] + KEj[p][q]) + tb[j] * (EiK[p][q] + KEi[p][q])) +
                              b * (EiEj[p][q] + EjEi[p][q]);
            cx w[3];
            axial_of(M, w);
            for (int c = 0; c < 3; ++c) {
                d2[c][i][j] = w[c];
                d2[c][j][i] = w[c];
            }
        }
    // local displacements (corner nodes and bubbles) and rotations
    cx disp[3][10], rot[3][4];
    for (int c = 0; c < 3; ++c)
        for (int n = 4; n < 10; ++n) disp[c][n] = mkc(0, 0);
    for (int nd = 0; nd < 4; ++nd) {
        cx ux[3], th[3];
        for (int c = 0; c < 3; ++c) {
            ux[c] = ue[4 * c + nd] + sX[nd][c];
            th[c] = ue[12 + 4 * c + nd] - tb[c];
        }
        for (int c = 0; c < 3; ++c) {
            disp[c][nd] = I[c][0] * ux[0] + I[c][1] * ux[1] + I[c][2] * ux[2] -
                          mkc(sAl[c][0] * sX[nd][0] + sAl[c][1] * sX[nd][1] + sAl[c][2] * sX[nd][2], 0.0);
            cx r = d1[c][0] * th[0] + d1[c][1] * th[1] + d1[c][2] * th[2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) r = r + 0.5 * (d2[c][i][j] * th[i] * th[j]);
            rot[c][nd] = r;
        }
    }
    // internal DOF k: (component, bubble function) = (0, 4) (0, 5) (1, 4) (1, 5) (2, 6..9)
    for (int k = 0; k < 8; ++k) {
        const int c = k < 2 ? 0 : (k < 4 ? 1 : 2);
        const int fn = k < 4 ? 4 + (k & 1) : 6 + (k - 4);
        disp[c][fn] = disp[c][fn] + ue[24 + k];
    }
    for (int g = 0; g < 4; ++g) {
        cx ux[3][2], rx[3][2], pr[3];
        for (int c = 0; c < 3; ++c) {
            for (int j = 0; j < 2; ++j) {
                cx acc = mkc(0, 0);
                for (int n = 0; n < 10; ++n) acc = acc + sNx[g][n][j] * disp[c][n];
                ux[c][j] = acc;
                cx ar = mkc(0, 0);
                for (int n = 0; n < 4; ++n) ar = ar + sNx[g][n][j] * rot[c][n];
                rx[c][j] = ar;
            }
            cx ap = mkc(0, 0);
            for (int n = 0; n < 4; ++n) ap = ap + sN[g][n] * rot[c][n];
            pr[c] = ap;
        }
        ep[g][0] = ux[0][0];
        ep[g][1] = ux[1][1];
        ep[g][2] = ux[0][1] + ux[1][0];
        ep[g][3] = rx[1][0];
        ep[g][4] = -rx[0][1];
        ep[g][5] = rx[1][1] - rx[0][0];
        ep[g][6] = ux[2][0] + pr[1];
        ep[g][7] = ux[2][1] - pr[0];
        ep[g][8] = 2.0 * pr[2] - ux[1][0] + ux[0][1];
    }
}

// LDS state of one element in the co-rotational shell4 kernels.
struct Shell4CorotLds {
    double X[4][3], Al[3][3], D[9][9];
    double N[4][10], Nx[4][10][2], W[4];
    double B[4][9][32], DB[4][9][32];
    double K[32][33];
    double U[32], R[32], En;
    int Node[4], Dof[24];
};

// dE/dU_d at the element DOFs U by complex step along DOF d (step h), and
// en = E = 1/2 sum over the Gauss points of w e^T ABD e.
__device__ __forceinline__ double shell4_corot_grad(int d, double h, const double (&U)[32], const Shell4CorotLds& L,
                                                    double& en) {
    cx ep[4][9];
    shell4_corot_strains(d, h, U, L.X, L.Al, L.N, L.Nx, ep);
    double r = 0.0;
    en = 0.0;
    for (int g = 0; g < 4; ++g)
        for (int m = 0; m < 9; ++m) {
            double s = 0.0;
            for (int k = 0; k < 9; ++k) s += L.D[m][k] * ep[g][k].r;
            r += L.W[g] * (ep[g][m].i / h) * s;
            en += 0.5 * L.W[g] * ep[g][m].r * s;
        }
    return r;
}

// Co-rotational state of shell4 element e at the displacement u (the shell
// branch of r_getElRuk, Working/AStrO_r_elementEqns.f90:2245-2423): the rest
// tangent K (internal block LU-factored in place), the element DOFs U (24
// from u, the 8 internal ones from the linear recovery -K_ii^-1 K_ie u_e
// brought to element equilibrium R_i = 0 by `iters` chord steps with K_ii at
// rest; the reference never recovers them, D7), R = dE/dU at U (lane d of
// 0..31 by complex step) and the strain energy En.  Ends with a barrier.
__device__ void shell4_corot_state(int e, int lane, const int* __restrict__ conn, const double* __restrict__ X,
                                   const double* __restrict__ elABD, const double* __restrict__ elAlpha,
                                   const int* __restrict__ rank, int nSolid, int nSec, const double* __restrict__ u,
                                   double h, int iters, Shell4CorotLds& L) {
    for (int k = lane; k < 81; k += 64) L.D[k % 9][k / 9] = elABD[81 * (size_t)e + k];
    shell4_stage_B(e, lane, conn, X, elAlpha, L.X, L.Al, L.N, L.Nx, L.W, L.B, L.Node);
    __syncthreads();
    for (int k = lane; k < 4 * 9 * 32; k += 64) {
        const int g = k / 288, r = (k / 32) % 9, d = k % 32;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 9; ++m) acc += L.D[r][m] * L.B[g][m][d];
        L.DB[g][r][d] = L.W[g] * acc;
    }
    __syncthreads();
    for (int k = lane; k < 32 * 32; k += 64) {
        const int i = k >> 5, j = k & 31;
        double acc = 0.0;
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int m = 0; m < 9; ++m) acc += L.B[g][m][i] * L.DB[g][m][j];
        L.K[i][j] = acc;
    }
    if (lane < 24) {
        const int vi = lane >> 2, nd = lane & 3;
        const int dof = rank[L.Node[nd]] + vi * nSec + 2 * nSolid;
        L.Dof[lane] = dof;
        L.U[lane] = u[dof - 1];
    }
    __syncthreads();
    if (lane == 0) {
        shell4_lu_internal(L.K);
        double rhs[8], x[8];
        for (int i = 0; i < 8; ++i) {
            double v = 0.0;
            for (int j = 0; j < 24; ++j) v -= L.K[24 + i][j] * L.U[j];
            rhs[i] = v;
        }
        shell4_solve_internal(L.K, rhs, x);
        for (int i = 0; i < 8; ++i) L.U[24 + i] = x[i];
    }
    __syncthreads();
    for (int it = 0; it <= iters; ++it) {
        if (lane < 32) {
            double en;
            L.R[lane] = shell4_corot_grad(lane, h, L.U, L, en);
            if (lane == 0) L.En = en;
        }
        __syncthreads();
        if (it == iters) break;
        if (lane == 0) {
            double x[8];
            shell4_solve_internal(L.K, &L.R[24], x);
            for (int i = 0; i < 8; ++i) L.U[24 + i] -= x[i];
        }
        __syncthreads();
    }
}

// Co-rotational internal force of shell4 elements at the displacement u: the
// condensed force R_e - K_ei K_ii^-1 R_i of shell4_corot_state's element
// state added into force at the element's DOFs (elements of one launch share
// no node) and E into energy[e].  At rest the derivative of this force is the
// condensed tangent of k_shell4_tangent; a finite rigid rotation gives zero
// force.
__global__ __launch_bounds__(64) void k_shell4_corot(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ elABD, const double* __restrict__ elAlpha, const int* __restrict__ rank, int nSolid,
    int nSec, const double* __restrict__ u, double h, int iters, double* __restrict__ force,
    double* __restrict__ energy) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ Shell4CorotLds L;
    __shared__ double sY[8];
    shell4_corot_state(e, lane, conn, X, elABD, elAlpha, rank, nSolid, nSec, u, h, iters, L);
    if (lane == 0) shell4_solve_internal(L.K, &L.R[24], sY);
    __syncthreads();
    if (lane < 24) {
        double v = L.R[lane];
        for (int m = 0; m < 8; ++m) v -= L.K[lane][24 + m] * sY[m];
        force[L.Dof[lane] - 1] += v;
    }
    if (lane == 0 && energy) energy[e] = L.En;
}

// Co-rotational tangent of shell4 elements at the displacement u: the dRdU
// of r_getElRuk's shell branch that getElasticSolnLoad(1) assembles at every
// Newton step of a nonlinear solve (Working/AStrO_r_elementEqns.f90:2245-2423,
// Working/AStrO_commandFunctions.f90:260-351).  With the internal DOFs at
// element equilibrium (shell4_corot_state), column j of the 32 x 32 Hessian of
// E is the central difference of the complex-step gradient,
// (R(U + s e_j) - R(U - s e_j)) / 2 s (lanes 0..31 take +, 32..63 take -),
// s = eps for the rotations and eps times the longest element edge for the
// translations and internal DOFs; symmetrised and condensed over the internal
// DOFs it is the derivative of k_shell4_corot's force (internal equilibrium
// makes the condensation exact), added into the CSR values as
// k_shell4_tangent's.  Elements of one launch share no node.
__global__ __launch_bounds__(64) void k_shell4_corot_tangent(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ elABD, const double* __restrict__ elAlpha, const int* __restrict__ rank, int nSolid,
    int nSec, const double* __restrict__ u, double h, int iters, double eps, const int* __restrict__ csrRange,
    const uint8_t* __restrict__ elPos, double* __restrict__ val) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ Shell4CorotLds L;
    __shared__ double sUp[2][32], sRp[2][32], sH[32][33], sXs[24][8], sLen;
    shell4_corot_state(e, lane, conn, X, elABD, elAlpha, rank, nSolid, nSec, u, h, iters, L);
    if (lane == 0) {
        double m = 0.0;
        for (int a = 0; a < 4; ++a) {
            const int b = (a + 1) & 3;
            double d2 = 0.0;
            for (int c = 0; c < 3; ++c) d2 += (L.X[b][c] - L.X[a][c]) * (L.X[b][c] - L.X[a][c]);
            m = fmax(m, d2);
        }
        sLen = sqrt(m);
    }
    __syncthreads();
    const int d = lane & 31, sg = lane >> 5;
    for (int j = 0; j < 32; ++j) {
        const double s = (j >= 12 && j < 24) ? eps : eps * sLen;
        if (lane < 32) {
            const double p = lane == j ? s : 0.0;
            sUp[0][lane] = L.U[lane] + p;
            sUp[1][lane] = L.U[lane] - p;
        }
        __syncthreads();
        double en;
        sRp[sg][d] = shell4_corot_grad(d, h, sUp[sg], L, en);
        __syncthreads();
        if (lane < 32) sH[lane][j] = (sRp[0][lane] - sRp[1][lane]) / (sUp[0][j] - sUp[1][j]);
    }
    __syncthreads();
    for (int k = lane; k < 32 * 32; k += 64) {
        const int i = k >> 5, j = k & 31;
        L.K[i][j] = 0.5 * (sH[i][j] + sH[j][i]);
    }
    shell4_condense_scatter(e, lane, L.K, sXs, L.Node, rank, nSolid, nSec, csrRange, elPos, val);
}

// ---------------------------------------------------------------------------
// Design sensitivities by complex step, shell4 layer thickness: every element
// evaluates its condensed tangent with the thickness of its design
// variable's layer perturbed by i h coef (ABD real and imaginary parts from
// the host), in complex arithmetic through the condensation, and contributes
// Im(Kc u_e) / h = coef dKc/dt u_e to column elDV[e] of dRdD (n x nDV,
// column-major) and -u_e . that to gPart[e] (the adjoint compliance term).
// kim (may be null): Im(Kc) of the element, 24 x 24 row-major at 576 e, and
// its DOFs (1-based) at kimDof[24 e], for the batched complex solves.
// Elements of one launch share no node; elements of different design
// variables write different columns.  Working/AStrO_commandFunctions.f90:
// 1172-1375 (getdRelasticdD) and the AStrO_c_elementEqns twin of the shell
// tangent.
struct cdbl {
    double r, i;
};
__device__ __forceinline__ cdbl cmul(cdbl a, cdbl b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cdbl csub(cdbl a, cdbl b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cdbl cdiv(cdbl a, cdbl b) {
    const double d = 1.0 / (b.r * b.r + b.i * b.i);
    return {(a.r * b.r + a.i * b.i) * d, (a.i * b.r - a.r * b.i) * d};
}

__global__ __launch_bounds__(64) void k_shell4_dtangent_cs(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ elABD, const double* __restrict__ elABDim, const double* __restrict__ elAlpha,
    const int* __restrict__ rank, int nSolid, int nSec, const double* __restrict__ u, const int* __restrict__ elDV,
    double invH, int64_t n, double* __restrict__ dRdD, double* __restrict__ gPart, double* __restrict__ kim,
    int* __restrict__ kimDof) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ double sX[4][3], sAl[3][3], sDr[9][9], sDi[9][9];
    __shared__ double sN[4][10], sNx[4][10][2], sW[4];
    __shared__ double sB[4][9][32], sDBr[4][9][32], sDBi[4][9][32];
    __shared__ cdbl sK[32][33];
    __shared__ cdbl sXs[24][8];
    __shared__ double sU[24], sVh[24], sVl[24], sIm[24][24];
    __shared__ int sNode[4], sDof[24];
    for (int k = lane; k < 81; k += 64) {
        sDr[k % 9][k / 9] = elABD[81 * (size_t)e + k];
        sDi[k % 9][k / 9] = elABDim[81 * (size_t)e + k];
    }
    shell4_stage_B(e, lane, conn, X, elAlpha, sX, sAl, sN, sNx, sW, sB, sNode);
    for (int k = lane; k < 4 * 9 * 32; k += 64) {
        const int g = k / 288, r = (k / 32) % 9, d = k % 32;
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            ar += sDr[r][m] * sB[g][m][d];
            ai += sDi[r][m] * sB[g][m][d];
        }
        sDBr[g][r][d] = sW[g] * ar;
        sDBi[g][r][d] = sW[g] * ai;
    }
    if (lane < 24) {
        const int var = lane >> 2, a = lane & 3;
        const int row = rank[sNode[a]] + var * nSec + 2 * nSolid;
        sDof[lane] = row;
        sU[lane] = u[row - 1];
    }
    __syncthreads();
    for (int k = lane; k < 32 * 32; k += 64) {
        const int i = k >> 5, j = k & 31;
        double ar = 0.0, ai = 0.0;
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int m = 0; m < 9; ++m) {
                ar += sB[g][m][i] * sDBr[g][m][j];
                ai += sB[g][m][i] * sDBi[g][m][j];
            }
        sK[i][j] = {ar, ai};
    }
    __syncthreads();
    if (lane == 0) {  // complex LU (no pivoting) of the internal block
        for (int i = 24; i < 32; ++i) {
            for (int j = i; j < 32; ++j)
                for (int k = 24; k < i; ++k) sK[i][j] = csub(sK[i][j], cmul(sK[i][k], sK[k][j]));
            const cdbl piv = sK[i][i];
            for (int j = i + 1; j < 32; ++j) {
                for (int k = 24; k < i; ++k) sK[j][i] = csub(sK[j][i], cmul(sK[j][k], sK[k][i]));
                sK[j][i] = cdiv(sK[j][i], piv);
            }
        }
    }
    __syncthreads();
    if (lane < 24) {
        cdbl z[8], x[8];
        for (int i = 0; i < 8; ++i) {
            cdbl v = sK[24 + i][lane];
            for (int k = 0; k < i; ++k) v = csub(v, cmul(sK[24 + i][24 + k], z[k]));
            z[i] = v;
        }
        for (int i = 7; i >= 0; --i) {
            cdbl v = z[i];
            for (int k = i + 1; k < 8; ++k) v = csub(v, cmul(sK[24 + i][24 + k], x[k]));
            x[i] = cdiv(v, sK[24 + i][24 + i]);
        }
        for (int i = 0; i < 8; ++i) sXs[lane][i] = x[i];
    }
    __syncthreads();
    const int dv = elDV[e];
    // Im(Kc) with Kc = Kcc - Kci X, once per entry
    for (int k = lane; k < 576; k += 64) {
        const int i = k / 24, j = k - (k / 24) * 24;
        double im = sK[i][j].i;
#pragma unroll
        for (int m = 0; m < 8; ++m) im -= cmul(sK[i][24 + m], sXs[j][m]).i;
        sIm[i][j] = im;
        if (kim && dv >= 0) kim[576 * (size_t)e + k] = im;
    }
    if (kim && dv >= 0 && lane < 24) kimDof[24 * (size_t)e + lane] = sDof[lane];
    __syncthreads();
    // Im(Kc) u and u . Im(Kc) u in double-double: the displacement's
    // rigid-body part cancels in both, terms much larger than the result
    if (lane < 24) {
        const int i = lane;
        double vh = 0.0, vl = 0.0;
        for (int j = 0; j < 24; ++j) dd_add_prod(vh, vl, sIm[i][j], sU[j]);
        sVh[i] = vh * invH;
        sVl[i] = vl * invH;
        if (dv >= 0) dRdD[(size_t)dv * n + sDof[i] - 1] += (vh + vl) * invH;
    }
    __syncthreads();
    if (lane == 0) {
        double gh = 0.0, gl = 0.0;
        for (int i = 0; i < 24; ++i) {
            dd_add_prod(gh, gl, -sU[i], sVh[i]);
            dd_add_prod(gh, gl, -sU[i], sVl[i]);
        }
        gPart[e] = dv >= 0 ? gh + gl : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Stress and strain recovery, brick8 (r_getSolidStress,
// Working/AStrO_r_elementEqns.f90:924-976): at the natural point sp of every
// element, ux = sum_n u_n Nx_n of the element's own nodes (the reference's
// gather indexes by local node number, D4), the Green-Lagrange strain of
// r_getNLStrain (:622-648, orient = identity) and stress = C (strain -
// T alpha) with T the nodal temperatures (rank-indexed) interpolated to the
// point (:962-973; tExp null: no expansion).  Voigt order (11, 22, 33, 12,
// 13, 23), engineering shear.  One thread per element.
__global__ __launch_bounds__(256) void k_brick8_stress(int E, const int* __restrict__ conn,
                                                       const double* __restrict__ X, const int* __restrict__ elMat,
                                                       const double* __restrict__ elScale,
                                                       const double* __restrict__ matC, const int* __restrict__ rank,
                                                       int nS, const double* __restrict__ u,
                                                       const double* __restrict__ tExp,
                                                       const double* __restrict__ T, double s0, double s1,
                                                       double s2, double* __restrict__ stress,
                                                       double* __restrict__ strain) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double xe[8][3], ue[3][8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int nd = conn[8 * (size_t)e + n] - 1;
        const int r = rank[nd];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xe[n][c] = X[3 * (size_t)nd + c];
            ue[c][n] = u[r + c * nS - 1];
        }
    }
    double N[8], Nx[8][3], det;
    brick8_point(s0, s1, s2, xe, N, Nx, det);
    double ux[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int n = 0; n < 8; ++n) acc += ue[i][n] * Nx[n][j];
            ux[i][j] = acc;
        }
    double Ev[6];
    Ev[0] = ux[0][0] + 0.5 * (ux[0][0] * ux[0][0] + ux[1][0] * ux[1][0] + ux[2][0] * ux[2][0]);
    Ev[1] = ux[1][1] + 0.5 * (ux[0][1] * ux[0][1] + ux[1][1] * ux[1][1] + ux[2][1] * ux[2][1]);
    Ev[2] = ux[2][2] + 0.5 * (ux[0][2] * ux[0][2] + ux[1][2] * ux[1][2] + ux[2][2] * ux[2][2]);
    Ev[3] = ux[0][1] + ux[1][0] + ux[0][0] * ux[0][1] + ux[1][0] * ux[1][1] + ux[2][0] * ux[2][1];
    Ev[4] = ux[0][2] + ux[2][0] + ux[0][0] * ux[0][2] + ux[1][0] * ux[1][2] + ux[2][0] * ux[2][2];
    Ev[5] = ux[1][2] + ux[2][1] + ux[0][2] * ux[0][1] + ux[1][2] * ux[1][1] + ux[2][2] * ux[2][1];
    const double sc = elScale[e];
    const double* Cm = matC + 36 * (size_t)elMat[e];
    double Em[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) Em[k] = Ev[k];
    if (tExp) {
        double Tp = 0.0;
#pragma unroll
        for (int n = 0; n < 8; ++n) Tp += N[n] * T[rank[conn[8 * (size_t)e + n] - 1] - 1];
#pragma unroll
        for (int k = 0; k < 6; ++k) Em[k] -= Tp * tExp[6 * (size_t)e + k];
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) acc += Cm[r + 6 * k] * Em[k];
        stress[6 * (size_t)e + r] = sc * acc;
        strain[6 * (size_t)e + r] = Ev[r];
    }
}

// ---------------------------------------------------------------------------
// Thermal expansion of brick8 elements at u = 0 (the solid branch of
// r_getElRuk with temperatures, Working/AStrO_r_elementEqns.f90:2433-2447,
// stress = C (strain - T_g alpha), T_g = sum_n N_n(g) T_n): the residual is
// R = -f with f_(a,c) = sum_g detJ T_g (Nx_a . S)_c, S = C alpha in Voigt
// form, so that
//  - mode 0 adds scale * f to load at the element's DOFs (the thermal load
//    of the static right-hand side, F - R(0, T));
//  - mode 1 forms out[rank(b) - 1] += (dR/dT)^T lam at node b, i.e.
//    -sum_g detJ N_b(g) sum_(a,c) lam_(a,c) (Nx_a . S)_c (the intended dRdT
//    of r_getElRuk, whose loop writes a stale row, D10).
// One thread per element; elements of one launch share no node.
__global__ __launch_bounds__(128) void k_brick8_thermal_strain(int mode, const int* __restrict__ perm, int count,
                                                              const int* __restrict__ conn,
                                                              const double* __restrict__ X,
                                                              const int* __restrict__ elMat,
                                                              const double* __restrict__ elScale,
                                                              const double* __restrict__ matC,
                                                              const double* __restrict__ tExp,
                                                              const int* __restrict__ rank, int nS,
                                                              const double* __restrict__ T,
                                                              const double* __restrict__ lam, double scale,
                                                              double* __restrict__ out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    double xe[8][3], tn[8], lm[3][8];
    int rk[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int nd = conn[8 * (size_t)e + n] - 1;
        rk[n] = rank[nd];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xe[n][c] = X[3 * (size_t)nd + c];
            lm[c][n] = mode == 1 ? lam[rk[n] + c * nS - 1] : 0.0;
        }
        tn[n] = mode == 0 ? T[rk[n] - 1] : 0.0;
    }
    const double sc = elScale[e];
    const double* Cm = matC + 36 * (size_t)elMat[e];
    double S[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) acc += Cm[r + 6 * k] * tExp[6 * (size_t)e + k];
        S[r] = sc * acc;
    }
    double f[3][8], g[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        g[n] = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c][n] = 0.0;
    }
    for (int gp = 0; gp < 8; ++gp) {
        double N[8], Nx[8][3], w;
        brick8_gauss_point(gp, xe, N, Nx, w);
        double Tg = 0.0;
#pragma unroll
        for (int n = 0; n < 8; ++n) Tg += N[n] * tn[n];
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            // (Nx_n . S)_c: the B0 column of DOF (n, c) against S
            const double b0 = Nx[n][0] * S[0] + Nx[n][1] * S[3] + Nx[n][2] * S[4];
            const double b1 = Nx[n][1] * S[1] + Nx[n][0] * S[3] + Nx[n][2] * S[5];
            const double b2 = Nx[n][2] * S[2] + Nx[n][0] * S[4] + Nx[n][1] * S[5];
            f[0][n] += w * Tg * b0;
            f[1][n] += w * Tg * b1;
            f[2][n] += w * Tg * b2;
            s += lm[0][n] * b0 + lm[1][n] * b1 + lm[2][n] * b2;
        }
#pragma unroll
        for (int n = 0; n < 8; ++n) g[n] -= w * N[n] * s;
    }
    if (mode == 0) {
#pragma unroll
        for (int n = 0; n < 8; ++n)
#pragma unroll
            for (int c = 0; c < 3; ++c) out[rk[n] + c * nS - 1] += scale * f[c][n];
    } else {
#pragma unroll
        for (int n = 0; n < 8; ++n) out[rk[n] - 1] += g[n];
    }
}

// ---------------------------------------------------------------------------
// Element residuals of the Newmark step, brick8: R_e = Ruk(u_e) + M_e a_e with
// the Green-Lagrange internal force (Working/AStrO_r_elementEqns.f90:2424-2458,
// strain and derivative r_getNLStrain :622-662, orient = identity, T = 0) and
// the consistent-mass inertia of a_e = c1 (c2 (u - pu - dt pv) - (1 - 2 beta) pa)
// (r_getElRu :2684-2690, r_getElRum :2628-2653).  Every element gathers its own
// nodes' values (the reference's r_getElementSolution indexes by local node
// number, D4).  tExp (may be null) and T: the thermal strain subtracted from
// E before the stress (brick8_subtract_thermal).  load[dof] -= R_e; elements
// of one launch share no node.
// One wave per element: lanes 0..7 own a Gauss point, lanes 0..23 a DOF
// (component c = d / 8, node d % 8).
__global__ __launch_bounds__(64) void k_brick8_residual(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const int* __restrict__ elMat, const double* __restrict__ elScale, const double* __restrict__ matC,
    const double* __restrict__ elRho, const int* __restrict__ rank, int nS, const double* __restrict__ u,
    const double* __restrict__ pu, const double* __restrict__ pv, const double* __restrict__ pa, double c1, double c2,
    double dt, double beta, const double* __restrict__ tExp, const double* __restrict__ T, double* __restrict__ load,
    const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ double sX[8][3];
    __shared__ double sNx[8][8][3];
    __shared__ double sN[8][8];
    __shared__ double sW[8];
    __shared__ int sNode[8];
    __shared__ double sU[3][8], sA[3][8];
    __shared__ double sUx[8][3][3], sS[8][6], sAb[8][3];
    if (lane < 8) {
        int nd = conn[8 * (size_t)e + lane] - 1;
        sNode[lane] = nd;
        sX[lane][0] = X[3 * (size_t)nd];
        sX[lane][1] = X[3 * (size_t)nd + 1];
        sX[lane][2] = X[3 * (size_t)nd + 2];
    }
    __syncthreads();
    const int comp = lane >> 3, nd = lane & 7;
    int dof = 0;
    if (lane < 24) {
        dof = rank[sNode[nd]] + comp * nS;  // 1-based
        const size_t k = dof - 1;
        const double uk = u[k];
        sU[comp][nd] = uk;
        sA[comp][nd] = c1 * (c2 * (uk - pu[k] - dt * pv[k]) - (1.0 - 2.0 * beta) * pa[k]);
    }
    if (lane < 8) brick8_gauss_point(lane, sX, sN[lane], sNx[lane], sW[lane]);
    __syncthreads();
    if (lane < 8) {
        const int g = lane;
        double ux[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
#pragma unroll
                for (int n = 0; n < 8; ++n) acc += sU[i][n] * sNx[g][n][j];
                ux[i][j] = acc;
                sUx[g][i][j] = acc;
            }
        double E[6];
        E[0] = ux[0][0] + 0.5 * (ux[0][0] * ux[0][0] + ux[1][0] * ux[1][0] + ux[2][0] * ux[2][0]);
        E[1] = ux[1][1] + 0.5 * (ux[0][1] * ux[0][1] + ux[1][1] * ux[1][1] + ux[2][1] * ux[2][1]);
        E[2] = ux[2][2] + 0.5 * (ux[0][2] * ux[0][2] + ux[1][2] * ux[1][2] + ux[2][2] * ux[2][2]);
        E[3] = ux[0][1] + ux[1][0] + ux[0][0] * ux[0][1] + ux[1][0] * ux[1][1] + ux[2][0] * ux[2][1];
        E[4] = ux[0][2] + ux[2][0] + ux[0][0] * ux[0][2] + ux[1][0] * ux[1][2] + ux[2][0] * ux[2][2];
        E[5] = ux[1][2] + ux[2][1] + ux[0][2] * ux[0][1] + ux[1][2] * ux[1][1] + ux[2][2] * ux[2][1];
        brick8_subtract_thermal(e, sN[g], sNode, rank, tExp, T, E);
        const double sc = elScale[e];
        const double* Cm = matC + 36 * (size_t)elMat[e];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc += Cm[r + 6 * k] * E[k];
            sS[g][r] = sc * acc;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int n = 0; n < 8; ++n) acc += sN[g][n] * sA[c][n];
            sAb[g][c] = acc;
        }
    }
    __syncthreads();
    if (lane >= 24) return;
    const double rho = elRho[e];
    const double o0 = comp == 0 ? 1.0 : 0.0, o1 = comp == 1 ? 1.0 : 0.0, o2 = comp == 2 ? 1.0 : 0.0;
    double R = 0.0;
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const double x0 = sNx[g][nd][0], x1 = sNx[g][nd][1], x2 = sNx[g][nd][2];
        const double u0 = sUx[g][comp][0], u1 = sUx[g][comp][1], u2 = sUx[g][comp][2];
        const double dE0 = o0 * x0 + x0 * u0;
        const double dE1 = o1 * x1 + x1 * u1;
        const double dE2 = o2 * x2 + x2 * u2;
        const double dE3 = o0 * x1 + o1 * x0 + x0 * u1 + x1 * u0;
        const double dE4 = o0 * x2 + o2 * x0 + x0 * u2 + x2 * u0;
        const double dE5 = o1 * x2 + o2 * x1 + x1 * u2 + x2 * u1;
        const double* S = sS[g];
        const double w = sW[g];
        R += w * (S[0] * dE0 + S[1] * dE1 + S[2] * dE2 + S[3] * dE3 + S[4] * dE4 + S[5] * dE5);
        R += rho * w * sN[g][nd] * sAb[g][comp];
    }
    load[dof - 1] -= R;
}

// Newmark update after the step's increment du: u += du; the new acceleration
// and velocity from the previous state; then the previous state takes the new
// one.  Working/AStrO_commandFunctions.f90:792-806.
__global__ void k_newmark_update(int n, double* __restrict__ u, const double* __restrict__ du,
                                 double* __restrict__ pu, double* __restrict__ pv, double* __restrict__ pa,
                                 double* __restrict__ v, double* __restrict__ a, double c1, double c2, double dt,
                                 double beta, double gamma, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double un = u[i] + du[i];
        const double an = c1 * (c2 * (un - pu[i] - dt * pv[i]) - (1.0 - 2.0 * beta) * pa[i]);
        const double vn = pv[i] + dt * ((1.0 - gamma) * pa[i] + gamma * an);
        u[i] = un;
        a[i] = an;
        v[i] = vn;
        pu[i] = un;
        pv[i] = vn;
        pa[i] = an;
    }
}

// Transient thermal vectors (r_getElRt, Working/AStrO_r_elementEqns.f90:
// 2190-2200, and the update of solve, Working/AStrO_commandFunctions.f90:791-803):
// w = Tdot(T) - cm T with Tdot(T) = (1 / gamma) ((T - Tp) / dt - (1 - gamma) Tdp),
// so that K T + C Tdot = A T + C w for the tangent A = K + cm C, cm = 1 / (gamma dt).
__global__ void k_thermal_rate(int n, const double* __restrict__ T, const double* __restrict__ Tp,
                               const double* __restrict__ Tdp, double gamma, double dt, double cm,
                               double* __restrict__ w) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double td = (1.0 / gamma) * ((1.0 / dt) * (T[i] - Tp[i]) - (1.0 - gamma) * Tdp[i]);
        w[i] = td - cm * T[i];
    }
}
// After the step: T += dT, Tdot(T), then Tp = T, Tdp = Tdot.
__global__ void k_thermal_update(int n, double* __restrict__ T, const double* __restrict__ dT,
                                 double* __restrict__ Tp, double* __restrict__ Tdp, double* __restrict__ Tdot,
                                 double gamma, double dt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double t = T[i] + dT[i];
        const double td = (1.0 / gamma) * ((1.0 / dt) * (t - Tp[i]) - (1.0 - gamma) * Tdp[i]);
        T[i] = t;
        Tdot[i] = td;
        Tp[i] = t;
        Tdp[i] = td;
    }
}

// Brick8 conductivity (Working/AStrO_r_elementEqns.f90:2062-2124) times
// condFactor plus capFactor times the capacitance, scattered into the
// thermal CSR by node rank; slot 255 marks a column the reference's probe
// never stores (its `.lt.` bound, Working/AStrO_commandFunctions.f90:57).
// elCap: rho c per element (read only when capFactor != 0).
__global__ __launch_bounds__(64) void k_brick8_conduct(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ elCond, double condFactor, const double* __restrict__ elCap, double capFactor,
    const int* __restrict__ rank, const int* __restrict__ csrRange, const uint8_t* __restrict__ thPos,
    double* __restrict__ val) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ double sX[8][3];
    __shared__ double sNx[8][8][3];
    __shared__ double sN[8][8];
    __shared__ double sW[8];
    __shared__ int sNode[8];
    if (lane < 8) {
        int nd = conn[8 * (size_t)e + lane] - 1;
        sNode[lane] = nd;
        for (int c = 0; c < 3; ++c) sX[lane][c] = X[3 * (size_t)nd + c];
    }
    __syncthreads();
    if (lane < 8) brick8_gauss_point(lane, sX, sN[lane], sNx[lane], sW[lane]);
    __syncthreads();
    const int a = lane >> 3, b = lane & 7;
    double acc = 0.0;
    if (condFactor != 0.0) {
        const double* kc = elCond + 6 * (size_t)e;
        const double k00 = kc[0], k11 = kc[1], k22 = kc[2], k01 = kc[3], k02 = kc[4], k12 = kc[5];
        for (int g = 0; g < 8; ++g) {
            const double ax = sNx[g][a][0], ay = sNx[g][a][1], az = sNx[g][a][2];
            const double bx = sNx[g][b][0], by = sNx[g][b][1], bz = sNx[g][b][2];
            const double kbx = k00 * bx + k01 * by + k02 * bz;
            const double kby = k01 * bx + k11 * by + k12 * bz;
            const double kbz = k02 * bx + k12 * by + k22 * bz;
            acc += sW[g] * (ax * kbx + ay * kby + az * kbz);
        }
        acc *= condFactor;
    }
    if (capFactor != 0.0) {
        // the capacitance rho c N N^T of r_getElRtm (Working/AStrO_r_elementEqns.f90:2126-2165)
        double m = 0.0;
        for (int g = 0; g < 8; ++g) m += sW[g] * sN[g][a] * sN[g][b];
        acc += capFactor * elCap[e] * m;
    }
    const int slot = thPos[64 * (size_t)e + lane];
    if (slot != 255) {
        const int row = rank[sNode[a]];
        val[csrRange[row - 1] + slot] += acc;
    }
}

// The transient thermal load -Rt of getThermalSolnLoad, element by element:
// load_a -= sum_b K_ab T_b + C_ab Tdot_b with the whole element matrices of
// r_getElRt (Working/AStrO_r_elementEqns.f90:2166-2200), whatever the CSR
// stores (the reference's probe drops entries of the assembled matrix, not of
// the residual).  T, Tdot, load: by node rank.  One colour per launch.
__global__ __launch_bounds__(64) void k_brick8_thermal_residual(
    const int* __restrict__ perm, int count, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ elCond, const double* __restrict__ elCap, const int* __restrict__ rank,
    const double* __restrict__ T, const double* __restrict__ Tdot, double* __restrict__ load) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int e = perm[idx];
    const int lane = threadIdx.x;
    __shared__ double sX[8][3];
    __shared__ double sNx[8][8][3];
    __shared__ double sN[8][8];
    __shared__ double sW[8];
    __shared__ int sRow[8];
    if (lane < 8) {
        int nd = conn[8 * (size_t)e + lane] - 1;
        sRow[lane] = rank[nd] - 1;
        for (int c = 0; c < 3; ++c) sX[lane][c] = X[3 * (size_t)nd + c];
    }
    __syncthreads();
    if (lane < 8) brick8_gauss_point(lane, sX, sN[lane], sNx[lane], sW[lane]);
    __syncthreads();
    const int a = lane >> 3, b = lane & 7;
    const double* kc = elCond + 6 * (size_t)e;
    const double k00 = kc[0], k11 = kc[1], k22 = kc[2], k01 = kc[3], k02 = kc[4], k12 = kc[5];
    double kab = 0.0, mab = 0.0;
    for (int g = 0; g < 8; ++g) {
        const double ax = sNx[g][a][0], ay = sNx[g][a][1], az = sNx[g][a][2];
        const double bx = sNx[g][b][0], by = sNx[g][b][1], bz = sNx[g][b][2];
        const double kbx = k00 * bx + k01 * by + k02 * bz;
        const double kby = k01 * bx + k11 * by + k12 * bz;
        const double kbz = k02 * bx + k12 * by + k22 * bz;
        kab += sW[g] * (ax * kbx + ay * kby + az * kbz);
        mab += sW[g] * sN[g][a] * sN[g][b];
    }
    const int rb = sRow[b];
    double r = kab * T[rb] + elCap[e] * mab * Tdot[rb];
    r += __shfl_xor(r, 1);
    r += __shfl_xor(r, 2);
    r += __shfl_xor(r, 4);
    if (b == 0) load[sRow[a]] -= r;
}

// max |v| over a vector, combined into *out as the bit pattern of a
// non-negative double (ordered like an unsigned integer).
__global__ void k_max_abs(const double* __restrict__ v, int64_t n, unsigned long long* __restrict__ out) {
    double m = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(v[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    __shared__ double sm[16];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sm[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = fmax(r, sm[i]);
        atomicMax(out, (unsigned long long)__double_as_longlong(r));
    }
}

// Constraint rows normalised, then scaled by 100 * max|K|.
// Working/AStrO_commandFunctions.f90:537-566.
__global__ void k_scale_mpc(double* __restrict__ cmat, double* __restrict__ rhs, const int* __restrict__ crange,
                            int cdim, const unsigned long long* __restrict__ maxBits) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cdim) return;
    const double maxMat = __longlong_as_double((long long)*maxBits);
    const int a = crange[r], b = crange[r + 1];
    double mag = 0.0;
    for (int s = a; s < b; ++s) mag += cmat[s] * cmat[s];
    mag = sqrt(mag);
    const double inv = 1.0 / mag, f = 100.0 * maxMat;
    for (int s = a; s < b; ++s) cmat[s] = f * (inv * cmat[s]);
    rhs[r] = f * (inv * rhs[r]);
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_sentinel(double* __restrict__ p, int64_t n, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const double v = __longlong_as_double(ASTRO_SENTINEL);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_i32(int* __restrict__ p, int64_t n, int v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// In-profile CSR entries into the (zeroed) profile: row r keeps columns
// s[r] <= c <= r.  Working/AStrO_bookKeeping.f90:764-774.  32 lanes per row.
template <typename LT>
__global__ __launch_bounds__(256) void k_csr_to_profile(const double* __restrict__ val, const int* __restrict__ cols,
                                                        const int* __restrict__ range, int n,
                                                        const int* __restrict__ sfirst,
                                                        const int64_t* __restrict__ rowoff, LT* __restrict__ lt) {
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const int l = threadIdx.x & 31;
    if (row >= n) return;
    const int s = sfirst[row];
    const int64_t off = rowoff[row];
    for (int q = range[row] + l; q < range[row + 1]; q += 32) {
        const int c = cols[q] - 1;
        if (c >= s && c <= row) lt[off + c] += (LT)val[q];
    }
}
template __global__ void k_csr_to_profile<double>(const double*, const int*, const int*, int, const int*,
                                                  const int64_t*, double*);
template __global__ void k_csr_to_profile<float>(const double*, const int*, const int*, int, const int*,
                                                 const int64_t*, float*);

// Penalty C^T C into the profile (Working/AStrO_bookKeeping.f90:776-790), one
// thread per profile entry that a constraint row reaches: entry k is
// (row entRow[k], column entCol[k]), its products cmat[pa[q]] * cmat[pb[q]]
// for q in [start[k], start[k + 1]) in the order convertToLTri adds them
// (constraint row, then term pairs), added one by one to the entry.  No two
// threads share an entry, so the sum does not depend on scheduling; the
// host drops pairs outside the profile.
template <typename LT>
__global__ void k_penalty_plan(const double* __restrict__ cmat, int nEnt, const int* __restrict__ entRow,
                               const int* __restrict__ entCol, const int* __restrict__ start,
                               const int* __restrict__ pa, const int* __restrict__ pb,
                               const int64_t* __restrict__ rowoff, LT* __restrict__ lt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nEnt) return;
    const int64_t idx = rowoff[entRow[k]] + entCol[k];
    LT v = lt[idx];
    for (int q = start[k]; q < start[k + 1]; ++q) v += (LT)(cmat[pa[q]] * cmat[pb[q]]);
    lt[idx] = v;
}
template __global__ void k_penalty_plan<double>(const double*, int, const int*, const int*, const int*, const int*,
                                                const int*, const int64_t*, double*);
template __global__ void k_penalty_plan<float>(const double*, int, const int*, const int*, const int*, const int*,
                                               const int*, const int64_t*, float*);

// Fill of a profile array (either storage precision).
template <typename T>
__global__ void k_fill_t(T* __restrict__ p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
template __global__ void k_fill_t<double>(double*, int64_t, double);
template __global__ void k_fill_t<float>(float*, int64_t, float);

// gfx950 cross-lane moves on doubles.  swap_halves32(a, c): lanes 32..63 of
// a take lanes 0..31 of c and lanes 0..31 of c take lanes 32..63 of a, so
// a + c is, in lanes 0..31, a summed with its l ^ 32 partner and, in lanes
// 32..63, c summed with its partner.  swap_rows16 does the same between the
// odd 16-lane rows of a and the even rows of c (partner l ^ 16).
__device__ __forceinline__ void swap_halves32(double& a, double& c) {
    const long long x = __double_as_longlong(a), y = __double_as_longlong(c);
    const auto lo = __builtin_amdgcn_permlane32_swap((int)x, (int)y, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((int)(x >> 32), (int)(y >> 32), false, false);
    a = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
    c = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
}
__device__ __forceinline__ void swap_rows16(double& a, double& c) {
    const long long x = __double_as_longlong(a), y = __double_as_longlong(c);
    const auto lo = __builtin_amdgcn_permlane16_swap((int)x, (int)y, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((int)(x >> 32), (int)(y >> 32), false, false);
    a = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
    c = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
}
// DPP controls: row_shl:n reads lane l + n, row_shr:n lane l - n (within a
// 16-lane row); quad_perm 2,3,0,1 and 1,0,3,2 read lanes l ^ 2 and l ^ 1.
enum { DPP_ROW_SHL8 = 0x108, DPP_ROW_SHR8 = 0x118, DPP_ROW_SHL4 = 0x104, DPP_ROW_SHR4 = 0x114,
       DPP_QUAD_XOR2 = 0x4E, DPP_QUAD_XOR1 = 0xB1 };
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---------------------------------------------------------------------------
// Blocked left-looking LDL^T on the row-skyline profile.
//
// Panel P = columns [p, pe), pe <= p + 64, of one independent block.  Rows
// [p, iend) may reach into P.  Step U: W(i, j) = sum_{k < p} L(i,k) D(k) L(j,k)
// for 64-row tiles of those rows and j in P, split over splitK ranges of k,
// partial sums to scratch (deterministic).  Step FT: every tile re-forms the
// updated diagonal block A(P,P) - W, factors it in LDS (tile 0 writes it
// back), then solves its own rows: L(i,P) = (A(i,P) - W(i,P)) L_PP^-T D_P^-1.
// Profile masking keeps every entry with c < s[r] an implicit zero; the LDL^T
// of an envelope matrix stays inside the envelope, as the reference's
// in-place row factor (Working/AStrO_solvers.f90:10-78) relies on.

typedef double double4_t __attribute__((ext_vector_type(4)));

// Step U on FP64 matrix cores: each of the 4 waves owns 16 tile rows x 64
// panel columns as four 16x16 MFMA accumulators (v_mfma_f64_16x16x4f64);
// k is staged through LDS in chunks of 32.  Operand lane layouts (measured on
// gfx950): A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k, D(i, j) in
// lane j + 16 (i % 4), register i / 4.
constexpr int LDL_KC = 32;  // k per staged chunk of step U
constexpr int LDL_U_SMEM_BYTES = 2 * LDL_KC * 65 * 8;  // step U alone
// step FT: two 64 x 65 blocks (factored diagonal block, then the tile rows;
// its inverse), the column double buffer, per-wave 16 x 16 products, 1/D and
// the tile rows' offsets
constexpr int LDL_FT_A = 0, LDL_FT_X = 64 * 65 * 8, LDL_FT_COL = 2 * 64 * 65 * 8,
              LDL_FT_T = LDL_FT_COL + 2 * 4 * 64 * 8, LDL_FT_RD = LDL_FT_T + 4 * 16 * 17 * 8,
              LDL_FT_OFF = LDL_FT_RD + 64 * 8,
              LDL_FT_SI = LDL_FT_OFF + 64 * 8, LDL_STEP_SMEM_BYTES = LDL_FT_SI + 64 * 4;

// X = L^-1 for the unit-lower 64 x 64 block whose strictly lower part is in
// L (entries on and above the diagonal are not read), by 16 x 16 blocks: the
// four diagonal blocks by forward substitution (thread = column), then block
// rows 1..3 as X_ij = -X_ii sum_{j <= k < i} L_ik X_kj on the matrix cores;
// the strictly upper blocks are set to zero.  256 threads; Tw is this wave's
// 16 x 17 scratch.  Ends with a barrier.
__device__ __forceinline__ void unit_lower_inverse(const double (*L)[65], double (*X)[65], double (*Tw)[17]) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid < 64) {
        const int bb = tid >> 4, c = tid & 15, o = 16 * bb;
        double xv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            double acc2 = r == c ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < r; ++k) acc2 -= L[o + r][o + k] * xv[k];
            xv[r] = r < c ? 0.0 : acc2;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) X[o + r][o + c] = xv[r];
    }
    for (int id = tid; id < 6 * 256; id += 256) {
        // upper blocks (bi, bj), bi < bj: (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
        const int q = id >> 8, e = id & 255;
        const int bi = q < 3 ? 0 : (q < 5 ? 1 : 2), bj = q < 3 ? q + 1 : (q < 5 ? q - 1 : 3);
        X[16 * bi + (e >> 4)][16 * bj + (e & 15)] = 0.0;
    }
    __syncthreads();
    for (int bi = 1; bi < 4; ++bi) {
        const int bj = w;  // wave w computes block (bi, w) for w < bi
        if (bj < bi) {
            double4_t t = double4_t{0.0, 0.0, 0.0, 0.0};
            for (int bk = bj; bk < bi; ++bk) {
#pragma unroll
                for (int ks = 0; ks < 16; ks += 4) {
                    const int kk = ks + (lane >> 4);
                    const double av = L[16 * bi + (lane & 15)][16 * bk + kk];
                    const double bv = X[16 * bk + kk][16 * bj + (lane & 15)];
                    t = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, t, 0, 0, 0);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) Tw[4 * v + (lane >> 4)][lane & 15] = t[v];
            __builtin_amdgcn_wave_barrier();
            double4_t xb = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < 16; ks += 4) {
                const int kk = ks + (lane >> 4);
                const double av = X[16 * bi + (lane & 15)][16 * bi + kk];
                const double bv = Tw[kk][lane & 15];
                xb = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, xb, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) X[16 * bi + 4 * v + (lane >> 4)][16 * bj + (lane & 15)] = -xb[v];
        }
        __syncthreads();
    }
}

// C = op(A) op(B) for 64 x 64 blocks in LDS on the matrix cores, written
// column-major to dst (minus E, if given, also column-major): wave w owns
// rows 16 w .. 16 w + 15.  TA / TB: use A^T / B^T.
template <bool TA, bool TB>
__device__ __forceinline__ void mfma_block64(const double (*A)[65], const double (*B)[65], const double* E,
                                             double* __restrict__ dst) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < 64; ks += 4) {
        const int kk = ks + (lane >> 4), i = w * 16 + (lane & 15);
        const double av = TA ? A[kk][i] : A[i][kk];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int j = cb * 16 + (lane & 15);
            const double bv = TB ? B[j][kk] : B[kk][j];
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = w * 16 + 4 * v + (lane >> 4), j = cb * 16 + (lane & 15);
            dst[j * 64 + i] = E ? acc[cb][v] - E[j * 64 + i] : acc[cb][v];
        }
}

typedef float float4_t __attribute__((ext_vector_type(4)));

// Step U for a profile stored in FP32 (the mixed-precision factor): the same
// tiling, k range and staging as ldl_update_body below, with FP32 operands
// staged in LDS, v_mfma_f32_16x16x4f32 with FP32 accumulation, and FP32
// partial sums.  B = L(P, k) D_k is formed in FP64 (D is FP64) and rounded
// once.  Result layout (measured on gfx950): D(i, j) in lane j + 16 (i / 4),
// register i % 4.
__device__ __forceinline__ void ldl_update_body_f32(const PanelInfo P, int blk, int nb, int tile, int sp, int nSplit,
                                                    int kMode, int slot, const float* __restrict__ lt,
                                                    const int64_t* __restrict__ rowoff,
                                                    const int* __restrict__ sfirst, const int* __restrict__ winMin,
                                                    const double* __restrict__ diag, float* __restrict__ scratch,
                                                    int scratchRows, char* smem) {
    const int r0 = P.p + 64 * tile;
    if (r0 >= P.iend) return;
    const int nr = min(64, P.iend - r0);
    const int np = P.pe - P.p;
    constexpr int KC = LDL_KC;
    float(*As)[65] = reinterpret_cast<float(*)[65]>(smem);
    float(*Bs)[65] = reinterpret_cast<float(*)[65]>(smem + KC * 65 * 4);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int kLo = max(winMin[r0], P.smin), kHi = P.p;
    if (kMode == LDL_K_OLD) kHi = P.p - 64;
    if (kMode == LDL_K_NEW) kLo = max(kLo, P.p - 64);
    kLo = P.p - ((P.p - kLo + 31) / 32) * 32;
    const int total = max(0, kHi - kLo);
    int chunk = (total + nSplit - 1) / nSplit;
    chunk = (chunk + KC - 1) / KC * KC;
    const int kb = kLo + sp * chunk, ke = min(kHi, kb + chunk);
    const int kl = lane & 31;
    int aR[8], aS0[8], aS[8], bR[8], bS0[8], bS[8];
    int64_t aOff[8], bOff[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int sr = 16 * w + 2 * i + (lane >> 5);
        const bool aLive = sr < nr, bLive = sr < np;
        aR[i] = aLive ? r0 + sr : r0;
        bR[i] = bLive ? P.p + sr : P.p;
        aS0[i] = sfirst[aR[i]];
        bS0[i] = sfirst[bR[i]];
        aOff[i] = rowoff[aR[i]];
        bOff[i] = rowoff[bR[i]];
        aS[i] = aLive ? aS0[i] : 0x7fffffff;
        bS[i] = bLive ? bS0[i] : 0x7fffffff;
    }
    float4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
    float va0[8], vb0[8], va1[8], vb1[8];
    double vd0 = 0.0, vd1 = 0.0;
    auto issue = [&](int k0, float (&va)[8], float (&vb)[8], double& vd) {
        const int k = k0 + kl;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            va[i] = lt[aOff[i] + min(max(k, aS0[i]), aR[i])];
            vb[i] = lt[bOff[i] + min(max(k, bS0[i]), bR[i])];
        }
        vd = diag[min(k0 + kl, ke - 1)];
    };
    auto step = [&](int k0, float (&va)[8], float (&vb)[8], double& vd) {
        __syncthreads();
        {
            const int k = k0 + kl;
            const bool kv = k < ke;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int sr = 16 * w + 2 * i + (lane >> 5);
                As[kl][sr] = (kv && k >= aS[i]) ? va[i] : 0.0f;
                Bs[kl][sr] = (kv && k >= bS[i]) ? (float)((double)vb[i] * vd) : 0.0f;
            }
        }
        __syncthreads();
        if (k0 + 2 * KC < ke) issue(k0 + 2 * KC, va, vb, vd);
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int kk = ks + (lane >> 4);
            const float av = As[kk][w * 16 + (lane & 15)];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const float bv = Bs[kk][cb * 16 + (lane & 15)];
                acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[cb], 0, 0, 0);
            }
        }
    };
    if (kb < ke) issue(kb, va0, vb0, vd0);
    if (kb + KC < ke) issue(kb + KC, va1, vb1, vd1);
    for (int k0 = kb; k0 < ke; k0 += 2 * KC) {
        step(k0, va0, vb0, vd0);
        if (k0 + KC < ke) step(k0 + KC, va1, vb1, vd1);
    }
    float* out = scratch + (((size_t)slot * nb + blk) * scratchRows + (r0 - P.p)) * 64;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = w * 16 + 4 * (lane >> 4) + v;
            const int j = cb * 16 + (lane & 15);
            if (i < nr) out[(size_t)i * 64 + j] = acc[cb][v];
        }
}

__device__ __forceinline__ void ldl_update_body(const PanelInfo P, int blk, int nb, int tile, int sp, int nSplit,
                                                int kMode, int slot, const float* __restrict__ lt,
                                                const int64_t* __restrict__ rowoff, const int* __restrict__ sfirst,
                                                const int* __restrict__ winMin,
                                                const double* __restrict__ diag, float* __restrict__ scratch,
                                                int scratchRows, char* smem) {
    ldl_update_body_f32(P, blk, nb, tile, sp, nSplit, kMode, slot, lt, rowoff, sfirst, winMin, diag, scratch,
                        scratchRows, smem);
}

__device__ __forceinline__ void ldl_update_body(const PanelInfo P, int blk, int nb, int tile, int sp, int nSplit,
                                                int kMode, int slot, const double* __restrict__ lt,
                                                const int64_t* __restrict__ rowoff, const int* __restrict__ sfirst,
                                                const int* __restrict__ winMin,
                                                const double* __restrict__ diag, double* __restrict__ scratch,
                                                int scratchRows, char* smem) {
    const int r0 = P.p + 64 * tile;
    if (r0 >= P.iend) return;
    const int nr = min(64, P.iend - r0);
    const int np = P.pe - P.p;
    constexpr int KC = LDL_KC;
    double(*As)[65] = reinterpret_cast<double(*)[65]>(smem);
    double(*Bs)[65] = reinterpret_cast<double(*)[65]>(smem + KC * 65 * 8);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int kLo = max(winMin[r0], P.smin), kHi = P.p;
    if (kMode == LDL_K_OLD) kHi = P.p - 64;
    if (kMode == LDL_K_NEW) kLo = max(kLo, P.p - 64);
    // start on a column = P.p (mod 16): with rows laid out aligned per block
    // every 32-column chunk is two whole 128-byte lines (the columns added
    // lie left of every row's first stored column and are masked)
    kLo = P.p - ((P.p - kLo + 15) / 16) * 16;
    const int total = max(0, kHi - kLo);
    int chunk = (total + nSplit - 1) / nSplit;
    chunk = (chunk + KC - 1) / KC * KC;
    const int kb = kLo + sp * chunk, ke = min(kHi, kb + chunk);
    // staging: wave w loads rows 16 w .. 16 w + 15 of both operands, two
    // rows per instruction with lane l on k = k0 + (l & 31) of row
    // 16 w + 2 i + (l >> 5), so every load covers 2 x 256 contiguous bytes
    static_assert(KC == 32, "the staging layout assumes 32 k per chunk");
    const int kl = lane & 31;
    int aR[8], aS0[8], aS[8], bR[8], bS0[8], bS[8];
    int64_t aOff[8], bOff[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int sr = 16 * w + 2 * i + (lane >> 5);
        const bool aLive = sr < nr, bLive = sr < np;
        // dead staging rows read a valid row (r0, P.p) and are masked to zero
        aR[i] = aLive ? r0 + sr : r0;
        bR[i] = bLive ? P.p + sr : P.p;
        aS0[i] = sfirst[aR[i]];
        bS0[i] = sfirst[bR[i]];
        aOff[i] = rowoff[aR[i]];
        bOff[i] = rowoff[bR[i]];
        aS[i] = aLive ? aS0[i] : 0x7fffffff;
        bS[i] = bLive ? bS0[i] : 0x7fffffff;
    }
    double4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = double4_t{0.0, 0.0, 0.0, 0.0};
    // register double buffer: the next chunk's operands are loaded while the
    // current one runs on the matrix cores; they are masked only when staged
    // into LDS, so nothing waits on them before the MFMAs.  Loads go to
    // indices clamped into each row's stored range (all in flight together,
    // no per-entry branch and memory wait).
    // two register sets, so the loads of the next two chunks are in flight
    double va0[8], vb0[8], vd0 = 0.0, va1[8], vb1[8], vd1 = 0.0;
    auto issue = [&](int k0, double (&va)[8], double (&vb)[8], double& vd) {
        const int k = k0 + kl;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            va[i] = lt[aOff[i] + min(max(k, aS0[i]), aR[i])];
            vb[i] = lt[bOff[i] + min(max(k, bS0[i]), bR[i])];
        }
        vd = diag[min(k0 + kl, ke - 1)];
    };
    // stage chunk k0 from its register set into LDS, then, once every wave
    // has staged, refill that set with chunk k0 + 2 KC and multiply
    auto step = [&](int k0, double (&va)[8], double (&vb)[8], double& vd) {
        __syncthreads();
        {
            // the panel rows are staged as L(P, k) D_k (vd = D of this lane's k)
            const int k = k0 + kl;
            const bool kv = k < ke;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int sr = 16 * w + 2 * i + (lane >> 5);
                As[kl][sr] = (kv && k >= aS[i]) ? va[i] : 0.0;
                Bs[kl][sr] = (kv && k >= bS[i]) ? vb[i] * vd : 0.0;
            }
        }
        __syncthreads();
        if (k0 + 2 * KC < ke) issue(k0 + 2 * KC, va, vb, vd);
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int kk = ks + (lane >> 4);
            const double av = As[kk][w * 16 + (lane & 15)];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const double bv = Bs[kk][cb * 16 + (lane & 15)];
                acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[cb], 0, 0, 0);
            }
        }
    };
    if (kb < ke) issue(kb, va0, vb0, vd0);
    if (kb + KC < ke) issue(kb + KC, va1, vb1, vd1);
    for (int k0 = kb; k0 < ke; k0 += 2 * KC) {
        step(k0, va0, vb0, vd0);
        if (k0 + KC < ke) step(k0 + KC, va1, vb1, vd1);
    }
    double* out = scratch + (((size_t)slot * nb + blk) * scratchRows + (r0 - P.p)) * 64;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = w * 16 + 4 * v + (lane >> 4);
            const int j = cb * 16 + (lane & 15);
            if (i < nr) out[(size_t)i * 64 + j] = acc[cb][v];
        }
}

// Step U alone (the first panel's zero partials, and the "new" part): grid
// (tiles x nSplit, blocks); partial sp of the k range goes to slot slot0 + sp.
template <typename LT>
__global__ __launch_bounds__(256) void k_ldl_update_mfma(const PanelInfo* __restrict__ panels, int panelIdx,
                                                         int numPanelsPerBlock, int nSplit, int kMode, int slot0,
                                                         const LT* __restrict__ lt,
                                                         const int64_t* __restrict__ rowoff,
                                                         const int* __restrict__ sfirst,
                                                         const int* __restrict__ winMin,
                                                         const double* __restrict__ diag,
                                                         LT* __restrict__ scratch, int scratchRows) {
    __shared__ __align__(16) char smem[LDL_U_SMEM_BYTES];
    const int blk = blockIdx.y;
    const PanelInfo P = panels[blk * numPanelsPerBlock + panelIdx];
    if (P.p < 0) return;
    const int tile = blockIdx.x / nSplit, sp = blockIdx.x % nSplit;
    if (P.p + 64 * tile >= P.iend) return;
    ldl_update_body(P, blk, gridDim.y, tile, sp, nSplit, kMode, slot0 + sp, lt, rowoff, sfirst, winMin, diag,
                    scratch, scratchRows, smem);
}
template __global__ void k_ldl_update_mfma<double>(const PanelInfo*, int, int, int, int, int, const double*,
                                                   const int64_t*, const int*, const int*, const double*, double*,
                                                   int);
template __global__ void k_ldl_update_mfma<float>(const PanelInfo*, int, int, int, int, int, const float*,
                                                  const int64_t*, const int*, const int*, const double*, float*, int);

// Step FT of a panel P, in two parts.  In the diagonal part, thread (r = tid / 4, q = tid % 4)
// holds 16 entries of one row in registers: columns q, q+4, ..., q+60 of the
// panel.  gather: A(row, P) minus the split-K partial sums of step U; loads
// are issued at indices clamped into the row's stored range and masked after
// (no per-entry branch and memory wait).
//  - ldl_panel_diag (one workgroup per block): factors the 64x64 diagonal
//    block right-looking in blocks of four columns, writes it back with D in
//    `diag`, and stores X = L_PP^-1 and 1 / D in Xg for the row tiles;
//  - ldl_panel_rows (two workgroups per further 64-row tile, in the next
//    launch): L(i, P) = rows X^T D^-1 on the matrix cores.
template <typename LT>
__device__ __forceinline__ void ldl_panel_diag(const PanelInfo P, int blk, int nb, int nSlots,
                                               LT* __restrict__ lt, const int64_t* __restrict__ rowoff,
                                               const int* __restrict__ sfirst, const LT* __restrict__ scratch,
                                               int scratchRows, double* __restrict__ diag,
                                               int* __restrict__ zeroPivot, double* __restrict__ Xg, char* smem) {
    const int tile = 0;
    const int r0 = P.p + 64 * tile;
    if (r0 >= P.iend) return;
    const int nr = min(64, P.iend - r0);
    const int np = P.pe - P.p;
    double(*Lc)[65] = reinterpret_cast<double(*)[65]>(smem + LDL_FT_A);  // factored diagonal block: L below, D on the diagonal
    double(*Xi)[65] = reinterpret_cast<double(*)[65]>(smem + LDL_FT_X);  // L_PP^-1 (unit lower)
    double(*colb)[4 * 64] = reinterpret_cast<double(*)[4 * 64]>(smem + LDL_FT_COL);
    double* rD = reinterpret_cast<double*>(smem + LDL_FT_RD);
    int64_t* rowOff = reinterpret_cast<int64_t*>(smem + LDL_FT_OFF);
    int* rowS = reinterpret_cast<int*>(smem + LDL_FT_SI);
    const int tid = threadIdx.x, fr = tid >> 2, fq = tid & 3;
    const size_t qstride = (size_t)nb * scratchRows * 64;

    // gather row `rowInPanel` of the updated panel block; entry m is column
    // k = fq + 4 m of the panel, kept where live && k < kmax && inside the profile
    auto gather = [&](int gi, int rowInPanel, bool live, int kmax, double (&a)[16], int& si, int64_t& off) {
        const int gr = live ? gi : P.p;
        si = sfirst[gr];
        off = rowoff[gr];
        const int rp = live ? rowInPanel : 0;
        double raw[16], ws[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int c = P.p + fq + 4 * m;
            raw[m] = lt[off + min(max(c, si), gr)];
            ws[m] = 0.0;
        }
        const LT* sp = scratch + ((size_t)blk * scratchRows + rp) * 64 + fq;
        for (int q = 0; q < nSlots; ++q) {
#pragma unroll
            for (int m = 0; m < 16; ++m) ws[m] += (double)sp[(size_t)q * qstride + 4 * m];
        }
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int k = fq + 4 * m;
            a[m] = (live && k < kmax && P.p + k >= si) ? raw[m] - ws[m] : 0.0;
        }
    };

    // --- 1-2: the diagonal block, row fr
    double a[16];
    int dsi;
    int64_t doff;
    gather(P.p + fr, fr, fr < np, fr + 1, a, dsi, doff);
    // right-looking in blocks of four columns, one barrier per block: the 4x4
    // diagonal block is factored by every thread, then each row forms its
    // four L entries and applies the rank-4 update to its later columns
#pragma unroll
    for (int jb = 0; jb < 16; ++jb) {
        const int j0 = 4 * jb;
        if (j0 < np) {
            double(*cb)[64] = reinterpret_cast<double(*)[64]>(&colb[jb & 1][0]);  // cb[c][row] = A(row, j0 + c)
            cb[fq][fr] = a[jb];  // thread (fr, fq) holds column j0 + fq
            __syncthreads();
            // 4x4 block: d[c], l[r][c] = L(j0 + r, j0 + c); columns past np are inert
            double d[4], rd[4], l[4][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const bool valid = j0 + c < np;
                double dc = cb[c][j0 + c];
#pragma unroll
                for (int c2 = 0; c2 < c; ++c2) dc -= l[c][c2] * l[c][c2] * d[c2];
                d[c] = valid ? dc : 1.0;
                {
                    // 1 / d by v_rcp_f64 and two Newton steps: a shorter
                    // dependency chain than the IEEE division sequence
                    double r = __builtin_amdgcn_rcp(d[c]);
                    r = __builtin_fma(r, __builtin_fma(-d[c], r, 1.0), r);
                    r = __builtin_fma(r, __builtin_fma(-d[c], r, 1.0), r);
                    rd[c] = r;
                }
                if (valid && dc == 0.0 && tid == 0 && tile == 0) atomicMin(zeroPivot, P.p + j0 + c + 1);
#pragma unroll
                for (int r = c + 1; r < 4; ++r) {
                    double v = cb[c][j0 + r];
#pragma unroll
                    for (int c2 = 0; c2 < c; ++c2) v -= l[r][c2] * l[c][c2] * d[c2];
                    l[r][c] = valid ? v * rd[c] : 0.0;
                }
            }
            // this row's L(fr, j0 + c)
            double lf[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                double v = cb[c][fr];
#pragma unroll
                for (int c2 = 0; c2 < c; ++c2) v -= lf[c2] * d[c2] * l[c][c2];
                lf[c] = j0 + c < np ? v * rd[c] : 0.0;
            }
            // rank-4 update of the later columns k <= fr: u_k[c] = D_c L(k, j0 + c)
#pragma unroll
            for (int m = jb + 1; m < 16; ++m) {
                const int k = 4 * m + fq;
                double u[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    double v = cb[c][k];
#pragma unroll
                    for (int c2 = 0; c2 < c; ++c2) v -= u[c2] * l[c][c2];
                    u[c] = j0 + c < np ? v : 0.0;
                }
                const double upd = a[m] - (lf[0] * u[0] + lf[1] * u[1] + lf[2] * u[2] + lf[3] * u[3]);
                a[m] = k <= fr ? upd : a[m];
            }
            // this thread's own column of the block: L below the diagonal, D on it
            const int jc = j0 + fq;
            const double lmine = fq == 0 ? lf[0] : (fq == 1 ? lf[1] : (fq == 2 ? lf[2] : lf[3]));
            const double dmine = fq == 0 ? d[0] : (fq == 1 ? d[1] : (fq == 2 ? d[2] : d[3]));
            a[jb] = jc < fr ? lmine : (jc == fr && jc < np ? dmine : a[jb]);
        }
    }
    {
        const int gi = P.p + fr;
        if (fr < np) {
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int k = fq + 4 * m;
                if (k <= fr && P.p + k >= dsi) lt[doff + P.p + k] = (LT)a[m];
                if (k == fr) diag[gi] = a[m];
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) Lc[fr][fq + 4 * m] = a[m];
    __syncthreads();
    const int w = tid >> 6;
    // X = L_PP^-1 (rows past np are the identity: L is zero there) and 1 / D,
    // for the row tiles of the next launch
    if (tid < 64) rD[tid] = tid < np ? 1.0 / Lc[tid][tid] : 0.0;
    unit_lower_inverse(Lc, Xi, reinterpret_cast<double(*)[17]>(smem + LDL_FT_T + w * 16 * 17 * 8));
    __syncthreads();
    for (int id = tid; id < 64 * 64; id += 256) Xg[id] = Xi[id >> 6][id & 63];
    if (tid < 64) Xg[64 * 64 + tid] = rD[tid];
}

// Rows [32 half, 32 half + 32) of a further row tile (half 0 or 1, one
// workgroup each: twice the workgroups of one per tile, each with half the
// gather and product; factor 179.6 -> 177.1 ms on the 64x64x32 brick).
// Thread (r = tid / 8, q = tid % 8) gathers 8 entries of one row (columns q,
// q+8, ..., q+56); wave w forms the 16 x 32 product block at rows 16 (w & 1),
// columns 32 (w >> 1).
template <typename LT>
__device__ __forceinline__ void ldl_panel_rows(const PanelInfo P, int blk, int nb, int tile, int half, int nSlots,
                                                    LT* __restrict__ lt, const int64_t* __restrict__ rowoff,
                                                    const int* __restrict__ sfirst, const LT* __restrict__ scratch,
                                                    int scratchRows, const double* __restrict__ Xg, char* smem) {
    const int r0 = P.p + 64 * tile + 32 * half;
    if (r0 >= P.iend) return;
    const int nr = min(32, P.iend - r0);
    const int np = P.pe - P.p;
    double(*Lc)[65] = reinterpret_cast<double(*)[65]>(smem + LDL_FT_A);
    double(*Xi)[65] = reinterpret_cast<double(*)[65]>(smem + LDL_FT_X);
    double* rD = reinterpret_cast<double*>(smem + LDL_FT_RD);
    int64_t* rowOff = reinterpret_cast<int64_t*>(smem + LDL_FT_OFF);
    int* rowS = reinterpret_cast<int*>(smem + LDL_FT_SI);
    const int tid = threadIdx.x, fr = tid >> 3, fq = tid & 7;
    const size_t qstride = (size_t)nb * scratchRows * 64;
    const int lane = tid & 63, w = tid >> 6;
    const bool live = fr < nr;
    const int gr = live ? r0 + fr : P.p;
    const int si = sfirst[gr];
    const int64_t off = rowoff[gr];
    const int rp = live ? (r0 - P.p) + fr : 0;
    double raw[8], ws[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int c = P.p + fq + 8 * m;
        raw[m] = lt[off + min(max(c, si), gr)];
        ws[m] = 0.0;
    }
    const LT* sp = scratch + ((size_t)blk * scratchRows + rp) * 64 + fq;
    for (int q = 0; q < nSlots; ++q) {
#pragma unroll
        for (int m = 0; m < 8; ++m) ws[m] += (double)sp[(size_t)q * qstride + 8 * m];
    }
    for (int id = tid; id < 64 * 64; id += 256) Xi[id >> 6][id & 63] = Xg[id];
    if (tid < 64) rD[tid] = Xg[64 * 64 + tid];
    if (fq == 0) {
        rowOff[fr] = off;
        rowS[fr] = live ? si : 0x7fffffff;
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int k = fq + 8 * m;
        Lc[fr][k] = (live && k < np && P.p + k >= si) ? raw[m] - ws[m] : 0.0;
    }
    __syncthreads();
    const int rb = w & 1, cb0 = 2 * (w >> 1);
    double4_t acc[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) acc[q] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < 64; ks += 4) {
        const int kk = ks + (lane >> 4);
        const double av = Lc[rb * 16 + (lane & 15)][kk];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            const double bv = Xi[(cb0 + cb) * 16 + (lane & 15)][kk];
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = rb * 16 + 4 * v + (lane >> 4);
        if (i < nr) {
            const int64_t ro = rowOff[i];
            const int s0 = rowS[i];
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int j = (cb0 + cb) * 16 + (lane & 15);
                if (j < np && P.p + j >= s0) lt[ro + P.p + j] = (LT)(acc[cb][v] * rD[j]);
            }
        }
    }
}

// Factor launch L_q (one per panel index q, every block's panel q):
//  - x < numBlocks: the "new" update of panel q's diagonal tile (the
//    previous panel's columns), then ldl_panel_diag, which factors the
//    diagonal block and stores L_PP^-1 and 1 / D in ftX;
//  - the next numBlocks * (numTiles - 1): the "new" update of panel q's
//    other tiles into scratch slot nSplit of bufQ;
//  - the rest: the "old" update of panel q + 1 (every column before panel
//    q, nSplit partials into bufQ1), which needs only panels finished by
//    earlier launches, so it overlaps the diagonal factor.  Tile-major
//    order: the tiles nearest the diagonal have the longest column ranges
//    and are dispatched first.
// The row tiles of panel q follow in k_ldl_rows.
template <typename LT>
__global__ __launch_bounds__(256) void k_ldl_lead(const PanelInfo* __restrict__ panels, int panelIdx,
                                                  int numPanelsPerBlock, int numBlocks, int numTiles, int nSplit,
                                                  LT* __restrict__ lt, const int64_t* __restrict__ rowoff,
                                                  const int* __restrict__ sfirst, const int* __restrict__ winMin,
                                                  double* __restrict__ diag, LT* __restrict__ bufQ,
                                                  LT* __restrict__ bufQ1, int scratchRows,
                                                  int* __restrict__ zeroPivot, double* __restrict__ ftX) {
    __shared__ __align__(16) char smem[LDL_STEP_SMEM_BYTES];
    const int x = blockIdx.x;
    const int nNew = numBlocks * (numTiles - 1);
    if (x < numBlocks) {
        const int blk = x;
        const PanelInfo P = panels[blk * numPanelsPerBlock + panelIdx];
        if (P.p < 0 || P.p >= P.iend) return;
        ldl_update_body(P, blk, numBlocks, 0, 0, 1, LDL_K_NEW, nSplit, lt, rowoff, sfirst, winMin, diag, bufQ,
                        scratchRows, smem);
        __syncthreads();  // its partial sums (global) are read back by the gather
        ldl_panel_diag(P, blk, numBlocks, nSplit + 1, lt, rowoff, sfirst, bufQ, scratchRows, diag, zeroPivot,
                       ftX + (size_t)blk * (64 * 64 + 64), smem);
    } else if (x < numBlocks + nNew) {
        const int id = x - numBlocks;
        const int blk = id / (numTiles - 1), tile = 1 + id % (numTiles - 1);
        const PanelInfo P = panels[blk * numPanelsPerBlock + panelIdx];
        if (P.p < 0 || P.p + 64 * tile >= P.iend) return;
        ldl_update_body(P, blk, numBlocks, tile, 0, 1, LDL_K_NEW, nSplit, lt, rowoff, sfirst, winMin, diag, bufQ,
                        scratchRows, smem);
    } else {
        if (panelIdx + 1 >= numPanelsPerBlock) return;
        const int id = x - numBlocks - nNew;
        const int tile = id / (numBlocks * nSplit), blk = (id / nSplit) % numBlocks, sp = id % nSplit;
        const PanelInfo P = panels[blk * numPanelsPerBlock + panelIdx + 1];
        if (P.p < 0 || P.p + 64 * tile >= P.iend) return;
        ldl_update_body(P, blk, numBlocks, tile, sp, nSplit, LDL_K_OLD, sp, lt, rowoff, sfirst, winMin, diag, bufQ1,
                        scratchRows, smem);
    }
}
template __global__ void k_ldl_lead<double>(const PanelInfo*, int, int, int, int, int, double*, const int64_t*,
                                            const int*, const int*, double*, double*, double*, int, int*, double*);
template __global__ void k_ldl_lead<float>(const PanelInfo*, int, int, int, int, int, float*, const int64_t*,
                                           const int*, const int*, double*, float*, float*, int, int*, double*);

// Factor launch R_q: the row tiles >= 1 of every block's panel q (grid
// numBlocks * (numTiles - 1)), against the L_PP^-1 that L_q stored.
template <typename LT>
__global__ __launch_bounds__(256) void k_ldl_rows(const PanelInfo* __restrict__ panels, int panelIdx,
                                                  int numPanelsPerBlock, int numBlocks, int numTiles, int nSplit,
                                                  LT* __restrict__ lt, const int64_t* __restrict__ rowoff,
                                                  const int* __restrict__ sfirst, const LT* __restrict__ bufQ,
                                                  int scratchRows, const double* __restrict__ ftX) {
    __shared__ __align__(16) char smem[LDL_STEP_SMEM_BYTES];
    const int half = blockIdx.x & 1, idx = blockIdx.x >> 1;
    const int blk = idx / (numTiles - 1), tile = 1 + idx % (numTiles - 1);
    const PanelInfo P = panels[blk * numPanelsPerBlock + panelIdx];
    if (P.p < 0 || P.p + 64 * tile + 32 * half >= P.iend) return;
    ldl_panel_rows(P, blk, numBlocks, tile, half, nSplit + 1, lt, rowoff, sfirst, bufQ, scratchRows,
                   ftX + (size_t)blk * (64 * 64 + 64), smem);
}
template __global__ void k_ldl_rows<double>(const PanelInfo*, int, int, int, int, int, double*, const int64_t*,
                                            const int*, const double*, int, const double*);
template __global__ void k_ldl_rows<float>(const PanelInfo*, int, int, int, int, int, float*, const int64_t*,
                                           const int*, const float*, int, const double*);

// ---------------------------------------------------------------------------
// Triangular sweeps on the factored profile, one 64-row tile per workgroup in
// dependency order.  Tiles are handed out by an atomic ticket, so a tile is
// only ever processed after every tile it depends on has been claimed by a
// running workgroup; waits on their completion flags are bounded (spinLimit),
// after which the kernel records a timeout and exits.
//
// Forward (z = L^-1 b, Working/AStrO_solvers.f90:95-107): tile rows need z of
// the earlier tiles their profiles reach.  Backward (x = D^-1 z - L^T x,
// :109-123): tile columns need x of the later tiles whose rows reach them.

// Broadcast lane `src` (wave-uniform) of a double with two v_readlane.
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Cross-workgroup hand-off without cache-maintenance fences: the values a
// tile publishes and its flag are written with agent-scope relaxed atomic
// stores and read with agent-scope relaxed atomic loads, which go to the
// device-coherent level instead of the per-XCD L2.  The producer waits for
// its value stores to complete (vmcnt 0) before it stores the flag, so a
// consumer that has seen the flag reads the values; 64-bit stores are
// single-copy atomic, so a polled value is either the sentinel or final.  Agent-scope
// acquire/release would instead write back and invalidate the whole L2 of the
// XCD at every hand-off.
__device__ __forceinline__ double coherent_load(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coherent_store(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stores_complete() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Each sweep starts with its output vector filled with ASTRO_SENTINEL: a
// consumer polls the values themselves (one 64-bit load per lane), so the
// neighbouring tile's hand-off needs no separate flag round trip.  Flags
// (stored after the tile's value stores have completed) still tell which
// whole tiles are done, for bulk catch-up.

// Poll v[base + lane] (lane < cnt) until every lane holds a produced value;
// lanes >= cnt return 0.  `backoff` (wave-uniform, 0..12) adds that many
// ~1000-cycle sleeps between polls: waiters whose own result is many tiles
// away poll rarely, which keeps the few freshly written lines every waiter
// wants from being hammered.  Returns false if the bounded wait expired.
// The timeout flag is shared by the whole launch: once any wave's bounded
// wait has expired, every other waiter sees it within 256 polls and gives up
// too, and tiles claimed afterwards exit at once, so one stuck tile ends the
// launch in about one spin limit instead of one per waiting tile.
__device__ __forceinline__ bool timed_out(const int* timeout) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0;
}

__device__ __forceinline__ bool poll_values(const double* v, int base, int cnt, int backoff, long spinLimit,
                                            int* timeout, double& out) {
    const int lane = threadIdx.x & 63;
    long it = 0;
    while (true) {
        out = lane < cnt ? coherent_load(&v[base + lane]) : 0.0;
        const bool pending = lane < cnt && __double_as_longlong(out) == ASTRO_SENTINEL;
        if (__ballot(pending) == 0ull) return true;
        if (++it > spinLimit) {
            if (lane == 0) atomicExch(timeout, 1);
            return false;
        }
        if ((it & 255) == 0 && timed_out(timeout)) return false;
        for (int k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(16);
        __builtin_amdgcn_s_sleep(1);
    }
}

// Non-spinning scans of the completion flags of a chain's tiles (indexed by
// blk).  first_pending_up: the first blk in [from, last) whose flag does not
// hold `epoch` (last if none); first_pending_down: the first in (last, from]
// going down (last if none).
__device__ __forceinline__ int first_pending_up(const int* flags, int from, int last, int epoch) {
    const int lane = threadIdx.x & 63;
    while (from < last) {
        const int J = from + lane;
        const bool done =
            J >= last || __hip_atomic_load(&flags[J], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
        const unsigned long long pend = __ballot(!done);
        if (pend) return from + (int)__builtin_ctzll(pend);
        from += 64;
    }
    return last;
}
__device__ __forceinline__ int first_pending_down(const int* flags, int from, int last, int epoch) {
    const int lane = threadIdx.x & 63;
    while (from > last) {
        const int J = from - lane;
        const bool done =
            J <= last || __hip_atomic_load(&flags[J], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
        const unsigned long long pend = __ballot(!done);
        if (pend) return from - (int)__builtin_ctzll(pend);
        from -= 64;
    }
    return last;
}

// Waiters closer than this many tiles (of their chain) to the tile they
// wait for follow the chain tile by tile without back-off; farther ones poll
// with sleeps and catch up in bulk.  Measured on the 64x64x32 brick with one
// workgroup per CU: forward 4 / 8 / 16 / 32 / 64 / 256 tiles: 3.26 / 3.09 /
// 2.96 / 2.87 / 2.85 / 2.86 ms; backward 2 / 4 / 8: 3.60 / 3.59 / 3.67 ms.
#define ASTRO_NEAR_FWD 64
#define ASTRO_NEAR_BWD 4

__device__ __forceinline__ int backoff_for(int chainDist, int nearTiles) {
    const int b = chainDist - nearTiles;
    return b < 0 ? 0 : (b > 12 ? 12 : b);
}

__device__ __forceinline__ void publish_flag(int* flag, int epoch) {
    __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// C[i][c] = L(rowStart + i, colStart + c) inside the profile (every column
// used here lies left of its row), zero elsewhere and beyond nRows x nCols.
template <typename LT>
__device__ __forceinline__ void load_coupling(double (*C)[65], int rowStart, int nRows, int colStart, int nCols,
                                              const LT* __restrict__ lt, const int64_t* __restrict__ rowoff,
                                              const int* __restrict__ sfirst) {
    for (int id = threadIdx.x; id < 64 * 64; id += 256) {
        const int i = id >> 6, c = id & 63;
        double u = 0.0;
        if (i < nRows && c < nCols) {
            const int gi = rowStart + i;
            if (colStart + c >= sfirst[gi]) u = lt[rowoff[gi] + colStart + c];
        }
        C[i][c] = u;
    }
}

// First pass over the tiles (one workgroup each): L_tt^-1 (blocked inverse
// on the matrix cores), then its products with the coupling blocks to the
// two preceding and two following tiles.  Rows beyond the tile's end are the
// identity in L_tt.
template <typename LT>
__global__ __launch_bounds__(256) void k_sweep_blocks(const SweepBlock* __restrict__ blocks,
                                                      const LT* __restrict__ lt,
                                                      const int64_t* __restrict__ rowoff,
                                                      const int* __restrict__ sfirst, double* __restrict__ out) {
    __shared__ double L[64][65];  // strictly lower L_tt, then one coupling block at a time
    __shared__ double X[64][65];  // L_tt^-1
    __shared__ double Tw[4][16][17];
    const SweepBlock S = blocks[blockIdx.x];
    const int tid = threadIdx.x;
    const int nr = S.r1 - S.r0;
    double* o = out + (size_t)blockIdx.x * SWEEP_KINDS * SWEEP_BLOCK;
    // L_tt strictly below the diagonal (zero on and above it)
    for (int id = tid; id < 64 * 64; id += 256) {
        const int i = id >> 6, c = id & 63;
        double l = 0.0;
        if (i < nr && c < i) {
            const int gi = S.r0 + i;
            if (S.r0 + c >= sfirst[gi]) l = lt[rowoff[gi] + S.r0 + c];
        }
        L[i][c] = l;
    }
    __syncthreads();
    unit_lower_inverse(L, X, Tw[tid >> 6]);
    for (int id = tid; id < 64 * 64; id += 256) {
        const int cc = id >> 6, ii = id & 63;
        o[SWEEP_DINV * SWEEP_BLOCK + cc * 64 + ii] = X[ii][cc];
        o[SWEEP_DINVT * SWEEP_BLOCK + cc * 64 + ii] = X[cc][ii];
    }
    // F = X L(t, t-1), H = X L(t, t-2), G = X^T L(t+1, t)^T, HB = X^T L(t+2, t)^T
    load_coupling(L, S.r0, nr, S.pm1, S.r0 - S.pm1, lt, rowoff, sfirst);
    __syncthreads();
    mfma_block64<false, false>(X, L, nullptr, o + SWEEP_F * SWEEP_BLOCK);
    __syncthreads();
    load_coupling(L, S.r0, nr, S.pm2, S.pm1 - S.pm2, lt, rowoff, sfirst);
    __syncthreads();
    mfma_block64<false, false>(X, L, nullptr, o + SWEEP_H * SWEEP_BLOCK);
    __syncthreads();
    load_coupling(L, S.r1, S.q1 - S.r1, S.r0, nr, lt, rowoff, sfirst);
    __syncthreads();
    mfma_block64<true, true>(X, L, nullptr, o + SWEEP_G * SWEEP_BLOCK);
    __syncthreads();
    load_coupling(L, S.q1, S.q2 - S.q1, S.r0, nr, lt, rowoff, sfirst);
    __syncthreads();
    mfma_block64<true, true>(X, L, nullptr, o + SWEEP_HB * SWEEP_BLOCK);
}
template __global__ void k_sweep_blocks<double>(const SweepBlock*, const double*, const int64_t*, const int*,
                                                double*);
template __global__ void k_sweep_blocks<float>(const SweepBlock*, const float*, const int64_t*, const int*, double*);

// A column-major 64 x 64 block from global memory into As[i][c] (zero if null).
__device__ __forceinline__ void load_block64(const double* __restrict__ src, double (*As)[65]) {
    for (int id = threadIdx.x; id < 64 * 64; id += 256) {
        const int c = id >> 6, i = id & 63;
        As[i][c] = src ? src[id] : 0.0;
    }
}

// Second pass: the products of a tile's blocks with its neighbour's,
// FH = F_t H_{t-1}, K = F_t F_{t-1} - H_t, GHB = G_t HB_{t+1},
// KB = G_t G_{t+1} - HB_t.
__global__ __launch_bounds__(256) void k_sweep_blocks_pair(const SweepBlock* __restrict__ blocks,
                                                           double* __restrict__ out) {
    __shared__ double As[64][65];
    __shared__ double Bs[64][65];
    const SweepBlock S = blocks[blockIdx.x];
    double* o = out + (size_t)blockIdx.x * SWEEP_KINDS * SWEEP_BLOCK;
    const double* p = S.prevBlk >= 0 ? out + (size_t)S.prevBlk * SWEEP_KINDS * SWEEP_BLOCK : nullptr;
    const double* n = S.nextBlk >= 0 ? out + (size_t)S.nextBlk * SWEEP_KINDS * SWEEP_BLOCK : nullptr;
    load_block64(o + SWEEP_F * SWEEP_BLOCK, As);
    load_block64(p ? p + SWEEP_H * SWEEP_BLOCK : nullptr, Bs);
    __syncthreads();
    mfma_block64<false, false>(As, Bs, nullptr, o + SWEEP_FH * SWEEP_BLOCK);
    __syncthreads();
    load_block64(p ? p + SWEEP_F * SWEEP_BLOCK : nullptr, Bs);
    __syncthreads();
    mfma_block64<false, false>(As, Bs, o + SWEEP_H * SWEEP_BLOCK, o + SWEEP_K * SWEEP_BLOCK);
    __syncthreads();
    load_block64(o + SWEEP_G * SWEEP_BLOCK, As);
    load_block64(n ? n + SWEEP_HB * SWEEP_BLOCK : nullptr, Bs);
    __syncthreads();
    mfma_block64<false, false>(As, Bs, nullptr, o + SWEEP_GHB * SWEEP_BLOCK);
    __syncthreads();
    load_block64(n ? n + SWEEP_G * SWEEP_BLOCK : nullptr, Bs);
    __syncthreads();
    mfma_block64<false, false>(As, Bs, o + SWEEP_HB * SWEEP_BLOCK, o + SWEEP_KB * SWEEP_BLOCK);
}

// This wave's share of a block product: sum over 16 columns k of
// m[k] v(lane lane0 + k), m[k] = M(lane, c(k)).  The 16 values go through
// this wave's LDS slot vs and come back as broadcast reads (the wave's own
// LDS accesses stay in order, so no barrier is needed).
__device__ __forceinline__ double block_product16(const double (&m)[16], double v, int lane0, double* vs) {
    const int lane = threadIdx.x & 63;
    if (lane >= lane0 && lane < lane0 + 16) vs[lane - lane0] = v;
    __builtin_amdgcn_wave_barrier();
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
        s0 += m[k] * vs[k];
        s1 += m[k + 1] * vs[k + 1];
    }
    __builtin_amdgcn_wave_barrier();
    return s0 + s1;
}

// Values of rows [base, base + cnt) in lanes 0..cnt-1 (others 0), loaded
// speculatively; near_wait completes them.
struct NearVals {
    int base, cnt;
    double v;
};
__device__ __forceinline__ NearVals near_load(const double* v, int base, int cnt) {
    const int lane = threadIdx.x & 63;
    NearVals n{base, cnt, 0.0};
    n.v = lane < cnt ? coherent_load(&v[base + lane]) : 0.0;
    return n;
}
__device__ __forceinline__ bool near_wait(const double* v, NearVals& n, long spinLimit, int* timeout) {
    const int lane = threadIdx.x & 63;
    if (!__ballot(lane < n.cnt && __double_as_longlong(n.v) == ASTRO_SENTINEL)) return true;
    return poll_values(v, n.base, n.cnt, 0, spinLimit, timeout, n.v);
}
// Waits for a, re-polling b alongside it whenever b is still pending, so b's
// arrival is seen while waiting for a.
__device__ __forceinline__ bool near_wait_with(const double* va, NearVals& a, const double* vb, NearVals& b,
                                               long spinLimit, int* timeout) {
    const int lane = threadIdx.x & 63;
    long it = 0;
    while (true) {
        const bool pa = __ballot(lane < a.cnt && __double_as_longlong(a.v) == ASTRO_SENTINEL) != 0ull;
        if (!pa) return true;
        const bool pb = __ballot(lane < b.cnt && __double_as_longlong(b.v) == ASTRO_SENTINEL) != 0ull;
        if (++it > spinLimit) {
            if (lane == 0) atomicExch(timeout, 1);
            return false;
        }
        if ((it & 255) == 0 && timed_out(timeout)) return false;
        __builtin_amdgcn_s_sleep(1);
        a.v = lane < a.cnt ? coherent_load(&va[a.base + lane]) : 0.0;
        if (pb) b.v = lane < b.cnt ? coherent_load(&vb[b.base + lane]) : 0.0;
    }
}

// The end of a sweep tile, after the far part (including the third near
// tile) has been reduced to h (lane = row, zero beyond the tile): g_t = D h is
// published to gout; then
//   out_t = g_t - B1 g_{near1} + B3 v_{near3} + B2 v_{near2}
// (v = the sweep's own output, gin = gout of the same launch).  Every product
// is split over the four waves (16 columns each: dv and kv in registers, B1
// and B3 from LDS; v3 is broadcast from lanes v3lane0 + k); only the B2
// product follows the wait on v_{near2}.
__device__ __forceinline__ bool lookahead_tail(int r0, int nr, int nNear, int flagIdx, double h,
                                               const double (&dv)[16], const double (&kv)[16],
                                               const double (*B1s)[64], const double (*B3s)[64], double v3,
                                               int v3lane0, double* gout, double* out, int* flags, int epoch,
                                               long spinLimit, int* timeout, NearVals& n1, NearVals& n2,
                                               double (*acc)[64], double (*acc2)[64], double* vs) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    acc[w][lane] = block_product16(dv, h, 16 * w, vs);
    __syncthreads();
    const double gt = (acc[0][lane] + acc[1][lane]) + (acc[2][lane] + acc[3][lane]);
    if (w == 0 && lane < nr) coherent_store(&gout[r0 + lane], gt);
    double pw = 0.0;
    if (nNear >= 3) {
        double m[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = B3s[k][lane];
        pw += block_product16(m, v3, v3lane0, vs);
    }
    if (nNear >= 1) {
        if (!near_wait_with(gout, n1, out, n2, spinLimit, timeout)) return false;
        double m[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = B1s[k][lane];
        pw -= block_product16(m, n1.v, 0, vs);
    }
    if (nNear >= 2) {
        if (!near_wait(out, n2, spinLimit, timeout)) return false;
        pw += block_product16(kv, n2.v, 0, vs);
    }
    acc2[w][lane] = pw;
    __syncthreads();
    if (w == 0) {
        const double a = gt + ((acc2[0][lane] + acc2[1][lane]) + (acc2[2][lane] + acc2[3][lane]));
        if (lane < nr) coherent_store(&out[r0 + lane], a);
        stores_complete();
        if (lane == 0) publish_flag(&flags[flagIdx], epoch);
    }
    return true;
}

// Forward sweep tile.  Phase 1 accumulates the contributions of the far
// tiles (five or more tiles back; wave w owns rows w*16..+15, lane =
// column) and of the tiles four and three back (their coupling entries
// prefetched at the start), partials reduced across lanes once; the tail
// combines the near tiles through the precomputed blocks.
template <typename LT>
__global__ __launch_bounds__(256, 1) void k_sweep_forward(const SweepTile* __restrict__ tiles, int numTiles,
                                                          const LT* __restrict__ lt,
                                                          const int64_t* __restrict__ rowoff,
                                                          const int* __restrict__ sfirst,
                                                          const double* __restrict__ blocks,
                                                          const double* __restrict__ b, double* __restrict__ z,
                                                          double* __restrict__ g, int* __restrict__ flags,
                                                          int* __restrict__ ticket, int epoch, long spinLimit,
                                                          int* __restrict__ timeout, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ int sT;
    __shared__ double part[4][16];
    __shared__ double acc[4][64];
    __shared__ double acc2[4][64];
    __shared__ double B1s[4][16][64];  // wave w's 16 columns of F_t
    __shared__ double B3s[4][16][64];  // and of F_t H_{t-1}
    __shared__ double vsh[4][16];      // per-wave broadcast slot
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) sT = timed_out(timeout) ? numTiles : atomicAdd(ticket, 1);
    __syncthreads();
    const int t = sT;
    if (t >= numTiles) return;
    const SweepTile T = tiles[t];
    const int r0 = T.r0, nr = T.r1 - T.r0, nNear = T.nNear;
    // this wave's 16 columns of K_t, coalesced; F_t and F_t H_{t-1} go to
    // LDS now, far from the critical path (each wave reads back only what it
    // wrote)
    const double* Bt = blocks + (size_t)T.blk * SWEEP_KINDS * SWEEP_BLOCK;
    double kv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) kv[k] = nNear >= 2 ? Bt[SWEEP_K * SWEEP_BLOCK + (16 * w + k) * 64 + lane] : 0.0;
    if (nNear >= 1) {
#pragma unroll
        for (int k = 0; k < 16; ++k) B1s[w][k][lane] = Bt[SWEEP_F * SWEEP_BLOCK + (16 * w + k) * 64 + lane];
    }
    if (nNear >= 3) {
#pragma unroll
        for (int k = 0; k < 16; ++k) B3s[w][k][lane] = Bt[SWEEP_FH * SWEEP_BLOCK + (16 * w + k) * 64 + lane];
    }
    const double bl = lane < nr ? b[r0 + lane] : 0.0;
    // the wave's rows as 32-bit offsets from its first row's (rows beyond
    // the tile read that row's entries: column c < r0 is inside every row's
    // addressable range, so the unselected loads stay in the profile)
    const int rb = min(r0 + 16 * w, T.r1 - 1);
    const LT* ltw = lt + rowoff[rb];
    int rs[16], rd[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = w * 16 + r;
        rs[r] = i < nr ? sfirst[r0 + i] : 0x7fffffff;
        rd[r] = i < nr ? (int)(rowoff[r0 + i] - rowoff[rb]) : 0;
    }
    // coupling entries to the columns of the third and fourth tiles back
    double lv3[16], lv4[16];
    {
        const int c = r0 - 192 + lane, c4 = r0 - 256 + lane;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            lv3[r] = nNear >= 3 ? ltw[(unsigned int)(rd[r] + c)] : 0.0;
            lv4[r] = nNear >= 4 ? ltw[(unsigned int)(rd[r] + c4)] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            lv3[r] = c >= rs[r] ? lv3[r] : 0.0;
            lv4[r] = c4 >= rs[r] ? lv4[r] : 0.0;
        }
    }
    double p[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) p[r] = 0.0;
    // Far tiles: blk farBlk .. blk-5, whose rows are the contiguous columns
    // [r0 - 64 (blk - farBlk), r0 - 256).  Each wave streams the columns of
    // the tiles already published and re-polls at the first pending one; no
    // block-wide barrier.
    if (T.farBlk >= 0) {
        const int bEnd = T.blk - 4;
        int J = T.farBlk;
        int c0 = r0 - 64 * (T.blk - J);
        bool front = false;
        while (J < bEnd) {
            if (!front) {
                // bulk: every column of the published prefix of far tiles
                const int pend = first_pending_up(flags, J, bEnd, epoch);
                const int cStop = r0 - 64 * (T.blk - pend);
                // two 64-column chunks per step, so 32 row loads are in
                // flight (one workgroup per CU leaves four waves to cover the
                // memory latency); a lone last chunk reloads the first one
                // and adds it times zero (profile entries are finite)
#pragma unroll 1
                for (int cb = c0; cb < cStop; cb += 128) {
                    const int c = cb + lane;
                    const bool two = cb + 128 <= cStop;
                    const int c2 = two ? c + 64 : c;
                    const double zc = coherent_load(&z[c]);
                    const double zc2 = two ? coherent_load(&z[c2]) : 0.0;
                    // unconditional loads (a column left of a row's first
                    // stored column addresses earlier rows' entries, still
                    // inside the profile), then a select: no per-row branch
                    // and wait
                    double lv[16], lv2[16];
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        lv[r] = ltw[(unsigned int)(rd[r] + c)];
                        lv2[r] = ltw[(unsigned int)(rd[r] + c2)];
                    }
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        p[r] += (c >= rs[r] ? lv[r] : 0.0) * zc;
                        p[r] += (c2 >= rs[r] ? lv2[r] : 0.0) * zc2;
                    }
                }
                c0 = cStop;
                J = pend;
                if (J >= bEnd) break;
            }
            // tile J is not known to be done: load this wave's coupling
            // entries to its columns first, then wait for its values
            const int c = c0 + lane;
            double lv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) lv[r] = ltw[(unsigned int)(rd[r] + c)];
#pragma unroll
            for (int r = 0; r < 16; ++r) lv[r] = c >= rs[r] ? lv[r] : 0.0;
            const int dist = T.blk - J;
            double zc;
            if (!poll_values(z, c0, 64, backoff_for(dist, ASTRO_NEAR_FWD), spinLimit, timeout, zc)) return;
#pragma unroll
            for (int r = 0; r < 16; ++r) p[r] += lv[r] * zc;
            c0 += 64;
            ++J;
            front = dist - 1 <= ASTRO_NEAR_FWD;
        }
    }
    // the values of the tiles four and three back (needed at once) and of
    // the two nearest, loaded speculatively (usually produced by now), then
    // this wave's 16 columns of L_tt^-1
    NearVals n1{0, 0, 0.0}, n2{0, 0, 0.0}, n3{0, 0, 0.0}, n4{0, 0, 0.0};
    if (nNear >= 4) n4 = near_load(z, r0 - 256, 64);
    if (nNear >= 3) n3 = near_load(z, r0 - 192, 64);
    if (nNear >= 1) n1 = near_load(g, r0 - 64 + 16 * w, 16);
    if (nNear >= 2) n2 = near_load(z, r0 - 128 + 16 * w, 16);
    double dv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) dv[k] = Bt[SWEEP_DINV * SWEEP_BLOCK + (16 * w + k) * 64 + lane];
    if (nNear >= 4) {
        if (!near_wait(z, n4, spinLimit, timeout)) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) p[r] += lv4[r] * n4.v;
    }
    if (nNear >= 3) {
        if (!near_wait(z, n3, spinLimit, timeout)) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) p[r] += lv3[r] * n3.v;
    }
    // transpose-reduce the 16 row partials across the wave: each stage
    // exchanges half of the rows a lane carries with its partner lane
    // (l ^ 32, l ^ 16, l ^ 8, l ^ 4) and adds, so lane l ends with row
    // (l >> 2) & 15 summed over all 64 lanes.  The exchanges are VALU
    // cross-lane moves (permlane swaps, DPP), not LDS permutes.
    {
        const bool b3 = lane & 8, b2 = lane & 4;
        double q1[8], q2[4], q3[2];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            double a = p[r], c = p[r + 8];
            swap_halves32(a, c);
            q1[r] = a + c;  // row r + 8 * bit5
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double a = q1[r], c = q1[r + 4];
            swap_rows16(a, c);
            q2[r] = a + c;  // row r + 8 * bit5 + 4 * bit4
        }
        // both shifts are taken in every lane before the select: a DPP move
        // inside a divergent branch would read disabled source lanes as 0
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const double send = b3 ? q2[r] : q2[r + 2];
            const double up = dpp_f64<DPP_ROW_SHL8>(send), down = dpp_f64<DPP_ROW_SHR8>(send);
            q3[r] = (b3 ? q2[r + 2] : q2[r]) + (b3 ? down : up);
        }
        const double send = b2 ? q3[0] : q3[1];
        const double up = dpp_f64<DPP_ROW_SHL4>(send), down = dpp_f64<DPP_ROW_SHR4>(send);
        double q4 = (b2 ? q3[1] : q3[0]) + (b2 ? down : up);
        q4 += dpp_f64<DPP_QUAD_XOR2>(q4);
        q4 += dpp_f64<DPP_QUAD_XOR1>(q4);
        if ((lane & 3) == 0) part[w][(lane >> 2) & 15] = q4;
    }
    __syncthreads();
    const double h = lane < nr ? bl - part[lane >> 4][lane & 15] : 0.0;
    lookahead_tail(r0, nr, nNear, T.blk, h, dv, kv, B1s[w], B3s[w], n3.v, 16 * w, g, z, flags, epoch, spinLimit,
                   timeout, n1, n2, acc, acc2, vsh[w]);
}
template __global__ void k_sweep_forward<double>(const SweepTile*, int, const double*, const int64_t*, const int*,
                                                 const double*, const double*, double*, double*, int*, int*, int,
                                                 long, int*, const int*);
template __global__ void k_sweep_forward<float>(const SweepTile*, int, const float*, const int64_t*, const int*,
                                                const double*, const double*, double*, double*, int*, int*, int,
                                                long, int*, const int*);

// Backward sweep tile (columns [r0, r1)): phase 1 takes the rows of the far
// tiles (five or more tiles on) and of the tiles three and four on (coupling
// entries prefetched), the tail combines the near tiles.
template <typename LT>
__global__ __launch_bounds__(256) void k_sweep_backward(const SweepTile* __restrict__ tiles, int numTiles,
                                                        const LT* __restrict__ lt,
                                                        const int64_t* __restrict__ rowoff,
                                                        const int* __restrict__ sfirst,
                                                        const double* __restrict__ blocks,
                                                        const double* __restrict__ z, double* __restrict__ x,
                                                        double* __restrict__ g, int* __restrict__ flags,
                                                        int* __restrict__ ticket, int epoch, long spinLimit,
                                                        int* __restrict__ timeout, const double* __restrict__ diag, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ int sT;
    __shared__ double part[4][64];
    __shared__ double acc[4][64];
    __shared__ double acc2[4][64];
    __shared__ double B1s[4][16][64];  // wave w's 16 columns of G_t
    __shared__ double B3s[4][16][64];  // and of G_t HB_{t+1}
    __shared__ double vsh[4][16];      // per-wave broadcast slot
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) sT = timed_out(timeout) ? numTiles : atomicAdd(ticket, 1);
    __syncthreads();
    const int t = sT;
    if (t >= numTiles) return;
    const SweepTile T = tiles[t];
    const int r0 = T.r0, r1 = T.r1, nr = T.r1 - T.r0, nNear = T.nNear, lim = T.lim;
    const double* Bt = blocks + (size_t)T.blk * SWEEP_KINDS * SWEEP_BLOCK;
    double dv[16], kv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        dv[k] = Bt[SWEEP_DINVT * SWEEP_BLOCK + (16 * w + k) * 64 + lane];
        kv[k] = nNear >= 2 ? Bt[SWEEP_KB * SWEEP_BLOCK + (16 * w + k) * 64 + lane] : 0.0;
    }
    if (nNear >= 1) {
#pragma unroll
        for (int k = 0; k < 16; ++k) B1s[w][k][lane] = Bt[SWEEP_G * SWEEP_BLOCK + (16 * w + k) * 64 + lane];
    }
    if (nNear >= 3) {
#pragma unroll
        for (int k = 0; k < 16; ++k) B3s[w][k][lane] = Bt[SWEEP_GHB * SWEEP_BLOCK + (16 * w + k) * 64 + lane];
    }
    // D^-1 z for this tile's rows (z is complete: the forward sweep has ended)
    double zd = 0.0;
    if (lane < nr) {
        const int gi = r0 + lane;
        // the pivot: from the profile's diagonal in FP64 storage (which may
        // hold a factor handed in from the host); an FP32 profile keeps its
        // pivots in FP64 in diag
        if constexpr (sizeof(LT) == sizeof(double))
            zd = z[gi] / lt[rowoff[gi] + gi];
        else
            zd = z[gi] / diag[gi];
    }
    const int ci = r0 + lane;
    // an in-range column for idle lanes' loads; their partial sums are unused
    const int cl = lane < nr ? ci : r0;
    // coupling entries lc[r] = L(jb + r, ci) of 16 rows (lanes 0..cnt-1
    // carry their first columns and offsets; rows beyond cnt have offset 0,
    // so their loads read lt[cl], inside the profile, and are masked)
    auto rows_meta = [&](int jb, int cnt, int& sjl, long long& rol) {
        sjl = 0x7fffffff;
        rol = 0;
        if (lane < cnt) {
            sjl = sfirst[jb + lane];
            rol = rowoff[jb + lane];
        }
    };
    auto row_entries = [&](int sjl, long long rol, double (&lc)[16]) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int sj = __builtin_amdgcn_readlane(sjl, r);
            const long long ro = ((long long)__builtin_amdgcn_readlane((int)(rol >> 32), r) << 32) |
                                 (unsigned int)__builtin_amdgcn_readlane((int)(rol & 0xffffffffLL), r);
            const double v = lt[ro + cl];
            lc[r] = ci >= sj ? v : 0.0;
        }
    };
    // this wave's 16 rows of the tiles three and four on
    const int jb3 = r1 + 128 + 16 * w, jb4 = r1 + 192 + 16 * w;
    const int cnt3 = nNear >= 3 ? max(0, min(16, lim - jb3)) : 0;
    const int cnt4 = nNear >= 4 ? max(0, min(16, lim - jb4)) : 0;
    double lc3[16], lc4[16];
    {
        int s3, s4;
        long long o3, o4;
        rows_meta(jb3, cnt3, s3, o3);
        rows_meta(jb4, cnt4, s4, o4);
        row_entries(s3, o3, lc3);
        row_entries(s4, o4, lc4);
    }
    double p = 0.0;
    // Far tiles: blk blk+5 .. farBlk, rows [r1 + 256, end of farBlk).
    // Published tiles form a top-down prefix of that range (tiles finish in
    // descending order); each wave streams the published rows and re-polls,
    // without block barriers.
    if (T.farBlk >= 0) {
        const int bLow = T.blk + 4;
        const int jLow = r1 + 256;
        int J = T.farBlk;
        int j1 = min(r0 + 64 * (J - T.blk) + 64, lim);  // rows [jStop, j1) still to do, top-down
        bool front = false;
        while (J > bLow) {
            if (!front) {
                // bulk: every row of the published prefix of far tiles; wave
                // w takes 16-row groups jb = jStop + 16 w + 64 q, lanes 0..15
                // fetch the group's first columns, offsets and x, broadcast
                const int pend = first_pending_down(flags, J, bLow, epoch);
                const int jStop = pend <= bLow ? jLow : min(r0 + 64 * (pend - T.blk) + 64, lim);
                // software-pipelined: the next group's first columns, offsets
                // and x are fetched while the current group's entries load
                auto group_meta = [&](int jb, int& sjl, long long& rol, double& xl) {
                    sjl = 0x7fffffff;
                    rol = 0;
                    xl = 0.0;
                    if (lane < 16 && jb + lane < j1) {
                        sjl = sfirst[jb + lane];
                        rol = rowoff[jb + lane];
                        xl = coherent_load(&x[jb + lane]);
                    }
                };
                // two groups (jb, jb + 64) per step, so 32 row loads are in flight
                int sjl0, sjl1, nsjl0, nsjl1;
                long long rol0, rol1, nrol0, nrol1;
                double xl0, xl1, nxl0, nxl1;
                int jb = jStop + 16 * w;
                group_meta(jb, sjl0, rol0, xl0);
                group_meta(jb + 64, sjl1, rol1, xl1);
#pragma unroll 1
                for (; jb < j1; jb += 128) {
                    group_meta(jb + 128, nsjl0, nrol0, nxl0);
                    group_meta(jb + 192, nsjl1, nrol1, nxl1);
                    double v0[16], v1[16];
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const long long ro0 = ((long long)__builtin_amdgcn_readlane((int)(rol0 >> 32), r) << 32) |
                                              (unsigned int)__builtin_amdgcn_readlane((int)(rol0 & 0xffffffffLL), r);
                        const long long ro1 = ((long long)__builtin_amdgcn_readlane((int)(rol1 >> 32), r) << 32) |
                                              (unsigned int)__builtin_amdgcn_readlane((int)(rol1 & 0xffffffffLL), r);
                        v0[r] = lt[ro0 + cl];
                        v1[r] = lt[ro1 + cl];
                    }
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int sj0 = __builtin_amdgcn_readlane(sjl0, r), sj1 = __builtin_amdgcn_readlane(sjl1, r);
                        p += (ci >= sj0 ? v0[r] : 0.0) * lane_bcast(xl0, r);
                        p += (ci >= sj1 ? v1[r] : 0.0) * lane_bcast(xl1, r);
                    }
                    sjl0 = nsjl0;
                    rol0 = nrol0;
                    xl0 = nxl0;
                    sjl1 = nsjl1;
                    rol1 = nrol1;
                    xl1 = nxl1;
                }
                j1 = jStop;
                J = pend;
                if (J <= bLow) break;
            }
            // tile J is not known to be done: this wave's 16 of its rows,
            // coupling entries loaded first, then wait for their values
            const int q0 = r0 + 64 * (J - T.blk);
            const int jb = q0 + 16 * w;
            const int cnt = max(0, min(16, min(q0 + 64, lim) - jb));
            int sjl;
            long long rol;
            rows_meta(jb, cnt, sjl, rol);
            double lc[16];
            row_entries(sjl, rol, lc);
            const int dist = J - T.blk;
            double xl;
            if (!poll_values(x, jb, cnt, backoff_for(dist, ASTRO_NEAR_BWD), spinLimit, timeout, xl)) return;
#pragma unroll
            for (int r = 0; r < 16; ++r) p += lc[r] * lane_bcast(xl, r);
            j1 = q0;
            --J;
            front = dist - 1 <= ASTRO_NEAR_BWD;
        }
    }
    NearVals n1{0, 0, 0.0}, n2{0, 0, 0.0}, n3{0, 0, 0.0}, n4{0, 0, 0.0};
    if (nNear >= 4) n4 = near_load(x, jb4, cnt4);
    if (nNear >= 3) n3 = near_load(x, jb3, cnt3);
    if (nNear >= 1) n1 = near_load(g, r1 + 16 * w, max(0, min(16, lim - (r1 + 16 * w))));
    if (nNear >= 2) n2 = near_load(x, r1 + 64 + 16 * w, max(0, min(16, lim - (r1 + 64 + 16 * w))));
    if (nNear >= 4) {
        if (!near_wait(x, n4, spinLimit, timeout)) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) p += lc4[r] * lane_bcast(n4.v, r);
    }
    if (nNear >= 3) {
        if (!near_wait(x, n3, spinLimit, timeout)) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) p += lc3[r] * lane_bcast(n3.v, r);
    }
    part[w][lane] = p;
    __syncthreads();
    const double h = lane < nr ? zd - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) : 0.0;
    lookahead_tail(r0, nr, nNear, T.blk, h, dv, kv, B1s[w], B3s[w], n3.v, 0, g, x, flags, epoch, spinLimit, timeout,
                   n1, n2, acc, acc2, vsh[w]);
}
template __global__ void k_sweep_backward<double>(const SweepTile*, int, const double*, const int64_t*, const int*,
                                                  const double*, const double*, double*, double*, int*, int*, int,
                                                  long, int*, const double*, const int*);
template __global__ void k_sweep_backward<float>(const SweepTile*, int, const float*, const int64_t*, const int*,
                                                 const double*, const double*, double*, double*, int*, int*, int,
                                                 long, int*, const double*, const int*);

__global__ void k_reset_ticket(int* ticket, const int* __restrict__ gate) {
    if (gate && *gate == 0) return; *ticket = 0; }

// ---------------------------------------------------------------------------
// Operator y = A x + C^T (C x) with the reference's CSR (1-based columns,
// 0 = unused slot).  Working/AStrO_solvers.f90:143-171.  32 lanes per row.
__global__ __launch_bounds__(256) void k_spmv(const double* __restrict__ val, const int* __restrict__ cols,
                                              const int* __restrict__ range, int n, const double* __restrict__ x,
                                              double* __restrict__ y, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const int l = threadIdx.x & 31;
    if (row >= n) return;
    // a brick row holds at most 81 slots: each lane's first three slots are
    // loaded together (columns, values, then x), any further ones in a loop.
    // The window starts at the 16-slot boundary below the row (slots of the
    // row before are masked), so every 32-lane load covers whole 128-byte
    // lines of the values: 96 -> 91 us per call on the 64x64x32 brick.
    const int s0 = range[row], e = range[row + 1];
    const int q0 = (s0 & ~15) + l, q1 = q0 + 32, q2 = q0 + 64;
    const bool i0 = q0 >= s0 && q0 < e, i1 = q1 >= s0 && q1 < e, i2 = q2 >= s0 && q2 < e;
    const int c0 = i0 ? cols[q0] : 0, c1 = i1 ? cols[q1] : 0, c2 = i2 ? cols[q2] : 0;
    const double v0 = i0 ? val[q0] : 0.0, v1 = i1 ? val[q1] : 0.0, v2 = i2 ? val[q2] : 0.0;
    const double x0 = c0 ? x[c0 - 1] : 0.0, x1 = c1 ? x[c1 - 1] : 0.0, x2 = c2 ? x[c2 - 1] : 0.0;
    double s = (c0 ? v0 * x0 : 0.0) + (c1 ? v1 * x1 : 0.0) + (c2 ? v2 * x2 : 0.0);
    for (int q = q0 + 96; q < e; q += 32) {
        const int c = cols[q];
        if (c) s += val[q] * x[c - 1];
    }
    for (int off = 16; off > 0; off >>= 1) s += __shfl_xor(s, off, 32);
    if (l == 0) y[row] = s;
}

// The same operator's A x in the node-block form of a brick8 model (three
// displacement components per node, DOF d of node slot b = b + d N, 0-based):
// one lane per node slot a forms its three rows a, a + N, a + 2 N.  Neighbour
// j of slot a is nb[j N + a] (-1: none); the entry (row a + c N, column
// b + d N) of neighbour b = nb[j N + a] sits at bv[(9 j + 3 c + d) N + a].
// Index and value loads are coalesced across lanes (slot-major), the column
// indices are shared by the node's three rows and nine entries (4 bytes per
// nine values instead of per value), and each neighbour's three x values
// serve all three rows.  Per row the products are summed in neighbour order.
__global__ __launch_bounds__(256) void k_spmv_nodes(const double* __restrict__ bv, const int* __restrict__ nb, int K,
                                                    int N, const double* __restrict__ x, double* __restrict__ y,
                                                    const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < K; ++j) {
        const int b = nb[(size_t)j * N + a];
        const double* v = bv + (size_t)9 * j * N + a;
        double w[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = v[(size_t)k * N];
        if (b >= 0) {
            const double x0 = x[b], x1 = x[b + N], x2 = x[b + 2 * N];
            s0 += w[0] * x0 + w[1] * x1 + w[2] * x2;
            s1 += w[3] * x0 + w[4] * x1 + w[5] * x2;
            s2 += w[6] * x0 + w[7] * x1 + w[8] * x2;
        }
    }
    y[a] = s0;
    y[a + N] = s1;
    y[a + 2 * N] = s2;
}

// bv[q] = val[map[q]] (0 where map[q] < 0): the node-block copy of the CSR
// values, refreshed after every assembly.
__global__ void k_gather_values(const double* __restrict__ val, const int* __restrict__ map, int64_t count,
                                double* __restrict__ bv) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
        const int m = map[q];
        bv[q] = m >= 0 ? val[m] : 0.0;
    }
}

// The node-block copy by tiles: workgroup (t, c) takes node slots
// a0 = 64 t .. a0 + 63 and component c, i.e. CSR rows a0 + c N .. a0 + 63 + c N,
// which are consecutive, so their slots form one contiguous range read once
// and coalesced.  Each slot's block position 3 j + d (jd[q], from the CSR
// column) and its row (a search over the tile's 65 row starts) place the
// value into an LDS tile T[3 j + d][a - a0]; the tile's rows are then
// written to bv[(9 j + 3 c + d) N + a] for 64 consecutive a.  3 K <= 96.
__global__ __launch_bounds__(256) void k_node_values(const double* __restrict__ val, const int* __restrict__ range,
                                                     const uint8_t* __restrict__ jd, int K, int N,
                                                     double* __restrict__ bv) {
    __shared__ double T[96][65];
    __shared__ int rs[65];
    const int a0 = 64 * blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
    const int na = min(64, N - a0);
    const int r0 = a0 + c * N;
    if (tid <= na) rs[tid] = range[r0 + tid];
    for (int id = tid; id < 3 * K * 64; id += 256) T[id >> 6][id & 63] = 0.0;
    __syncthreads();
    const int q0 = rs[0], q1 = rs[na];
    for (int q = q0 + tid; q < q1; q += 256) {
        const int k = jd[q];
        if (k == 0xff) continue;  // unused slot
        int lo = 0, hi = na;      // rs[lo] <= q < rs[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rs[mid] <= q) lo = mid;
            else hi = mid;
        }
        T[k][lo] = val[q];
    }
    __syncthreads();
    for (int id = tid; id < 3 * K * 64; id += 256) {
        const int k = id >> 6, a = id & 63;
        if (a < na) {
            const int j = k / 3, d = k - 3 * j;
            bv[(size_t)(9 * j + 3 * c + d) * N + a0 + a] = T[k][a];
        }
    }
}

// t[r] = (C x)_r as a double-double pair (t[2r], t[2r + 1]).
__global__ void k_mpc_dd(const double* __restrict__ cmat, const int* __restrict__ ccols,
                         const int* __restrict__ crange, int cdim, const double* __restrict__ x,
                         double* __restrict__ t, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cdim) return;
    double hi = 0.0, lo = 0.0;
    for (int s = crange[r]; s < crange[r + 1]; ++s)
        if (ccols[s]) dd_add_prod(hi, lo, cmat[s], x[ccols[s] - 1]);
    t[2 * r] = hi;
    t[2 * r + 1] = lo;
}

// True residual r = b - (A x + C^T (C x)) with every product and sum carried
// in double-double and rounded once at the end, so r is accurate even when it
// is 1e-16 of |A| |x| (the restarts of GMRES then refine the solution like
// iterative refinement with extended-precision residuals).  32 lanes per row;
// the penalty term reads C transposed: for row i, entries ctPos[k] of cmat
// in constraint rows ctRow[k], k in [ctRange[i], ctRange[i + 1]), with
// t = C x from k_mpc_dd.
__global__ __launch_bounds__(256) void k_residual_dd(const double* __restrict__ val, const int* __restrict__ cols,
                                                     const int* __restrict__ range, int n,
                                                     const int* __restrict__ ctRange, const int* __restrict__ ctPos,
                                                     const int* __restrict__ ctRow, const double* __restrict__ cmat,
                                                     const double* __restrict__ t, const double* __restrict__ b,
                                                     const double* __restrict__ x, double* __restrict__ r, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const int l = threadIdx.x & 31;
    if (row >= n) return;
    double hi = 0.0, lo = 0.0;
    for (int q = range[row] + l; q < range[row + 1]; q += 32) {
        const int c = cols[q];
        if (c) dd_add_prod(hi, lo, val[q], x[c - 1]);
    }
    if (ctRange) {
        for (int k = ctRange[row] + l; k < ctRange[row + 1]; k += 32) {
            const double a = cmat[ctPos[k]];
            const int cr = ctRow[k];
            dd_add_prod(hi, lo, a, t[2 * cr]);
            dd_add_prod(hi, lo, a, t[2 * cr + 1]);
        }
    }
    for (int off = 16; off > 0; off >>= 1) {
        const double oh = __shfl_xor(hi, off, 32), ol = __shfl_xor(lo, off, 32);
        dd_add(hi, lo, oh, ol);
    }
    if (l == 0) {
#pragma clang fp contract(off)
        double s, e;
        dd_two_sum(b[row], -hi, s, e);
        r[row] = s + (e - lo);
    }
}

__global__ void k_mpc_apply(const double* __restrict__ cmat, const int* __restrict__ ccols,
                            const int* __restrict__ crange, int cdim, const double* __restrict__ x,
                            double* __restrict__ y, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cdim) return;
    double t = 0.0;
    for (int s = crange[r]; s < crange[r + 1]; ++s)
        if (ccols[s]) t += cmat[s] * x[ccols[s] - 1];
    for (int s = crange[r]; s < crange[r + 1]; ++s)
        if (ccols[s]) atomicAdd(&y[ccols[s] - 1], cmat[s] * t);
}

// RHS pieces: nodal loads and the constraint load C^T (rhs - C u) (u null:
// C^T rhs).  Working/AStrO_commandFunctions.f90:368-392, 506-535.
__global__ void k_nodal_loads(const int* __restrict__ dofs, const double* __restrict__ vals, int count,
                              double scale, double* __restrict__ load, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    atomicAdd(&load[dofs[k] - 1], scale * vals[k]);
}

__global__ void k_constraint_load(const double* __restrict__ cmat, const int* __restrict__ ccols,
                                  const int* __restrict__ crange, const double* __restrict__ crhs, int cdim,
                                  const double* __restrict__ u, double* __restrict__ load, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cdim) return;
    double t = crhs[r];
    if (u)
        for (int s = crange[r]; s < crange[r + 1]; ++s)
            if (ccols[s]) t -= cmat[s] * u[ccols[s] - 1];
    for (int s = crange[r]; s < crange[r + 1]; ++s)
        if (ccols[s]) atomicAdd(&load[ccols[s] - 1], cmat[s] * t);
}

// ---------------------------------------------------------------------------
// Krylov vector kernels.  Multi-dot: partial[blk][q] = sum over the block's
// rows of V[q][i] * w[i] for q in [q0, q0 + nq); the host-side launcher sums
// the partials in a second fixed-order pass (deterministic).
__global__ __launch_bounds__(256) void k_multidot(const double* __restrict__ V, int64_t ldv, int nqTotal,
                                                  const double* __restrict__ w, int n, double* __restrict__ partial,
                                                  int pstride, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int tid = threadIdx.x;
    const int q0 = blockIdx.y * 8, nq = min(8, nqTotal - q0);
    double acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    const int rowsPerBlock = (n + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * rowsPerBlock, i1 = min(n, i0 + rowsPerBlock);
    for (int i = i0 + tid; i < i1; i += 256) {
        const double wi = w[i];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < nq) acc[k] += V[(size_t)(q0 + k) * ldv + i] * wi;
    }
    __shared__ double red[4][8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((tid & 63) == 0) red[tid >> 6][k] = v;
    }
    __syncthreads();
    if (tid < nq) partial[(size_t)blockIdx.x * pstride + q0 + tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// out[q] = sum over blocks of partial[b][q]: one wave per q (grid nq x 64
// threads), lanes take blocks b = lane, lane + 64, ..., then a fixed
// butterfly, so the order of the additions never changes.
__global__ __launch_bounds__(64) void k_sum_partials(const double* __restrict__ partial, int nblk, int pstride,
                                                     int nq, double* __restrict__ out, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int q = blockIdx.x, lane = threadIdx.x;
    if (q >= nq) return;
    double s = 0.0;
    for (int b = lane; b < nblk; b += 64) s += partial[(size_t)b * pstride + q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) out[q] = s;
}

// ---------------------------------------------------------------------------
// Device-side GMRES cycle for captured (hipGraph) time steps.  The cycle's
// scalars live on the device: gd = [res, target, res0, hnorm, 1/hnorm,
// 1/res, judge target, -, R (m+1 x m, column-major), cs (m), sn (m),
// g (m+1), y (m)], gi = [active, k, run, stepIdx, failedStep].  run = 0 after a step
// whose solve missed its target (every later gated kernel then skips until
// the host has finished that step); active = 1 while the cycle's Arnoldi
// steps are still needed.  The arithmetic is the host cycle's
// (Gmres::solve): CGS2 coefficients, Givens rotations on the new column,
// g[j + 1] = -sn g[j], the stop |g[j + 1]| <= target, back substitution.
__device__ __forceinline__ size_t gmd_cs(int m) { return GMD_R + (size_t)(m + 1) * m; }

// Cycle start from nrm2 = |P^-1 r|^2: target = max(rtol res, atol) (the
// solve's first cycle), g = res e1, active = run && res > target.
__global__ void k_gm_begin(const double* __restrict__ nrm2, double rtol, double atol, double* __restrict__ gd,
                           int* __restrict__ gi, int m) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!gi[GMI_RUN]) {
        gi[GMI_ACTIVE] = 0;
        return;
    }
    const double res = sqrt(*nrm2);
    gd[GMD_RES] = res;
    gd[GMD_RES0] = res;
    gd[GMD_TARGET] = fmax(rtol * res, atol);
    gd[GMD_INVRES] = res > 0.0 ? 1.0 / res : 0.0;
    double* g = gd + gmd_cs(m) + 2 * (size_t)m;
    for (int i = 0; i <= m; ++i) g[i] = 0.0;
    g[0] = res;
    gi[GMI_K] = 0;
    gi[GMI_ACTIVE] = res > gd[GMD_TARGET] ? 1 : 0;
}

// Arnoldi step j: the new column h1 + h2 (CGS2 passes) and |w| = sqrt(hn),
// rotated by the earlier rotations, a new rotation, g, and the stop test.
__global__ void k_gm_givens(int j, const double* __restrict__ h1, const double* __restrict__ h2,
                            const double* __restrict__ hn, double* __restrict__ gd, int* __restrict__ gi, int m) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!gi[GMI_ACTIVE]) return;
    double* R = gd + GMD_R + (size_t)(m + 1) * j;
    double* cs = gd + gmd_cs(m);
    double* sn = cs + m;
    double* g = sn + m;
    const double hnorm = sqrt(*hn);
    for (int i = 0; i <= j; ++i) R[i] = h1[i] + h2[i];
    for (int i = 0; i < j; ++i) {
        const double a = cs[i] * R[i] + sn[i] * R[i + 1];
        const double b = -sn[i] * R[i] + cs[i] * R[i + 1];
        R[i] = a;
        R[i + 1] = b;
    }
    const double rr = hypot(R[j], hnorm);
    cs[j] = rr == 0.0 ? 1.0 : R[j] / rr;
    sn[j] = rr == 0.0 ? 0.0 : hnorm / rr;
    R[j] = rr;
    R[j + 1] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    gi[GMI_K] = j + 1;
    gd[GMD_HNORM] = hnorm;
    gd[GMD_INVH] = hnorm > 0.0 ? 1.0 / hnorm : 0.0;
    if (hnorm == 0.0 || fabs(g[j + 1]) <= gd[GMD_TARGET]) gi[GMI_ACTIVE] = 0;
}

// Judge target rtol sqrt(nrm2) (the plain-preconditioned starting residual).
__global__ void k_gm_judge_target(const double* __restrict__ nrm2, double rtol, double* __restrict__ gd,
                                  const int* __restrict__ gi) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (gi[GMI_RUN]) gd[GMD_JUDGE] = rtol * sqrt(*nrm2);
}

// y = R^-1 g over the cycle's k columns (run only).
__global__ void k_gm_lsq(double* __restrict__ gd, const int* __restrict__ gi, int m) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!gi[GMI_RUN]) return;
    const int k = gi[GMI_K];
    const double* cs = gd + gmd_cs(m);
    const double* g = cs + 2 * (size_t)m;
    double* y = (double*)g + (m + 1);
    for (int i = k - 1; i >= 0; --i) {
        double v = g[i];
        for (int q = i + 1; q < k; ++q) v -= gd[GMD_R + (size_t)(m + 1) * q + i] * y[q];
        y[i] = v / gd[GMD_R + (size_t)(m + 1) * i + i];
    }
}

// w += sign sum_{q < *nq} V[q] h[q] (nq on the device).
__global__ void k_sub_combination_dn(const double* __restrict__ V, int64_t ldv, const int* __restrict__ nq,
                                     const double* __restrict__ h, double* __restrict__ w, int n, double sign,
                                     const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int k = *nq;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < k; ++q) s += V[(size_t)q * ldv + i] * h[q];
        w[i] += sign * s;
    }
}

// The step's judge: the true preconditioned residual after the cycle,
// sqrt(nrm2), against gd[slot] (the cycle's target or the judge target); a
// miss stops the pipeline.
__global__ void k_gm_check(const double* __restrict__ nrm2, double* __restrict__ gd, int* __restrict__ gi,
                           int slot) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!gi[GMI_RUN]) return;
    const double res = sqrt(*nrm2);
    gd[GMD_RES] = res;
    if (!(res <= gd[slot])) {
        gi[GMI_RUN] = 0;
        gi[GMI_FAILED] = gi[GMI_STEP];
    }
}

__global__ void k_gm_advance(int* __restrict__ gi) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (gi[GMI_RUN]) gi[GMI_STEP] += 1;
}

// y = x * (*s), gated.
__global__ void k_scale_by(const double* __restrict__ x, double* __restrict__ y, int n, const double* __restrict__ sp,
                           const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const double a = *sp;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a * x[i];
}

__global__ void k_fill_gated(double* __restrict__ p, int64_t n, double v, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Nodal loads times scales[*stepIdx] (the captured step reads its load
// factor from the device table).
__global__ void k_nodal_loads_dev(const int* __restrict__ dofs, const double* __restrict__ vals, int count,
                                  const double* __restrict__ scales, const int* __restrict__ stepIdx,
                                  double* __restrict__ load, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    atomicAdd(&load[dofs[k] - 1], scales[*stepIdx] * vals[k]);
}

// w -= sum_q V[q] h[q]   (h on the device)
__global__ void k_sub_combination(const double* __restrict__ V, int64_t ldv, int nq, const double* __restrict__ h,
                                  double* __restrict__ w, int n, double sign, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ double sh[256];
    for (int q = threadIdx.x; q < nq; q += blockDim.x) sh[q] = h[q];
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < nq; ++q) s += V[(size_t)q * ldv + i] * sh[q];
        w[i] += sign * s;
    }
}

// y = alpha * x  (alpha read from device memory as 1 / *den when den != null)
__global__ void k_scale_copy(const double* __restrict__ x, double* __restrict__ y, int n,
                             const double* __restrict__ den) {
    const double a = 1.0 / *den;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a * x[i];
}

__global__ void k_axpby(const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out,
                        int n, double a, double b, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = a * x[i] + b * y[i];
}

// CSR max |value| entry check helper: count entries with column 0 (unused slots)
__global__ void k_count_zero_cols(const int* __restrict__ cols, int64_t nnz, int* __restrict__ out) {
    int c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        c += cols[i] == 0;
    if (c) atomicAdd(out, c);
}

}  // namespace astro
