// astro_oracle.cpp -- CPU FP64 restatement of the AStrO finite-element hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
// compared against; it is never linked into, loaded by or called from the
// product path (misc_amd/).  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may use it.
//
// Every routine restates one reference routine (Fortran, andersone35/Misc,
// directory Working/) and cites it as file:line.  Index conventions follow the
// reference exactly: node ids, DOF numbers and CSR/profile column numbers are
// 1-based values, range arrays are (0:n) prefix offsets stored as n+1 ints, and
// a CSR column value of 0 marks an unused slot.  Arrays passed in from the
// caller are plain C pointers, column-major where the reference is.
//
// Pinning of each part (what checks it besides being a restatement) is listed
// in tests/test_oracle_*.py; the literal reference compiled from source
// (oracle/_ref, built by oracle/Makefile) supplies golden vectors for the
// defect-free linear-static brick path.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <complex>

namespace {

// ---------------------------------------------------------------------------
// Literal constants.  Working/AStrO_constantVals.f90:6-17 and the single
// precision literals of Working/AStrO_solvers.f90:184 (1e-12) and :336
// (1.570796326794897, read as a default-kind REAL).
const double kOneRt3 = 1.0 / std::sqrt(3.0);
const double kGmresTol = (double)1e-12f;
const double kHalfPiSingle = (double)1.570796326794897f;

// Element profile (integration points, weights, DOF table) for the element
// types this oracle evaluates.  Working/AStrO_r_elementEqns.f90:334-449.
struct ElProfile {
    int numNds = 0, dofPerNd = 0, numIntDof = 0, numIntPts = 0;
    int dofTable[2][33];  // [0]=variable (1-based), [1]=local node (1-based)
    double intPts[8][3];
    double ipWt[8];
};

void element_profile(int eType, ElProfile& p) {
    p = ElProfile{};
    if (eType == 8) {
        p.numNds = 8; p.dofPerNd = 3; p.numIntDof = 0; p.numIntPts = 8;
        // GP order: s1 fastest, s3 slowest; :363-376
        int k = 0;
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    p.intPts[k][0] = kOneRt3 * (a ? 1.0 : -1.0);
                    p.intPts[k][1] = kOneRt3 * (b ? 1.0 : -1.0);
                    p.intPts[k][2] = kOneRt3 * (c ? 1.0 : -1.0);
                    p.ipWt[k] = 1.0;
                    ++k;
                }
    } else if (eType == 41) {
        // :399-416: 4 Gauss points in the element plane, 8 internal DOFs on
        // the bubble functions (u1, u2 on "nodes" 5-6; u3 on 7-10)
        p.numNds = 4; p.dofPerNd = 6; p.numIntDof = 8; p.numIntPts = 4;
        const double pts[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
        for (int k = 0; k < 4; ++k) {
            p.intPts[k][0] = kOneRt3 * pts[k][0];
            p.intPts[k][1] = kOneRt3 * pts[k][1];
            p.intPts[k][2] = 0.0;
            p.ipWt[k] = 1.0;
        }
        const int tab[8][2] = {{1, 5}, {1, 6}, {2, 5}, {2, 6}, {3, 7}, {3, 8}, {3, 9}, {3, 10}};
        for (int k = 0; k < 8; ++k) {
            p.dofTable[0][24 + k] = tab[k][0];
            p.dofTable[1][24 + k] = tab[k][1];
        }
    } else {
        std::fprintf(stderr, "oracle: element type %d not supported\n", eType);
        std::abort();
    }
    // component-major DOF order: :440-447
    int i3 = 0;
    for (int v = 1; v <= p.dofPerNd; ++v)
        for (int n = 1; n <= p.numNds; ++n) {
            p.dofTable[0][i3] = v;
            p.dofTable[1][i3] = n;
            ++i3;
        }
}

// Trilinear brick shape functions and natural derivatives.
// Working/AStrO_r_elementEqns.f90:27-36 (N) and :110-135 (dN/ds).
void brick_shape(const double s[3], double N[8], double Ns[8][3]) {
    static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    for (int n = 0; n < 8; ++n) {
        double f0 = 1.0 + sg[n][0] * s[0];
        double f1 = 1.0 + sg[n][1] * s[1];
        double f2 = 1.0 + sg[n][2] * s[2];
        N[n] = 0.125 * f0 * f1 * f2;
        Ns[n][0] = 0.125 * sg[n][0] * f1 * f2;
        Ns[n][1] = 0.125 * sg[n][1] * f0 * f2;
        Ns[n][2] = 0.125 * sg[n][2] * f0 * f1;
    }
}

// Determinant and inverse by adjugate.  Working/AStrO_r_elementEqns.f90:212-238.
// A, I are 3x3 column-major (A[r + 3*c]).
double det_inv3(const double A[9], double I[9]) {
    auto a = [&](int r, int c) { return A[(r - 1) + 3 * (c - 1)]; };
    double adj11 = a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2);
    double adj12 = a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3);
    double adj13 = a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1);
    double D = a(1, 1) * adj11 + a(1, 2) * adj12 + a(1, 3) * adj13;
    double dinv = 1.0 / D;
    auto I_ = [&](int r, int c) -> double& { return I[(r - 1) + 3 * (c - 1)]; };
    I_(1, 1) = dinv * adj11;
    I_(2, 1) = dinv * adj12;
    I_(3, 1) = dinv * adj13;
    I_(1, 2) = dinv * (a(1, 3) * a(3, 2) - a(1, 2) * a(3, 3));
    I_(2, 2) = dinv * (a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1));
    I_(3, 2) = dinv * (a(1, 2) * a(3, 1) - a(1, 1) * a(3, 2));
    I_(1, 3) = dinv * (a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2));
    I_(2, 3) = dinv * (a(1, 3) * a(2, 1) - a(1, 1) * a(2, 3));
    I_(3, 3) = dinv * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1));
    return D;
}

// Integration point data for a brick: Jacobian dx/ds = locNds * Ns, its
// determinant and Nx = Ns * ds/dx.  Working/AStrO_r_elementEqns.f90:580-620.
// xyz is 3x8 column-major (node coordinates in the section frame).
double brick_ipdata(const double* xyz, const double s[3], double N[8], double Nx[8][3]) {
    double Ns[8][3];
    brick_shape(s, N, Ns);
    double J[9] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0;
            for (int n = 0; n < 8; ++n) acc += xyz[i + 3 * n] * Ns[n][j];
            J[i + 3 * j] = acc;
        }
    double Ji[9];
    double detJ = det_inv3(J, Ji);
    for (int n = 0; n < 8; ++n)
        for (int j = 0; j < 3; ++j) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += Ns[n][k] * Ji[k + 3 * j];
            Nx[n][j] = acc;
        }
    return detJ;
}

// Doolittle LU without pivoting on rows/cols [st, en] (0-based, inclusive) of
// a column-major n x n matrix.  Working/AStrO_r_designPropertyFunctions.f90:1015-1039.
template <typename T>
void lu_factor(T* A, int n, int st, int en) {
    auto a = [&](int r, int c) -> T& { return A[r + (size_t)n * c]; };
    for (int i = st; i <= en; ++i) {
        for (int j = i; j <= en; ++j)
            for (int k = st; k < i; ++k) a(i, j) -= a(i, k) * a(k, j);
        T f = T(1.0) / a(i, i);
        for (int j = i + 1; j <= en; ++j) {
            for (int k = st; k < i; ++k) a(j, i) -= a(j, k) * a(k, i);
            a(j, i) = f * a(j, i);
        }
    }
}

// Forward/back substitution with the factor of lu_factor.
// Working/AStrO_r_designPropertyFunctions.f90:1042-1073.
template <typename T>
void lu_solve(T* x, const T* LU, const T* b, int n, int st, int en) {
    std::vector<T> z(n, T(0.0));
    auto lu = [&](int r, int c) { return LU[r + (size_t)n * c]; };
    z[st] = b[st];
    for (int i = st + 1; i <= en; ++i) {
        z[i] = b[i];
        for (int k = st; k < i; ++k) z[i] -= lu(i, k) * z[k];
    }
    x[en] = z[en] / lu(en, en);
    for (int i = en - 1; i >= st; --i) {
        T f = T(1.0) / lu(i, i);
        T dp = T(0);
        for (int k = i + 1; k <= en; ++k) dp += lu(i, k) * x[k];
        x[i] = f * (z[i] - dp);
    }
}

// Shell4 shape functions: 4 bilinear and 6 bubbles (incompatible modes) with
// their natural derivatives.  Working/AStrO_r_elementEqns.f90:44-55, 143-162.
void shell4_shape(double s1, double s2, double N[10], double Ns[10][2]) {
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
    Ns[7][0] = (1 - s2 * s2);      Ns[7][1] = -2 * s2 * (1 + s1);
    Ns[8][0] = -2 * s1 * (1 + s2); Ns[8][1] = (1 - s1 * s1);
    Ns[9][0] = -(1 - s2 * s2);     Ns[9][1] = -2 * s2 * (1 - s1);
}

// Rotation matrix of the rotation vector v (Rodrigues).
void rotation_matrix(const double v[3], double R[3][3]) {
    const double th = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double k[3] = {0, 0, 0};
    if (th > 0) for (int i = 0; i < 3; ++i) k[i] = v[i] / th;
    const double c = std::cos(th), sn = std::sin(th), t = 1 - c;
    R[0][0] = c + k[0] * k[0] * t;        R[0][1] = k[0] * k[1] * t - k[2] * sn; R[0][2] = k[0] * k[2] * t + k[1] * sn;
    R[1][0] = k[1] * k[0] * t + k[2] * sn; R[1][1] = c + k[1] * k[1] * t;        R[1][2] = k[1] * k[2] * t - k[0] * sn;
    R[2][0] = k[2] * k[0] * t - k[1] * sn; R[2][1] = k[2] * k[1] * t + k[0] * sn; R[2][2] = c + k[2] * k[2] * t;
}

// Bodies of orc_shell_abd and orc_shell4_stiffness (documented there),
// generic in the scalar type so that the complex-step derivative runs the
// same arithmetic on std::complex<double> (the reference's AStrO_c_* twins).
template <typename T>
void shell_abd_t(int nL, const double* layers, const T* thick, double zOff, T* ABD) {
    for (int k = 0; k < 81; ++k) ABD[k] = T(0.0);
    T t = T(0);
    for (int l = 0; l < nL; ++l) t += thick[l];
    T z = -0.5 * t * (zOff + 1.0);
    auto A = [&](int r, int c) -> T& { return ABD[(r - 1) + 9 * (c - 1)]; };
    for (int l = 0; l < nL; ++l) {
        const double* L = layers + 6 * l;
        const T zn = z + thick[l];
        double S[9] = {1.0 / L[0], -L[2] / L[0], 0, -L[2] / L[0], 1.0 / L[1], 0, 0, 0, 1.0 / L[3]};
        double Q[9];
        lu_factor(S, 3, 0, 2);
        for (int c = 0; c < 3; ++c) {
            double f[3] = {0, 0, 0};
            f[c] = 1.0;
            lu_solve(Q + 3 * c, S, f, 3, 0, 2);
        }
        const double a11 = std::cos(L[5]), a21 = std::sin(L[5]), a12 = -a21, a22 = a11;
        double Ts[3][3] = {{a11 * a11, a12 * a12, 2 * a11 * a12},
                           {a21 * a21, a22 * a22, 2 * a22 * a21},
                           {a11 * a21, a12 * a22, a11 * a22 + a12 * a21}};
        double Te[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double v = Ts[r][c];
                if (c == 2) v *= 0.5;
                if (r == 2) v *= 2.0;
                Te[r + 3 * c] = v;
            }
        double TeInv[9];
        lu_factor(Te, 3, 0, 2);
        for (int c = 0; c < 3; ++c) {
            double f[3] = {0, 0, 0};
            f[c] = 1.0;
            lu_solve(TeInv + 3 * c, Te, f, 3, 0, 2);
        }
        double QT[3][3], Qb[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                QT[r][c] = 0;
                for (int k = 0; k < 3; ++k) QT[r][c] += Q[r + 3 * k] * TeInv[k + 3 * c];
            }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                Qb[r][c] = 0;
                for (int k = 0; k < 3; ++k) Qb[r][c] += Ts[r][k] * QT[k][c];
            }
        for (int r = 1; r <= 3; ++r)
            for (int c = 1; c <= 3; ++c) {
                A(r, c) += (zn - z) * Qb[r - 1][c - 1];
                A(r, c + 3) += 0.5 * (zn * zn - z * z) * Qb[r - 1][c - 1];
                A(r + 3, c + 3) += (1.0 / 3.0) * (zn * zn * zn - z * z * z) * Qb[r - 1][c - 1];
            }
        z = zn;
    }
    for (int r = 2; r <= 6; ++r)
        for (int c = 1; c < r; ++c) A(r, c) = A(c, r);
    A(7, 7) = A(3, 3);
    A(8, 8) = A(3, 3);
    A(9, 9) = A(3, 3);
}

template <typename T>
void shell4_stiffness_t(const double* xyz, const double* alpha, const T* ABD, T* K32, T* Kc) {
    ElProfile p;
    element_profile(41, p);
    double al[3][3], loc[3][4];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) al[i][j] = alpha[i + 3 * j];
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 4; ++a) {
            loc[i][a] = 0;
            for (int k = 0; k < 3; ++k) loc[i][a] += al[i][k] * xyz[k + 3 * a];
        }
    std::vector<T> K(33 * 33, T(0.0));  // r_condenseMat works on the 33x33 array
    auto Kf = [&](int r, int c) -> T& { return K[r + 33 * (size_t)c]; };
    for (int g = 0; g < 4; ++g) {
        double N[10], Ns[10][2];
        shell4_shape(p.intPts[g][0], p.intPts[g][1], N, Ns);
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 2; ++j)
                for (int a = 0; a < 4; ++a) J[i][j] += loc[i][a] * Ns[a][j];
        J[2][2] = 1.0;
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
        double Nx[10][2];
        for (int n = 0; n < 10; ++n)
            for (int j = 0; j < 2; ++j) Nx[n][j] = Ns[n][0] * I[0][j] + Ns[n][1] * I[1][j];
        double B[9][32];
        for (int d = 0; d < 32; ++d) {
            const int var = p.dofTable[0][d], nd = p.dofTable[1][d] - 1;
            double ux[3][2] = {{0, 0}, {0, 0}, {0, 0}}, rx[3][2] = {{0, 0}, {0, 0}, {0, 0}}, rot[3] = {0, 0, 0};
            if (d >= 24) {  // internal DOF: local component var on bubble nd
                for (int j = 0; j < 2; ++j) ux[var - 1][j] = Nx[nd][j];
            } else if (var <= 3) {  // nodal translation: u_L(nd) = alpha e_var
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 2; ++j) ux[i][j] = al[i][var - 1] * Nx[nd][j];
            } else {  // nodal rotation e_q at node nd
                const int q = var - 4;
                for (int a = 0; a < 4; ++a) {
                    // d u_L(a) = -alpha (e_q x X_a) / 4
                    const double* X = xyz + 3 * a;
                    double ex[3] = {0, 0, 0};
                    if (q == 0) { ex[1] = -X[2]; ex[2] = X[1]; }
                    if (q == 1) { ex[0] = X[2]; ex[2] = -X[0]; }
                    if (q == 2) { ex[0] = -X[1]; ex[1] = X[0]; }
                    for (int i = 0; i < 3; ++i) {
                        double v = 0;
                        for (int k = 0; k < 3; ++k) v += al[i][k] * ex[k];
                        v *= -0.25;
                        for (int j = 0; j < 2; ++j) ux[i][j] += v * Nx[a][j];
                    }
                    // d r_L(a) = alpha e_q (delta_a,nd - 1/4)
                    const double f = (a == nd ? 1.0 : 0.0) - 0.25;
                    for (int i = 0; i < 3; ++i) {
                        for (int j = 0; j < 2; ++j) rx[i][j] += al[i][q] * f * Nx[a][j];
                        rot[i] += al[i][q] * f * N[a];
                    }
                }
            }
            B[0][d] = ux[0][0];
            B[1][d] = ux[1][1];
            B[2][d] = ux[0][1] + ux[1][0];
            B[3][d] = rx[1][0];
            B[4][d] = -rx[0][1];
            B[5][d] = rx[1][1] - rx[0][0];
            B[6][d] = ux[2][0] + rot[1];
            B[7][d] = ux[2][1] - rot[0];
            B[8][d] = 2.0 * rot[2] - ux[1][0] + ux[0][1];
        }
        const double wd = det * p.ipWt[g];
        T DB[9][32];
        for (int r = 0; r < 9; ++r)
            for (int d = 0; d < 32; ++d) {
                T acc = T(0);
                for (int k = 0; k < 9; ++k) acc += ABD[r + 9 * k] * B[k][d];
                DB[r][d] = wd * acc;
            }
        for (int j = 0; j < 32; ++j)
            for (int i = 0; i < 32; ++i) {
                T acc = T(0);
                for (int k = 0; k < 9; ++k) acc += B[k][i] * DB[k][j];
                Kf(i, j) += acc;
            }
    }
    if (K32)
        for (int j = 0; j < 32; ++j)
            for (int i = 0; i < 32; ++i) K32[i + 32 * j] = Kf(i, j);
    // condensation: LU of rows/cols 24..31, then each nodal column
    lu_factor(K.data(), 33, 24, 31);
    std::vector<T> b(33), x(33);
    for (int c = 0; c < 24; ++c) {
        for (int r = 0; r < 33; ++r) b[r] = Kf(r, c);
        std::fill(x.begin(), x.end(), T(0.0));
        lu_solve(x.data(), K.data(), b.data(), 33, 24, 31);
        for (int r = 0; r < 24; ++r) {
            T acc = T(0);
            for (int k = 24; k < 32; ++k) acc += Kf(r, k) * x[k];
            Kf(r, c) -= acc;
        }
    }
    for (int j = 0; j < 24; ++j)
        for (int i = 0; i < 24; ++i) Kc[i + 24 * j] = Kf(i, j);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// a6: shell frame.  Rows of alpha are the local axes: x along the section's
// first orientation vector, z = x cross (second vector), y = z cross x, all
// normalised (Working/AStrO_input.f90:1071-1095); then correctAlpha
// (Working/AStrO_r_elementEqns.f90:240-286) turns the frame about
// z_sec cross n by asin(|z_sec cross n|) so that its z axis is the element
// normal n (diagonal cross product, oriented with z_sec).  The reference's
// r_asin / r_rotateAlpha run on r_sin/r_cos, which read an uninitialised angle
// (D1); here the rotation is exact.  xyz: 3x4 column-major; alpha[i + 3 j] =
// component j of axis i.
void orc_shell_frame(const double* or6, const double* xyz, double* alpha) {
    double x[3] = {or6[0], or6[1], or6[2]}, v[3] = {or6[3], or6[4], or6[5]};
    auto norm = [](double* a) {
        double m = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (int i = 0; i < 3; ++i) a[i] /= m;
    };
    auto cross = [](const double* a, const double* b, double* c) {
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
    };
    double z[3], y[3];
    norm(x);
    cross(x, v, z);
    norm(z);
    cross(z, x, y);
    double ax[3][3] = {{x[0], x[1], x[2]}, {y[0], y[1], y[2]}, {z[0], z[1], z[2]}};
    double d1[3], d2[3], n[3];
    for (int i = 0; i < 3; ++i) {
        d1[i] = xyz[i + 3 * 2] - xyz[i + 3 * 0];
        d2[i] = xyz[i + 3 * 3] - xyz[i + 3 * 1];
    }
    cross(d1, d2, n);
    if (n[0] * z[0] + n[1] * z[1] + n[2] * z[2] < 0) for (int i = 0; i < 3; ++i) n[i] = -n[i];
    norm(n);
    double w[3];
    cross(z, n, w);
    const double mag = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (mag > 1e-5) {
        const double th = std::asin(std::min(1.0, mag));
        for (int i = 0; i < 3; ++i) w[i] *= th / mag;
        double R[3][3];
        rotation_matrix(w, R);
        double out[3][3];
        for (int r = 0; r < 3; ++r)
            for (int i = 0; i < 3; ++i) {
                out[r][i] = 0;
                for (int k = 0; k < 3; ++k) out[r][i] += R[i][k] * ax[r][k];
            }
        std::memcpy(ax, out, sizeof ax);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) alpha[i + 3 * j] = ax[i][j];
}

// a5: shell ABD by layer integration, Working/AStrO_r_designPropertyFunctions.f90:
// 259-363.  layers: nL x 6 (E1, E2, nu12, G12, thickness, angle in radians);
// the stack starts at z = -t (zOff + 1) / 2.  Q = S^-1 of the in-plane
// compliance, rotated by the layer angle (Ts Q Te^-1), A += dz Q,
// B += dz2 / 2 Q, D += dz3 / 3 Q; transverse shear and drilling take A33
// (:359-361, no shear correction).  The reference's r_cos/r_sin read an
// uninitialised angle (D1) and carry single-precision coefficients (D2): here
// cos/sin are exact.  ABD: 9x9 column-major.
void orc_shell_abd(int nL, const double* layers, double zOff, double* ABD) {
    std::vector<double> th(nL);
    for (int l = 0; l < nL; ++l) th[l] = layers[6 * l + 4];
    shell_abd_t<double>(nL, layers, th.data(), zOff, ABD);
}

// a10 (shell branch at u = 0) + a9 + a14: shell4 tangent
// sum_gp detJ*w * dedU^T ABD dedU (Working/AStrO_r_elementEqns.f90:2245-2423,
// 2478-2494) with the 9 generalised strains of r_getShellDef (:800-814), then
// static condensation of the 8 internal DOFs (r_condenseMat :2722-2743).
// dedU follows r_getInstDof (:703-891) linearised at rest with exact
// derivatives: local displacement u_L(a) = alpha (u_a - rbar x X_a) and local
// rotation r_L(a) = alpha (r_a - rbar), rbar = mean nodal rotation (the
// reference's derivative branch drops d rbar / d r_b from the rotations and
// reads past instDofMat for the interpolated rotation, D3).  Interpolation:
// u_L over the 4 nodes plus the bubbles of the internal DOFs, r_L over the 4
// nodes (:2253-2262).  xyz: 3x4 global coordinates; alpha as orc_shell_frame;
// ABD 9x9.  K32 (optional): 32x32 before condensation; Kc: 24x24 condensed,
// both column-major in the element DOF order of dofTable.
void orc_shell4_stiffness(const double* xyz, const double* alpha, const double* ABD, double* K32, double* Kc) {
    shell4_stiffness_t<double>(xyz, alpha, ABD, K32, Kc);
}

// ---------------------------------------------------------------------------
// a4: material stiffness C = S^-1 from 9 engineering constants
// (E1,E2,E3,nu12,nu13,nu23,G12,G13,G23) by LU inverse without pivoting.
// Working/AStrO_r_designPropertyFunctions.f90:8-80 (compliance branch :43-77)
// and r_GetInvLU :996-1013.  C is 6x6 column-major, Voigt order
// (11,22,33,12,13,23) with engineering shears.
void orc_material_stiffness(const double* ep, double* C) {
    double S[36] = {0};
    auto s = [&](int r, int c) -> double& { return S[(r - 1) + 6 * (c - 1)]; };
    s(1, 1) = 1.0 / ep[0];
    s(1, 2) = -ep[3] / ep[0];
    s(1, 3) = -ep[4] / ep[0];
    s(2, 2) = 1.0 / ep[1];
    s(2, 3) = -ep[5] / ep[1];
    s(3, 3) = 1.0 / ep[2];
    s(4, 4) = 1.0 / ep[6];
    s(5, 5) = 1.0 / ep[7];
    s(6, 6) = 1.0 / ep[8];
    for (int i = 2; i <= 6; ++i)
        for (int j = 1; j < i; ++j) s(i, j) = s(j, i);
    lu_factor(S, 6, 0, 5);
    double f[6];
    for (int c = 0; c < 6; ++c) {
        for (int k = 0; k < 6; ++k) f[k] = 0.0;
        f[c] = 1.0;
        lu_solve(C + 6 * c, S, f, 6, 0, 5);
    }
}

// ---------------------------------------------------------------------------
// a10 (solid branch at u = 0): brick8 tangent dR/dU = sum_gp detJ*w * B^T C B.
// Working/AStrO_r_elementEqns.f90:2424-2475 (B columns from r_getNLStrain
// :649-662 with orient = identity) and :2512-2529 (contraction).  The
// geometric-stress term (:2466-2473) is zero at u = 0, T = 0 and is omitted.
// xyz: 3x8 column-major section-frame coordinates.  Ke: 24x24 column-major,
// DOF order component-major [u1..u8, v1..v8, w1..w8].
void orc_brick8_stiffness(const double* xyz, const double* C, double* Ke) {
    ElProfile p;
    element_profile(8, p);
    std::memset(Ke, 0, sizeof(double) * 576);
    for (int g = 0; g < 8; ++g) {
        double N[8], Nx[8][3];
        double detJ = brick_ipdata(xyz, p.intPts[g], N, Nx);
        double B[6][24];
        for (int d = 0; d < 24; ++d) {
            int var = p.dofTable[0][d], nd = p.dofTable[1][d] - 1;
            double o[3] = {var == 1 ? 1.0 : 0.0, var == 2 ? 1.0 : 0.0, var == 3 ? 1.0 : 0.0};
            B[0][d] = o[0] * Nx[nd][0];
            B[1][d] = o[1] * Nx[nd][1];
            B[2][d] = o[2] * Nx[nd][2];
            B[3][d] = o[0] * Nx[nd][1] + o[1] * Nx[nd][0];
            B[4][d] = o[0] * Nx[nd][2] + o[2] * Nx[nd][0];
            B[5][d] = o[1] * Nx[nd][2] + o[2] * Nx[nd][1];
        }
        double wd = detJ * p.ipWt[g];
        double CB[6][24];
        for (int r = 0; r < 6; ++r)
            for (int d = 0; d < 24; ++d) {
                double acc = 0;
                for (int k = 0; k < 6; ++k) acc += C[r + 6 * k] * B[k][d];
                CB[r][d] = wd * acc;
            }
        for (int j = 0; j < 24; ++j)
            for (int i = 0; i < 24; ++i) {
                double acc = 0;
                for (int k = 0; k < 6; ++k) acc += B[k][i] * CB[k][j];
                Ke[i + 24 * j] += acc;
            }
    }
}

// a11 (solid branch): consistent mass rho * sum_gp detJ*w * N^T N per
// component.  Working/AStrO_r_elementEqns.f90:2628-2653.  Me: 24x24.
void orc_brick8_mass(const double* xyz, double rho, double* Me) {
    ElProfile p;
    element_profile(8, p);
    std::memset(Me, 0, sizeof(double) * 576);
    for (int g = 0; g < 8; ++g) {
        double N[8], Nx[8][3];
        double detJ = brick_ipdata(xyz, p.intPts[g], N, Nx);
        double f = rho * detJ * p.ipWt[g];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 8; ++a)
                for (int b = 0; b < 8; ++b) Me[(c * 8 + a) + 24 * (c * 8 + b)] += f * N[a] * N[b];
    }
}

// a10 + a8 (solid branch at any u): internal force with the Green-Lagrange
// strain, R_i = sum_gp detJ*w * S . dE/du_i with S = C E(u) (T = 0).
// Working/AStrO_r_elementEqns.f90:2424-2458 (ux :2426-2432, stress :2437-2447,
// Ruk :2448-2458); strain and its first derivative r_getNLStrain :622-662 with
// orient = identity.  ue: 24 DOFs component-major; R: 24.
void orc_brick8_internal_force(const double* xyz, const double* C, const double* ue, double* R) {
    ElProfile p;
    element_profile(8, p);
    for (int d = 0; d < 24; ++d) R[d] = 0.0;
    for (int g = 0; g < 8; ++g) {
        double N[8], Nx[8][3];
        double detJ = brick_ipdata(xyz, p.intPts[g], N, Nx);
        double ux[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int d = 0; d < 24; ++d) {
            int var = p.dofTable[0][d] - 1, nd = p.dofTable[1][d] - 1;
            for (int j = 0; j < 3; ++j) ux[var][j] += ue[d] * Nx[nd][j];
        }
        double E[6];
        E[0] = ux[0][0] + 0.5 * (ux[0][0] * ux[0][0] + ux[1][0] * ux[1][0] + ux[2][0] * ux[2][0]);
        E[1] = ux[1][1] + 0.5 * (ux[0][1] * ux[0][1] + ux[1][1] * ux[1][1] + ux[2][1] * ux[2][1]);
        E[2] = ux[2][2] + 0.5 * (ux[0][2] * ux[0][2] + ux[1][2] * ux[1][2] + ux[2][2] * ux[2][2]);
        E[3] = ux[0][1] + ux[1][0] + ux[0][0] * ux[0][1] + ux[1][0] * ux[1][1] + ux[2][0] * ux[2][1];
        E[4] = ux[0][2] + ux[2][0] + ux[0][0] * ux[0][2] + ux[1][0] * ux[1][2] + ux[2][0] * ux[2][2];
        E[5] = ux[1][2] + ux[2][1] + ux[0][2] * ux[0][1] + ux[1][2] * ux[1][1] + ux[2][2] * ux[2][1];
        double S[6];
        for (int r = 0; r < 6; ++r) {
            S[r] = 0.0;
            for (int k = 0; k < 6; ++k) S[r] += C[r + 6 * k] * E[k];
        }
        double wd = detJ * p.ipWt[g];
        for (int d = 0; d < 24; ++d) {
            int v = p.dofTable[0][d] - 1, nd = p.dofTable[1][d] - 1;
            double o[3] = {v == 0 ? 1.0 : 0.0, v == 1 ? 1.0 : 0.0, v == 2 ? 1.0 : 0.0};
            const double* x = Nx[nd];
            double dE[6];
            dE[0] = o[0] * x[0] + x[0] * ux[v][0];
            dE[1] = o[1] * x[1] + x[1] * ux[v][1];
            dE[2] = o[2] * x[2] + x[2] * ux[v][2];
            dE[3] = o[0] * x[1] + o[1] * x[0] + x[0] * ux[v][1] + x[1] * ux[v][0];
            dE[4] = o[0] * x[2] + o[2] * x[0] + x[0] * ux[v][2] + x[2] * ux[v][0];
            dE[5] = o[1] * x[2] + o[2] * x[1] + x[1] * ux[v][2] + x[2] * ux[v][1];
            double acc = 0.0;
            for (int k = 0; k < 6; ++k) acc += S[k] * dE[k];
            R[d] += wd * acc;
        }
    }
}

// f4: brick8 stress and strain at the natural point s (r_getSolidStress,
// Working/AStrO_r_elementEqns.f90:924-976): ux of the element DOFs ue
// (component-major) at s, the Green-Lagrange strain of r_getNLStrain
// (:622-648, orient = identity) and stress = C (strain - T alpha) with T alpha
// = 0 (the models carry no expansion coefficients).  C: 6x6 column-major.
void orc_brick8_stress(const double* xyz, const double* C, const double* ue, const double* s, double* stress,
                       double* strain) {
    ElProfile p;
    element_profile(8, p);
    double N[8], Nx[8][3];
    brick_ipdata(xyz, s, N, Nx);
    double ux[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int d = 0; d < 24; ++d) {
        int var = p.dofTable[0][d] - 1, nd = p.dofTable[1][d] - 1;
        for (int j = 0; j < 3; ++j) ux[var][j] += ue[d] * Nx[nd][j];
    }
    strain[0] = ux[0][0] + 0.5 * (ux[0][0] * ux[0][0] + ux[1][0] * ux[1][0] + ux[2][0] * ux[2][0]);
    strain[1] = ux[1][1] + 0.5 * (ux[0][1] * ux[0][1] + ux[1][1] * ux[1][1] + ux[2][1] * ux[2][1]);
    strain[2] = ux[2][2] + 0.5 * (ux[0][2] * ux[0][2] + ux[1][2] * ux[1][2] + ux[2][2] * ux[2][2]);
    strain[3] = ux[0][1] + ux[1][0] + ux[0][0] * ux[0][1] + ux[1][0] * ux[1][1] + ux[2][0] * ux[2][1];
    strain[4] = ux[0][2] + ux[2][0] + ux[0][0] * ux[0][2] + ux[1][0] * ux[1][2] + ux[2][0] * ux[2][2];
    strain[5] = ux[1][2] + ux[2][1] + ux[0][2] * ux[0][1] + ux[1][2] * ux[1][1] + ux[2][2] * ux[2][1];
    for (int r = 0; r < 6; ++r) {
        stress[r] = 0.0;
        for (int k = 0; k < 6; ++k) stress[r] += C[r + 6 * k] * strain[k];
    }
}

// a13: brick8 conductivity sum_gp detJ*w * Nx k Nx^T.
// Working/AStrO_r_elementEqns.f90:2062-2124.  k: 3x3 column-major; Kt: 8x8.
void orc_brick8_conductivity(const double* xyz, const double* k, double* Kt) {
    ElProfile p;
    element_profile(8, p);
    std::memset(Kt, 0, sizeof(double) * 64);
    for (int g = 0; g < 8; ++g) {
        double N[8], Nx[8][3];
        double detJ = brick_ipdata(xyz, p.intPts[g], N, Nx);
        double wd = detJ * p.ipWt[g];
        double kNx[8][3];
        for (int a = 0; a < 8; ++a)
            for (int j = 0; j < 3; ++j) {
                double acc = 0;
                for (int m = 0; m < 3; ++m) acc += Nx[a][m] * k[m + 3 * j];
                kNx[a][j] = wd * acc;
            }
        for (int b = 0; b < 8; ++b)
            for (int a = 0; a < 8; ++a) {
                double acc = 0;
                for (int m = 0; m < 3; ++m) acc += Nx[a][m] * kNx[b][m];
                Kt[a + 8 * b] += acc;
            }
    }
}

// f3: brick8 capacitance sum_gp detJ*w * rhoc N N^T, the dRdTdot of r_getElRtm
// (Working/AStrO_r_elementEqns.f90:2126-2165).  Mt: 8x8.
void orc_brick8_capacitance(const double* xyz, double rhoc, double* Mt) {
    ElProfile p;
    element_profile(8, p);
    std::memset(Mt, 0, sizeof(double) * 64);
    for (int g = 0; g < 8; ++g) {
        double N[8], Nx[8][3];
        double wd = brick_ipdata(xyz, p.intPts[g], N, Nx) * p.ipWt[g] * rhoc;
        for (int b = 0; b < 8; ++b)
            for (int a = 0; a < 8; ++a) Mt[a + 8 * b] += N[a] * wd * N[b];
    }
}

// Volume sum_gp detJ*w of one brick (used by the closed-form checks).
double orc_brick8_volume(const double* xyz) {
    ElProfile p;
    element_profile(8, p);
    double v = 0;
    for (int g = 0; g < 8; ++g) {
        double N[8], Nx[8][3];
        v += brick_ipdata(xyz, p.intPts[g], N, Nx) * p.ipWt[g];
    }
    return v;
}

// ---------------------------------------------------------------------------
// a20: profile from CSR + penalty.  Working/AStrO_bookKeeping.f90:754-792.
// lTri is zeroed first (:764).
void orc_convert_ltri(double* lTri, const int* lTRange, long long lTSize, const double* mainMat,
                      const int* mainCols, const int* mainRange, int mainDim, const double* cMat,
                      const int* cCols, const int* cRange, int cDim) {
    std::memset(lTri, 0, sizeof(double) * (size_t)lTSize);
    for (int i1 = 1; i1 <= mainDim; ++i1) {
        long long minCol = (long long)i1 - lTRange[i1] + lTRange[i1 - 1] + 1;
        for (int i2 = mainRange[i1 - 1] + 1; i2 <= mainRange[i1]; ++i2) {
            int i3 = mainCols[i2 - 1];
            if (i3 >= minCol && i3 <= i1) {
                long long i4 = (long long)lTRange[i1] - i1 + i3;
                lTri[i4 - 1] += mainMat[i2 - 1];
            }
        }
    }
    for (int i1 = 1; i1 <= cDim; ++i1)
        for (int i2 = cRange[i1 - 1] + 1; i2 <= cRange[i1]; ++i2) {
            int i4 = cCols[i2 - 1];
            if (i4 == 0) continue;
            long long minCol = (long long)i4 - lTRange[i4] + lTRange[i4 - 1] + 1;
            for (int i3 = cRange[i1 - 1] + 1; i3 <= cRange[i1]; ++i3) {
                int i5 = cCols[i3 - 1];
                if (i5 >= minCol && i5 <= i4) {
                    long long i6 = (long long)lTRange[i4] - i4 + i5;
                    lTri[i6 - 1] += cMat[i2 - 1] * cMat[i3 - 1];
                }
            }
        }
}

// a21: in-place row-skyline LDL^T.  Working/AStrO_solvers.f90:10-78.
// ARange is the (0:n) prefix array; row i holds columns i-len+1..i with the
// diagonal last.  After the call the strict lower part holds L (unit
// diagonal implicit) and the diagonal holds D.  Returns 0, or the 1-based row
// of the first zero pivot (the reference divides by it regardless).
int orc_ldl_factor(double* A, long long ASize, const int* ARange, int n) {
    (void)ASize;
    std::vector<long long> rowSt(n + 1);
    int maxBW = 0;
    for (int i = 1; i <= n; ++i) {
        int len = ARange[i] - ARange[i - 1];
        rowSt[i] = (long long)i - len + 1;
        if (len > maxBW) maxBW = len;
    }
    std::vector<double> diag(n + 1, 0.0), dL(maxBW + 1, 0.0);
    int bad = 0;
    for (int i1 = 1; i1 <= n; ++i1) {
        long long i3 = ARange[i1];          // diagonal (1-based)
        long long i2 = ARange[i1 - 1] + 1;  // first entry of row
        // dL(maxBW - k) = D(col) * L(i1, col), right-aligned: :42-48
        int i5 = maxBW;
        long long col = i1 - 1;
        for (long long i4 = i3 - 1; i4 >= i2; --i4) {
            dL[i5] = diag[col] * A[i4 - 1];
            --i5;
            --col;
        }
        // pivot: :50-57
        long long c = i1 - i3 + i2;
        double d = A[i3 - 1];
        for (long long i4 = i2; i4 <= i3 - 1; ++i4) {
            d -= diag[c] * A[i4 - 1] * A[i4 - 1];
            ++c;
        }
        diag[i1] = d;
        A[i3 - 1] = d;
        if (d == 0.0 && !bad) bad = i1;
        double fact = 1.0 / d;
        // column i1 of L below the pivot: :58-73
        int imax = std::min(n, i1 + maxBW);
        for (int r = i1 + 1; r <= imax; ++r) {
            long long stCol = std::max(rowSt[i1], rowSt[r]);
            long long rowLen = i1 - stCol;
            if (rowLen >= 0) {
                long long j3 = maxBW - rowLen + 1;
                long long j4 = ARange[r] - (r - stCol);
                long long j5 = j4 + rowLen;
                double v = A[j5 - 1];
                for (long long k = 0; k <= rowLen - 1; ++k) v -= dL[j3 + k] * A[j4 + k - 1];
                A[j5 - 1] = v * fact;
            }
        }
    }
    return bad;
}

// a22: z = L^-1 b (row dots), then x = D^-1 z - L^T x by probing later rows.
// Working/AStrO_solvers.f90:83-126.
void orc_ldl_solve(double* soln, const double* rhs, const double* L, const int* LRange, int n,
                   double* z) {
    int maxBW = 0;
    for (int i1 = 1; i1 <= n; ++i1) {
        long long i2 = LRange[i1 - 1] + 1;
        long long rowLen = LRange[i1] - i2;
        if (rowLen > maxBW) maxBW = (int)rowLen;
        long long stCol = i1 - rowLen;
        double v = rhs[i1 - 1];
        for (long long k = 0; k <= rowLen - 1; ++k) v -= L[i2 + k - 1] * z[stCol + k - 1];
        z[i1 - 1] = v;
    }
    for (int i1 = n; i1 >= 1; --i1) {
        double dt = 0;
        double d = L[LRange[i1] - 1];
        int imax = std::min(n, i1 + maxBW);
        for (int i2 = i1 + 1; i2 <= imax; ++i2) {
            long long rowLen = LRange[i2] - LRange[i2 - 1];
            long long stCol = i2 - rowLen + 1;
            if (stCol <= i1) {
                long long i3 = LRange[i2] + i1 - i2;
                dt += L[i3 - 1] * soln[i2 - 1];
            }
        }
        dt *= d;
        soln[i1 - 1] = (z[i1 - 1] - dt) / d;
    }
}

// Givens QR of the (m+1) x m Hessenberg matrix with angles stored in the
// sub-diagonal (triDiag = 2 branch).  Working/AStrO_solvers.f90:299-353.
static void r_factor_hessenberg(double* H, int rowDim, int colDim) {
    auto h = [&](int r, int c) -> double& { return H[(r - 1) + (size_t)rowDim * (c - 1)]; };
    for (int i1 = 1; i1 <= colDim; ++i1) {
        int maxRow = std::min(i1 + 1, rowDim);
        for (int i2 = i1 + 1; i2 <= maxRow; ++i2) {
            if (std::fabs(h(i2, i1)) > 1e-10) {
                double theta = (h(i1, i1) == 0.0) ? kHalfPiSingle : std::atan(h(i2, i1) / h(i1, i1));
                double st = std::sin(theta), ct = std::cos(theta);
                for (int i3 = i1; i3 <= colDim; ++i3) {
                    double a1 = ct * h(i1, i3) + st * h(i2, i3);
                    double a2 = -st * h(i1, i3) + ct * h(i2, i3);
                    h(i1, i3) = a1;
                    h(i2, i3) = a2;
                }
                h(i2, i1) = theta;
            }
        }
    }
}

// Apply the stored rotations to b, then back-substitute R.
// Working/AStrO_solvers.f90:355-404 (triDiag = 2).
static void r_solve_hessenberg(double* soln, const double* H, double* b, int rowDim, int colDim) {
    auto h = [&](int r, int c) { return H[(r - 1) + (size_t)rowDim * (c - 1)]; };
    for (int i1 = 1; i1 <= colDim; ++i1) {
        int maxRow = std::min(i1 + 1, rowDim);
        for (int i2 = i1 + 1; i2 <= maxRow; ++i2) {
            if (std::fabs(h(i2, i1)) > 1e-10) {
                double st = std::sin(h(i2, i1)), ct = std::cos(h(i2, i1));
                double r1 = ct * b[i1 - 1] + st * b[i2 - 1];
                double r2 = -st * b[i1 - 1] + ct * b[i2 - 1];
                b[i1 - 1] = r1;
                b[i2 - 1] = r2;
            }
        }
    }
    for (int i1 = colDim; i1 >= 1; --i1) {
        double v = b[i1 - 1];
        for (int i2 = i1 + 1; i2 <= colDim; ++i2) v -= h(i1, i2) * soln[i2 - 1];
        soln[i1 - 1] = v / h(i1, i1);
    }
}

// y = A x + C^T (C x) with the reference's 0-means-empty CSR convention.
// Working/AStrO_solvers.f90:143-171 (same loops at :189-215, :253-281).
static void apply_operator(double* y, const double* x, const double* Amat, const int* Acols,
                           const int* ARange, int n, const double* Cmat, const int* Ccols,
                           const int* CRange, int cDim, std::vector<double>& t1) {
    t1.assign(std::max(cDim, 1), 0.0);
    for (int i1 = 1; i1 <= cDim; ++i1)
        for (int i2 = CRange[i1 - 1] + 1; i2 <= CRange[i1]; ++i2) {
            int c = Ccols[i2 - 1];
            if (c != 0) t1[i1 - 1] += Cmat[i2 - 1] * x[c - 1];
        }
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int i1 = 1; i1 <= cDim; ++i1)
        for (int i2 = CRange[i1 - 1] + 1; i2 <= CRange[i1]; ++i2) {
            int c = Ccols[i2 - 1];
            if (c != 0) y[c - 1] += Cmat[i2 - 1] * t1[i1 - 1];
        }
    for (int i1 = 1; i1 <= n; ++i1) {
        double acc = y[i1 - 1];
        for (int i2 = ARange[i1 - 1] + 1; i2 <= ARange[i1]; ++i2) {
            int c = Acols[i2 - 1];
            if (c != 0) acc += Amat[i2 - 1] * x[c - 1];
        }
        y[i1 - 1] = acc;
    }
}

// a23: left-preconditioned restarted GMRES(m) with modified Gram-Schmidt.
// Working/AStrO_solvers.f90:128-297.  xVec is the initial guess on entry and
// the solution on exit.  Returns the number of restart cycles run; *finalRes
// receives the last preconditioned residual norm |P^-1 (b - A x)|, *stepsOut
// the number of Arnoldi steps.
//
// mode 0 (literal): every cycle runs all m Arnoldi steps, the least-squares
//   right-hand side is V^T r (:238-243), the stop test is the absolute
//   |P^-1 r| > 1e-12 (single-precision literal, :184), at most maxIt/m + 1
//   cycles (:181).
// mode 1 (intended): the same operator, preconditioner, MGS and Givens
//   rotations, but the least-squares right-hand side is beta*e1 and a cycle ends
//   as soon as the rotated residual estimate drops below rtol * |P^-1 r0|; the
//   solve stops when the true preconditioned residual does.  Once a literal
//   cycle converges before step m its Arnoldi basis loses orthogonality and the
//   V^T r projection turns the update into rounding noise; mode 1 is the
//   numerically stable form of the same method.
int orc_gmres_mode(double* xVec, const double* Amat, const int* Acols, const int* ARange, int n,
                   const double* Cmat, const int* Ccols, const int* CRange, int cDim, const double* Pmat,
                   const int* PRange, const double* bVec, int numVecs, int maxIt, int mode, double rtol,
                   double* finalRes, int* stepsOut) {
    int m = numVecs;
    std::vector<double> z(n), r(n), t1, t2(n), tmp(n);
    std::vector<double> V((size_t)n * (m + 1));
    std::vector<double> H((size_t)(m + 1) * m);
    auto hh = [&](int rr, int cc) -> double& { return H[(rr - 1) + (size_t)(m + 1) * (cc - 1)]; };
    apply_operator(t2.data(), xVec, Amat, Acols, ARange, n, Cmat, Ccols, CRange, cDim, t1);
    for (int i = 0; i < n; ++i) tmp[i] = bVec[i] - t2[i];
    orc_ldl_solve(r.data(), tmp.data(), Pmat, PRange, n, z.data());
    double res = 0;
    for (int i = 0; i < n; ++i) res += r[i] * r[i];
    res = std::sqrt(res);
    const double res0 = res;
    const double stopAt = mode == 0 ? kGmresTol : rtol * res0;
    int maxOutIt = maxIt / numVecs + 1;
    int cycles = 0, steps = 0;
    while (res > stopAt && cycles < maxOutIt) {
        for (int i = 0; i < n; ++i) V[i] = (1.0 / res) * r[i];
        std::fill(H.begin(), H.end(), 0.0);
        // running Givens state for the mode-1 residual estimate
        std::vector<double> cs(m + 1), sn(m + 1), gg(m + 1, 0.0);
        gg[0] = res;
        int k = 0;
        for (int i5 = 2; i5 <= m + 1; ++i5) {
            double* vprev = &V[(size_t)n * (i5 - 2)];
            double* vnew = &V[(size_t)n * (i5 - 1)];
            apply_operator(t2.data(), vprev, Amat, Acols, ARange, n, Cmat, Ccols, CRange, cDim, t1);
            orc_ldl_solve(vnew, t2.data(), Pmat, PRange, n, z.data());
            for (int i1 = 1; i1 <= i5 - 1; ++i1) {
                const double* vi = &V[(size_t)n * (i1 - 1)];
                double dp = 0;
                for (int q = 0; q < n; ++q) dp += vi[q] * vnew[q];
                for (int q = 0; q < n; ++q) vnew[q] -= dp * vi[q];
                hh(i1, i5 - 1) = dp;
            }
            double dp = 0;
            for (int q = 0; q < n; ++q) dp += vnew[q] * vnew[q];
            dp = std::sqrt(dp);
            for (int q = 0; q < n; ++q) vnew[q] = (1.0 / dp) * vnew[q];
            hh(i5, i5 - 1) = dp;
            ++steps;
            k = i5 - 1;
            if (mode == 1) {
                // estimate only (H itself is rotated below exactly as the reference does)
                int j = k;
                double hcol_j = hh(j, j), hsub = hh(j + 1, j);
                std::vector<double> col(j + 1);
                for (int q = 1; q <= j + 1; ++q) col[q - 1] = hh(q, j);
                for (int q = 1; q < j; ++q) {
                    double a = cs[q] * col[q - 1] + sn[q] * col[q];
                    double b = -sn[q] * col[q - 1] + cs[q] * col[q];
                    col[q - 1] = a;
                    col[q] = b;
                }
                (void)hcol_j;
                (void)hsub;
                double rr = std::hypot(col[j - 1], col[j]);
                cs[j] = rr == 0.0 ? 1.0 : col[j - 1] / rr;
                sn[j] = rr == 0.0 ? 0.0 : col[j] / rr;
                gg[j] = -sn[j] * gg[j - 1];
                gg[j - 1] = cs[j] * gg[j - 1];
                if (std::fabs(gg[j]) <= stopAt) break;
            }
        }
        int mk = (mode == 1) ? k : m;  // columns actually built
        // QR of the leading (mk+1) x mk block, angles stored in the sub-diagonal
        std::vector<double> Hk((size_t)(mk + 1) * mk);
        for (int c = 1; c <= mk; ++c)
            for (int rr = 1; rr <= mk + 1; ++rr) Hk[(rr - 1) + (size_t)(mk + 1) * (c - 1)] = hh(rr, c);
        r_factor_hessenberg(Hk.data(), mk + 1, mk);
        std::vector<double> g(mk + 1, 0.0), y(mk, 0.0);
        if (mode == 0) {
            for (int i2 = 1; i2 <= mk + 1; ++i2) {
                const double* vi = &V[(size_t)n * (i2 - 1)];
                double acc = 0;
                for (int i1 = 0; i1 < n; ++i1) acc += vi[i1] * r[i1];
                g[i2 - 1] = acc;
            }
        } else {
            g[0] = res;
        }
        r_solve_hessenberg(y.data(), Hk.data(), g.data(), mk + 1, mk);
        for (int i1 = 0; i1 < n; ++i1) {
            double acc = xVec[i1];
            for (int i2 = 0; i2 < mk; ++i2) acc += V[i1 + (size_t)n * i2] * y[i2];
            xVec[i1] = acc;
        }
        apply_operator(t2.data(), xVec, Amat, Acols, ARange, n, Cmat, Ccols, CRange, cDim, t1);
        for (int i = 0; i < n; ++i) tmp[i] = bVec[i] - t2[i];
        orc_ldl_solve(r.data(), tmp.data(), Pmat, PRange, n, z.data());
        res = 0;
        for (int i = 0; i < n; ++i) res += r[i] * r[i];
        res = std::sqrt(res);
        ++cycles;
    }
    if (finalRes) *finalRes = res;
    if (stepsOut) *stepsOut = steps;
    return cycles;
}

int orc_gmres(double* xVec, const double* Amat, const int* Acols, const int* ARange, int n,
              const double* Cmat, const int* Ccols, const int* CRange, int cDim, const double* Pmat,
              const int* PRange, const double* bVec, int numVecs, int maxIt, double* finalRes,
              int* stepsOut) {
    return orc_gmres_mode(xVec, Amat, Acols, ARange, n, Cmat, Ccols, CRange, cDim, Pmat, PRange, bVec,
                          numVecs, maxIt, 0, 0.0, finalRes, stepsOut);
}

// Plain CSR operator y = A x + C^T C x (exported for the SpMV parity tests).
void orc_apply_operator(double* y, const double* x, const double* Amat, const int* Acols,
                        const int* ARange, int n, const double* Cmat, const int* Ccols,
                        const int* CRange, int cDim) {
    std::vector<double> t1;
    apply_operator(y, x, Amat, Acols, ARange, n, Cmat, Ccols, CRange, cDim, t1);
}

}  // extern "C"

// ===========================================================================
// Model-level bookkeeping and assembly (a16, a17, a19, a24).
// ===========================================================================

struct OrcModel {
    int numNodes = 0, numEls = 0;
    std::vector<double> nodeList;   // 3*N
    std::vector<int> elementList;   // 8*E (1-based node ids, 0 = unused)
    std::vector<int> elementType;   // E
    std::vector<double> elC;        // 36*E
    std::vector<double> elDen;      // E
    std::vector<double> elCond;     // 9*E
    std::vector<double> elCap;      // E: density * specific heat
    std::vector<double> elABD;      // 81*E (shell elements)
    std::vector<double> elOrient;   // 6*E section orientation vectors (shell elements)
    std::vector<double> elLayers;   // nLayers*6*E (shell elements)
    std::vector<double> elZOff;     // E
    int nLayers = 0;
    // design variables: layer dvLayer's thickness of element e += elCoef[e] * d[elDV[e]]
    int nDV = 0, dvLayer = 0;
    std::vector<int> elDV;
    std::vector<double> elCoef;
    // constraints (Working/AStrO_input.f90 constraints block, after set expansion
    // of the node names): term t has equation mpcEqn, dof, coef and either a
    // plain node (isSet = 0, one node) or a node set (isSet = 1, node list)
    std::vector<int> mpcEqn, mpcDof, mpcIsSet, mpcNodeStart, mpcNodes;
    std::vector<double> mpcCoef, mpcRHS;
    int numMPC = 0;
    // nodal loads after set expansion: node id and 7 components
    std::vector<int> loadNode;
    std::vector<double> loadVal;
    // bookkeeping
    int solverMaxBW = 0, numSolidNds = 0, numSectionNds = 0;
    int coupledProfile = 0;  // 1: profile rows reach every component of the neighbours
    std::vector<int> nCRange, nodalConn, currentRank, originalRank, nDofIndex;
    int elMatDim = 0;
    std::vector<int> elMatRange, elMatCols, elMatLTRange;
    std::vector<double> elasticMat, elMatLT, elasticLoad, delDisp, nodeDisp;
    std::vector<int> thermMatRange, thermMatCols, thermMatLTRange;
    std::vector<double> thermMat, thermMatLT, thermalLoad, delTemp;
    int elMPCDim = 0;
    std::vector<int> elMPCMatRange, elMPCMatCols;
    std::vector<double> elMPCMat, elMPCRHS;
    int thermMPCDim = 0;
    std::vector<int> thermMPCMatRange, thermMPCMatCols;
    std::vector<double> thermMPCMat, thermMPCRHS;
    int gmresCycles = 0, gmresSteps = 0;
    double gmresRes = 0;
    // Newmark state (solve's dynamic branch)
    double nmBeta = 0.25, nmGamma = 0.5, delT = 0.0;
    // transient thermal: capacitance on the thermal CSR's columns, the state
    std::vector<double> capMat, nodeTemp, prevTemp, prevTdot, nodeTdot;
    std::vector<int> capCols;
    double thGamma = 0.5, thDt = 0.0;
    std::vector<double> prevDisp, prevVel, prevAcc, nodeVel, nodeAcc;
    std::vector<double> timings;  // seconds per phase of the last static solve
};

namespace {

int ndof_of_type(int eT) { return (eT == 41 || eT == 3 || eT == 2) ? 6 : 3; }

// levelSortNodes: breadth-first ordering from the first node of minimum
// connectivity, neighbours taken in nodalConn insertion order; when the queue
// empties, restart at the nearest unvisited node of the same DOF type.
// Working/AStrO_bookKeeping.f90:233-331.  (The block-split branch at :272-296
// needs nextAvail - blockSt == solverBlockDim == numNodes and cannot fire.)
void level_sort(OrcModel& m, int& nextAvail, int stopLength, int ndof, std::vector<int>& inserted) {
    const int N = m.numNodes;
    auto ndofOf = [&](int nd) { return m.nDofIndex[6 * (nd - 1)]; };
    auto nearest = [&](int from) {
        double best = 1e100;
        int st = 0;
        for (int i4 = 1; i4 <= N; ++i4) {
            if (inserted[i4] == 0 && ndofOf(i4) == ndof) {
                double v0 = m.nodeList[3 * (i4 - 1)] - m.nodeList[3 * (from - 1)];
                double v1 = m.nodeList[3 * (i4 - 1) + 1] - m.nodeList[3 * (from - 1) + 1];
                double v2 = m.nodeList[3 * (i4 - 1) + 2] - m.nodeList[3 * (from - 1) + 2];
                double d = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                if (d < best) {
                    best = d;
                    st = i4;
                }
            }
        }
        return st;
    };
    int minCon = 1000, stNd = 0;
    for (int i1 = 1; i1 <= N; ++i1)
        if (ndofOf(i1) == ndof) {
            int c = m.nCRange[i1] - m.nCRange[i1 - 1];
            if (c < minCon) {
                stNd = i1;
                minCon = c;
            }
        }
    m.originalRank[nextAvail] = stNd;
    int currentNd = stNd;
    int currentInd = nextAvail;
    ++nextAvail;
    inserted[stNd] = 1;
    while (nextAvail <= stopLength) {
        for (int i1 = m.nCRange[currentNd - 1] + 1; i1 <= m.nCRange[currentNd]; ++i1) {
            int i2 = m.nodalConn[i1 - 1];
            if (i2 != 0 && inserted[i2] == 0) {
                m.originalRank[nextAvail] = i2;
                ++nextAvail;
                inserted[i2] = 1;
            }
        }
        ++currentInd;
        if (currentInd == nextAvail && nextAvail <= stopLength) {
            int from = m.originalRank[nextAvail - 1];
            int st = nearest(from);
            m.originalRank[nextAvail] = st;
            currentInd = nextAvail;
            inserted[st] = 1;
            ++nextAvail;
        }
        currentNd = m.originalRank[currentInd];
    }
}

// Linear-probe insert of (col, val) into CSR row `row` (1-based), adding to an
// existing slot with the same column.  `strictLast` reproduces the thermal
// scatter's `.lt.` bound that never fills a row's last slot
// (Working/AStrO_commandFunctions.f90:57); the elastic scatter uses `.le.`
// (:334).
inline void probe_add(std::vector<double>& val, std::vector<int>& cols, const std::vector<int>& range,
                      int row, int col, double v, bool strictLast) {
    int end = strictLast ? range[row] - 1 : range[row];
    for (int s = range[row - 1] + 1; s <= end; ++s) {
        int c = cols[s - 1];
        if (c == 0) {
            val[s - 1] = v;
            cols[s - 1] = col;
            return;
        }
        if (c == col) {
            val[s - 1] += v;
            return;
        }
    }
}

// buildElasticMPC / buildThermalMPC share this expansion: equation blocks sized
// by the largest node set among their terms, one row per set member.
// Working/AStrO_bookKeeping.f90:474-612 (elastic) and :614-752 (thermal).
void build_mpc(OrcModel& m, bool thermal, int& dim, std::vector<int>& range, std::vector<int>& cols,
               std::vector<double>& mat, std::vector<double>& rhs) {
    int T = (int)m.mpcEqn.size();
    std::vector<int> nEqns(m.numMPC + 1, 0), nTrms(m.numMPC + 1, 0);
    auto use = [&](int t) { return thermal ? m.mpcDof[t] == 7 : m.mpcDof[t] <= 6; };
    for (int t = 0; t < T; ++t) {
        if (!use(t)) continue;
        int e = m.mpcEqn[t];
        nTrms[e] += 1;
        if (!m.mpcIsSet[t]) {
            if (nEqns[e] < 1) nEqns[e] = 1;
        } else {
            int sz = m.mpcNodeStart[t + 1] - m.mpcNodeStart[t];
            if (sz > nEqns[e]) nEqns[e] = sz;
        }
    }
    dim = 0;
    long long size = 0;
    for (int e = 1; e <= m.numMPC; ++e) {
        dim += nEqns[e];
        size += (long long)nEqns[e] * nTrms[e];
    }
    mat.assign(size, 0.0);
    cols.assign(size, 0);
    rhs.assign(dim, 0.0);
    range.assign(dim + 1, 0);
    for (int e = 2; e <= m.numMPC; ++e) nEqns[e] += nEqns[e - 1];
    for (int e = 1; e <= m.numMPC; ++e)
        for (int r = nEqns[e - 1] + 1; r <= nEqns[e]; ++r) range[r] = nTrms[e];
    for (int r = 2; r <= dim; ++r) range[r] += range[r - 1];
    auto dofcol = [&](int node, int dof) {
        return thermal ? m.currentRank[node - 1] : m.nDofIndex[6 * (node - 1) + dof - 1];
    };
    auto insert = [&](int row, int col, double coef) {
        for (int s = range[row - 1] + 1; s <= range[row]; ++s) {
            if (cols[s - 1] == 0) {
                mat[s - 1] = coef;
                cols[s - 1] = col;
                return;
            }
            if (cols[s - 1] == col) return;
        }
    };
    for (int t = 0; t < T; ++t) {
        if (!use(t)) continue;
        int e = m.mpcEqn[t];
        int i3 = nEqns[e - 1] + 1;
        int a = m.mpcNodeStart[t], b = m.mpcNodeStart[t + 1];
        if (!m.mpcIsSet[t] || b - a == 1) {
            int col = dofcol(m.mpcNodes[a], m.mpcDof[t]);
            for (; i3 <= nEqns[e]; ++i3) {
                insert(i3, col, m.mpcCoef[t]);
                rhs[i3 - 1] = m.mpcRHS[e - 1];
            }
        } else {
            for (int k = a; k < b; ++k) {
                int col = dofcol(m.mpcNodes[k], m.mpcDof[t]);
                insert(i3, col, m.mpcCoef[t]);
                rhs[i3 - 1] = m.mpcRHS[e - 1];
                ++i3;
            }
        }
    }
}

}  // namespace

extern "C" {

OrcModel* orc_create(int numNodes, const double* nodeList, int numEls, const int* elementList,
                     const int* elementType) {
    auto* m = new OrcModel;
    m->numNodes = numNodes;
    m->numEls = numEls;
    m->nodeList.assign(nodeList, nodeList + 3 * (size_t)numNodes);
    m->elementList.assign(elementList, elementList + 8 * (size_t)numEls);
    m->elementType.assign(elementType, elementType + numEls);
    m->elC.assign(36 * (size_t)numEls, 0.0);
    m->elDen.assign(numEls, 0.0);
    m->elCond.assign(9 * (size_t)numEls, 0.0);
    m->elCap.assign(numEls, 0.0);
    return m;
}

void orc_free(OrcModel* m) { delete m; }

// Profile option set before orc_analysis_prep: 0 = the reference's rows (same
// component only, Working/AStrO_bookKeeping.f90:380-402), 1 = rows reaching
// every component of the node and its neighbours (the factor is then a direct
// solve instead of a component-block preconditioner).
void orc_set_coupled_profile(OrcModel* m, int on) { m->coupledProfile = on; }

// Shell sections per element: nL layers of (E1, E2, nu12, G12, thickness,
// angle) each, the z offset and the two orientation vectors of the section.
void orc_set_shell_sections(OrcModel* m, int nL, const double* layers, const double* zOff, const double* orient6) {
    m->elABD.assign(81 * (size_t)m->numEls, 0.0);
    m->elOrient.assign(6 * (size_t)m->numEls, 0.0);
    m->nLayers = nL;
    m->elLayers.assign(layers, layers + 6 * (size_t)nL * m->numEls);
    m->elZOff.assign(m->numEls, 0.0);
    if (zOff) m->elZOff.assign(zOff, zOff + m->numEls);
    for (int e = 0; e < m->numEls; ++e) {
        orc_shell_abd(nL, layers + 6 * (size_t)nL * e, zOff ? zOff[e] : 0.0, &m->elABD[81 * (size_t)e]);
        for (int k = 0; k < 6; ++k) m->elOrient[6 * (size_t)e + k] = orient6[6 * (size_t)e + k];
    }
}

// Per-element properties: eprops (9*E engineering constants), den (E),
// cond6 (6*E: k11,k22,k33,k12,k13,k23).  Any pointer may be null.
// Working/AStrO_r_designPropertyFunctions.f90:8-80, 83-127, 129-162.
void orc_set_properties(OrcModel* m, const double* eprops, const double* den, const double* cond6) {
    for (int e = 0; e < m->numEls; ++e) {
        if (eprops) orc_material_stiffness(eprops + 9 * (size_t)e, &m->elC[36 * (size_t)e]);
        if (den) m->elDen[e] = den[e];
        if (cond6) {
            const double* k = cond6 + 6 * (size_t)e;
            double* K = &m->elCond[9 * (size_t)e];
            K[0] = k[0]; K[4] = k[1]; K[8] = k[2];
            K[3] = K[1] = k[3];
            K[6] = K[2] = k[4];
            K[7] = K[5] = k[5];
        }
    }
}

// Specific heat per element (after orc_set_properties): capacitance rho c.
void orc_set_spec_heat(OrcModel* m, const double* sh) {
    for (int e = 0; e < m->numEls; ++e) m->elCap[e] = m->elDen[e] * sh[e];
}

void orc_set_mpc(OrcModel* m, int nTerms, const int* eqn, const int* dof, const double* coef,
                 const int* isSet, const int* nodeStart, const int* nodes, int numMPC,
                 const double* rhs) {
    m->mpcEqn.assign(eqn, eqn + nTerms);
    m->mpcDof.assign(dof, dof + nTerms);
    m->mpcCoef.assign(coef, coef + nTerms);
    m->mpcIsSet.assign(isSet, isSet + nTerms);
    m->mpcNodeStart.assign(nodeStart, nodeStart + nTerms + 1);
    m->mpcNodes.assign(nodes, nodes + nodeStart[nTerms]);
    m->numMPC = numMPC;
    m->mpcRHS.assign(rhs, rhs + numMPC);
}

void orc_set_nodal_loads(OrcModel* m, int nLoads, const int* node, const double* val7) {
    m->loadNode.assign(node, node + nLoads);
    m->loadVal.assign(val7, val7 + 7 * (size_t)nLoads);
}

// solverMaxBW as the *solve command sets it: smallest b with b^3 >= 27*N*N,
// the product formed in 32-bit two's-complement arithmetic as the reference's
// default INTEGER does (it wraps above N = 8918 nodes, and a negative result
// leaves b = 1).  Working/AStrO_runJob.f90:107-114.
int orc_solver_max_bw(int numNodes) {
    uint32_t p = (uint32_t)27u * (uint32_t)numNodes;
    p = p * (uint32_t)numNodes;
    int32_t nnd2 = (int32_t)p;
    int bw = 1;
    long long b3 = 1;
    while (b3 < nnd2) {
        ++bw;
        b3 = (long long)bw * bw * bw;
    }
    return bw;
}

// analysisPrep: adjacency, level ordering, DOF numbering, CSR capacity and
// profile lengths, MPC matrices.  Working/AStrO_bookKeeping.f90:794-1039,
// allocateElasticMat :333-411, allocateThermalMat :413-472.
// maxBW <= 0 selects the reference's own rule (orc_solver_max_bw).
void orc_analysis_prep(OrcModel* mp, int maxBW) {
    OrcModel& m = *mp;
    const int N = m.numNodes, E = m.numEls;
    m.solverMaxBW = maxBW > 0 ? maxBW : orc_solver_max_bw(N);
    m.currentRank.assign(N, 0);
    m.originalRank.assign(N + 2, 0);  // 1-based, one spare slot
    m.nDofIndex.assign(6 * (size_t)N, 0);
    m.nCRange.assign(N + 1, 0);
    for (int e = 0; e < E; ++e) {
        int nd = ndof_of_type(m.elementType[e]);
        for (int a = 0; a < 8; ++a) {
            int i3 = m.elementList[8 * (size_t)e + a];
            if (i3 == 0) continue;
            for (int b = 0; b < 8; ++b) {
                int i5 = m.elementList[8 * (size_t)e + b];
                if (i5 != 0 && i5 != i3) m.nCRange[i3] += 1;
            }
            m.nDofIndex[6 * (size_t)(i3 - 1)] = nd;
        }
    }
    m.numSolidNds = 0;
    for (int i = 1; i <= N; ++i)
        if (m.nDofIndex[6 * (size_t)(i - 1)] == 3) ++m.numSolidNds;
    m.numSectionNds = N - m.numSolidNds;
    for (int i = 2; i <= N; ++i) m.nCRange[i] += m.nCRange[i - 1];
    m.nodalConn.assign(m.nCRange[N], 0);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 8; ++a) {
            int i3 = m.elementList[8 * (size_t)e + a];
            if (i3 == 0) continue;
            for (int b = 0; b < 8; ++b) {
                int i5 = m.elementList[8 * (size_t)e + b];
                if (i5 == 0 || i5 == i3) continue;
                for (int s = m.nCRange[i3 - 1] + 1; s <= m.nCRange[i3]; ++s) {
                    if (m.nodalConn[s - 1] == 0) {
                        m.nodalConn[s - 1] = i5;
                        break;
                    }
                    if (m.nodalConn[s - 1] == i5) break;
                }
            }
        }
    std::vector<int> inserted(N + 1, 0);
    int nextAvail = 1;
    if (m.numSolidNds > 0) level_sort(m, nextAvail, m.numSolidNds, 3, inserted);
    if (nextAvail <= N) level_sort(m, nextAvail, N, 6, inserted);
    for (int i = 1; i <= N; ++i) m.currentRank[m.originalRank[i] - 1] = i;
    for (int i = 1; i <= N; ++i) {
        int* nd = &m.nDofIndex[6 * (size_t)(i - 1)];
        int r = m.currentRank[i - 1];
        if (nd[0] == 3) {
            for (int c = 0; c < 3; ++c) nd[c] = r + c * m.numSolidNds;
        } else if (nd[0] == 6) {
            for (int c = 0; c < 6; ++c) nd[c] = r + c * m.numSectionNds + 2 * m.numSolidNds;
        }
    }
    m.elMatDim = 3 * m.numSolidNds + 6 * m.numSectionNds;
    const int n = m.elMatDim;
    // allocateElasticMat
    m.elMatRange.assign(n + 1, 0);
    m.elMatLTRange.assign(n + 1, 0);
    auto ndi = [&](int c, int node) { return m.nDofIndex[6 * (size_t)(node - 1) + c - 1]; };
    for (int i1 = 1; i1 <= N; ++i1)
        for (int i2 = 1; i2 <= 6; ++i2) {
            int i3 = ndi(i2, i1);
            if (i3 == 0) continue;
            for (int i4 = 1; i4 <= 6; ++i4)
                if (ndi(i4, i1) != 0) m.elMatRange[i3] += 1;
            for (int s = m.nCRange[i1 - 1] + 1; s <= m.nCRange[i1]; ++s) {
                int i5 = m.nodalConn[s - 1];
                if (i5 == 0) continue;
                for (int i6 = 1; i6 <= 6; ++i6)
                    if (ndi(i6, i5) != 0) m.elMatRange[i3] += 1;
            }
        }
    for (int i = 2; i <= n; ++i) m.elMatRange[i] += m.elMatRange[i - 1];
    m.elasticMat.assign(m.elMatRange[n], 0.0);
    m.elMatCols.assign(m.elMatRange[n], 0);
    for (int i1 = 1; i1 <= N; ++i1)
        for (int i2 = 1; i2 <= 6; ++i2) {
            int i3 = ndi(i2, i1);
            if (i3 == 0) continue;
            m.elMatLTRange[i3] = 1;
            if (m.coupledProfile)  // the node's own lower components
                for (int c = 1; c <= 6; ++c) {
                    int i7 = ndi(c, i1);
                    if (i7 != 0 && i7 <= i3) m.elMatLTRange[i3] = std::max(m.elMatLTRange[i3], i3 - i7 + 1);
                }
            for (int s = m.nCRange[i1 - 1] + 1; s <= m.nCRange[i1]; ++s) {
                int i5 = m.nodalConn[s - 1];
                if (i5 == 0) continue;
                int i6 = m.currentRank[i1 - 1] - m.currentRank[i5 - 1];
                if (i6 >= 0 && i6 <= m.solverMaxBW) {
                    for (int c = 1; c <= 6; ++c) {
                        if (!m.coupledProfile && c != i2) continue;
                        int i7 = ndi(c, i5);
                        if (i7 != 0 && i7 <= i3) {
                            int i8 = i3 - i7 + 1;
                            if (i8 > m.elMatLTRange[i3]) m.elMatLTRange[i3] = i8;
                        }
                    }
                }
            }
        }
    for (int i = 2; i <= n; ++i) m.elMatLTRange[i] += m.elMatLTRange[i - 1];
    // allocateThermalMat (indices by node rank)
    m.thermMatRange.assign(N + 1, 0);
    m.thermMatLTRange.assign(N + 1, 0);
    for (int i1 = 1; i1 <= N; ++i1) {
        int i3 = m.currentRank[i1 - 1];
        m.thermMatRange[i3] += 1;
        for (int s = m.nCRange[i1 - 1] + 1; s <= m.nCRange[i1]; ++s)
            if (m.nodalConn[s - 1] != 0) m.thermMatRange[i3] += 1;
    }
    for (int i = 2; i <= N; ++i) m.thermMatRange[i] += m.thermMatRange[i - 1];
    m.thermMat.assign(m.thermMatRange[N], 0.0);
    m.thermMatCols.assign(m.thermMatRange[N], 0);
    for (int i1 = 1; i1 <= N; ++i1) {
        int i3 = m.currentRank[i1 - 1];
        m.thermMatLTRange[i3] = 1;
        for (int s = m.nCRange[i1 - 1] + 1; s <= m.nCRange[i1]; ++s) {
            int i5 = m.nodalConn[s - 1];
            if (i5 == 0) continue;
            int i6 = m.currentRank[i1 - 1] - m.currentRank[i5 - 1];
            if (i6 >= 0 && i6 <= m.solverMaxBW && i6 + 1 > m.thermMatLTRange[i3])
                m.thermMatLTRange[i3] = i6 + 1;
        }
    }
    for (int i = 2; i <= N; ++i) m.thermMatLTRange[i] += m.thermMatLTRange[i - 1];
    build_mpc(m, false, m.elMPCDim, m.elMPCMatRange, m.elMPCMatCols, m.elMPCMat, m.elMPCRHS);
    build_mpc(m, true, m.thermMPCDim, m.thermMPCMatRange, m.thermMPCMatCols, m.thermMPCMat,
              m.thermMPCRHS);
    m.elasticLoad.assign(n, 0.0);
    m.delDisp.assign(n, 0.0);
    m.nodeDisp.assign(n, 0.0);
    m.thermalLoad.assign(N, 0.0);
    m.delTemp.assign(N, 0.0);
}

// Element coordinates (3x8 column-major) of element e (0-based).
// Working/AStrO_r_designPropertyFunctions.f90:945-968 with no DV offsets,
// rotated by the identity section frame (solids with default orientation).
void orc_element_coords(const OrcModel* m, int e, double* xyz) {
    for (int a = 0; a < 8; ++a) {
        int nd = m->elementList[8 * (size_t)e + a];
        for (int c = 0; c < 3; ++c) xyz[c + 3 * a] = nd > 0 ? m->nodeList[3 * (size_t)(nd - 1) + c] : 1.0;
    }
}

// a16: element loop + linear-probe CSR scatter (buildMat = 1 path; at u = 0
// the internal force is zero, so the residual scatter adds nothing).
// Working/AStrO_commandFunctions.f90:260-351.  `massFactor` adds
// massFactor * M_e to each tangent (the dynamic branch of r_getElRu,
// Working/AStrO_r_elementEqns.f90:2684-2696, c1*c2 = 1/(beta*dt^2)).
void orc_assemble_elastic(OrcModel* mp, double massFactor) {
    OrcModel& m = *mp;
    std::fill(m.elasticMat.begin(), m.elasticMat.end(), 0.0);
    std::vector<double> Ke(576), Me(576);
    double xyz[24];
    for (int e = 0; e < m.numEls; ++e) {
        if (m.elementType[e] == 41) {
            if (massFactor != 0.0 || m.elABD.empty()) {
                std::fprintf(stderr, "oracle: shell4 needs sections and supports statics only\n");
                std::abort();
            }
            double alpha[9];
            orc_element_coords(&m, e, xyz);
            orc_shell_frame(&m.elOrient[6 * (size_t)e], xyz, alpha);
            orc_shell4_stiffness(xyz, alpha, &m.elABD[81 * (size_t)e], nullptr, Ke.data());
            ElProfile p;
            element_profile(41, p);
            for (int i = 0; i < 24; ++i) {
                int kGi = m.nDofIndex[6 * (size_t)(m.elementList[8 * (size_t)e + p.dofTable[1][i] - 1] - 1) +
                                      p.dofTable[0][i] - 1];
                for (int j = 0; j < 24; ++j) {
                    int kGj = m.nDofIndex[6 * (size_t)(m.elementList[8 * (size_t)e + p.dofTable[1][j] - 1] - 1) +
                                          p.dofTable[0][j] - 1];
                    probe_add(m.elasticMat, m.elMatCols, m.elMatRange, kGi, kGj, Ke[i + 24 * j], false);
                }
            }
            continue;
        }
        if (m.elementType[e] != 8) {
            std::fprintf(stderr, "oracle: assemble supports brick8 and shell4 only\n");
            std::abort();
        }
        orc_element_coords(&m, e, xyz);
        orc_brick8_stiffness(xyz, &m.elC[36 * (size_t)e], Ke.data());
        if (massFactor != 0.0) {
            orc_brick8_mass(xyz, m.elDen[e], Me.data());
            for (int k = 0; k < 576; ++k) Ke[k] += massFactor * Me[k];
        }
        ElProfile p;
        element_profile(8, p);
        for (int i = 0; i < 24; ++i) {
            int kGi = m.nDofIndex[6 * (size_t)(m.elementList[8 * (size_t)e + p.dofTable[1][i] - 1] - 1) +
                                  p.dofTable[0][i] - 1];
            for (int j = 0; j < 24; ++j) {
                int kGj = m.nDofIndex[6 * (size_t)(m.elementList[8 * (size_t)e + p.dofTable[1][j] - 1] - 1) +
                                      p.dofTable[0][j] - 1];
                probe_add(m.elasticMat, m.elMatCols, m.elMatRange, kGi, kGj, Ke[i + 24 * j], false);
            }
        }
    }
}

// a17: thermal element loop + scatter, indices by node rank.
// Working/AStrO_commandFunctions.f90:16-74.  literalD8 = 1 keeps the
// reference's `.lt.` probe bound that never fills the last slot of a row (:57).
void orc_assemble_thermal(OrcModel* mp, int literalD8) {
    OrcModel& m = *mp;
    std::fill(m.thermMat.begin(), m.thermMat.end(), 0.0);
    std::fill(m.thermMatCols.begin(), m.thermMatCols.end(), 0);
    double xyz[24], Kt[64];
    for (int e = 0; e < m.numEls; ++e) {
        orc_element_coords(&m, e, xyz);
        orc_brick8_conductivity(xyz, &m.elCond[9 * (size_t)e], Kt);
        for (int a = 0; a < 8; ++a) {
            int kGi = m.currentRank[m.elementList[8 * (size_t)e + a] - 1];
            for (int b = 0; b < 8; ++b) {
                int kGj = m.currentRank[m.elementList[8 * (size_t)e + b] - 1];
                probe_add(m.thermMat, m.thermMatCols, m.thermMatRange, kGi, kGj, Kt[a + 8 * b],
                          literalD8 != 0);
            }
        }
    }
}

// a19: nodal loads (Working/AStrO_commandFunctions.f90:368-392), then the
// constraint load C^T (rhs - C u) (:506-535).  Resets elasticLoad first, as
// stepSolve does (:718-723), with the residual term zero at u = 0.
void orc_elastic_rhs(OrcModel* mp) {
    OrcModel& m = *mp;
    std::fill(m.elasticLoad.begin(), m.elasticLoad.end(), 0.0);
    for (size_t l = 0; l < m.loadNode.size(); ++l)
        for (int c = 0; c < 6; ++c) {
            int d = m.nDofIndex[6 * (size_t)(m.loadNode[l] - 1) + c];
            if (d != 0) m.elasticLoad[d - 1] += m.loadVal[7 * l + c];
        }
    std::vector<double> t(m.elMPCDim);
    for (int i = 1; i <= m.elMPCDim; ++i) {
        double v = m.elMPCRHS[i - 1];
        for (int s = m.elMPCMatRange[i - 1] + 1; s <= m.elMPCMatRange[i]; ++s) {
            int c = m.elMPCMatCols[s - 1];
            if (c != 0) v -= m.elMPCMat[s - 1] * m.nodeDisp[c - 1];
        }
        t[i - 1] = v;
    }
    for (int i = 1; i <= m.elMPCDim; ++i)
        for (int s = m.elMPCMatRange[i - 1] + 1; s <= m.elMPCMatRange[i]; ++s) {
            int c = m.elMPCMatCols[s - 1];
            if (c != 0) m.elasticLoad[c - 1] += m.elMPCMat[s - 1] * t[i - 1];
        }
}

// Thermal right-hand side: nodal heat loads (component 7) and the constraint
// load; Working/AStrO_commandFunctions.f90:76-251 (nodal branch) and the
// thermal twin of getElasticConstraintLoad.
void orc_thermal_rhs(OrcModel* mp) {
    OrcModel& m = *mp;
    std::fill(m.thermalLoad.begin(), m.thermalLoad.end(), 0.0);
    for (size_t l = 0; l < m.loadNode.size(); ++l) {
        int r = m.currentRank[m.loadNode[l] - 1];
        m.thermalLoad[r - 1] += m.loadVal[7 * l + 6];
    }
    std::vector<double> t(m.thermMPCDim);
    for (int i = 1; i <= m.thermMPCDim; ++i) t[i - 1] = m.thermMPCRHS[i - 1];  // T = 0 initially
    for (int i = 1; i <= m.thermMPCDim; ++i)
        for (int s = m.thermMPCMatRange[i - 1] + 1; s <= m.thermMPCMatRange[i]; ++s) {
            int c = m.thermMPCMatCols[s - 1];
            if (c != 0) m.thermalLoad[c - 1] += m.thermMPCMat[s - 1] * t[i - 1];
        }
}

// scaleElasticMPC / scaleThermalMPC: normalise every constraint row, then
// scale rows and rhs by 100 * max|K_ij|.  Working/AStrO_commandFunctions.f90:537-566
// (thermal twin :228-258).
static void scale_mpc(const std::vector<double>& K, std::vector<double>& C, const std::vector<int>& range,
                      std::vector<double>& rhs, int dim) {
    double maxMat = 0;
    for (double v : K) maxMat = std::max(maxMat, std::fabs(v));
    for (int i = 1; i <= dim; ++i) {
        double mag = 0;
        for (int s = range[i - 1] + 1; s <= range[i]; ++s) mag += C[s - 1] * C[s - 1];
        mag = std::sqrt(mag);
        for (int s = range[i - 1] + 1; s <= range[i]; ++s) C[s - 1] = (1.0 / mag) * C[s - 1];
        rhs[i - 1] = (1.0 / mag) * rhs[i - 1];
    }
    for (double& v : C) v = 100.0 * maxMat * v;
    for (double& v : rhs) v = 100.0 * maxMat * v;
}

void orc_scale_elastic_mpc(OrcModel* m) {
    scale_mpc(m->elasticMat, m->elMPCMat, m->elMPCMatRange, m->elMPCRHS, m->elMPCDim);
}
void orc_scale_thermal_mpc(OrcModel* m) {
    scale_mpc(m->thermMat, m->thermMPCMat, m->thermMPCMatRange, m->thermMPCRHS, m->thermMPCDim);
}

void orc_build_elastic_profile(OrcModel* mp) {
    OrcModel& m = *mp;
    m.elMatLT.assign(m.elMatLTRange[m.elMatDim], 0.0);
    orc_convert_ltri(m.elMatLT.data(), m.elMatLTRange.data(), (long long)m.elMatLT.size(), m.elasticMat.data(),
                     m.elMatCols.data(), m.elMatRange.data(), m.elMatDim, m.elMPCMat.data(),
                     m.elMPCMatCols.data(), m.elMPCMatRange.data(), m.elMPCDim);
}
void orc_build_thermal_profile(OrcModel* mp) {
    OrcModel& m = *mp;
    m.thermMatLT.assign(m.thermMatLTRange[m.numNodes], 0.0);
    orc_convert_ltri(m.thermMatLT.data(), m.thermMatLTRange.data(), (long long)m.thermMatLT.size(),
                     m.thermMat.data(), m.thermMatCols.data(), m.thermMatRange.data(), m.numNodes,
                     m.thermMPCMat.data(), m.thermMPCMatCols.data(), m.thermMPCMatRange.data(),
                     m.thermMPCDim);
}

int orc_factor_elastic(OrcModel* m) {
    return orc_ldl_factor(m->elMatLT.data(), (long long)m->elMatLT.size(), m->elMatLTRange.data(), m->elMatDim);
}
int orc_factor_thermal(OrcModel* m) {
    return orc_ldl_factor(m->thermMatLT.data(), (long long)m->thermMatLT.size(), m->thermMatLTRange.data(),
                          m->numNodes);
}

// numVecs rule of stepSolve: n/2 below 200 unknowns, else 100.
// Working/AStrO_commandFunctions.f90:711-715 (thermal :690-694).
int orc_num_vecs(int n) { return n < 200 ? n / 2 : 100; }

// stepSolve (linear static elastic branch): RHS, GMRES from the stored delDisp,
// nodeDisp += delDisp.  Working/AStrO_commandFunctions.f90:681-749.  mode 0
// keeps the reference's GMRES (maxIt = n); mode 1 is the stable form with a
// relative tolerance rtol and no practical cycle cap.
void orc_solve_elastic(OrcModel* mp, int numVecs, int mode, double rtol) {
    OrcModel& m = *mp;
    orc_elastic_rhs(mp);
    if (numVecs <= 0) numVecs = orc_num_vecs(m.elMatDim);
    int maxIt = mode == 0 ? m.elMatDim : 1000 * numVecs;
    m.gmresCycles = orc_gmres_mode(m.delDisp.data(), m.elasticMat.data(), m.elMatCols.data(), m.elMatRange.data(),
                              m.elMatDim, m.elMPCMat.data(), m.elMPCMatCols.data(), m.elMPCMatRange.data(),
                              m.elMPCDim, m.elMatLT.data(), m.elMatLTRange.data(), m.elasticLoad.data(), numVecs,
                              maxIt, mode, rtol, &m.gmresRes, &m.gmresSteps);
    for (int i = 0; i < m.elMatDim; ++i) m.nodeDisp[i] += m.delDisp[i];
}

void orc_solve_thermal(OrcModel* mp, int numVecs, int mode, double rtol) {
    OrcModel& m = *mp;
    orc_thermal_rhs(mp);
    if (numVecs <= 0) numVecs = orc_num_vecs(m.numNodes);
    int maxIt = mode == 0 ? m.numNodes : 1000 * numVecs;
    m.gmresCycles = orc_gmres_mode(m.delTemp.data(), m.thermMat.data(), m.thermMatCols.data(), m.thermMatRange.data(),
                              m.numNodes, m.thermMPCMat.data(), m.thermMPCMatCols.data(),
                              m.thermMPCMatRange.data(), m.thermMPCDim, m.thermMatLT.data(),
                              m.thermMatLTRange.data(), m.thermalLoad.data(), numVecs, maxIt, mode, rtol,
                              &m.gmresRes, &m.gmresSteps);
}

// ---------------------------------------------------------------------------
// Newmark dynamics, the dynamic branch of solve (Working/AStrO_commandFunctions.f90:
// 762-810) with the intended element gather (D4 corrected: every element sees
// its own nodes' solution, velocity and acceleration not swapped).
// c1 = 0.5/beta, c2 = 2/dt^2 (:787-788, Working/AStrO_r_elementEqns.f90:2685-2686).

// Assemble the tangent K + c1*c2*M at u = 0 (getElasticSolnLoad(1) with
// dynamic = 1), scale the penalty rows, build the profile; zero the state.
void orc_dyn_init(OrcModel* mp, double beta, double gamma, double dt) {
    OrcModel& m = *mp;
    m.nmBeta = beta;
    m.nmGamma = gamma;
    m.delT = dt;
    const double c1 = 0.5 / beta, c2 = 2.0 / (dt * dt);
    orc_assemble_elastic(mp, c1 * c2);
    orc_scale_elastic_mpc(mp);
    orc_build_elastic_profile(mp);
    const size_t n = m.elMatDim;
    m.nodeDisp.assign(n, 0.0);
    m.delDisp.assign(n, 0.0);
    m.prevDisp.assign(n, 0.0);
    m.prevVel.assign(n, 0.0);
    m.prevAcc.assign(n, 0.0);
    m.nodeVel.assign(n, 0.0);
    m.nodeAcc.assign(n, 0.0);
}

// Right-hand side of one step (stepSolve :718-723): minus the element
// residuals Ru = Ruk(u) + M a_next(u) (r_getElRu :2659-2696), plus the nodal
// loads times loadScale (the caller's time function; the reference switches
// whole loads on and off by their active-time windows, :366), plus the
// constraint load C^T (rhs - C u) (:506-535).
void orc_dyn_residual_load(OrcModel* mp, double loadScale) {
    OrcModel& m = *mp;
    const double c1 = 0.5 / m.nmBeta, c2 = 2.0 / (m.delT * m.delT), dt = m.delT, beta = m.nmBeta;
    std::fill(m.elasticLoad.begin(), m.elasticLoad.end(), 0.0);
    ElProfile p;
    element_profile(8, p);
    double xyz[24], ue[24], ae[24], Ri[24], Me[576];
    int dof[24];
    for (int e = 0; e < m.numEls; ++e) {
        orc_element_coords(&m, e, xyz);
        for (int d = 0; d < 24; ++d) {
            int node = m.elementList[8 * (size_t)e + p.dofTable[1][d] - 1];
            dof[d] = m.nDofIndex[6 * (size_t)(node - 1) + p.dofTable[0][d] - 1];
            const size_t k = dof[d] - 1;
            ue[d] = m.nodeDisp[k];
            ae[d] = c1 * (c2 * (m.nodeDisp[k] - m.prevDisp[k] - dt * m.prevVel[k]) - (1.0 - 2.0 * beta) * m.prevAcc[k]);
        }
        orc_brick8_internal_force(xyz, &m.elC[36 * (size_t)e], ue, Ri);
        orc_brick8_mass(xyz, m.elDen[e], Me);
        for (int i = 0; i < 24; ++i) {
            double rm = 0.0;
            for (int j = 0; j < 24; ++j) rm += Me[i + 24 * j] * ae[j];
            m.elasticLoad[dof[i] - 1] -= Ri[i] + rm;
        }
    }
    for (size_t l = 0; l < m.loadNode.size(); ++l)
        for (int c = 0; c < 6; ++c) {
            int d = m.nDofIndex[6 * (size_t)(m.loadNode[l] - 1) + c];
            if (d != 0) m.elasticLoad[d - 1] += loadScale * m.loadVal[7 * l + c];
        }
    std::vector<double> t(m.elMPCDim);
    for (int i = 1; i <= m.elMPCDim; ++i) {
        double v = m.elMPCRHS[i - 1];
        for (int s = m.elMPCMatRange[i - 1] + 1; s <= m.elMPCMatRange[i]; ++s) {
            int c = m.elMPCMatCols[s - 1];
            if (c != 0) v -= m.elMPCMat[s - 1] * m.nodeDisp[c - 1];
        }
        t[i - 1] = v;
    }
    for (int i = 1; i <= m.elMPCDim; ++i)
        for (int s = m.elMPCMatRange[i - 1] + 1; s <= m.elMPCMatRange[i]; ++s) {
            int c = m.elMPCMatCols[s - 1];
            if (c != 0) m.elasticLoad[c - 1] += m.elMPCMat[s - 1] * t[i - 1];
        }
}

// One time step: RHS, GMRES from the previous increment (stepSolve passes
// delDisp as the initial guess), u += du, then the Newmark updates of
// acceleration and velocity and the shift to the previous state (:792-806).
// mode/rtol as orc_solve_elastic.
void orc_dyn_step(OrcModel* mp, double loadScale, int numVecs, int mode, double rtol) {
    OrcModel& m = *mp;
    orc_dyn_residual_load(mp, loadScale);
    if (numVecs <= 0) numVecs = orc_num_vecs(m.elMatDim);
    int maxIt = mode == 0 ? m.elMatDim : 1000 * numVecs;
    m.gmresCycles = orc_gmres_mode(m.delDisp.data(), m.elasticMat.data(), m.elMatCols.data(), m.elMatRange.data(),
                                   m.elMatDim, m.elMPCMat.data(), m.elMPCMatCols.data(), m.elMPCMatRange.data(),
                                   m.elMPCDim, m.elMatLT.data(), m.elMatLTRange.data(), m.elasticLoad.data(),
                                   numVecs, maxIt, mode, rtol, &m.gmresRes, &m.gmresSteps);
    const double c1 = 0.5 / m.nmBeta, c2 = 2.0 / (m.delT * m.delT), dt = m.delT;
    const double beta = m.nmBeta, gam = m.nmGamma;
    for (int i = 0; i < m.elMatDim; ++i) {
        m.nodeDisp[i] += m.delDisp[i];
        m.nodeAcc[i] = c1 * (c2 * (m.nodeDisp[i] - m.prevDisp[i] - dt * m.prevVel[i]) - (1.0 - 2.0 * beta) * m.prevAcc[i]);
        m.nodeVel[i] = m.prevVel[i] + dt * ((1.0 - gam) * m.prevAcc[i] + gam * m.nodeAcc[i]);
        m.prevDisp[i] = m.nodeDisp[i];
        m.prevVel[i] = m.nodeVel[i];
        m.prevAcc[i] = m.nodeAcc[i];
    }
}

// ---------------------------------------------------------------------------
// a15 + a18: design sensitivities of shell4 layer thickness by complex step.
// Design variable k adds coef_e * d_k to the thickness of layer dvLayer
// (0-based) of every element e with elDV[e] = k (the reference's 'layup'
// 'thickness' category, Working/AStrO_r_designPropertyFunctions.f90:283-302).
void orc_set_design(OrcModel* m, int nDV, const int* elDV, const double* elCoef, int dvLayer) {
    m->nDV = nDV;
    m->dvLayer = dvLayer;
    m->elDV.assign(elDV, elDV + m->numEls);
    m->elCoef.assign(elCoef, elCoef + m->numEls);
}

// Complex-step derivative of one shell4's condensed tangent with respect to
// the thickness of layer `layer`: Im(Kc(t + i h)) / h, the whole element
// evaluation (ABD, tangent, condensation) run in complex arithmetic as the
// reference's AStrO_c_* twins do.  K32 (optional) gets the derivative before
// condensation.  dKc: 24x24 column-major.
void orc_shell4_dstiffness_cs(const double* xyz, const double* alpha, int nL, const double* layers, double zOff,
                              int layer, double h, double* K32, double* dKc) {
    using cd = std::complex<double>;
    std::vector<cd> th(nL), abd(81), k32(32 * 32), kc(24 * 24);
    for (int l = 0; l < nL; ++l) th[l] = cd(layers[6 * l + 4], l == layer ? h : 0.0);
    shell_abd_t<cd>(nL, layers, th.data(), zOff, abd.data());
    shell4_stiffness_t<cd>(xyz, alpha, abd.data(), k32.data(), kc.data());
    if (K32)
        for (int k = 0; k < 32 * 32; ++k) K32[k] = k32[k].imag() / h;
    for (int k = 0; k < 24 * 24; ++k) dKc[k] = kc[k].imag() / h;
}

// dR/dd at the current nodeDisp, the per-variable loop of getdRelasticdD
// (Working/AStrO_commandFunctions.f90:1172-1375, with its element-id and DOF
// indexing corrected, D9): for each design variable k, the elements on k
// evaluate R_e = Kc_e(t + i h coef_e) u_e in complex arithmetic and column k
// of dRdD takes Im(R_e) / h at the element's DOFs.  The analysis is linear
// (R = K(d) u - F with F independent of d), so dR/dd_k = (dK/dd_k) u.
// dRdD: elMatDim x nDV column-major.
void orc_dRdD(OrcModel* mp, double h, double* dRdD) {
    OrcModel& m = *mp;
    const size_t n = m.elMatDim;
    std::fill(dRdD, dRdD + n * m.nDV, 0.0);
    ElProfile p;
    element_profile(41, p);
    double xyz[24], alpha[9], dK[576];
    for (int k = 0; k < m.nDV; ++k)
        for (int e = 0; e < m.numEls; ++e) {
            if (m.elDV[e] != k || m.elementType[e] != 41) continue;
            orc_element_coords(&m, e, xyz);
            orc_shell_frame(&m.elOrient[6 * (size_t)e], xyz, alpha);
            orc_shell4_dstiffness_cs(xyz, alpha, m.nLayers, &m.elLayers[6 * (size_t)m.nLayers * e], m.elZOff[e],
                                     m.dvLayer, h * m.elCoef[e], nullptr, dK);
            int dof[24];
            double ue[24];
            for (int d = 0; d < 24; ++d) {
                int node = m.elementList[8 * (size_t)e + p.dofTable[1][d] - 1];
                dof[d] = m.nDofIndex[6 * (size_t)(node - 1) + p.dofTable[0][d] - 1];
                ue[d] = m.nodeDisp[dof[d] - 1];
            }
            for (int i = 0; i < 24; ++i) {
                double acc = 0;
                for (int j = 0; j < 24; ++j) acc += dK[i + 24 * j] * ue[j];
                dRdD[(size_t)k * n + dof[i] - 1] += m.elCoef[e] * acc;
            }
        }
}

// Tip-compliance (or any work-conjugate compliance C = F . u) gradient by the
// adjoint: K symmetric and F independent of d give lambda = u, so
// dC/dd_k = -u . dR/dd_k (getTotaldLdD, Working/AStrO_commandFunctions.f90:
// 1377-1466, for this objective).
void orc_compliance_gradient(OrcModel* mp, double h, double* grad) {
    OrcModel& m = *mp;
    const size_t n = m.elMatDim;
    std::vector<double> R(n * m.nDV);
    orc_dRdD(mp, h, R.data());
    for (int k = 0; k < m.nDV; ++k) {
        double acc = 0;
        for (size_t i = 0; i < n; ++i) acc += m.nodeDisp[i] * R[(size_t)k * n + i];
        grad[k] = -acc;
    }
}

// Every element of a model at s with its own nodes' displacements (the
// reference's r_getElementSolution gathers by local node number, D4).
// stress, strain: 6 x numEls.
void orc_element_stress(OrcModel* mp, const double* s, double* stress, double* strain) {
    OrcModel& m = *mp;
    ElProfile p;
    element_profile(8, p);
    double xyz[24], ue[24];
    for (int e = 0; e < m.numEls; ++e) {
        orc_element_coords(&m, e, xyz);
        for (int d = 0; d < 24; ++d) {
            int node = m.elementList[8 * (size_t)e + p.dofTable[1][d] - 1];
            int dof = m.nDofIndex[6 * (size_t)(node - 1) + p.dofTable[0][d] - 1];
            ue[d] = m.nodeDisp[dof - 1];
        }
        orc_brick8_stress(xyz, &m.elC[36 * (size_t)e], ue, s, stress + 6 * (size_t)e, strain + 6 * (size_t)e);
    }
}

// f3: transient conduction, the thermal branch of solve() with dynamic = 1
// (Working/AStrO_commandFunctions.f90:763-769, 689-703, 791-803).  Set-up:
// getThermalSolnLoad(1), whose element tangent r_getElRt adds
// dRdTdot / (gamma dt) to the conductivity (Working/AStrO_r_elementEqns.f90:
// 2190-2200), probed into the thermal CSR in the same order; the capacitance
// alone goes into capMat on the same columns.  Then scaleThermalMPC, the
// profile and the factor; the state starts at T = Tdot = 0.
void orc_thermal_dyn_init(OrcModel* mp, double gamma, double dt, int literalD8) {
    OrcModel& m = *mp;
    const int N = m.numNodes;
    const double cm = 1.0 / (gamma * dt);
    m.thGamma = gamma;
    m.thDt = dt;
    std::fill(m.thermMat.begin(), m.thermMat.end(), 0.0);
    std::fill(m.thermMatCols.begin(), m.thermMatCols.end(), 0);
    m.capMat.assign(m.thermMat.size(), 0.0);
    m.capCols.assign(m.thermMat.size(), 0);
    double xyz[24], Kt[64], Mt[64];
    for (int e = 0; e < m.numEls; ++e) {
        orc_element_coords(&m, e, xyz);
        orc_brick8_conductivity(xyz, &m.elCond[9 * (size_t)e], Kt);
        orc_brick8_capacitance(xyz, m.elCap[e], Mt);
        for (int a = 0; a < 8; ++a) {
            int kGi = m.currentRank[m.elementList[8 * (size_t)e + a] - 1];
            for (int b = 0; b < 8; ++b) {
                int kGj = m.currentRank[m.elementList[8 * (size_t)e + b] - 1];
                probe_add(m.thermMat, m.thermMatCols, m.thermMatRange, kGi, kGj, Kt[a + 8 * b] + cm * Mt[a + 8 * b],
                          literalD8 != 0);
                probe_add(m.capMat, m.capCols, m.thermMatRange, kGi, kGj, Mt[a + 8 * b], literalD8 != 0);
            }
        }
    }
    orc_scale_thermal_mpc(mp);
    orc_build_thermal_profile(mp);
    orc_factor_thermal(mp);
    for (auto* v : {&m.nodeTemp, &m.prevTemp, &m.prevTdot, &m.nodeTdot, &m.delTemp}) v->assign(N, 0.0);
}

// One step (stepSolve's thermal branch): thermalLoad = -Rt(T) element by
// element, Rt = K_e T_e + C_e Tdot_e with Tdot = (1 / gamma) ((T - Tp) / dt
// - (1 - gamma) Tdp) per node (r_getElRt), then the constraint load
// C^T (rhs - C T); GMRES from the previous increment; T += dT; Tdot and the
// previous state (:791-803).  mode / rtol as orc_solve_elastic.
void orc_thermal_dyn_step(OrcModel* mp, int numVecs, int mode, double rtol) {
    OrcModel& m = *mp;
    const int N = m.numNodes;
    const double g = m.thGamma, dt = m.thDt;
    std::vector<double> Td(N);
    for (int i = 0; i < N; ++i)
        Td[i] = (1.0 / g) * ((1.0 / dt) * (m.nodeTemp[i] - m.prevTemp[i]) - (1.0 - g) * m.prevTdot[i]);
    std::fill(m.thermalLoad.begin(), m.thermalLoad.end(), 0.0);
    double xyz[24], Kt[64], Mt[64];
    for (int e = 0; e < m.numEls; ++e) {
        orc_element_coords(&m, e, xyz);
        orc_brick8_conductivity(xyz, &m.elCond[9 * (size_t)e], Kt);
        orc_brick8_capacitance(xyz, m.elCap[e], Mt);
        int r[8];
        for (int a = 0; a < 8; ++a) r[a] = m.currentRank[m.elementList[8 * (size_t)e + a] - 1];
        for (int a = 0; a < 8; ++a) {
            double acc = 0.0;
            for (int b = 0; b < 8; ++b) acc += Kt[a + 8 * b] * m.nodeTemp[r[b] - 1] + Mt[a + 8 * b] * Td[r[b] - 1];
            m.thermalLoad[r[a] - 1] -= acc;
        }
    }
    std::vector<double> t(m.thermMPCDim);
    for (int i = 1; i <= m.thermMPCDim; ++i) {
        double v = m.thermMPCRHS[i - 1];
        for (int s = m.thermMPCMatRange[i - 1] + 1; s <= m.thermMPCMatRange[i]; ++s) {
            int c = m.thermMPCMatCols[s - 1];
            if (c != 0) v -= m.thermMPCMat[s - 1] * m.nodeTemp[c - 1];
        }
        t[i - 1] = v;
    }
    for (int i = 1; i <= m.thermMPCDim; ++i)
        for (int s = m.thermMPCMatRange[i - 1] + 1; s <= m.thermMPCMatRange[i]; ++s) {
            int c = m.thermMPCMatCols[s - 1];
            if (c != 0) m.thermalLoad[c - 1] += m.thermMPCMat[s - 1] * t[i - 1];
        }
    if (numVecs <= 0) numVecs = orc_num_vecs(N);
    int maxIt = mode == 0 ? N : 1000 * numVecs;
    m.gmresCycles = orc_gmres_mode(m.delTemp.data(), m.thermMat.data(), m.thermMatCols.data(), m.thermMatRange.data(),
                                   N, m.thermMPCMat.data(), m.thermMPCMatCols.data(), m.thermMPCMatRange.data(),
                                   m.thermMPCDim, m.thermMatLT.data(), m.thermMatLTRange.data(), m.thermalLoad.data(),
                                   numVecs, maxIt, mode, rtol, &m.gmresRes, &m.gmresSteps);
    for (int i = 0; i < N; ++i) {
        const double T = m.nodeTemp[i] + m.delTemp[i];
        const double td = (1.0 / g) * ((1.0 / dt) * (T - m.prevTemp[i]) - (1.0 - g) * m.prevTdot[i]);
        m.nodeTemp[i] = T;
        m.nodeTdot[i] = td;
        m.prevTemp[i] = T;
        m.prevTdot[i] = td;
    }
}

// Array getters for the Python side of the tests.
const int* orc_geti(OrcModel* m, const char* name, long long* len) {
    std::string s(name);
    const std::vector<int>* v = nullptr;
    if (s == "currentRank") v = &m->currentRank;
    else if (s == "originalRank") v = &m->originalRank;
    else if (s == "nDofIndex") v = &m->nDofIndex;
    else if (s == "nCRange") v = &m->nCRange;
    else if (s == "nodalConn") v = &m->nodalConn;
    else if (s == "elMatRange") v = &m->elMatRange;
    else if (s == "elMatCols") v = &m->elMatCols;
    else if (s == "elMatLTRange") v = &m->elMatLTRange;
    else if (s == "elMPCMatRange") v = &m->elMPCMatRange;
    else if (s == "elMPCMatCols") v = &m->elMPCMatCols;
    else if (s == "thermMatRange") v = &m->thermMatRange;
    else if (s == "thermMatCols") v = &m->thermMatCols;
    else if (s == "thermMatLTRange") v = &m->thermMatLTRange;
    else if (s == "thermMPCMatRange") v = &m->thermMPCMatRange;
    else if (s == "thermMPCMatCols") v = &m->thermMPCMatCols;
    if (!v) { *len = -1; return nullptr; }
    *len = (long long)v->size();
    return v->data();
}

double* orc_getd(OrcModel* m, const char* name, long long* len) {
    std::string s(name);
    std::vector<double>* v = nullptr;
    if (s == "elasticMat") v = &m->elasticMat;
    else if (s == "elMatLT") v = &m->elMatLT;
    else if (s == "elasticLoad") v = &m->elasticLoad;
    else if (s == "delDisp") v = &m->delDisp;
    else if (s == "nodeDisp") v = &m->nodeDisp;
    else if (s == "elMPCMat") v = &m->elMPCMat;
    else if (s == "elMPCRHS") v = &m->elMPCRHS;
    else if (s == "thermMat") v = &m->thermMat;
    else if (s == "thermMatLT") v = &m->thermMatLT;
    else if (s == "thermalLoad") v = &m->thermalLoad;
    else if (s == "delTemp") v = &m->delTemp;
    else if (s == "nodeTemp") v = &m->nodeTemp;
    else if (s == "nodeTdot") v = &m->nodeTdot;
    else if (s == "capMat") v = &m->capMat;
    else if (s == "thermMPCMat") v = &m->thermMPCMat;
    else if (s == "thermMPCRHS") v = &m->thermMPCRHS;
    else if (s == "elC") v = &m->elC;
    else if (s == "nodeVel") v = &m->nodeVel;
    else if (s == "nodeAcc") v = &m->nodeAcc;
    else if (s == "prevDisp") v = &m->prevDisp;
    if (!v) { *len = -1; return nullptr; }
    *len = (long long)v->size();
    return v->data();
}

int orc_geti_scalar(OrcModel* m, const char* name) {
    std::string s(name);
    if (s == "elMatDim") return m->elMatDim;
    if (s == "numEls") return m->numEls;
    if (s == "solverMaxBW") return m->solverMaxBW;
    if (s == "numSolidNds") return m->numSolidNds;
    if (s == "elMPCDim") return m->elMPCDim;
    if (s == "thermMPCDim") return m->thermMPCDim;
    if (s == "gmresCycles") return m->gmresCycles;
    if (s == "gmresSteps") return m->gmresSteps;
    return -1;
}

double orc_getd_scalar(OrcModel* m, const char* name) {
    std::string s(name);
    if (s == "gmresRes") return m->gmresRes;
    return NAN;
}

}  // extern "C"
