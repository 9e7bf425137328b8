// Host-side shell section data of the device path: the section frame and the
// layup stiffness ABD of every shell4 element, computed once per model.

#include <cmath>
#include <complex>

#include "astro_prep.h"

namespace astro {

namespace {

void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

double unit3(double* a) {
    const double m = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int i = 0; i < 3; ++i) a[i] /= m;
    return m;
}

// inverse of a 3x3 (row-major) by the adjugate
void inv3(const double A[3][3], double I[3][3]) {
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    const double d = 1.0 / det;
    I[0][0] = d * c00;
    I[1][0] = d * c01;
    I[2][0] = d * c02;
    I[0][1] = d * (A[0][2] * A[2][1] - A[0][1] * A[2][2]);
    I[1][1] = d * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
    I[2][1] = d * (A[0][1] * A[2][0] - A[0][0] * A[2][1]);
    I[0][2] = d * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
    I[1][2] = d * (A[0][2] * A[1][0] - A[0][0] * A[1][2]);
    I[2][2] = d * (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
}

}  // namespace

// Section frame, rows = local axes: x along the first orientation vector,
// z = x cross (second vector), y = z cross x (Working/AStrO_input.f90:1071-1095),
// then turned about z cross n by asin|z cross n| so that z is the element
// normal n (the diagonals' cross product, oriented with z), as correctAlpha
// does (Working/AStrO_r_elementEqns.f90:240-286).  xyz: 3 x 4 column-major;
// alpha[i + 3 j] = component j of axis i.
void shell_frame(const double* or6, const double* xyz, double* alpha) {
    double ax[3][3];
    double x[3] = {or6[0], or6[1], or6[2]}, v[3] = {or6[3], or6[4], or6[5]}, z[3], y[3];
    unit3(x);
    cross3(x, v, z);
    unit3(z);
    cross3(z, x, y);
    for (int j = 0; j < 3; ++j) {
        ax[0][j] = x[j];
        ax[1][j] = y[j];
        ax[2][j] = z[j];
    }
    double d1[3], d2[3], n[3], w[3];
    for (int i = 0; i < 3; ++i) {
        d1[i] = xyz[i + 6] - xyz[i];
        d2[i] = xyz[i + 9] - xyz[i + 3];
    }
    cross3(d1, d2, n);
    if (n[0] * z[0] + n[1] * z[1] + n[2] * z[2] < 0.0)
        for (double& c : n) c = -c;
    unit3(n);
    cross3(z, n, w);
    const double s = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (s > 1e-5) {
        // rotate each axis about k = w / |w| by th = asin(s) (Rodrigues)
        const double th = std::asin(s < 1.0 ? s : 1.0), c = std::cos(th), sn = std::sin(th);
        const double k[3] = {w[0] / s, w[1] / s, w[2] / s};
        for (int r = 0; r < 3; ++r) {
            double kxa[3], a[3] = {ax[r][0], ax[r][1], ax[r][2]};
            cross3(k, a, kxa);
            const double kd = k[0] * a[0] + k[1] * a[1] + k[2] * a[2];
            for (int j = 0; j < 3; ++j) ax[r][j] = a[j] * c + kxa[j] * sn + k[j] * kd * (1.0 - c);
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) alpha[i + 3 * j] = ax[i][j];
}

// Layup stiffness by integration through the thickness
// (Working/AStrO_r_designPropertyFunctions.f90:259-363): per layer the plane
// stress Q = S^-1, rotated by the layer angle to Qbar = Ts Q Te^-1; A, B, D
// from the layer bounds, the stack starting at z = -t (zOff + 1) / 2;
// transverse shear and drilling stiffness A33.  layers: nL x (E1, E2, nu12,
// G12, thickness, angle in radians).  ABD: 9 x 9 column-major.
void shell_section_abd(int nL, const double* layers, double zOff, double* ABD) {
    for (int k = 0; k < 81; ++k) ABD[k] = 0.0;
    double t = 0.0;
    for (int l = 0; l < nL; ++l) t += layers[6 * l + 4];
    double z0 = -0.5 * t * (zOff + 1.0);
    for (int l = 0; l < nL; ++l) {
        const double* L = layers + 6 * l;
        const double z1 = z0 + L[4];
        const double S[3][3] = {{1.0 / L[0], -L[2] / L[0], 0.0}, {-L[2] / L[0], 1.0 / L[1], 0.0}, {0.0, 0.0, 1.0 / L[3]}};
        double Q[3][3];
        inv3(S, Q);
        const double c = std::cos(L[5]), s = std::sin(L[5]);
        const double Ts[3][3] = {{c * c, s * s, -2 * c * s}, {s * s, c * c, 2 * c * s}, {c * s, -s * c, c * c - s * s}};
        double Te[3][3];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) Te[r][q] = Ts[r][q] * (q == 2 ? 0.5 : 1.0) * (r == 2 ? 2.0 : 1.0);
        double TeI[3][3], QT[3][3], Qb[3][3];
        inv3(Te, TeI);
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) {
                QT[r][q] = 0.0;
                for (int k = 0; k < 3; ++k) QT[r][q] += Q[r][k] * TeI[k][q];
            }
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) {
                Qb[r][q] = 0.0;
                for (int k = 0; k < 3; ++k) Qb[r][q] += Ts[r][k] * QT[k][q];
            }
        const double fA = z1 - z0, fB = 0.5 * (z1 * z1 - z0 * z0), fD = (z1 * z1 * z1 - z0 * z0 * z0) / 3.0;
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) {
                ABD[r + 9 * q] += fA * Qb[r][q];
                ABD[r + 9 * (q + 3)] += fB * Qb[r][q];
                ABD[(r + 3) + 9 * q] += fB * Qb[q][r];
                ABD[(r + 3) + 9 * (q + 3)] += fD * Qb[r][q];
            }
        z0 = z1;
    }
    ABD[6 + 9 * 6] = ABD[7 + 9 * 7] = ABD[8 + 9 * 8] = ABD[2 + 9 * 2];
}

}  // namespace astro

namespace astro {

// Complex-step layup stiffness: the thickness of layer `layer` carries the
// imaginary part hc; the integration of shell_section_abd runs in complex
// arithmetic (the reference's complex twin c_getShellStiffness).  The real
// and imaginary parts of ABD go to ABDre / ABDim (9 x 9 column-major).
void shell_section_abd_cs(int nL, const double* layers, double zOff, int layer, double hc, double* ABDre,
                          double* ABDim) {
    using cd = std::complex<double>;
    cd A[81];
    for (cd& v : A) v = 0.0;
    cd t = 0.0;
    for (int l = 0; l < nL; ++l) t += cd(layers[6 * l + 4], l == layer ? hc : 0.0);
    cd z0 = -0.5 * t * (zOff + 1.0);
    for (int l = 0; l < nL; ++l) {
        const double* L = layers + 6 * l;
        const cd z1 = z0 + cd(L[4], l == layer ? hc : 0.0);
        // the real layer matrix Qbar, from the real routine with a unit layer
        double lay[6] = {L[0], L[1], L[2], L[3], 1.0, L[5]};
        double Q1[81];
        shell_section_abd(1, lay, -1.0, Q1);  // zOff = -1: z in [0, 1], A block = Qbar
        const cd fA = z1 - z0, fB = 0.5 * (z1 * z1 - z0 * z0), fD = (z1 * z1 * z1 - z0 * z0 * z0) / 3.0;
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) {
                const double Qb = Q1[r + 9 * q];
                A[r + 9 * q] += fA * Qb;
                A[r + 9 * (q + 3)] += fB * Qb;
                A[(r + 3) + 9 * q] += fB * Q1[q + 9 * r];
                A[(r + 3) + 9 * (q + 3)] += fD * Qb;
            }
        z0 = z1;
    }
    A[6 + 9 * 6] = A[7 + 9 * 7] = A[8 + 9 * 8] = A[2 + 9 * 2];
    for (int k = 0; k < 81; ++k) {
        ABDre[k] = A[k].real();
        ABDim[k] = A[k].imag();
    }
}

}  // namespace astro
