// Output-time Cauchy stress of a triangle (calc_stresses.cpp:45-179) and
// the eigen-decomposition the output file reports (write_output.cpp:136-156).
//
// The reference's tensor is, per triangle, with W = W_s + W_b the energy
// density per reference area (bend stiffening included), F = [R_x R_y] the
// surface gradient and J = |R_x x R_y| the area ratio,
//     sigma_ij = (A / J) sum_alpha (dW/dR^i_alpha) R^j_alpha,
// i.e. the reference area A times the membrane Cauchy stress J^-1 dW/dF F^T.
// The derivatives are those of the force path (dW/dR_x, dW/dR_y at fixed
// R_xx, R_xy, R_yy; the bending part enters through the normal), so the
// tensor is not symmetric in general.  A symmetric eigensolver given it reads
// its lower triangle only, as the reference's does; the smallest-magnitude
// eigenvalue (the normal direction, zero up to the bending part) is dropped
// and the other two are kept in the reference's cyclic order.
#pragma once

#include <math.h>

#include "tri_math.h"

namespace ms {

// Lower triangle (s00, s10, s11, s20, s21, s22) of the tensor above.  Rx, Ry,
// R2 (R2[k][c], k = xx, xy, yy), n and J = |R_x x R_y| of the current state;
// ab, bb the programmed metric at the shell's dial factor; h, mu, A the
// triangle's thickness, shear modulus and reference area.
__device__ __forceinline__ void cauchy_lower(double nu, double gamma, double h, double mu, double A,
                                             const double ab[3], const double bb[3], const double Rx[3],
                                             const double Ry[3], const double R2[3][3], const double n[3],
                                             const double b[3], double J, double S[6]) {
    const double P[3] = {ab[2], -ab[1], ab[0]};
    const double det = ab[0] * ab[2] - ab[1] * ab[1];
    const double idet32 = 1.0 / (det * sqrt(det));
    const double cs = h * mu / (1.0 - nu) * idet32;
    const double cb = mu * h * h * h / (12.0 * (1.0 - nu)) * idet32;
    const double a[3] = {Rx[0] * Rx[0] + Rx[1] * Rx[1] + Rx[2] * Rx[2], Rx[0] * Ry[0] + Rx[1] * Ry[1] + Rx[2] * Ry[2],
                         Ry[0] * Ry[0] + Ry[1] * Ry[1] + Ry[2] * Ry[2]};
    const double Da[3] = {a[0] - ab[0], a[1] - ab[1], a[2] - ab[2]};
    const double Db[3] = {b[0] - bb[0], b[1] - bb[1], b[2] - bb[2]};
    double ga[3], gbv[3];
    phi_grad(nu, P, Da, ga);
    const double Wb0 = cb * phi_grad(nu, P, Db, gbv);
    const double stiff = 1.0 + 2.0 * Wb0 / (gamma * mu * h);  // calc_deformation_forces.cpp:141
    const double Wa[3] = {cs * ga[0], cs * ga[1], cs * ga[2]};
    const double Wb[3] = {cb * stiff * gbv[0], cb * stiff * gbv[1], cb * stiff * gbv[2]};
    // bending reaches R_x, R_y through n = m/|m|, m = R_x x R_y:
    // dW/dR_x gets R_y x u, dW/dR_y gets u x R_x, u = (I - n n^T) w / |m|
    double w[3];
    for (int c = 0; c < 3; ++c) w[c] = -(Wb[0] * R2[0][c] + Wb[1] * R2[1][c] + Wb[2] * R2[2][c]);
    const double nw = n[0] * w[0] + n[1] * w[1] + n[2] * w[2];
    double u[3];
    for (int c = 0; c < 3; ++c) u[c] = (w[c] - nw * n[c]) / J;
    const double cx[3] = {Ry[1] * u[2] - Ry[2] * u[1], Ry[2] * u[0] - Ry[0] * u[2], Ry[0] * u[1] - Ry[1] * u[0]};
    const double cy[3] = {u[1] * Rx[2] - u[2] * Rx[1], u[2] * Rx[0] - u[0] * Rx[2], u[0] * Rx[1] - u[1] * Rx[0]};
    double gx[3], gy[3];
    for (int c = 0; c < 3; ++c) {
        gx[c] = 2.0 * Wa[0] * Rx[c] + Wa[1] * Ry[c] + cx[c];
        gy[c] = Wa[1] * Rx[c] + 2.0 * Wa[2] * Ry[c] + cy[c];
    }
    const double f = A / J;
    auto sig = [&](int i, int j) { return f * (gx[i] * Rx[j] + gy[i] * Ry[j]); };
    S[0] = sig(0, 0);
    S[1] = sig(1, 0);
    S[2] = sig(1, 1);
    S[3] = sig(2, 0);
    S[4] = sig(2, 1);
    S[5] = sig(2, 2);
}

// Eigenvalues (ascending) and orthonormal eigenvectors (columns of V) of the
// symmetric 3x3 matrix with lower triangle S, by cyclic Jacobi rotations
// (Jacobi 1846; the rotation of Golub and Van Loan, Matrix Computations,
// sec. 8.5: tan(phi) = sgn(theta)/(|theta| + sqrt(1 + theta^2)),
// theta = (a_qq - a_pp)/(2 a_pq), which zeroes a_pq), run until the
// off-diagonal part vanishes to rounding.
__device__ __forceinline__ void sym3_eigen(const double S[6], double lam[3], double V[3][3]) {
    double a[3][3] = {{S[0], S[1], S[3]}, {S[1], S[2], S[4]}, {S[3], S[4], S[5]}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (!(off > 1e-36 * diag)) break;  // also ends on an all-zero matrix
        for (int k = 0; k < 3; ++k) {
            const int p = k == 2 ? 1 : 0, q = k == 0 ? 1 : 2;
            const double apq = a[p][q];
            if (apq == 0.0) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
            const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
            a[p][p] -= t * apq;
            a[q][q] += t * apq;
            a[p][q] = a[q][p] = 0.0;
            const int r = 3 - p - q;
            const double arp = a[r][p], arq = a[r][q];
            a[r][p] = a[p][r] = c * arp - s * arq;
            a[r][q] = a[q][r] = s * arp + c * arq;
            for (int i = 0; i < 3; ++i) {
                const double vp = V[i][p], vq = V[i][q];
                V[i][p] = c * vp - s * vq;
                V[i][q] = s * vp + c * vq;
            }
        }
    }
    int ord[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (a[ord[j]][ord[j]] < a[ord[i]][ord[i]]) {
                const int tmp = ord[i];
                ord[i] = ord[j];
                ord[j] = tmp;
            }
    double W[3][3];
    for (int k = 0; k < 3; ++k) {
        lam[k] = a[ord[k]][ord[k]];
        for (int i = 0; i < 3; ++i) W[i][k] = V[i][ord[k]];
    }
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) V[i][k] = W[i][k];
}

// The reference's pick (calc_stresses.cpp:165-176): drop the first eigenvalue
// of least magnitude, keep the next two in cyclic order as (eigval_1,
// eigvec_1, eigval_2, eigvec_2).  An eigenvector's sign is not fixed by the
// reference (its eigensolver's is arbitrary); here its component of largest
// magnitude (the first such) is positive.
__device__ __forceinline__ void stress_pack(const double lam[3], const double V[3][3], double out[8]) {
    int m = 0;
    for (int j = 0; j < 3; ++j)
        if (fabs(lam[m]) > fabs(lam[j])) m = j;
    for (int r = 0; r < 2; ++r) {
        const int k = (m + 1 + r) % 3;
        int big = 0;
        for (int i = 1; i < 3; ++i)
            if (fabs(V[i][k]) > fabs(V[big][k])) big = i;
        const double sgn = V[big][k] < 0.0 ? -1.0 : 1.0;
        out[4 * r] = lam[k];
        for (int i = 0; i < 3; ++i) out[4 * r + 1 + i] = sgn * V[i][k];
    }
}

}  // namespace ms
