/*
 * oracle.c -- FP64 CPU restatement of MorphoShell's per-step
 * force-and-dynamics path.  TEST INFRASTRUCTURE ONLY: it is the checker the
 * parity tests and the CPU-baseline leg of bench.py compare against; no
 * product code links or calls it.
 *
 * It follows the reference's dataflow, not the GPU design:
 *   q = M . dofs          (calc_surface_derivs.cpp:27-34; M built as a CSR
 *                          matrix of 72 non-zeros per triangle exactly as
 *                          calc_deriv_mats.cpp:64-119 lays it out)
 *   g = d(A_t W_t)/dq_t   (calc_deformation_forces.cpp:53-176, restated from
 *                          the energy by the chain rule, NOT from the
 *                          generated expressions)
 *   F = -M^T g            (calc_deformation_forces.cpp:185, with an explicit
 *                          row-major transpose summed in ascending q order)
 *   damping, clamp, blow-up check, symplectic Euler
 *                         (calc_non_deformation_forces.cpp:44-65,
 *                          impose_constraints.cpp:56-68,
 *                          advance_dynamics.cpp:58-93)
 *   equilibrium max norms (is_at_equilibrium.cpp:58-98)
 *   energies, serial Kahan (calc_energy_densities.cpp:46-82,
 *                          calc_energies.cpp:44-64, kahan_sum.cpp:32-49)
 * plus its own restatement of the patch choice and of G/H
 * (set_up_patch_fitting.cpp:57-297, calc_deriv_mats.cpp:70-91), written
 * independently of the product's host code (Gauss-Jordan with partial
 * pivoting instead of a full-pivot LU, power iteration instead of Jacobi).
 *
 * Parity pins (tests/test_oracle_*.py): G exact on affine maps and H exact on
 * quadratic maps (moment identities), a/b/n and W_s/W_b closed forms
 * (flat, uniform stretch, quadratic graph, isotropic abar), forces against
 * central finite differences of the total energy, translation/rotation
 * invariance, zero force at the programmed state, Kahan against math.fsum,
 * the symplectic-Euler trajectory of a single oscillator, scalar formulas on
 * an equilateral mesh.  No function is left "parity unpinned".
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------- utilities */

/* kahan_sum.cpp:32-49, serial, in index order. */
EXPORT double orc_kahan_sum(const double* v, int64_t n) {
    volatile double sum = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double y = v[i] - c;
        double t = sum + y;
        volatile double d = t - sum;
        c = d - y;
        sum = t;
    }
    return sum;
}

/* ---------------------------------------------------------- patch fitting */

/* Gauss-Jordan inverse with partial pivoting of an n x n row-major matrix.
 * Returns 0 if a pivot is below 1e-14 of the largest |entry| (singular). */
static int gj_inverse(int n, const double* A_in, double* inv) {
    double A[36], B[36];
    memcpy(A, A_in, sizeof(double) * n * n);
    double amax = 0.0;
    for (int i = 0; i < n * n; ++i) amax = fmax(amax, fabs(A[i]));
    if (amax == 0.0) return 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) B[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
        if (!(fabs(A[p * n + k]) > 1e-14 * amax)) return 0;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t;
                t = B[k * n + j]; B[k * n + j] = B[p * n + j]; B[p * n + j] = t;
            }
        double d = A[k * n + k];
        for (int j = 0; j < n; ++j) { A[k * n + j] /= d; B[k * n + j] /= d; }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            double l = A[i * n + k];
            if (l == 0.0) continue;
            for (int j = 0; j < n; ++j) { A[i * n + j] -= l * A[k * n + j]; B[i * n + j] -= l * B[k * n + j]; }
        }
    }
    memcpy(inv, B, sizeof(double) * n * n);
    return 1;
}

/* Largest singular value of a 6x3 row-major matrix by power iteration on
 * the Gram matrix (converged to ~1e-15 relative). */
static double sigma_max(const double* H) {
    double S[9] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int n = 0; n < 6; ++n) S[i * 3 + j] += H[n * 3 + i] * H[n * 3 + j];
    double x[3] = {1.0, 0.7071067811865476, 0.5773502691896258}, lam = 0.0;
    for (int it = 0; it < 2000; ++it) {
        double y[3];
        for (int i = 0; i < 3; ++i) y[i] = S[i * 3] * x[0] + S[i * 3 + 1] * x[1] + S[i * 3 + 2] * x[2];
        double nrm = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        if (nrm == 0.0) return 0.0;
        for (int i = 0; i < 3; ++i) x[i] = y[i] / nrm;
        if (fabs(nrm - lam) <= 1e-16 * nrm && it > 20) { lam = nrm; break; }
        lam = nrm;
    }
    return sqrt(lam);
}

/* Restatement of set_up_patch_fitting.cpp:57-297 and calc_deriv_mats.cpp:70-91.
 * Serial.  Returns -1 on success, else the first triangle whose search was
 * exhausted.  Ties in the distance ordering keep first-seen order. */
EXPORT int64_t orc_patch_fit(int32_t N, int32_t T, const double* pos, const int32_t* tv, double thresh,
                             int32_t* pid, double* G, double* H) {
    const double *X = pos, *Y = pos + N, *Z = pos + 2 * (int64_t)N;
    /* vertex incidence, ascending triangle id */
    int64_t* off = calloc((size_t)N + 1, sizeof(int64_t));
    int32_t* inc = malloc(sizeof(int32_t) * 3 * (size_t)T);
    for (int64_t i = 0; i < 3LL * T; ++i) off[tv[i] + 1]++;
    for (int32_t n = 0; n < N; ++n) off[n + 1] += off[n];
    int64_t* cur = malloc(sizeof(int64_t) * (size_t)N);
    for (int32_t n = 0; n < N; ++n) cur[n] = off[n];
    for (int32_t t = 0; t < T; ++t)
        for (int v = 0; v < 3; ++v) inc[cur[tv[3 * (int64_t)t + v]]++] = t;
    free(cur);
    int64_t failed = -1;
    int32_t cand[512];
    double dist[512];
    int32_t ord[512];
    for (int32_t t = 0; t < T && failed < 0; ++t) {
        const int32_t* v = tv + 3 * (int64_t)t;
        double c[3] = {(X[v[0]] + X[v[1]] + X[v[2]]) / 3.0, (Y[v[0]] + Y[v[1]] + Y[v[2]]) / 3.0,
                       (Z[v[0]] + Z[v[1]] + Z[v[2]]) / 3.0};
        int nc = 0;
        for (int a = 0; a < 3; ++a)
            for (int64_t k = off[v[a]]; k < off[v[a] + 1]; ++k)
                for (int w = 0; w < 3; ++w) {
                    int32_t m = tv[3 * (int64_t)inc[k] + w];
                    int dup = (m == v[0] || m == v[1] || m == v[2]);
                    for (int u = 0; u < nc && !dup; ++u) dup = (cand[u] == m);
                    if (dup || nc >= 512) continue;
                    cand[nc] = m;
                    double dx = X[m] - c[0], dy = Y[m] - c[1], dz = Z[m] - c[2];
                    dist[nc] = sqrt(dx * dx + dy * dy + dz * dz);
                    nc++;
                }
        /* insertion sort by distance: stable */
        for (int i = 0; i < nc; ++i) ord[i] = i;
        for (int i = 1; i < nc; ++i) {
            int x = ord[i], j = i - 1;
            while (j >= 0 && dist[ord[j]] > dist[x]) { ord[j + 1] = ord[j]; --j; }
            ord[j + 1] = x;
        }
        int ok = 0;
        int32_t p3 = -1, p4 = -1, p5 = -1;
        double Hc[18];
        if (nc >= 3) {
            p3 = cand[ord[0]];
            for (int p = 1; p < (nc < 4 ? nc : 4) && !ok; ++p) {
                p4 = cand[ord[p]];
                for (int q = 2; q < nc; ++q) {
                    p5 = cand[ord[q]];
                    int32_t ids[6] = {v[0], v[1], v[2], p3, p4, p5};
                    double size2 = 0.0;
                    for (int n = 0; n < 6; ++n) {
                        double dx = X[ids[n]] - c[0], dy = Y[ids[n]] - c[1], dz = Z[ids[n]] - c[2];
                        size2 += dx * dx + dy * dy + dz * dz;
                    }
                    size2 /= 6.0;
                    if (p4 == p5) continue;
                    double M6[36], inv[36];
                    for (int n = 0; n < 6; ++n) {
                        double dx = X[ids[n]] - c[0], dy = Y[ids[n]] - c[1];
                        M6[0 * 6 + n] = 1.0;
                        M6[1 * 6 + n] = dx;
                        M6[2 * 6 + n] = dy;
                        M6[3 * 6 + n] = 0.5 * dx * dx;
                        M6[4 * 6 + n] = dx * dy;
                        M6[5 * 6 + n] = 0.5 * dy * dy;
                    }
                    if (!gj_inverse(6, M6, inv)) continue;
                    for (int n = 0; n < 6; ++n)
                        for (int k = 0; k < 3; ++k) Hc[n * 3 + k] = inv[n * 6 + 3 + k];
                    if (!(sigma_max(Hc) * size2 / thresh < 1.0)) continue;
                    ok = 1;
                    break;
                }
            }
        }
        if (!ok) { failed = t; break; }
        int32_t* P = pid + 6 * (int64_t)t;
        P[0] = v[0]; P[1] = v[1]; P[2] = v[2]; P[3] = p3; P[4] = p4; P[5] = p5;
        memcpy(H + 18 * (int64_t)t, Hc, sizeof(Hc));
        double T3[9], I3[9];
        for (int i = 0; i < 3; ++i) {
            T3[i * 3 + 0] = 1.0;
            T3[i * 3 + 1] = X[v[i]] - c[0];
            T3[i * 3 + 2] = Y[v[i]] - c[1];
        }
        if (!gj_inverse(3, T3, I3)) { failed = t; break; }
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < 3; ++i) G[6 * (int64_t)t + 3 * k + i] = I3[(1 + k) * 3 + i];
    }
    free(off);
    free(inc);
    return failed;
}

/* ------------------------------------------------------------ the system */

typedef struct orc_sys {
    int32_t N, T;
    /* M: 15T x 3N CSR; Mt: 3N x 15T CSR */
    int64_t *M_off, *Mt_off;
    int32_t *M_col, *Mt_col;
    double *M_val, *Mt_val;
    double *area, *h, *mu, *abar, *bbar, *mass; /* abar/bbar 3T SoA */
    uint8_t* constrained;
    double nu, gamma, rho, hmin, dt, eta, fchar, cchar;
    int gradient_descent;
    double *q, *g; /* scratch 15T */
} orc_sys;

EXPORT void orc_destroy(orc_sys* s) {
    if (!s) return;
    free(s->M_off); free(s->Mt_off); free(s->M_col); free(s->Mt_col); free(s->M_val); free(s->Mt_val);
    free(s->area); free(s->h); free(s->mu); free(s->abar); free(s->bbar); free(s->mass);
    free(s->constrained); free(s->q); free(s->g);
    free(s);
}

static double* dupd(const double* p, size_t n) {
    double* r = malloc(sizeof(double) * n);
    memcpy(r, p, sizeof(double) * n);
    return r;
}

/* Build M exactly as calc_deriv_mats.cpp:64-119 lays out its triplets:
 * row 15t + 5c + {0,1} = first derivatives over the 3 vertices (G),
 * rows 15t + 5c + {2,3,4} = second derivatives over the 6 patch nodes (H),
 * column = node + c*N.  Entries of a row are sorted by column. */
EXPORT orc_sys* orc_create(int32_t N, int32_t T, const int32_t* pid, const double* G, const double* H,
                           const double* area, const double* h, const double* mu, const double* abar,
                           const double* bbar, const double* mass, const uint8_t* constrained, double nu,
                           double gamma, double rho, double hmin, double dt, double eta, double fchar,
                           double cchar, int gradient_descent) {
    orc_sys* s = calloc(1, sizeof(orc_sys));
    s->N = N; s->T = T;
    s->nu = nu; s->gamma = gamma; s->rho = rho; s->hmin = hmin; s->dt = dt; s->eta = eta;
    s->fchar = fchar; s->cchar = cchar; s->gradient_descent = gradient_descent;
    s->area = dupd(area, T); s->h = dupd(h, T); s->mu = dupd(mu, T);
    s->abar = dupd(abar, 3 * (size_t)T); s->bbar = dupd(bbar, 3 * (size_t)T);
    s->mass = dupd(mass, N);
    s->constrained = calloc((size_t)N, 1);
    if (constrained) memcpy(s->constrained, constrained, (size_t)N);
    s->q = malloc(sizeof(double) * 15 * (size_t)T);
    s->g = malloc(sizeof(double) * 15 * (size_t)T);
    int64_t rows = 15LL * T;
    s->M_off = malloc(sizeof(int64_t) * (size_t)(rows + 1));
    s->M_col = malloc(sizeof(int32_t) * 72 * (size_t)T);
    s->M_val = malloc(sizeof(double) * 72 * (size_t)T);
    int64_t nz = 0;
    for (int32_t t = 0; t < T; ++t) {
        const int32_t* P = pid + 6 * (int64_t)t;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 5; ++r) {
                int64_t row = 15LL * t + 5 * c + r;
                s->M_off[row] = nz;
                int cnt = (r < 2) ? 3 : 6;
                int32_t cols[6];
                double vals[6];
                for (int n = 0; n < cnt; ++n) {
                    cols[n] = P[n] + c * N;
                    vals[n] = (r < 2) ? G[6 * (int64_t)t + 3 * r + n] : H[18 * (int64_t)t + 3 * n + (r - 2)];
                }
                /* sort by column */
                for (int i = 1; i < cnt; ++i) {
                    int32_t cc = cols[i]; double vv = vals[i]; int j = i - 1;
                    while (j >= 0 && cols[j] > cc) { cols[j + 1] = cols[j]; vals[j + 1] = vals[j]; --j; }
                    cols[j + 1] = cc; vals[j + 1] = vv;
                }
                for (int n = 0; n < cnt; ++n) { s->M_col[nz] = cols[n]; s->M_val[nz] = vals[n]; ++nz; }
            }
    }
    s->M_off[rows] = nz;
    /* explicit transpose (main.cpp:168-174), rows in ascending q order */
    int64_t dofs = 3LL * N;
    s->Mt_off = calloc((size_t)dofs + 1, sizeof(int64_t));
    s->Mt_col = malloc(sizeof(int32_t) * (size_t)nz);
    s->Mt_val = malloc(sizeof(double) * (size_t)nz);
    for (int64_t k = 0; k < nz; ++k) s->Mt_off[s->M_col[k] + 1]++;
    for (int64_t d = 0; d < dofs; ++d) s->Mt_off[d + 1] += s->Mt_off[d];
    int64_t* cur = malloc(sizeof(int64_t) * (size_t)dofs);
    memcpy(cur, s->Mt_off, sizeof(int64_t) * (size_t)dofs);
    for (int64_t row = 0; row < rows; ++row)
        for (int64_t k = s->M_off[row]; k < s->M_off[row + 1]; ++k) {
            int64_t d = s->M_col[k];
            s->Mt_col[cur[d]] = (int32_t)row;
            s->Mt_val[cur[d]] = s->M_val[k];
            cur[d]++;
        }
    free(cur);
    return s;
}

EXPORT void orc_set_targets(orc_sys* s, const double* abar, const double* bbar, const double* h, const double* mu) {
    if (abar) memcpy(s->abar, abar, sizeof(double) * 3 * (size_t)s->T);
    if (bbar) memcpy(s->bbar, bbar, sizeof(double) * 3 * (size_t)s->T);
    if (h) memcpy(s->h, h, sizeof(double) * (size_t)s->T);
    if (mu) memcpy(s->mu, mu, sizeof(double) * (size_t)s->T);
}

/* q = M . dofs (calc_surface_derivs.cpp:33). */
EXPORT void orc_surface_derivs(const orc_sys* s, const double* dofs, double* q) {
    int64_t rows = 15LL * s->T;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; ++r) {
        double acc = 0.0;
        for (int64_t k = s->M_off[r]; k < s->M_off[r + 1]; ++k) acc += s->M_val[k] * dofs[s->M_col[k]];
        q[r] = acc;
    }
}

/* --------------------------------------------------------- the mechanics
 * Symmetric 2x2 tensors are stored as (xx, xy, yy).  With P = adj(abar),
 * det = det(abar) and the quadratic form
 *     Phi(D) = nu tr(P D)^2 + (1 - nu) tr((P D)^2)
 * the reference's energy densities per reference area are
 *     W_s  = h mu / (1 - nu) det^-3/2 Phi(a - abar)
 *     W_b0 = mu h^3 / (12 (1 - nu)) det^-3/2 Phi(b - bbar)
 *     W_b  = W_b0 (1 + W_b0 / (gamma mu h))
 * (calc_energy_densities.cpp:62-77: prefactors 2h mu/(1-nu) det^-3/2 and
 * mu h^3/(6(1-nu)) det^-3/2, each times one half).
 */
typedef struct { double s, b0, b, dPhia[3], dPhib[3]; } tri_energy;

static void phi_and_grad(double nu, const double P[3], const double D[3], double* phi, double grad[3]) {
    /* P D (general 2x2) */
    double m00 = P[0] * D[0] + P[1] * D[1], m01 = P[0] * D[1] + P[1] * D[2];
    double m10 = P[1] * D[0] + P[2] * D[1], m11 = P[1] * D[1] + P[2] * D[2];
    double tr = m00 + m11;
    double tr2 = m00 * m00 + 2.0 * m01 * m10 + m11 * m11;
    *phi = nu * tr * tr + (1.0 - nu) * tr2;
    if (grad) {
        /* dPhi/dD = 2 nu tr P + 2 (1 - nu) P D P ; stored-component gradient
         * (D00, 2 D01, D11) since D01 appears twice in the tensor */
        double pdp00 = m00 * P[0] + m01 * P[1];
        double pdp01 = m00 * P[1] + m01 * P[2];
        double pdp11 = m10 * P[1] + m11 * P[2];
        grad[0] = 2.0 * nu * tr * P[0] + 2.0 * (1.0 - nu) * pdp00;
        grad[1] = 2.0 * (2.0 * nu * tr * P[1] + 2.0 * (1.0 - nu) * pdp01);
        grad[2] = 2.0 * nu * tr * P[2] + 2.0 * (1.0 - nu) * pdp11;
    }
}

/* calc_a_comps_and_b_comps_and_normals.cpp:49-60 for one triangle. */
static void geometry_one(const double* q, double a[3], double b[3], double n[3], double* mnorm) {
    double Rx[3] = {q[0], q[5], q[10]}, Ry[3] = {q[1], q[6], q[11]};
    double Rxx[3] = {q[2], q[7], q[12]}, Rxy[3] = {q[3], q[8], q[13]}, Ryy[3] = {q[4], q[9], q[14]};
    a[0] = Rx[0] * Rx[0] + Rx[1] * Rx[1] + Rx[2] * Rx[2];
    a[1] = Rx[0] * Ry[0] + Rx[1] * Ry[1] + Rx[2] * Ry[2];
    a[2] = Ry[0] * Ry[0] + Ry[1] * Ry[1] + Ry[2] * Ry[2];
    double m[3] = {Rx[1] * Ry[2] - Rx[2] * Ry[1], Rx[2] * Ry[0] - Rx[0] * Ry[2], Rx[0] * Ry[1] - Rx[1] * Ry[0]};
    double mn = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    for (int i = 0; i < 3; ++i) n[i] = m[i] / mn;
    b[0] = -(n[0] * Rxx[0] + n[1] * Rxx[1] + n[2] * Rxx[2]);
    b[1] = -(n[0] * Rxy[0] + n[1] * Rxy[1] + n[2] * Rxy[2]);
    b[2] = -(n[0] * Ryy[0] + n[1] * Ryy[1] + n[2] * Ryy[2]);
    if (mnorm) *mnorm = mn;
}

static void energy_one(const orc_sys* s, int32_t t, const double a[3], const double b[3], tri_energy* e,
                       int want_grad) {
    int32_t T = s->T;
    double ab[3] = {s->abar[t], s->abar[T + t], s->abar[2 * (int64_t)T + t]};
    double bb[3] = {s->bbar[t], s->bbar[T + t], s->bbar[2 * (int64_t)T + t]};
    double P[3] = {ab[2], -ab[1], ab[0]};
    double det = ab[0] * ab[2] - ab[1] * ab[1];
    double idet32 = 1.0 / (det * sqrt(det));
    double h = s->h[t], mu = s->mu[t], nu = s->nu;
    double cs = h * mu / (1.0 - nu) * idet32;
    double cb = mu * h * h * h / (12.0 * (1.0 - nu)) * idet32;
    double Da[3] = {a[0] - ab[0], a[1] - ab[1], a[2] - ab[2]};
    double Db[3] = {b[0] - bb[0], b[1] - bb[1], b[2] - bb[2]};
    double pa, pb;
    phi_and_grad(nu, P, Da, &pa, want_grad ? e->dPhia : NULL);
    phi_and_grad(nu, P, Db, &pb, want_grad ? e->dPhib : NULL);
    e->s = cs * pa;
    e->b0 = cb * pb;
    double scale = s->gamma * mu * h; /* bend-stiffening energy-density scale */
    e->b = e->b0 * (1.0 + e->b0 / scale);
    if (want_grad) {
        double stiff = 1.0 + 2.0 * e->b0 / scale; /* calc_deformation_forces.cpp:141 */
        for (int k = 0; k < 3; ++k) {
            e->dPhia[k] *= cs;          /* now dW_s/da_k */
            e->dPhib[k] *= cb * stiff;  /* now dW_b/db_k */
        }
    }
}

/* g_t = d(A_t (W_s + W_b))/dq_t, 15 values in q order. */
static void dE_dq_one(const orc_sys* s, int32_t t, const double* q, double* g) {
    double a[3], b[3], n[3], mn;
    geometry_one(q, a, b, n, &mn);
    tri_energy e;
    energy_one(s, t, a, b, &e, 1);
    const double* Wa = e.dPhia;
    const double* Wb = e.dPhib;
    double Rx[3] = {q[0], q[5], q[10]}, Ry[3] = {q[1], q[6], q[11]};
    double Rxx[3] = {q[2], q[7], q[12]}, Rxy[3] = {q[3], q[8], q[13]}, Ryy[3] = {q[4], q[9], q[14]};
    /* bending depends on R_x, R_y through n = m/|m|, m = R_x x R_y:
     * dW/dn = w = -(Wb0 Rxx + Wb1 Rxy + Wb2 Ryy); dW/dm = (I - n n^T) w / |m|;
     * dW/dR_x = R_y x u, dW/dR_y = u x R_x */
    double w[3], u[3];
    for (int i = 0; i < 3; ++i) w[i] = -(Wb[0] * Rxx[i] + Wb[1] * Rxy[i] + Wb[2] * Ryy[i]);
    double nw = n[0] * w[0] + n[1] * w[1] + n[2] * w[2];
    for (int i = 0; i < 3; ++i) u[i] = (w[i] - nw * n[i]) / mn;
    double gx[3] = {Ry[1] * u[2] - Ry[2] * u[1], Ry[2] * u[0] - Ry[0] * u[2], Ry[0] * u[1] - Ry[1] * u[0]};
    double gy[3] = {u[1] * Rx[2] - u[2] * Rx[1], u[2] * Rx[0] - u[0] * Rx[2], u[0] * Rx[1] - u[1] * Rx[0]};
    double A = s->area[t];
    for (int c = 0; c < 3; ++c) {
        g[5 * c + 0] = A * (2.0 * Wa[0] * Rx[c] + Wa[1] * Ry[c] + gx[c]);
        g[5 * c + 1] = A * (Wa[1] * Rx[c] + 2.0 * Wa[2] * Ry[c] + gy[c]);
        g[5 * c + 2] = -A * Wb[0] * n[c];
        g[5 * c + 3] = -A * Wb[1] * n[c];
        g[5 * c + 4] = -A * Wb[2] * n[c];
    }
}

EXPORT void orc_dE_dq(const orc_sys* s, const double* q, double* g) {
#pragma omp parallel for schedule(static)
    for (int32_t t = 0; t < s->T; ++t) dE_dq_one(s, t, q + 15 * (int64_t)t, g + 15 * (int64_t)t);
}

/* forces = -M^T g (calc_deformation_forces.cpp:185); ascending q per row. */
EXPORT void orc_deformation_forces(orc_sys* s, const double* dofs, double* forces) {
    orc_surface_derivs(s, dofs, s->q);
    orc_dE_dq(s, s->q, s->g);
    int64_t D = 3LL * s->N;
#pragma omp parallel for schedule(static)
    for (int64_t d = 0; d < D; ++d) {
        double acc = 0.0;
        for (int64_t k = s->Mt_off[d]; k < s->Mt_off[d + 1]; ++k) acc += s->Mt_val[k] * s->g[s->Mt_col[k]];
        forces[d] = -acc;
    }
}

/* a, b, n as 3T SoA (any may be NULL). */
EXPORT void orc_geometry(orc_sys* s, const double* dofs, double* a, double* b, double* n) {
    orc_surface_derivs(s, dofs, s->q);
    int32_t T = s->T;
#pragma omp parallel for schedule(static)
    for (int32_t t = 0; t < T; ++t) {
        double aa[3], bb[3], nn[3];
        geometry_one(s->q + 15 * (int64_t)t, aa, bb, nn, NULL);
        for (int k = 0; k < 3; ++k) {
            if (a) a[(int64_t)k * T + t] = aa[k];
            if (b) b[(int64_t)k * T + t] = bb[k];
            if (n) n[(int64_t)k * T + t] = nn[k];
        }
    }
}

/* Energy densities (T x 2 SoA: stretch then bend, calc_energy_densities.cpp). */
EXPORT void orc_energy_densities(orc_sys* s, const double* dofs, double* dens) {
    orc_surface_derivs(s, dofs, s->q);
    int32_t T = s->T;
#pragma omp parallel for schedule(static)
    for (int32_t t = 0; t < T; ++t) {
        double a[3], b[3], n[3];
        geometry_one(s->q + 15 * (int64_t)t, a, b, n, NULL);
        tri_energy e;
        energy_one(s, t, a, b, &e, 0);
        dens[t] = e.s;
        dens[T + t] = e.b;
    }
}

/* calc_energies.cpp:44-64: serial Kahan sums of A W_s, A W_b, 1/2 m |v|^2. */
EXPORT void orc_energies(orc_sys* s, const double* dofs, const double* vel, double out[3]) {
    int32_t T = s->T, N = s->N;
    double* dens = malloc(sizeof(double) * 2 * (size_t)T);
    double* tmp = malloc(sizeof(double) * (size_t)(T > N ? T : N));
    orc_energy_densities(s, dofs, dens);
    for (int32_t t = 0; t < T; ++t) tmp[t] = s->area[t] * dens[t];
    out[0] = orc_kahan_sum(tmp, T);
    for (int32_t t = 0; t < T; ++t) tmp[t] = s->area[t] * dens[T + t];
    out[1] = orc_kahan_sum(tmp, T);
    for (int32_t n = 0; n < N; ++n) {
        double vx = vel[n], vy = vel[N + n], vz = vel[2 * (int64_t)N + n];
        tmp[n] = 0.5 * s->mass[n] * (vx * vx + vy * vy + vz * vz);
    }
    out[2] = orc_kahan_sum(tmp, N);
    free(dens);
    free(tmp);
}

/* Damping (calc_non_deformation_forces.cpp:44-65), Newtonian mode only. */
EXPORT void orc_damping(const orc_sys* s, const double* vel, double* forces) {
    if (s->gradient_descent) return;
    int32_t N = s->N;
    double den = s->rho * s->hmin * s->hmin;
    for (int64_t d = 0; d < 3LL * N; ++d) forces[d] += -s->eta * s->mass[d % N] * vel[d] / den;
}

/* impose_constraints.cpp:56-68 (full clamp). */
EXPORT void orc_constraints(const orc_sys* s, double* forces) {
    int32_t N = s->N;
    for (int32_t n = 0; n < N; ++n)
        if (s->constrained[n]) forces[n] = forces[N + n] = forces[2 * (int64_t)N + n] = 0.0;
}

/* advance_dynamics.cpp:58-93.  Returns the first offending dof (and its force
 * in *bad) without updating anything, or -1 after the update. */
EXPORT int64_t orc_advance(const orc_sys* s, double* dofs, double* vel, const double* forces, double* bad) {
    int32_t N = s->N;
    int64_t D = 3LL * N;
    double lim = 50000.0 * s->fchar;
    for (int64_t d = 0; d < D; ++d)
        if (!(fabs(forces[d]) < lim)) {
            if (bad) *bad = forces[d];
            return d;
        }
    for (int64_t d = 0; d < D; ++d) {
        double m = s->mass[d % N];
        vel[d] += (s->dt / m) * forces[d];
        if (s->gradient_descent) vel[d] = s->rho * s->hmin * s->hmin * forces[d] / (s->eta * m);
        dofs[d] += s->dt * vel[d];
    }
    return -1;
}

/* is_at_equilibrium.cpp:58-98: max |F_n| / F_char and max |v_n| / c_char,
 * first argmax (std::max_element). */
EXPORT void orc_equilibrium(const orc_sys* s, const double* forces, const double* vel, double* fratio,
                            int32_t* fnode, double* vratio, int32_t* vnode) {
    int32_t N = s->N;
    double fm = -1.0, vm = -1.0;
    int32_t fi = 0, vi = 0;
    for (int32_t n = 0; n < N; ++n) {
        double fx = forces[n], fy = forces[N + n], fz = forces[2 * (int64_t)N + n];
        double vx = vel[n], vy = vel[N + n], vz = vel[2 * (int64_t)N + n];
        double f = sqrt(fx * fx + fy * fy + fz * fz), v = sqrt(vx * vx + vy * vy + vz * vz);
        if (f > fm) { fm = f; fi = n; }
        if (v > vm) { vm = v; vi = n; }
    }
    *fratio = fm / s->fchar;
    *fnode = fi;
    *vratio = vm / s->cchar;
    *vnode = vi;
}

/* nsteps of the reference step (forces, damping, clamp, advance).  Returns -1
 * or the (step << 32 | dof) of the first blow-up.  forces is 3N scratch. */
EXPORT int64_t orc_run_steps(orc_sys* s, double* dofs, double* vel, double* forces, int64_t nsteps) {
    for (int64_t k = 0; k < nsteps; ++k) {
        orc_deformation_forces(s, dofs, forces);
        orc_damping(s, vel, forces);
        orc_constraints(s, forces);
        double bad;
        int64_t d = orc_advance(s, dofs, vel, forces, &bad);
        if (d >= 0) return (k << 32) | d;
    }
    return -1;
}

/* Dialling restated (do_dialling.cpp:83-84 protocol 1; 104-124 protocol 3;
 * 194-228 dial-from-ansatz blend).  info arrays are T x cols row-major;
 * a0/b0 3T SoA or NULL (= (1,0,1) and 0).  Outputs 3T SoA. */
EXPORT void orc_dial(int32_t T, int protocol, int from_ansatz, double f, const double* abar_info, int acols,
                     const double* bbar_info, int bcols, const double* a0, const double* b0, double* abar,
                     double* bbar) {
    double fin = from_ansatz ? 1.0 : f;
    for (int32_t t = 0; t < T; ++t) {
        double A0[3] = {1.0, 0.0, 1.0}, B0[3] = {0.0, 0.0, 0.0};
        if (a0) for (int k = 0; k < 3; ++k) A0[k] = a0[(int64_t)k * T + t];
        if (b0) for (int k = 0; k < 3; ++k) B0[k] = b0[(int64_t)k * T + t];
        double ab[3], bb[3];
        if (protocol == 3) {
            double ang = abar_info[(int64_t)t * acols + 0];
            double lam = (1.0 - fin) * 1.0 + fin * abar_info[(int64_t)t * acols + 1];
            double nu_ot = abar_info[(int64_t)t * acols + 2];
            double c = cos(ang), sn = sin(ang), l2 = lam * lam, lp = pow(lam, -2.0 * nu_ot);
            ab[0] = l2 * c * c + lp * sn * sn;
            ab[1] = (l2 - lp) * c * sn;
            ab[2] = lp * c * c + l2 * sn * sn;
            bb[0] = bb[1] = bb[2] = 0.0;
        } else {
            /* protocol 1 uses the zero-dial-factor a, b as its start */
            for (int k = 0; k < 3; ++k) {
                ab[k] = (1.0 - fin) * A0[k] + fin * abar_info[(int64_t)t * acols + k];
                bb[k] = (1.0 - fin) * B0[k] + fin * (bbar_info ? bbar_info[(int64_t)t * bcols + k] : 0.0);
            }
        }
        if (from_ansatz) {
            for (int k = 0; k < 3; ++k) {
                ab[k] = (1.0 - f) * A0[k] + f * ab[k];
                bb[k] = (1.0 - f) * B0[k] + f * bb[k];
            }
        }
        for (int k = 0; k < 3; ++k) {
            abar[(int64_t)k * T + t] = ab[k];
            bbar[(int64_t)k * T + t] = bb[k];
        }
    }
}

EXPORT int orc_num_threads(void) { return omp_get_max_threads(); }
EXPORT void orc_set_num_threads(int n) { omp_set_num_threads(n); }

/* Energy densities and dW/dq of one triangle from its 15 surface derivatives
 * (q order Xx, Xy, Xxx, Xxy, Xyy, Yx, ...): for closed-form tests.
 * out[0] = W_s, out[1] = W_b (per reference area); g (15, may be NULL) =
 * dW/dq with unit area. */
EXPORT void orc_tri_energy_q(const double* q, const double* abar, const double* bbar, double h, double mu, double nu,
                             double gamma, double* out, double* g) {
    orc_sys s;
    memset(&s, 0, sizeof(s));
    double A = 1.0;
    s.T = 1;
    s.abar = (double*)abar;
    s.bbar = (double*)bbar;
    s.h = &h;
    s.mu = &mu;
    s.area = &A;
    s.nu = nu;
    s.gamma = gamma;
    double a[3], b[3], n[3];
    geometry_one(q, a, b, n, NULL);
    tri_energy e;
    energy_one(&s, 0, a, b, &e, 0);
    out[0] = e.s;
    out[1] = e.b;
    if (g) dE_dq_one(&s, 0, q, g);
}

/* Change the damping factor (the reference recomputes it at phase switches,
 * main.cpp:301, 505). */
EXPORT void orc_set_eta(orc_sys* s, double eta) { s->eta = eta; }
