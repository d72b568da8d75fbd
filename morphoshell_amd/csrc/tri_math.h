// Per-triangle mechanics shared by the triangle launch (tri_kernels.hip) and
// the fused tile launch (tile_kernels.hip), so both produce the same bits.
//
// For triangle t: gather its six patch nodes, form R_x, R_y (G applied to
// the three vertices) and R_xx, R_xy, R_yy (H applied to the six patch
// nodes), both as weights of the edge vectors to patch node 0 (the weights
// of each derivative sum to zero; in FP64 G and the vertex rows of H come
// from the reference positions, see RecF64), the metric
// a, the normal n and the curvature b, the programmed
// abar, bbar at the shell's dial factor, and the 18 force contributions
// -A dW/dx of the six patch nodes.  The energy densities and gradient are
// the reference's (calc_energy_densities.cpp:62-77,
// calc_deformation_forces.cpp:93-171) restated through the quadratic form
// Phi(D) = nu tr(PD)^2 + (1-nu) tr((PD)^2), P = adj(abar):
//   W_s = h mu/(1-nu) det(abar)^-3/2 Phi(a - abar)
//   W_b = W_b0 (1 + W_b0/(gamma mu h)),  W_b0 = mu h^3/(12(1-nu)) det^-3/2 Phi(b - bbar)
//   dW/dR_x = 2 W_a0 R_x + W_a1 R_y + R_y x u,  dW/dR_y = W_a1 R_x + 2 W_a2 R_y + u x R_x
//   dW/dR_k'' = -W_bk n,  u = (I - n n^T) w / |R_x x R_y|,  w = -sum_k W_bk R_k''
// with W_ak = dW_s/da_k and W_bk = dW_b/db_k (stored-component gradients).
#pragma once

#include <math.h>

#include "kernels.h"

namespace ms {

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of chunks so neighbouring patches share an L2.
__device__ __forceinline__ int xcd_remap(int i, int n) {
    const int X = 8;
    const int q = n / X, r = n % X;
    const int j = i % X, local = i / X;
    return j * q + (j < r ? j : r) + local;
}

// ---- uniform values and loads ------------------------------------------------
// A block works on one shell (chunks never straddle shells), so everything
// read from the shell record, and every stream base, is the same for all
// lanes.  Broadcasting those values from lane 0 keeps them in scalar
// registers; loads then take a scalar base plus a 32-bit per-lane byte offset
// instead of a 64-bit vector address per stream.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename P>
__device__ __forceinline__ P* uni_ptr(P* p) {
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return (P*)(((unsigned long long)hi << 32) | lo);
}
// base[i] for a uniform base and i * sizeof(E) < 2^32.
template <typename E>
__device__ __forceinline__ E ld(const E* base, unsigned i) {
    return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(E));
}
// The same for data a step streams once (records, tile-local ids): a
// non-temporal load when NT, which leaves the caches to the node positions the
// neighbouring tiles re-read.  Measured on the 1.8M-triangle shell: the fused
// FP64 step 7% faster; the mixed step (3.8%) and both two-launch steps slower,
// so only the fused FP64 path sets it.
template <bool NT, typename E>
__device__ __forceinline__ E ld_stream(const E* base, unsigned i) {
    const E* p = reinterpret_cast<const E*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(E));
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// The shell quantities the per-triangle mechanics needs, block-uniform.
struct TriParams {
    double nu, gamma, f;  // Poisson ratio, stiffening scale, dial factor
    int protocol, from_ansatz, has_a0, has_b0;
    double h, mu, b[3];   // uniform h, mu, bbar_info (UNI kernels only)
    int lce_cached;       // fused step, LCE shell: the records' metric cache holds f's metric
};

// Factor the programmed metric is evaluated at: under the device-run protocol
// the last dialled one; otherwise the dial clock's.
__device__ __forceinline__ double dial_factor(const ShellDev& s) {
    return s.auto_protocol ? s.f_targets : shell_dial_factor(s);
}

// UNI: every shell of the context has uniform thickness, shear modulus and
// (blend protocol) bbar_info, which then come from the shell record; the
// per-triangle streams are not read.  A compile-time choice: a runtime switch
// keeps both the uniform values and the stream bases live.
template <bool UNI = false>
__device__ __forceinline__ TriParams tri_params(const ShellDev& s) {
    TriParams p;
    p.nu = uni(s.nu);
    p.gamma = uni(s.gamma);
    p.f = uni(dial_factor(s));
    p.protocol = uni(s.protocol);
    p.from_ansatz = uni(s.from_ansatz);
    p.has_a0 = uni(s.has_a0);
    p.has_b0 = uni(s.has_b0);
    p.lce_cached = 0;
    if (UNI) {
        p.h = uni(s.h_u);
        p.mu = uni(s.mu_u);
#pragma unroll
        for (int k = 0; k < 3; ++k) p.b[k] = uni(s.b_u[k]);
    }
    return p;
}

// The per-triangle dial data of triangle t: dial info (blend: abar_info and
// bbar_info; LCE: c2, cs, s2, lambda, -2nu) and the ansatz a0, b0 when
// dialling from it.  The force kernels load it next to the patch
// coefficients, ahead of every branch and use, so that a triangle's record
// costs one memory round trip.
struct DialRec {
    double ai[5], bi[3], A0[3], B0[3];
    // fused step, LCE shell without a valid cache: where to keep the dialled
    // metric (3 values at a stride of 64, the entry's bbar-info record slots)
    double* wb = nullptr;
};

// CS: component stride of the record streams when known at compile time (the
// tile launch's blocked records, 64), else 0 for td.T.
template <bool LCE, bool UNI, int CS = 0, bool NT = false>
__device__ __forceinline__ void load_dial(const TriData& td, const TriParams& s, unsigned t, DialRec& d) {
    const size_t T = CS ? (size_t)CS : (size_t)td.T;
    if (!(LCE && s.protocol == 3 && s.lce_cached)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) d.ai[k] = ld_stream<NT>(td.ainfo + k * T, t);
    }
    if (LCE && s.protocol == 3) {
        if (s.lce_cached) {
            // the metric this dial factor gives, kept by an earlier step
#pragma unroll
            for (int k = 0; k < 3; ++k) d.ai[k] = ld_stream<NT>(td.binfo + k * T, t);
        } else {
            d.ai[3] = ld_stream<NT>(td.ainfo + 3 * T, t);
            d.ai[4] = ld_stream<NT>(td.ainfo + 4 * T, t);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) d.bi[k] = UNI ? s.b[k] : ld_stream<NT>(td.binfo + k * T, t);
    }
    d.A0[0] = 1.0;
    d.A0[1] = 0.0;
    d.A0[2] = 1.0;
    d.B0[0] = d.B0[1] = d.B0[2] = 0.0;
    if (s.has_a0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) d.A0[k] = ld_stream<NT>(td.a0 + k * T, t);
    }
    if (s.has_b0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) d.B0[k] = ld_stream<NT>(td.b0 + k * T, t);
    }
}

// Programmed abar, bbar from the dial data (do_dialling.cpp:83-84 protocol 1,
// 104-124 protocol 3, 194-228 blend from the ansatz a, b).  LCE = false
// compiles only the blend protocol: the LCE branch's pow() costs the hot
// kernels some 60 registers, so contexts without LCE shells do without it.
template <bool LCE>
__device__ __forceinline__ void dial_targets(const DialRec& d, const TriParams& s, double ab[3], double bb[3]) {
#pragma clang fp contract(on)
    const double f = s.f;
    const double fin = s.from_ansatz ? 1.0 : f;
    if (LCE && s.protocol == 3) {
        if (s.lce_cached) {
#pragma unroll
            for (int k = 0; k < 3; ++k) ab[k] = d.ai[k];
        } else {
            const double c2 = d.ai[0], cs = d.ai[1], s2 = d.ai[2];
            const double lamf = d.ai[3], m2nu = d.ai[4];
            const double lam = (1.0 - fin) * 1.0 + fin * lamf;
            const double l2 = lam * lam, lp = pow(lam, m2nu);
            ab[0] = l2 * c2 + lp * s2;
            ab[1] = (l2 - lp) * cs;
            ab[2] = lp * c2 + l2 * s2;
            if (d.wb) {
#pragma unroll
                for (int k = 0; k < 3; ++k) d.wb[64 * k] = ab[k];
            }
        }
        bb[0] = bb[1] = bb[2] = 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ab[k] = (1.0 - fin) * d.A0[k] + fin * d.ai[k];
            bb[k] = (1.0 - fin) * d.B0[k] + fin * d.bi[k];
        }
    }
    if (s.from_ansatz) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ab[k] = (1.0 - f) * d.A0[k] + f * ab[k];
            bb[k] = (1.0 - f) * d.B0[k] + f * bb[k];
        }
    }
}

// Programmed abar, bbar of triangle t at the shell's dial factor.
template <bool LCE = true, bool UNI = false, int CS = 0>
__device__ __forceinline__ void targets(const TriData& td, const TriParams& s, unsigned t, double ab[3],
                                        double bb[3]) {
    DialRec d;
    load_dial<LCE, UNI, CS>(td, s, t, d);
    dial_targets<LCE>(d, s, ab, bb);
}

// Phi(D) and its stored-component gradient (dPhi/dD00, dPhi/dD01 counted
// twice, dPhi/dD11) for symmetric D and P given as (xx, xy, yy).
template <typename C>
__device__ __forceinline__ C phi_grad(C nu, const C P[3], const C D[3], C g[3]) {
#pragma clang fp contract(on)
    const C m00 = P[0] * D[0] + P[1] * D[1], m01 = P[0] * D[1] + P[1] * D[2];
    const C m10 = P[1] * D[0] + P[2] * D[1], m11 = P[1] * D[1] + P[2] * D[2];
    const C tr = m00 + m11;
    const C one_m_nu = C(1) - nu;
    const C q00 = m00 * P[0] + m01 * P[1];
    const C q01 = m00 * P[1] + m01 * P[2];
    const C q11 = m10 * P[1] + m11 * P[2];
    const C two_nu_tr = C(2) * nu * tr;
    g[0] = two_nu_tr * P[0] + C(2) * one_m_nu * q00;
    g[1] = C(2) * (two_nu_tr * P[1] + C(2) * one_m_nu * q01);
    g[2] = two_nu_tr * P[2] + C(2) * one_m_nu * q11;
    return nu * tr * tr + one_m_nu * (m00 * m00 + C(2) * m01 * m10 + m11 * m11);
}

template <typename C>
__device__ __forceinline__ C phi_only(C nu, const C P[3], const C D[3]) {
#pragma clang fp contract(on)
    const C m00 = P[0] * D[0] + P[1] * D[1], m01 = P[0] * D[1] + P[1] * D[2];
    const C m10 = P[1] * D[0] + P[2] * D[1], m11 = P[1] * D[1] + P[2] * D[2];
    const C tr = m00 + m11;
    return nu * tr * tr + (C(1) - nu) * (m00 * m00 + C(2) * m01 * m10 + m11 * m11);
}

// The six patch node ids of triangle t (loaded ahead of the mechanics, so a
// caller can fetch the next triangle's ids while this one computes).
__device__ __forceinline__ void load_patch_ids(const TriData& td, unsigned t, unsigned p[6]) {
    const size_t T = (size_t)td.T;
#pragma unroll
    for (int s = 0; s < 6; ++s) p[s] = (unsigned)ld(td.pid + s * T, t);
}

// Position fetchers: pos(s, x, y, z) gives patch node s of the triangle.
// GlobalPos reads the SoA state (3N, stride N) at global patch ids;
// TilePos reads the tile's LDS copy (x, y, z per tile-local node).
struct GlobalPos {
    const double* X;
    size_t N;
    const unsigned* p;
    __device__ __forceinline__ void operator()(int s, double& x, double& y, double& z) const {
        x = ld(X, p[s]);
        y = ld(X + N, p[s]);
        z = ld(X + 2 * N, p[s]);
    }
};
struct TilePos {
    const double* lds;        // 3 doubles per tile-local node
    const unsigned* p;        // tile-local ids
    __device__ __forceinline__ void operator()(int s, double& x, double& y, double& z) const {
        const double* q = lds + 3 * p[s];
        x = q[0];
        y = q[1];
        z = q[2];
    }
};
// Reference x, y of patch node s: GlobalRef from the 2N SoA stream,
// TileRef from the tile's LDS copy (2 doubles per tile-local node).
struct GlobalRef {
    const double* R;
    size_t N;
    const unsigned* p;
    __device__ __forceinline__ void operator()(int s, double& x, double& y) const {
        x = ld(R, p[s]);
        y = ld(R + N, p[s]);
    }
};
struct TileRef {
    const double* lds;
    const unsigned* p;
    __device__ __forceinline__ void operator()(int s, double& x, double& y) const {
        const double* q = lds + 2 * p[s];
        x = q[0];
        y = q[1];
    }
};

// The FP64 record of triangle t: the second-derivative weights of its three
// non-vertex patch nodes, dial data, h, mu and area.  Loaded whole before any
// use, so that a triangle's record costs one memory round trip (uniform
// branches, e.g. dial from ansatz, would otherwise split the loads into
// serial round trips), and so that a caller can have the next triangle's
// record in flight while it evaluates this one.
//
// The other 15 of the 24 patch coefficients follow from the reference
// positions, which the kernels stage per node.  G is rows 1, 2 of
// inv([1, x_v - c_x, y_v - c_y]) (calc_deriv_mats.cpp:70-91): in edge-vector
// form (weights of P_1 - P_0 and P_2 - P_0) it is the inverse of the 2x2
// matrix E of reference edge vectors (xi_v, eta_v) = (X_v - X_0, Y_v - Y_0),
// v = 1, 2.  H interpolates quadratics on the six patch nodes
// (set_up_patch_fitting.cpp:206-232), so each of its columns annihilates 1,
// X and Y: sum_n H[n] = 0 and sum_n H[n] (xi_n, eta_n) = 0.  Given the rows of
// nodes 3..5 these fix the vertex rows: (H_1, H_2) = -E^-T (t1, t2) with
// t = sum_{n>=3} H[n] (xi_n, eta_n), and H_0 = -(H_1 + H_2 + sum_{n>=3} H[n]).
// The reconstruction agrees with the host's G and H to ~4e-16 relative.
struct RecF64 {
    double hc[9];  // H[n][k] of patch nodes n = 3..5 at 3 (n - 3) + k
    double h, mu, A;
    DialRec dr;
};

// load_area = false: the area is formed from the reference positions
// (tri_forces_fp64_rec's area_ref) and not read.
template <bool LCE, bool UNI, int CS = 0>
__device__ __forceinline__ void load_rec_fp64(const TriData& td, const TriParams& sp, unsigned t, RecF64& r,
                                              bool load_area = true) {
    const size_t T = CS ? (size_t)CS : (size_t)td.T;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.hc[k] = ld_stream<CS != 0>(td.H + (9 + k) * T, t);
    r.h = UNI ? sp.h : ld_stream<CS != 0>(td.h, t);
    r.mu = UNI ? sp.mu : ld_stream<CS != 0>(td.mu, t);
    r.A = load_area ? ld_stream<CS != 0>(td.area, t) : 0.0;
    load_dial<LCE, UNI, CS, CS != 0>(td, sp, t, r.dr);  // tile records (CS = 64): non-temporal
}

// 1/x to about an ulp for normal, finite x: the hardware reciprocal refined
// by two Newton steps (no final rounding correction).
__device__ __forceinline__ double fast_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// 0.5 |x1 y2 - x2 y1|, each product rounded (no fused multiply-add): the
// area of a triangle in the plane from its edge vectors (x1, y1), (x2, y2).
__device__ __forceinline__ double planar_area(double x1, double y1, double x2, double y2) {
#pragma clang fp contract(off)
    return 0.5 * fabs(x1 * y2 - x2 * y1);
}

// FP64 force contributions (out[3*slot + c]) of the triangle whose record is
// r; pos fetches its six patch nodes, ref their reference x, y.  area_ref:
// the reference area is 0.5 |x1 y2 - x2 y1| of the reference edge vectors to
// patch nodes 1, 2 (each product rounded; the host checked that this equals
// the area stream), else r.A.
// UNI: h and mu are the shell's (sp), so the factors built from them alone
// are loop invariants of the kernels' triangle loops.
template <bool LCE, bool UNI, class Pos, class Ref>
__device__ __forceinline__ void tri_forces_fp64_rec(const RecF64& r, const TriParams& sp, const Pos& pos,
                                                    const Ref& ref, double out[18], bool area_ref = false) {
#pragma clang fp contract(on)
    const double* hc = r.hc;
    const double h = UNI ? sp.h : r.h, mu = UNI ? sp.mu : r.mu;
    const double nu = sp.nu;
    // reference edge vectors to patch node 0, and the first-derivative
    // weights gc of P_1 - P_0, P_2 - P_0: (G[0][1], G[0][2], G[1][1], G[1][2])
    double xi[6], eta[6];
    {
        double X0, Y0;
        ref(0, X0, Y0);
#pragma unroll
        for (int s = 1; s < 6; ++s) {
            double X, Y;
            ref(s, X, Y);
            xi[s] = X - X0;
            eta[s] = Y - Y0;
        }
    }
    const double idet = fast_rcp(xi[1] * eta[2] - xi[2] * eta[1]);
    const double A = area_ref ? planar_area(xi[1], eta[1], xi[2], eta[2]) : r.A;
    const double gc[4] = {eta[2] * idet, -eta[1] * idet, -xi[2] * idet, xi[1] * idet};
    double x0, y0, z0;
    pos(0, x0, y0, z0);
    // First derivatives from the edge vectors e_1, e_2 to patch node 0.  The
    // vertex rows of H follow from the moment identities (its columns
    // annihilate 1, X and Y): (H_1, H_2) = -E^-T sum_n H_n (xi_n, eta_n) over
    // the non-vertex nodes n = 3..5, H_0 = -(H_1 + H_2 + sum_n H_n).  So the
    // second derivatives are R2 = sum_n H_n d_n with d_n = e_n - xi_n Rx -
    // eta_n Ry, node n's offset from the tangent-plane prediction.  R2 enters
    // only through b = -n . R2 and Wb . R2, which are formed from the d_n
    // directly; neither R2 nor the vertex rows are formed.
    double Rx[3], Ry[3], d[3][3];  // d[n - 3][c]
    {
        double p1[3], p2[3];
        pos(1, p1[0], p1[1], p1[2]);
        pos(2, p2[0], p2[1], p2[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double e1 = p1[c] - (c == 0 ? x0 : c == 1 ? y0 : z0);
            const double e2 = p2[c] - (c == 0 ? x0 : c == 1 ? y0 : z0);
            Rx[c] = gc[0] * e1 + gc[1] * e2;
            Ry[c] = gc[2] * e1 + gc[3] * e2;
        }
    }
#pragma unroll
    for (int s = 3; s < 6; ++s) {
        double p[3];
        pos(s, p[0], p[1], p[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            d[s - 3][c] = (p[c] - (c == 0 ? x0 : c == 1 ? y0 : z0)) - xi[s] * Rx[c] - eta[s] * Ry[c];
    }
    const double a0 = Rx[0] * Rx[0] + Rx[1] * Rx[1] + Rx[2] * Rx[2];
    const double a1 = Rx[0] * Ry[0] + Rx[1] * Ry[1] + Rx[2] * Ry[2];
    const double a2 = Ry[0] * Ry[0] + Ry[1] * Ry[1] + Ry[2] * Ry[2];
    const double m0 = Rx[1] * Ry[2] - Rx[2] * Ry[1];
    const double m1 = Rx[2] * Ry[0] - Rx[0] * Ry[2];
    const double m2 = Rx[0] * Ry[1] - Rx[1] * Ry[0];
    // reciprocal square roots (a few ulp) and reciprocals of per-shell values
    // in place of per-triangle FP64 divisions and square roots
    const double imn = rsqrt(m0 * m0 + m1 * m1 + m2 * m2);
    const double n[3] = {m0 * imn, m1 * imn, m2 * imn};
    // b = -n . R2 = -sum_n H_n (n . d_n)
    double dn[3], b[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) dn[j] = n[0] * d[j][0] + n[1] * d[j][1] + n[2] * d[j][2];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) b[kk] = -(hc[kk] * dn[0] + hc[3 + kk] * dn[1] + hc[6 + kk] * dn[2]);
    double ab[3], bb[3];
    dial_targets<LCE>(r.dr, sp, ab, bb);
    const double P[3] = {ab[2], -ab[1], ab[0]};
    const double det = ab[0] * ab[2] - ab[1] * ab[1];
    const double rdet = rsqrt(det);
    const double idet32 = rdet * rdet * rdet;  // det^(-3/2)
    const double i1mnu = 1.0 / (1.0 - nu);
    const double cs = h * mu * i1mnu * idet32;
    const double cb = mu * h * h * h * (i1mnu * (1.0 / 12.0)) * idet32;
    const double Da[3] = {a0 - ab[0], a1 - ab[1], a2 - ab[2]};
    const double Db[3] = {b[0] - bb[0], b[1] - bb[1], b[2] - bb[2]};
    double ga[3], gbv[3];
    phi_grad(nu, P, Da, ga);
    const double Wb0 = cb * phi_grad(nu, P, Db, gbv);
    const double stiff = 1.0 + Wb0 * (2.0 / (sp.gamma * mu * h));  // calc_deformation_forces.cpp:141
    // the stresses times the reference area (every force term carries A once)
    const double csA = cs * A, cbsA = cb * stiff * A;
    const double Wa[3] = {csA * ga[0], csA * ga[1], csA * ga[2]};
    const double Wb[3] = {cbsA * gbv[0], cbsA * gbv[1], cbsA * gbv[2]};
    // H_s . Wb per patch node: the non-vertex nodes' directly, the vertex
    // nodes' through the moment identities above
    double ws[6];
#pragma unroll
    for (int s = 3; s < 6; ++s) {
        const double* hs = hc + 3 * (s - 3);
        ws[s] = hs[0] * Wb[0] + hs[1] * Wb[1] + hs[2] * Wb[2];
    }
    // w = -Wb . R2 = -sum_n (H_n . Wb) d_n
    double w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) w[c] = -(ws[3] * d[0][c] + ws[4] * d[1][c] + ws[5] * d[2][c]);
    const double nw = n[0] * w[0] + n[1] * w[1] + n[2] * w[2];
    double u[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c] = (w[c] - nw * n[c]) * imn;
    const double cx[3] = {Ry[1] * u[2] - Ry[2] * u[1], Ry[2] * u[0] - Ry[0] * u[2], Ry[0] * u[1] - Ry[1] * u[0]};
    const double cy[3] = {u[1] * Rx[2] - u[2] * Rx[1], u[2] * Rx[0] - u[0] * Rx[2], u[0] * Rx[1] - u[1] * Rx[0]};
    double Agx[3], Agy[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Agx[c] = 2.0 * Wa[0] * Rx[c] + Wa[1] * Ry[c] + cx[c];
        Agy[c] = Wa[1] * Rx[c] + 2.0 * Wa[2] * Ry[c] + cy[c];
    }
    // node 0's first-derivative weights: minus the sums of the others'
    const double g0[2] = {-(gc[0] + gc[1]), -(gc[2] + gc[3])};
    {
        const double T1 = ws[3] * xi[3] + ws[4] * xi[4] + ws[5] * xi[5];
        const double T2 = ws[3] * eta[3] + ws[4] * eta[4] + ws[5] * eta[5];
        ws[1] = -(gc[0] * T1 + gc[2] * T2);
        ws[2] = -(gc[1] * T1 + gc[3] * T2);
        ws[0] = -((ws[1] + ws[2]) + ((ws[3] + ws[4]) + ws[5]));
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double f = ws[s] * n[c];
            if (s == 0) f -= g0[0] * Agx[c] + g0[1] * Agy[c];
            if (s == 1 || s == 2) f -= gc[s - 1] * Agx[c] + gc[s + 1] * Agy[c];
            out[3 * s + c] = f;
        }
    }
}

// FP64 force contributions of triangle t: its record, then the mechanics.
template <bool LCE, bool UNI, int CS, class Pos, class Ref>
__device__ __forceinline__ void tri_forces_fp64(const TriData& td, const TriParams& sp, unsigned t, const Pos& pos,
                                                const Ref& ref, double out[18]) {
    RecF64 r;
    load_rec_fp64<LCE, UNI, CS>(td, sp, t, r);
    tri_forces_fp64_rec<LCE, UNI>(r, sp, pos, ref, out);
}

// The mixed-mode record of triangle t: 19 of the 24 FP32 patch coefficients
// (node 0's weights are minus the sums of the others'), h, mu, area and the
// dial data.  Loaded whole before any use, as the FP64 record.
struct RecMixed {
    float g[4];       // G[0][1], G[0][2], G[1][1], G[1][2]
    float hc[5][3];   // H[s][k] of patch nodes s = 1..5
    float h, mu, A;
    DialRec dr;
};

template <bool LCE, bool UNI, int CS = 0>
__device__ __forceinline__ void load_rec_mixed(const TriData& td, const TriParams& sp, unsigned t, RecMixed& r) {
    const size_t T = CS ? (size_t)CS : (size_t)td.T;
    r.g[0] = ld(td.G32 + 1 * T, t);
    r.g[1] = ld(td.G32 + 2 * T, t);
    r.g[2] = ld(td.G32 + 4 * T, t);
    r.g[3] = ld(td.G32 + 5 * T, t);
#pragma unroll
    for (int s = 1; s < 6; ++s)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) r.hc[s - 1][kk] = ld(td.H32 + (s * 3 + kk) * T, t);
    r.h = UNI ? (float)sp.h : ld(td.hmuA32, t);
    r.mu = UNI ? (float)sp.mu : ld(td.hmuA32 + T, t);
    r.A = ld(td.hmuA32 + 2 * T, t);
    load_dial<LCE, UNI, CS>(td, sp, t, r.dr);
}

// Mixed mode: edge vectors to vertex 0 and the strains a - abar, b - bbar in
// double, the patch fit, stresses and contributions in float.
template <bool LCE, class Pos>
__device__ __forceinline__ void tri_forces_mixed_rec(const RecMixed& r, const TriParams& sp, const Pos& pos,
                                                     float out[18]) {
#pragma clang fp contract(on)
    // Edge vectors, the surface derivatives and the strains a - abar, b - bbar
    // are formed in FP64 (the FP32 patch-fit coefficients are exact in
    // double).  Rounding the tangent vectors to FP32 before forming the
    // metric makes a - abar a noisy function of the positions (~1e-7 per
    // step, uncorrelated between steps); under the weak damping that noise
    // holds the residual force far above its static error.  The stresses
    // and the force assembly below run in FP32.
    const float* g = r.g;
    const float(*hc)[3] = r.hc;
    const float h = r.h, mu = r.mu, A = r.A;
    const DialRec& dr = r.dr;
    double x0, y0, z0;
    pos(0, x0, y0, z0);
    // one edge vector at a time into the first and second derivatives
    double rx[3], ry[3], r2[3][3];
#pragma unroll
    for (int s = 1; s < 6; ++s) {
        double x, y, z;
        pos(s, x, y, z);
        const double e[3] = {x - x0, y - y0, z - z0};
        if (s == 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                rx[c] = (double)g[0] * e[c];
                ry[c] = (double)g[2] * e[c];
            }
        } else if (s == 2) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                rx[c] += (double)g[1] * e[c];
                ry[c] += (double)g[3] * e[c];
            }
        }
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (s == 1)
                    r2[kk][c] = (double)hc[0][kk] * e[c];
                else
                    r2[kk][c] += (double)hc[s - 1][kk] * e[c];
            }
    }
    double ab[3], bb[3];
    dial_targets<LCE>(dr, sp, ab, bb);
    float Rx[3], Ry[3], R2[3][3], Da[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Rx[c] = (float)rx[c];
        Ry[c] = (float)ry[c];
    }
    Da[0] = (float)((rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2]) - ab[0]);
    Da[1] = (float)((rx[0] * ry[0] + rx[1] * ry[1] + rx[2] * ry[2]) - ab[1]);
    Da[2] = (float)((ry[0] * ry[0] + ry[1] * ry[1] + ry[2] * ry[2]) - ab[2]);
    const float m0 = Rx[1] * Ry[2] - Rx[2] * Ry[1];
    const float m1 = Rx[2] * Ry[0] - Rx[0] * Ry[2];
    const float m2 = Rx[0] * Ry[1] - Rx[1] * Ry[0];
    const float imn = rsqrtf(m0 * m0 + m1 * m1 + m2 * m2);
    const float n[3] = {m0 * imn, m1 * imn, m2 * imn};
    float Db[3];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
#pragma unroll
        for (int c = 0; c < 3; ++c) R2[kk][c] = (float)r2[kk][c];
        const double bk = -((double)n[0] * r2[kk][0] + (double)n[1] * r2[kk][1] + (double)n[2] * r2[kk][2]);
        Db[kk] = (float)(bk - bb[kk]);
    }
    const double det = ab[0] * ab[2] - ab[1] * ab[1];
    const double idet32 = 1.0 / (det * sqrt(det));
    const float nu = (float)sp.nu;
    const float cs = (float)((double)h * (double)mu / (1.0 - sp.nu) * idet32);
    const float cb = (float)((double)mu * h * h * h / (12.0 * (1.0 - sp.nu)) * idet32);
    const float P[3] = {(float)ab[2], (float)-ab[1], (float)ab[0]};
    float ga[3], gbv[3];
    phi_grad(nu, P, Da, ga);
    const float Wb0 = cb * phi_grad(nu, P, Db, gbv);
    const float stiff = 1.0f + 2.0f * Wb0 / (float)(sp.gamma * (double)mu * (double)h);
    const float Wa[3] = {cs * ga[0], cs * ga[1], cs * ga[2]};
    const float cbs = cb * stiff;
    const float Wb[3] = {cbs * gbv[0], cbs * gbv[1], cbs * gbv[2]};
    float w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) w[c] = -(Wb[0] * R2[0][c] + Wb[1] * R2[1][c] + Wb[2] * R2[2][c]);
    const float nw = n[0] * w[0] + n[1] * w[1] + n[2] * w[2];
    float u[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c] = (w[c] - nw * n[c]) * imn;
    const float cx[3] = {Ry[1] * u[2] - Ry[2] * u[1], Ry[2] * u[0] - Ry[0] * u[2], Ry[0] * u[1] - Ry[1] * u[0]};
    const float cy[3] = {u[1] * Rx[2] - u[2] * Rx[1], u[2] * Rx[0] - u[0] * Rx[2], u[0] * Rx[1] - u[1] * Rx[0]};
    float Agx[3], Agy[3], An[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Agx[c] = A * (2.0f * Wa[0] * Rx[c] + Wa[1] * Ry[c] + cx[c]);
        Agy[c] = A * (Wa[1] * Rx[c] + 2.0f * Wa[2] * Ry[c] + cy[c]);
        An[c] = A * n[c];
    }
    // node 0's weights: minus the sums of the others' (G and H annihilate
    // constants), so the record's node-0 coefficients are never read
    const float g0[2] = {-(g[0] + g[1]), -(g[2] + g[3])};
    float h0[3];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) h0[kk] = -(((hc[0][kk] + hc[1][kk]) + (hc[2][kk] + hc[3][kk])) + hc[4][kk]);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const float* hs = s == 0 ? h0 : hc[s - 1];
        const float ws = hs[0] * Wb[0] + hs[1] * Wb[1] + hs[2] * Wb[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float f = ws * An[c];
            if (s == 0) f -= g0[0] * Agx[c] + g0[1] * Agy[c];
            if (s == 1 || s == 2) f -= g[s - 1] * Agx[c] + g[s + 1] * Agy[c];
            out[3 * s + c] = f;
        }
    }
}

// Mixed-mode force contributions of triangle t: its record, then the mechanics.
template <bool LCE, bool UNI, int CS, class Pos>
__device__ __forceinline__ void tri_forces_mixed(const TriData& td, const TriParams& sp, unsigned t, const Pos& pos,
                                                 float out[18]) {
    RecMixed r;
    load_rec_mixed<LCE, UNI, CS>(td, sp, t, r);
    tri_forces_mixed_rec<LCE>(r, sp, pos, out);
}

}  // namespace ms
