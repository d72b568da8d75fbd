// Per-triangle launches of the per-step path (gfx950).
//
// One thread per triangle, one chunk of triangles (of one shell) per block.
// The force launch streams each triangle's patch record (6 node ids, G, H,
// dial info, h, mu, area), gathers its six patch nodes and writes the 18
// force contributions -A dW/dx of its patch nodes (tri_math.h); they are
// gathered per node, in a fixed order, by node_kernels.hip: no atomics.
// The output-time launches evaluate the energy densities (with per-chunk
// compensated partial sums) and the geometry a, b, n.

#include <math.h>

#include "kernels.h"
#include "stress_math.h"
#include "tri_math.h"

namespace ms {
namespace {

// ---- compensated block reduction (two-sum, fixed tree: deterministic) ----
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}
__device__ __forceinline__ void dd_add(double& hi, double& lo, double h2, double l2) {
#pragma clang fp contract(off)
    double s, e;
    two_sum(hi, h2, s, e);
    e = e + (lo + l2);
    two_sum(s, e, hi, lo);
}

// Reduce (hi, lo) pairs of NV values across the block; result in thread 0.
template <int NV>
__device__ __forceinline__ void block_dd_reduce(double hi[NV], double lo[NV]) {
    __shared__ double sh_hi[NV][16], sh_lo[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double h2 = __shfl_down(hi[k], off, 64);
            const double l2 = __shfl_down(lo[k], off, 64);
            if (lane < off) dd_add(hi[k], lo[k], h2, l2);
        }
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            sh_hi[k][wave] = hi[k];
            sh_lo[k][wave] = lo[k];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            hi[k] = sh_hi[k][0];
            lo[k] = sh_lo[k][0];
            for (int w = 1; w < nw; ++w) dd_add(hi[k], lo[k], sh_hi[k][w], sh_lo[k][w]);
        }
    }
    __syncthreads();
}

struct TriKArgs {
    TriData td;
    const double* pos;   // buffer 0 (3N)
    const double* pos1;  // buffer 1
    double* contrib;
    float* contrib32;
    const Chunk* chunks;
    int nchunks;
    ShellDev* shells;
    int skip_halted;
    double* dens;
    double* part;
    double* ga;
    double* gb;
    double* gn;
    int obegin;   // geometry outputs: triangle t at [k * ostride + t - obegin]
    int ostride;
};

// ------------------------------------------------------------------ forces
template <bool MIXED, bool LCE, bool UNI>
__global__ __launch_bounds__(256) void tri_forces_kernel(TriKArgs k) {
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const int ci = xcd_remap(vi, k.nchunks);
        const Chunk ch = make_int4(uni(k.chunks[ci].x), uni(k.chunks[ci].y), uni(k.chunks[ci].z), 0);
        const ShellDev& sh = k.shells[ch.z];
        if (k.skip_halted && uni(sh.halted)) continue;  // block-uniform
        const TriParams tp = tri_params<UNI>(sh);
        const double* X = uni(sh.par) ? k.pos1 : k.pos;
        const int t = ch.x + threadIdx.x;
        if (t >= ch.y) continue;
        unsigned p[6];
        load_patch_ids(k.td, (unsigned)t, p);
        const GlobalPos pos{X, (size_t)k.td.N, p};
        if (MIXED) {
            float out[18];
            tri_forces_mixed<LCE, UNI, 0>(k.td, tp, (unsigned)t, pos, out);
            float2* dst = reinterpret_cast<float2*>(k.contrib32 + (size_t)t * 18);
#pragma unroll
            for (int i = 0; i < 9; ++i) dst[i] = make_float2(out[2 * i], out[2 * i + 1]);
        } else {
            double out[18];
            const GlobalRef ref{k.td.ref, (size_t)k.td.N, p};
            tri_forces_fp64<LCE, UNI, 0>(k.td, tp, (unsigned)t, pos, ref, out);
            double2* dst = reinterpret_cast<double2*>(k.contrib + (size_t)t * 18);
#pragma unroll
            for (int i = 0; i < 9; ++i) dst[i] = make_double2(out[2 * i], out[2 * i + 1]);
        }
    }
}

// ---------------------------------------------- energies and geometry (FP64)
template <int MODE>
__global__ __launch_bounds__(256) void tri_output_kernel(TriKArgs k) {
    const TriData& td = k.td;
    const int T = td.T, N = td.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const int ci = xcd_remap(vi, k.nchunks);
        const Chunk ch = k.chunks[ci];
        const ShellDev& sh = k.shells[ch.z];
        const int t = ch.x + threadIdx.x;
        double Es = 0.0, Eb = 0.0;
        if (t < ch.y) {
            const double* X = sh.par ? k.pos1 : k.pos;
            double Rx[3] = {0, 0, 0}, Ry[3] = {0, 0, 0}, R2[3][3] = {};
            for (int s = 0; s < 6; ++s) {
                const int p = td.pid[s * T + t];
                const double P[3] = {X[p], X[N + p], X[2 * N + p]};
                for (int c = 0; c < 3; ++c) {
                    if (s < 3) {
                        Rx[c] += td.G[s * T + t] * P[c];
                        Ry[c] += td.G[(3 + s) * T + t] * P[c];
                    }
                    for (int kk = 0; kk < 3; ++kk) R2[kk][c] += td.H[(s * 3 + kk) * T + t] * P[c];
                }
            }
            const double a[3] = {Rx[0] * Rx[0] + Rx[1] * Rx[1] + Rx[2] * Rx[2],
                                 Rx[0] * Ry[0] + Rx[1] * Ry[1] + Rx[2] * Ry[2],
                                 Ry[0] * Ry[0] + Ry[1] * Ry[1] + Ry[2] * Ry[2]};
            const double m0 = Rx[1] * Ry[2] - Rx[2] * Ry[1];
            const double m1 = Rx[2] * Ry[0] - Rx[0] * Ry[2];
            const double m2 = Rx[0] * Ry[1] - Rx[1] * Ry[0];
            const double mn = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
            const double imn = 1.0 / mn;
            const double n[3] = {m0 * imn, m1 * imn, m2 * imn};
            double b[3];
            for (int kk = 0; kk < 3; ++kk) b[kk] = -(n[0] * R2[kk][0] + n[1] * R2[kk][1] + n[2] * R2[kk][2]);
            if (MODE == TRI_GEOMETRY) {
                const size_t o = (size_t)(t - k.obegin);
                for (int kk = 0; kk < 3; ++kk) {
                    k.ga[kk * (size_t)k.ostride + o] = a[kk];
                    k.gb[kk * (size_t)k.ostride + o] = b[kk];
                    k.gn[kk * (size_t)k.ostride + o] = n[kk];
                }
            } else if (MODE == TRI_STRESS) {
                double ab[3], bb[3], S[6], lam[3], V[3][3], o[8];
                targets(td, tri_params(sh), (unsigned)t, ab, bb);
                cauchy_lower(sh.nu, sh.gamma, td.h[t], td.mu[t], td.area[t], ab, bb, Rx, Ry, R2, n, b, mn, S);
                sym3_eigen(S, lam, V);
                stress_pack(lam, V, o);
                for (int kk = 0; kk < 8; ++kk) k.ga[kk * (size_t)T + t] = o[kk];
            } else if (MODE == TRI_DIAGNOSTICS) {
                // the shape operator a^-1 b and the strain a^-1 (a - abar):
                // invariants of the reference's QR and pseudo-inverse forms
                double ab[3], bb[3];
                targets(td, tri_params(sh), (unsigned)t, ab, bb);
                const double det = a[0] * a[2] - a[1] * a[1];
                const double gauss = (b[0] * b[2] - b[1] * b[1]) / det;
                const double mean = 0.5 * (a[2] * b[0] - 2.0 * a[1] * b[1] + a[0] * b[2]) / det;
                const double D0 = a[0] - ab[0], D1 = a[1] - ab[1], D2 = a[2] - ab[2];
                const double m00 = (a[2] * D0 - a[1] * D1) / det, m01 = (a[2] * D1 - a[1] * D2) / det;
                const double m10 = (a[0] * D1 - a[1] * D0) / det, m11 = (a[0] * D2 - a[1] * D1) / det;
                const double tr2 = m00 * m00 + 2.0 * m01 * m10 + m11 * m11;
                k.ga[t] = gauss;
                k.ga[T + t] = mean;
                k.ga[2 * T + t] = sqrt(tr2 / 8.0);
            } else {
                double ab[3], bb[3];
                targets(td, tri_params(sh), (unsigned)t, ab, bb);
                const double P[3] = {ab[2], -ab[1], ab[0]};
                const double det = ab[0] * ab[2] - ab[1] * ab[1];
                const double idet32 = 1.0 / (det * sqrt(det));
                const double h = td.h[t], mu = td.mu[t], nu = sh.nu;
                const double cs = h * mu / (1.0 - nu) * idet32;
                const double cb = mu * h * h * h / (12.0 * (1.0 - nu)) * idet32;
                const double Da[3] = {a[0] - ab[0], a[1] - ab[1], a[2] - ab[2]};
                const double Db[3] = {b[0] - bb[0], b[1] - bb[1], b[2] - bb[2]};
                const double Ws = cs * phi_only(nu, P, Da);
                const double Wb0 = cb * phi_only(nu, P, Db);
                const double Wb = Wb0 * (1.0 + Wb0 / (sh.gamma * mu * h));
                if (k.dens) {
                    k.dens[t] = Ws;
                    k.dens[T + t] = Wb;
                }
                Es = td.area[t] * Ws;
                Eb = td.area[t] * Wb;
            }
        }
        if (MODE == TRI_ENERGY) {
            double hi[2] = {Es, Eb}, lo[2] = {0.0, 0.0};
            block_dd_reduce<2>(hi, lo);
            if (threadIdx.x == 0) {
                k.part[4 * ci + 0] = hi[0];
                k.part[4 * ci + 1] = lo[0];
                k.part[4 * ci + 2] = hi[1];
                k.part[4 * ci + 3] = lo[1];
            }
        }
    }
}

TriKArgs make_args(const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks, ShellDev* shells) {
    TriKArgs k{};
    k.td = td;
    k.pos = b.x[0];
    k.pos1 = b.x[1];
    k.contrib = b.contrib;
    k.contrib32 = b.contrib32;
    k.chunks = chunks;
    k.nchunks = nchunks;
    k.shells = shells;
    k.part = b.part;
    k.obegin = 0;
    k.ostride = td.T;
    return k;
}

}  // namespace

void launch_tri_forces(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, bool mixed, bool lce, bool uni_data, bool skip_halted) {
    TriKArgs k = make_args(td, b, chunks, nchunks, shells);
    k.skip_halted = skip_halted ? 1 : 0;
    const dim3 g(L.grid), bl(L.block);
    const int v = (mixed ? 4 : 0) | (lce ? 2 : 0) | (uni_data ? 1 : 0);
#define MS_TRI_CASE(M, C, U)                                                             \
    case (M ? 4 : 0) | (C ? 2 : 0) | (U ? 1 : 0):                                        \
        hipLaunchKernelGGL((tri_forces_kernel<M, C, U>), g, bl, 0, L.stream, k); \
        break;
    switch (v) {
        MS_TRI_CASE(false, false, false)
        MS_TRI_CASE(false, false, true)
        MS_TRI_CASE(false, true, false)
        MS_TRI_CASE(false, true, true)
        MS_TRI_CASE(true, false, false)
        MS_TRI_CASE(true, false, true)
        MS_TRI_CASE(true, true, false)
        MS_TRI_CASE(true, true, true)
    }
#undef MS_TRI_CASE
}

void launch_tri_energy(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, double* dens) {
    TriKArgs k = make_args(td, b, chunks, nchunks, shells);
    k.dens = dens;
    hipLaunchKernelGGL(tri_output_kernel<TRI_ENERGY>, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_tri_diagnostics(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                            ShellDev* shells, double* out) {
    TriKArgs k = make_args(td, b, chunks, nchunks, shells);
    k.ga = out;
    hipLaunchKernelGGL(tri_output_kernel<TRI_DIAGNOSTICS>, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_tri_stresses(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, double* out) {
    TriKArgs k = make_args(td, b, chunks, nchunks, shells);
    k.ga = out;
    hipLaunchKernelGGL(tri_output_kernel<TRI_STRESS>, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_tri_geometry(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, double* ga, double* gb, double* gn, int out_begin, int out_stride) {
    TriKArgs k = make_args(td, b, chunks, nchunks, shells);
    k.ga = ga;
    k.gb = gb;
    k.gn = gn;
    k.obegin = out_begin;
    k.ostride = out_stride;
    hipLaunchKernelGGL(tri_output_kernel<TRI_GEOMETRY>, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

}  // namespace ms
