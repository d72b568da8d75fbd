// Fused owner-computes step (gfx950): one workgroup per node tile.
//
// A tile is a node chunk of one shell.  Its workgroup evaluates every
// triangle whose six-node patch touches one of its nodes (the tile's own
// triangles plus a halo of about 1.3-1.5x as many, recomputed by each
// neighbouring tile), writes the contributions its nodes need into LDS at
// positions fixed by the patch-incidence CSR, and then, one thread per node,
// sums them in CSR order -- the order of the two-launch path, so the forces
// are bit-identical to it -- and applies damping, clamp, the equilibrium
// maxima, the blow-up flag and the update.  The per-slot contributions never
// go to HBM.

#include <limits.h>

#include <algorithm>

#include "kernels.h"
#include "node_math.h"
#include "tri_math.h"

namespace ms {
namespace {

struct TileKArgs {
    NodeData nd;
    Buffers b;
    TileData tl;
    const Chunk* chunks;
    int nchunks;
    ShellDev* shells;
    StepFlags fl;
};

template <int NV>
__device__ __forceinline__ void block_argmax(double v[NV], int idx[NV]) {
    __shared__ double sv[NV][16];
    __shared__ int si[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double v2 = __shfl_down(v[k], off, 64);
            const int i2 = __shfl_down(idx[k], off, 64);
            if (lane < off && (v2 > v[k] || (v2 == v[k] && i2 < idx[k]))) {
                v[k] = v2;
                idx[k] = i2;
            }
        }
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            sv[k][wave] = v[k];
            si[k][wave] = idx[k];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            v[k] = sv[k][0];
            idx[k] = si[k][0];
            for (int w = 1; w < nw; ++w)
                if (sv[k][w] > v[k] || (sv[k][w] == v[k] && si[k][w] < idx[k])) {
                    v[k] = sv[k][w];
                    idx[k] = si[k][w];
                }
        }
    }
    __syncthreads();
}

template <bool MIXED, bool CHECK>
__global__ __launch_bounds__(256) void tile_step_kernel(TileKArgs k) {
    typedef typename std::conditional<MIXED, float, double>::type CT;
    extern __shared__ unsigned char smem_raw[];
    CT* lds = reinterpret_cast<CT*>(smem_raw);
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const int ci = xcd_remap(vi, k.nchunks);
        const Chunk ch = make_int4(uni(k.chunks[ci].x), uni(k.chunks[ci].y), uni(k.chunks[ci].z), 0);
        const ShellDev& sh = k.shells[ch.z];
        if (uni(sh.halted)) continue;  // block-uniform
        const bool do_check = CHECK || uni(sh.check_due);
        const int par = uni(sh.par);
        const double* X = par ? k.b.x[1] : k.b.x[0];
        const TriParams tp = tri_params(sh);
        // phase A: contributions of the tile's triangles into LDS, from the
        // tile's contiguous copy of their records
        const int j0 = uni(k.tl.tri_off[ci]), j1 = uni(k.tl.tri_off[ci + 1]);
        for (int j = j0 + (int)threadIdx.x; j < j1; j += blockDim.x) {
            // opaque to loop strength reduction: the record loads keep one
            // 32-bit lane offset against scalar stream bases instead of a
            // 64-bit induction pointer per stream
            unsigned t = (unsigned)j;
            asm volatile("" : "+v"(t));
            unsigned pc[6];
            load_patch_ids(k.tl.td, t, pc);
            CT out[18];
            if (MIXED)
                tri_forces_mixed(k.tl.td, tp, t, pc, X, reinterpret_cast<float*>(out));
            else
                tri_forces_fp64(k.tl.td, tp, t, pc, X, reinterpret_cast<double*>(out));
            const unsigned short* d = k.tl.dst + (size_t)j * 6;
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const unsigned int q = d[s];
                if (q != 0xFFFFu) {
                    lds[3 * q] = out[3 * s];
                    lds[3 * q + 1] = out[3 * s + 1];
                    lds[3 * q + 2] = out[3 * s + 2];
                }
            }
        }
        __syncthreads();
        // phase B: one thread per node
        const int n = ch.x + threadIdx.x;
        double fnorm = -1.0, vnorm = -1.0;
        int fidx = INT_MAX, vidx = INT_MAX;
        if (n < ch.y) {
            const long long base = k.nd.csr_off[ch.x];
            double fx, fy, fz;
            ordered_sum3(lds, [](long long e) { return e; }, k.nd.csr_off[n] - base, k.nd.csr_off[n + 1] - base, fx,
                         fy, fz);
            const double* V = k.b.v[par];
            double* Xo = k.b.x[par ^ 1];
            double* Vo = k.b.v[par ^ 1];
            double vx = V[n], vy = V[N + n], vz = V[2 * N + n];
            const double m = k.nd.mass[n];
            if (!sh.gd) {
                fx = damped(fx, sh.eta, m, vx, sh.damp_den);
                fy = damped(fy, sh.eta, m, vy, sh.damp_den);
                fz = damped(fz, sh.eta, m, vz, sh.damp_den);
            }
            if (k.nd.clamp[n]) fx = fy = fz = 0.0;
            if (do_check) {
                fnorm = norm3(fx, fy, fz);
                vnorm = norm3(vx, vy, vz);
                fidx = vidx = n;
            }
            const double lim = sh.lim;
            if (!(fabs(fx) < lim) || !(fabs(fy) < lim) || !(fabs(fz) < lim)) atomicOr(&k.shells[ch.z].failed, 1);
            double x = X[n], y = X[N + n], z = X[2 * N + n];
            update(sh, m, fx, vx, x);
            update(sh, m, fy, vy, y);
            update(sh, m, fz, vz, z);
            Xo[n] = x;
            Xo[N + n] = y;
            Xo[2 * N + n] = z;
            Vo[n] = vx;
            Vo[N + n] = vy;
            Vo[2 * N + n] = vz;
        }
        if (do_check) {
            double v2[2] = {fnorm, vnorm};
            int i2[2] = {fidx, vidx};
            block_argmax<2>(v2, i2);
            if (threadIdx.x == 0) {
                k.b.part[2 * ci] = v2[0];
                k.b.part[2 * ci + 1] = v2[1];
                k.b.part_i[2 * ci] = i2[0];
                k.b.part_i[2 * ci + 1] = i2[1];
            }
        }
        __syncthreads();  // the next tile reuses the LDS
    }
}

template <typename E>
__global__ __launch_bounds__(256) void tile_gather_kernel(const E* __restrict__ in, E* __restrict__ out, int ncomp,
                                                          int T, int Tp, const int* __restrict__ tris) {
    const long long total = (long long)ncomp * Tp;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int kc = (int)(i / Tp), j = (int)(i - (long long)kc * Tp);
        out[i] = in[(size_t)kc * T + tris[j]];
    }
}

}  // namespace

void launch_tile_gather(hipStream_t s, const void* in, void* out, int elem_bytes, int ncomp, int T, int Tp,
                        const int* tris) {
    const long long total = (long long)ncomp * Tp;
    if (total <= 0) return;
    const int grid = (int)std::min<long long>(8192, (total + 255) / 256);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(tile_gather_kernel<double>, dim3(grid), dim3(256), 0, s, (const double*)in, (double*)out,
                           ncomp, T, Tp, tris);
    else
        hipLaunchKernelGGL(tile_gather_kernel<float>, dim3(grid), dim3(256), 0, s, (const float*)in, (float*)out,
                           ncomp, T, Tp, tris);
}

int tile_lds_bytes(int max_entries, bool mixed) { return max_entries * 3 * (mixed ? 4 : 8); }

template <bool MIXED>
hipError_t raise_lds(int bytes) {
    if (bytes <= 65536) return hipSuccess;
    const void* fns[2] = {(const void*)tile_step_kernel<MIXED, false>, (const void*)tile_step_kernel<MIXED, true>};
    for (const void* f : fns) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t prepare_tile_kernels(int lds_bytes_fp64, int lds_bytes_mixed) {
    const hipError_t e = raise_lds<false>(lds_bytes_fp64);
    return e != hipSuccess ? e : raise_lds<true>(lds_bytes_mixed);
}

template <bool MIXED>
void launch_tile_variant(const Launch& L, const TileKArgs& k, size_t lds, bool check) {
    if (check)
        hipLaunchKernelGGL((tile_step_kernel<MIXED, true>), dim3(L.grid), dim3(L.block), lds, L.stream, k);
    else
        hipLaunchKernelGGL((tile_step_kernel<MIXED, false>), dim3(L.grid), dim3(L.block), lds, L.stream, k);
}

void launch_tile_step(const Launch& L, const NodeData& nd, const Buffers& b, const TileData& tl, const Chunk* chunks,
                      int nchunks, ShellDev* shells, bool mixed, const StepFlags& fl) {
    TileKArgs k{};
    k.nd = nd;
    k.b = b;
    k.tl = tl;
    k.chunks = chunks;
    k.nchunks = nchunks;
    k.shells = shells;
    k.fl = fl;
    const size_t lds = (size_t)tile_lds_bytes(tl.max_entries, mixed);
    if (mixed)
        launch_tile_variant<true>(L, k, lds, fl.check_equil != 0);
    else
        launch_tile_variant<false>(L, k, lds, fl.check_equil != 0);
}

}  // namespace ms
