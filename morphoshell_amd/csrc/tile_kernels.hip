// Fused owner-computes step (gfx950): one workgroup per node tile.
//
// A tile is a node chunk of one shell.  Its workgroup
//   1. stages the positions of its nodes and of the halo nodes its triangles
//      reach into LDS (coalesced loads, once per tile);
//   2. evaluates every triangle whose six-node patch touches one of its nodes
//      (its own triangles plus a halo of about 1.3x as many, recomputed by
//      each neighbouring tile), reading the patch nodes from LDS through
//      16-bit tile-local ids, and writes the contributions its nodes need into
//      LDS in patch-incidence CSR order;
//   3. one thread per node sums them in that order -- the order of the
//      two-launch path, so forces are bit-identical to it -- and applies
//      damping, clamp, the equilibrium maxima, the blow-up flag and the update.
// Steps 2-3 run in passes of B triangles (B = block size) with each node's
// running sum carried over, so the LDS holds one pass's contributions (about
// a third of the tile's).  The per-slot contributions never go to HBM.

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
    ShellDev* shells;
    StepFlags fl;
};

// Block-uniform tables and the shell record, read through the constant
// address space so that they are scalar loads: those bypass the vector memory
// queue, which the record streams keep full, so the chain tile -> shell ->
// dial factor at the start of a tile does not wait behind them.  The tables
// are written only at set-up and the fields read here of the shell record
// only by the bookkeeping launch between steps (the scalar cache is
// invalidated at every launch); the atomically set failure flag is not read.
template <typename T>
__device__ __forceinline__ T kload(const T* p, long long i) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    typedef const __attribute__((address_space(4))) unsigned KWord;
    KWord* src = (KWord*)(p + i);
    T v;
    unsigned* d = reinterpret_cast<unsigned*>(&v);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(T) / 4); ++w) d[w] = src[w];
    return v;
}

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

// LDS layout: positions (3 doubles per tile-local node), in mixed mode the
// nodes' running force sums (3 doubles per own node, carried from pass to
// pass in LDS so that they hold no registers while triangles are evaluated;
// the FP64 kernel keeps them in registers, where its larger LDS tile would
// otherwise cost a workgroup per CU), then contributions (3 CT per pass
// entry), then the tile's per-pass entry offsets (staged with the positions:
// a global load in a pass's sums would wait, by the in-order count of
// outstanding vector loads, for the next entries' records already in flight).
// Waves per SIMD asked of the compiler (their LDS tiles allow them): 2 in
// FP64 and 4 in mixed mode (3 with the LCE protocol).  Each triangle's record
// is loaded whole before its first use; at 2 waves the FP64 kernel keeps
// every wave's record in flight without spilling, and measured faster than at
// 3 waves (168 registers, where the loads had to wait for registers).
template <bool MIXED, bool CHECK, bool LCE, bool UNI>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(LCE ? (MIXED ? 3 : 2) : (MIXED ? 4 : 2))))
void tile_step_kernel(TileKArgs k) {
    typedef typename std::conditional<MIXED, float, double>::type CT;
    extern __shared__ double smem[];
    const int N = k.nd.N;
    const int B = blockDim.x;
    const int stride = k.tl.poff_stride;
    double* lpos = smem;
    double* lref = smem + 3 * (size_t)k.tl.max_nodes;  // FP64: reference x, y
    double* lsum = smem + (MIXED ? 3 : 5) * (size_t)k.tl.max_nodes;
    CT* lcon = reinterpret_cast<CT*>(MIXED ? lsum + 3 * (size_t)stride : lsum);
    unsigned short* lpo = reinterpret_cast<unsigned short*>(lcon + 3 * (size_t)k.tl.max_entries);
    const int nt = k.tl.active ? k.tl.nactive : k.tl.ntiles;
    for (int vi = blockIdx.x; vi < nt; vi += gridDim.x) {
        const int vr = xcd_remap(vi, nt);
        const int ci = k.tl.active ? kload(k.tl.active, vr) : vr;
        // nodes [ch.x, ch.y) of shell ch.z; chunk ch.w + r holds nodes from ch.x + r * B
        const int4 ch = kload(k.tl.tiles, ci);
        const ShellDev sh = kload(k.shells, ch.z);
        if (uni(sh.halted)) continue;  // block-uniform
        const bool do_check = CHECK || uni(sh.check_due);
        const int par = uni(sh.par);
        const double* X = par ? k.b.x[1] : k.b.x[0];
        TriParams tp = tri_params<UNI>(sh);
        // LCE shells: the records hold the dialled metric of this dial factor
        // when an earlier fused step on these records used the same factor
        if (LCE) tp.lce_cached = (tp.protocol == 3 && sh.lce_gen == k.tl.gen && sh.lce_f == tp.f) ? 1 : 0;
        const int nn = ch.y - ch.x;
        const int h0 = kload(k.tl.halo_off, ci), h1 = kload(k.tl.halo_off, ci + 1);
        const int j0 = kload(k.tl.tri_off, ci), j1 = kload(k.tl.tri_end, ci);
        const int split = k.tl.pass_split;
        const int npass = (j1 - j0 + split - 1) / split;
        const unsigned short* poff = k.tl.poff + (size_t)ci * k.tl.max_passes * stride;
        // the record block of entry j (a wave's 64 entries start at a multiple of 64)
        auto view = [&](int j) {
            const size_t blk = (size_t)uni(j >> 6) * (REC_SLOTS * 64);
            const double* r64 = k.tl.rec64 + blk;
            const float* r32 = k.tl.rec32 + blk;
            TriData rv{};
            rv.T = 64;
            rv.G = r64 + RS_G * 64;
            rv.H = r64 + RS_H * 64;
            rv.ainfo = r64 + RS_AI * 64;
            rv.binfo = r64 + RS_BI * 64;
            rv.h = r64 + RS_THICK * 64;
            rv.mu = r64 + RS_MU * 64;
            rv.area = r64 + RS_AREA * 64;
            rv.a0 = r64 + RS_A0 * 64;
            rv.b0 = r64 + RS_B0 * 64;
            rv.G32 = r32 + RS_G * 64;
            rv.H32 = r32 + RS_H * 64;
            rv.hmuA32 = r32 + RS_HMUA * 64;
            return rv;
        };
        // six 16-bit values of entry j (read as three 32-bit words) from
        // stream w: its tile-local patch ids (lpid) or the pass-local LDS
        // positions of its contributions (dst)
        auto words = [&](const unsigned short* w, int j, unsigned q[3]) {
            // opaque to loop strength reduction: the loads keep one 32-bit
            // lane offset against scalar bases instead of a 64-bit induction
            // pointer per stream
            unsigned t = (unsigned)j;
            asm volatile("" : "+v"(t));
            const unsigned* p = reinterpret_cast<const unsigned*>(w);
            q[0] = ld_stream<!MIXED>(p, 3 * t);
            q[1] = ld_stream<!MIXED>(p, 3 * t + 1);
            q[2] = ld_stream<!MIXED>(p, 3 * t + 2);
        };
        auto unpack = [](const unsigned q[3], unsigned v[6]) {
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                v[2 * s] = q[s] & 0xFFFFu;
                v[2 * s + 1] = q[s] >> 16;
            }
        };
        // an entry's contributions to the tile's nodes, at their pass-local LDS positions
        auto store = [&](const unsigned dw[3], const CT out[18]) {
            unsigned dq[6];
            unpack(dw, dq);
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const unsigned q = dq[s];
                if (q != 0xFFFFu) {
                    lcon[3 * q] = out[3 * s];
                    lcon[3 * q + 1] = out[3 * s + 1];
                    lcon[3 * q + 2] = out[3 * s + 2];
                }
            }
        };
        // FP64: the thread's entries are j0 + tid + m B (m = 0, 1, ...) across
        // all passes, and each one's record, ids and LDS positions are loaded
        // as soon as the previous entry's contributions are in LDS: at the end
        // of a pass the loads are in flight through the barriers and the node
        // sums, and the first entry's are in flight with the position staging
        // below.  (Loading them during the previous entry's mechanics, or
        // before its LDS store, needs more than the 256 registers of 2 waves
        // per SIMD and spills.)  The LDS positions come with the ids, so a
        // store never waits on a global load.
        int jn = j0 + (int)threadIdx.x;
        RecF64 nrec;
        unsigned nlw[3], ndw[3];
        // entry j's bbar-info record slots, where an LCE shell's dialled
        // metric is kept (null when the records already hold it)
        auto keep = [&](int j) -> double* {
            if (!LCE || tp.lce_cached) return nullptr;
            return const_cast<double*>(k.tl.rec64) + (size_t)uni(j >> 6) * (REC_SLOTS * 64) + RS_BI * 64 + (j & 63);
        };
        auto fetch = [&](int j) {
            words(k.tl.lpid, j, nlw);
            words(k.tl.dst, j, ndw);
            load_rec_fp64<LCE, UNI, 64>(view(j), tp, (unsigned)j & 63u, nrec, !k.tl.area_ref);
            nrec.dr.wb = keep(j);
        };
        if (!MIXED && jn < j1) fetch(jn);
        // FP64: the node state of the thread's nodes (velocity, mass, clamp
        // flag) is loaded here, in flight through the tile, so that the
        // update at its end waits on no memory (FP64 step 3-5% faster on
        // discs of 18k to 1.8M triangles; the mixed kernel spills at its
        // register limit and loads them at the update)
        double pv[MAX_TILE_CHUNKS][3], pm[MAX_TILE_CHUNKS];
        unsigned char pc[MAX_TILE_CHUNKS];
        if constexpr (!MIXED) {
            const double* V = k.b.v[par];
#pragma unroll
            for (int r = 0; r < MAX_TILE_CHUNKS; ++r) {
                const int i = (int)threadIdx.x + r * B;
                if (i < nn) {
                    const int n = ch.x + i;
                    pv[r][0] = __builtin_nontemporal_load(V + n);
                    pv[r][1] = __builtin_nontemporal_load(V + N + n);
                    pv[r][2] = __builtin_nontemporal_load(V + 2 * N + n);
                    pm[r] = __builtin_nontemporal_load(k.nd.mass + n);
                    pc[r] = k.nd.clamp[n];
                }
            }
        }
        // 1. positions of the tile's nodes and its halo nodes, and the tile's
        // per-pass entry offsets
        for (int i = threadIdx.x; i < nn + (h1 - h0); i += B) {
            const unsigned g = i < nn ? (unsigned)(ch.x + i) : (unsigned)k.tl.halo[h0 + i - nn];
            lpos[3 * i] = ld(X, g);
            lpos[3 * i + 1] = ld(X + N, g);
            lpos[3 * i + 2] = ld(X + 2 * (size_t)N, g);
            if (!MIXED) {
                lref[2 * i] = ld(k.nd.ref, g);
                lref[2 * i + 1] = ld(k.nd.ref + N, g);
            }
        }
        {
            // by 32-bit words: the table's stride is even
            const unsigned* pw = reinterpret_cast<const unsigned*>(poff);
            unsigned* lw = reinterpret_cast<unsigned*>(lpo);
            for (int i = threadIdx.x; i < npass * (stride >> 1); i += B) lw[i] = pw[i];
        }
        __syncthreads();
        double fx[MAX_TILE_CHUNKS], fy[MAX_TILE_CHUNKS], fz[MAX_TILE_CHUNKS];
#pragma unroll
        for (int r = 0; r < MAX_TILE_CHUNKS; ++r) fx[r] = fy[r] = fz[r] = 0.0;
#pragma unroll 1
        for (int pass = 0; pass < npass; ++pass) {
            // 2. this pass's triangles into LDS
            const int ja = j0 + pass * split, jb = min(j1, ja + split);
            if constexpr (MIXED) {
                for (int j = ja + (int)threadIdx.x; j < jb; j += B) {
                    unsigned lw[3], lp[6], dw[3];
                    words(k.tl.lpid, j, lw);
                    unpack(lw, lp);
                    const TilePos pos{lpos, lp};
                    float out[18];
                    RecMixed rec;
                    load_rec_mixed<LCE, UNI, 64>(view(j), tp, (unsigned)j & 63u, rec);
                    rec.dr.wb = keep(j);
                    tri_forces_mixed_rec<LCE>(rec, tp, pos, out);
                    // the mixed kernel, at its 128-register limit, loads the
                    // LDS positions at the store (loaded with the ids: 5% slower)
                    words(k.tl.dst, j, dw);
                    store(dw, reinterpret_cast<const CT*>(out));
                }
            } else {
                for (; jn < jb; jn += B) {
                    unsigned lp[6], dw[3] = {ndw[0], ndw[1], ndw[2]};
                    unpack(nlw, lp);
                    double out[18];
                    tri_forces_fp64_rec<LCE, UNI>(nrec, tp, TilePos{lpos, lp}, TileRef{lref, lp}, out, k.tl.area_ref != 0);
                    store(dw, reinterpret_cast<const CT*>(out));
                    if (jn + B < j1) fetch(jn + B);
                }
            }
            __syncthreads();
            // 3. each node adds this pass's entries, in CSR order
            const unsigned short* po = lpo + pass * stride;
#pragma unroll
            for (int r = 0; r < MAX_TILE_CHUNKS; ++r) {
                const int i = (int)threadIdx.x + r * B;
                if (i >= nn) continue;
                if (MIXED) {
                    double* ls = lsum + 3 * i;
                    double sx = 0.0, sy = 0.0, sz = 0.0;
                    if (pass > 0) {
                        sx = ls[0];
                        sy = ls[1];
                        sz = ls[2];
                    }
                    ordered_sum3_acc(lcon, [](long long e) { return e; }, (long long)po[i], (long long)po[i + 1], sx,
                                     sy, sz);
                    ls[0] = sx;
                    ls[1] = sy;
                    ls[2] = sz;
                } else {
                    ordered_sum3_acc(lcon, [](long long e) { return e; }, (long long)po[i], (long long)po[i + 1],
                                     fx[r], fy[r], fz[r]);
                }
            }
            __syncthreads();  // the next pass (or tile) reuses the LDS
        }
        if (MIXED && npass > 0) {
#pragma unroll
            for (int r = 0; r < MAX_TILE_CHUNKS; ++r) {
                const int i = (int)threadIdx.x + r * B;
                if (i < nn) {
                    fx[r] = lsum[3 * i];
                    fy[r] = lsum[3 * i + 1];
                    fz[r] = lsum[3 * i + 2];
                }
            }
        }
        // each node's update, chunk by chunk (the equilibrium maxima are
        // per-chunk partials)
#pragma unroll
        for (int r = 0; r < MAX_TILE_CHUNKS; ++r) {
            if (r * B >= nn) break;  // block-uniform
            const int i = (int)threadIdx.x + r * B;
            const int n = ch.x + i;
            double fnorm = -1.0, vnorm = -1.0;
            int fidx = INT_MAX, vidx = INT_MAX;
            if (i < nn) {
                double gx = fx[r], gy = fy[r], gz = fz[r];
                const double* V = k.b.v[par];
                double* Xo = k.b.x[par ^ 1];
                double* Vo = k.b.v[par ^ 1];
                // the node state is read and written once per step: non-temporal
                // loads and stores leave the caches to the halo positions
                // (measured: mixed step 3.4% faster, FP64 unchanged)
                double vx, vy, vz, m;
                unsigned char clamped;
                if constexpr (MIXED) {
                    vx = __builtin_nontemporal_load(V + n);
                    vy = __builtin_nontemporal_load(V + N + n);
                    vz = __builtin_nontemporal_load(V + 2 * N + n);
                    m = __builtin_nontemporal_load(k.nd.mass + n);
                    clamped = k.nd.clamp[n];
                } else {
                    vx = pv[r][0];
                    vy = pv[r][1];
                    vz = pv[r][2];
                    m = pm[r];
                    clamped = pc[r];
                }
                if (!sh.gd) {
                    gx = damped(gx, sh.eta, m, vx, sh.damp_den);
                    gy = damped(gy, sh.eta, m, vy, sh.damp_den);
                    gz = damped(gz, sh.eta, m, vz, sh.damp_den);
                }
                if (clamped) gx = gy = gz = 0.0;
                if (do_check) {
                    fnorm = norm3(gx, gy, gz);
                    vnorm = norm3(vx, vy, vz);
                    fidx = vidx = n;
                }
                const double lim = sh.lim;
                if (!(fabs(gx) < lim) || !(fabs(gy) < lim) || !(fabs(gz) < lim))
                    atomicOr(&k.shells[ch.z].failed, 1);
                double x = lpos[3 * i], y = lpos[3 * i + 1], z = lpos[3 * i + 2];
                update(sh, m, gx, vx, x);
                update(sh, m, gy, vy, y);
                update(sh, m, gz, vz, z);
                __builtin_nontemporal_store(x, Xo + n);
                __builtin_nontemporal_store(y, Xo + N + n);
                __builtin_nontemporal_store(z, Xo + 2 * N + n);
                __builtin_nontemporal_store(vx, Vo + n);
                __builtin_nontemporal_store(vy, Vo + N + n);
                __builtin_nontemporal_store(vz, Vo + 2 * N + n);
            }
            if (do_check) {
                double v2[2] = {fnorm, vnorm};
                int i2[2] = {fidx, vidx};
                block_argmax<2>(v2, i2);
                if (threadIdx.x == 0) {
                    const int cc = ch.w + r;
                    k.b.part[2 * cc] = v2[0];
                    k.b.part[2 * cc + 1] = v2[1];
                    k.b.part_i[2 * cc] = i2[0];
                    k.b.part_i[2 * cc + 1] = i2[1];
                }
            }
        }
        __syncthreads();  // the next tile restages the LDS
    }
}

template <typename E>
__global__ __launch_bounds__(256) void tile_gather_kernel(const E* __restrict__ in, E* __restrict__ out, int ncomp,
                                                          int T, int Tp, int slot, const int* __restrict__ tris) {
    const long long total = (long long)ncomp * Tp;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int kc = (int)(i / Tp), j = (int)(i - (long long)kc * Tp);
        const int t = tris[j];
        if (t < 0) continue;  // padding between tiles
        out[((size_t)(j >> 6) * REC_SLOTS + slot + kc) * 64 + (j & 63)] = in[(size_t)kc * T + t];
    }
}

}  // namespace

void launch_tile_gather(hipStream_t s, const void* in, void* out, int elem_bytes, int ncomp, int T, int Tp, int slot,
                        const int* tris) {
    const long long total = (long long)ncomp * Tp;
    if (total <= 0) return;
    const int grid = (int)std::min<long long>(8192, (total + 255) / 256);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(tile_gather_kernel<double>, dim3(grid), dim3(256), 0, s, (const double*)in, (double*)out,
                           ncomp, T, Tp, slot, tris);
    else
        hipLaunchKernelGGL(tile_gather_kernel<float>, dim3(grid), dim3(256), 0, s, (const float*)in, (float*)out,
                           ncomp, T, Tp, slot, tris);
}

template <bool MIXED>
hipError_t raise_lds(int bytes) {
    if (bytes <= 65536) return hipSuccess;
    const void* fns[8] = {(const void*)tile_step_kernel<MIXED, false, false, false>,
                          (const void*)tile_step_kernel<MIXED, true, false, false>,
                          (const void*)tile_step_kernel<MIXED, false, true, false>,
                          (const void*)tile_step_kernel<MIXED, true, true, false>,
                          (const void*)tile_step_kernel<MIXED, false, false, true>,
                          (const void*)tile_step_kernel<MIXED, true, false, true>,
                          (const void*)tile_step_kernel<MIXED, false, true, true>,
                          (const void*)tile_step_kernel<MIXED, true, true, true>};
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

template <bool MIXED, bool LCE, bool UNI>
void launch_tile_variant(const Launch& L, const TileKArgs& k, size_t lds, bool check) {
    if (check)
        hipLaunchKernelGGL((tile_step_kernel<MIXED, true, LCE, UNI>), dim3(L.grid), dim3(L.block), lds, L.stream, k);
    else
        hipLaunchKernelGGL((tile_step_kernel<MIXED, false, LCE, UNI>), dim3(L.grid), dim3(L.block), lds, L.stream, k);
}

void launch_tile_step(const Launch& L, const NodeData& nd, const Buffers& b, const TileData& tl, ShellDev* shells,
                      bool mixed, bool lce, bool uni_data, const StepFlags& fl) {
    TileKArgs k{};
    k.nd = nd;
    k.b = b;
    k.tl = tl;
    k.shells = shells;
    k.fl = fl;
    const size_t lds = (size_t)tile_lds_bytes(tl, mixed);
    const bool check = fl.check_equil != 0;
    const int v = (mixed ? 4 : 0) | (lce ? 2 : 0) | (uni_data ? 1 : 0);
#define MS_TILE_CASE(M, C, U)                                \
    case (M ? 4 : 0) | (C ? 2 : 0) | (U ? 1 : 0):            \
        launch_tile_variant<M, C, U>(L, k, lds, check); \
        break;
    switch (v) {
        MS_TILE_CASE(false, false, false)
        MS_TILE_CASE(false, false, true)
        MS_TILE_CASE(false, true, false)
        MS_TILE_CASE(false, true, true)
        MS_TILE_CASE(true, false, false)
        MS_TILE_CASE(true, false, true)
        MS_TILE_CASE(true, true, false)
        MS_TILE_CASE(true, true, true)
    }
#undef MS_TILE_CASE
}

}  // namespace ms
