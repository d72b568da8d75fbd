// Per-node kernels of the per-step path (gfx950).  Compiled with
// -ffp-contract=off: the elementwise dynamics round exactly as the
// reference's C++ statements do (one rounding per operation), so given the
// same forces the update is bit-identical to the CPU path.
//
// Force assembly is a deterministic gather: node n sums the contributions
// of every (triangle, patch slot) whose patch contains n, in ascending
// (triangle, slot) order, from the patch-incidence CSR.  This is the row
// structure of the reference's explicit transpose M^T
// (calc_deformation_forces.cpp:185, main.cpp:168-174); vertex incidence
// alone (calc_triangles_incident_on_nodes.cpp) would miss the non-vertex
// patch nodes.
//
// The fused step launch then applies damping (calc_non_deformation_forces
// .cpp:55-57), the clamp (impose_constraints.cpp:61-66), optionally the
// equilibrium max norms (is_at_equilibrium.cpp:62-93), the blow-up test
// !(|F| < 5e4 F_char) (advance_dynamics.cpp:58-73) and the symplectic-Euler
// or gradient-descent update (advance_dynamics.cpp:76-93), writing the other
// of two position/velocity buffers.  A one-block launch after it does the
// per-step bookkeeping: buffer parity, step and dial clocks, equilibrium
// verdicts, halting of failed shells.

#include <limits.h>
#include <math.h>

#include <algorithm>

#include "kernels.h"
#include "node_math.h"

namespace ms {
namespace {

__device__ __forceinline__ int xcd_remap(int i, int n) {
    const int X = 8;
    const int q = n / X, r = n % X;
    const int j = i % X, local = i / X;
    return j * q + (j < r ? j : r) + local;
}

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}
__device__ __forceinline__ void dd_add(double& hi, double& lo, double h2, double l2) {
    double s, e;
    two_sum(hi, h2, s, e);
    e = e + (lo + l2);
    two_sum(s, e, hi, lo);
}

// (max value, first index) block reduction of NV pairs; result in thread 0.
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

struct NodeKArgs {
    NodeData nd;
    Buffers b;
    const Chunk* chunks;
    int nchunks;
    ShellDev* shells;
    int num_shells;
    ClockDev* clock;
    StepFlags fl;
};

// Sum of the patch contributions of node n (fixed CSR order).
template <typename CT>
__device__ __forceinline__ void gather(const NodeData& nd, const CT* contrib, int n, double& fx, double& fy,
                                       double& fz) {
    const int* ent = nd.csr_ent;
    ordered_sum3(contrib, [ent](long long e) { return ent[e]; }, nd.csr_off[n], nd.csr_off[n + 1], fx, fy, fz);
}

// Per-step bookkeeping, one small launch after a step-like node launch (the
// kernel boundary makes the node launch's partials and flags visible, so no
// grid-wide fences are needed).  MODE 0 = fused step, 1 = ops advance (shells
// with a blow-up are not advanced), 2 = equilibrium maxima only.
template <int MODE, bool CHECK>
__global__ __launch_bounds__(256) void bookkeeping_kernel(NodeKArgs k) {
    const long long step = k.clock->steps;
    __syncthreads();
    for (int s = threadIdx.x; s < k.num_shells; s += blockDim.x) {
        ShellDev* S = &k.shells[s];
        const int halted = S->halted;
        if (MODE == 0 && halted) continue;
        if (MODE == 0) {
            // the tile records of generation fl.lce_gen now hold the LCE
            // metric at the dial factor this step used (-1: none do)
            S->lce_f = S->auto_protocol ? S->f_targets : shell_dial_factor(*S);
            S->lce_gen = k.fl.lce_gen;
        }
        const bool proto_check = MODE == 0 && S->auto_protocol && S->check_due;
        if (CHECK || proto_check) {
            double fm = -1.0, vm = -1.0;
            int fi = INT_MAX, vi = INT_MAX;
            const int c0 = S->node_chunk_begin, c1 = c0 + S->node_chunk_count;
            for (int c = c0; c < c1; ++c) {
                const double f = k.b.part[2 * c], v = k.b.part[2 * c + 1];
                const int fii = k.b.part_i[2 * c], vii = k.b.part_i[2 * c + 1];
                if (f > fm) {
                    fm = f;
                    fi = fii;
                }
                if (v > vm) {
                    vm = v;
                    vi = vii;
                }
            }
            const double fr = fm / S->fchar, vr = vm / S->cchar;
            S->max_f = fr;
            S->max_v = vr;
            S->f_node = fi - S->node_begin;
            S->v_node = vi - S->node_begin;
            S->checked = 1;
            const double tf = CHECK ? k.fl.tol_f : S->tol_f, tv = CHECK ? k.fl.tol_v : S->tol_v;
            S->at_eq = (fr < tf && vr < tv) ? 1 : 0;
        }
        if (MODE == 2) continue;
        if (MODE == 0) {
            if (S->failed) {
                S->halted = 1;
                S->fail_step = step;
                S->halt_step = step;
                continue;
            }
            if (CHECK && k.fl.halt_on_eq && S->at_eq) {
                S->halted = 1;
                S->halt_step = step;
                continue;
            }
            if (S->auto_protocol) {
                // main.cpp:396-407: the verdict of this step's check
                const double f_used = shell_dial_factor(*S);
                S->f_override = -1.0;
                bool eq_event = false;
                if (proto_check) {
                    if (k.b.clog) {
                        CheckRec& e = k.b.clog[(size_t)s * k.b.clog_cap + S->num_logged % k.b.clog_cap];
                        e.step = S->own_steps;
                        e.dial_factor = f_used;
                        e.max_force_ratio = S->max_f;
                        e.max_speed_ratio = S->max_v;
                        e.max_force_node = S->f_node;
                        e.max_speed_node = S->v_node;
                        e.at_equilibrium = S->at_eq;
                        e.kind = 0;
                        S->num_logged += 1;
                    }
                    S->since_check = 0;
                    if (S->at_eq) {
                        eq_event = true;
                        S->last_eq_step = S->own_steps;  // in the shell's own step count
                        S->num_eq += 1;
                    }
                }
                if (eq_event) {
                    S->dial_last_eq = f_used;
                    if (f_used >= 1.0 || S->pause_on_eq) {
                        // keep the checked state: the input of this step
                        S->status = f_used >= 1.0 ? 2 : 1;
                        S->halted = 1;
                        S->halt_step = step;
                        S->check_due = 0;
                        continue;
                    }
                    // main.cpp:483-485 computes the next step's factor before
                    // main.cpp:501-507 restarts the phase counter
                    S->f_override = f_used + S->dial_spacing * ((double)S->phase_steps / S->dial_K);
                    S->status = 0;
                    S->dialling = 1;
                    S->phase_steps = 0;
                    S->eta = S->eta1;
                } else {
                    if (S->status == 0) S->phase_steps += 1;
                    if (S->status == 1) S->since_check += 1;
                }
                // main.cpp:293-304 at the start of the next step
                if (S->status == 0 && S->phase_steps >= (long long)S->dial_K) {
                    S->status = 1;
                    S->dialling = 0;
                    S->since_check = 0;
                    S->eta = S->eta2;
                    if (k.b.clog) {
                        CheckRec& e = k.b.clog[(size_t)s * k.b.clog_cap + S->num_logged % k.b.clog_cap];
                        e.step = S->own_steps + 1;
                        e.dial_factor = shell_dial_factor(*S);
                        e.max_force_ratio = e.max_speed_ratio = 0.0;
                        e.max_force_node = e.max_speed_node = -1;
                        e.at_equilibrium = 0;
                        e.kind = 1;
                        S->num_logged += 1;
                    }
                }
                S->check_due = (S->status == 1 && S->since_check > S->equil_interval) ? 1 : 0;
                if (S->status == 0) S->f_targets = shell_dial_factor(*S);
                S->par ^= 1;
                S->own_steps += 1;
                continue;
            }
        }
        if (MODE == 1 && k.b.bad_dof[s] != LLONG_MAX) continue;
        S->par ^= 1;
        S->own_steps += 1;
        S->f_override = -1.0;
        if (S->dialling) S->phase_steps += 1;
    }
    __syncthreads();
    if (threadIdx.x == 0 && MODE != 2) k.clock->steps = step + 1;
}

// ----------------------------------------------------------- fused step
template <bool MIXED, bool CHECK>
__global__ __launch_bounds__(512) void node_step_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const int ci = xcd_remap(vi, k.nchunks);
        const Chunk ch = k.chunks[ci];
        const ShellDev& sh = k.shells[ch.z];
        if (sh.halted) continue;  // block-uniform
        const bool do_check = CHECK || sh.check_due;  // block-uniform
        const int n = ch.x + threadIdx.x;
        double fnorm = -1.0, vnorm = -1.0;
        int fidx = INT_MAX, vidx = INT_MAX;
        if (n < ch.y) {
            const int par = sh.par;
            const double* X = k.b.x[par];
            const double* V = k.b.v[par];
            double* Xo = k.b.x[par ^ 1];
            double* Vo = k.b.v[par ^ 1];
            double fx, fy, fz;
            if (MIXED)
                gather(k.nd, k.b.contrib32, n, fx, fy, fz);
            else
                gather(k.nd, k.b.contrib, n, fx, fy, fz);
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
            if (!(fabs(fx) < lim) || !(fabs(fy) < lim) || !(fabs(fz) < lim))
                atomicOr(&k.shells[ch.z].failed, 1);
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
            if (threadIdx.x == 0) {  // at the chunk's place in the full table
                k.b.part[2 * ch.w] = v2[0];
                k.b.part[2 * ch.w + 1] = v2[1];
                k.b.part_i[2 * ch.w] = i2[0];
                k.b.part_i[2 * ch.w + 1] = i2[1];
            }
        }
    }
}

// ------------------------------------------------------------ ops kernels
template <bool MIXED>
__global__ __launch_bounds__(512) void node_gather_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const Chunk ch = k.chunks[xcd_remap(vi, k.nchunks)];
        const int n = ch.x + threadIdx.x;
        if (n >= ch.y) continue;
        double fx, fy, fz;
        if (MIXED)
            gather(k.nd, k.b.contrib32, n, fx, fy, fz);
        else
            gather(k.nd, k.b.contrib, n, fx, fy, fz);
        k.b.F[n] = fx;
        k.b.F[N + n] = fy;
        k.b.F[2 * N + n] = fz;
    }
}

__global__ __launch_bounds__(512) void node_damping_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const Chunk ch = k.chunks[xcd_remap(vi, k.nchunks)];
        const ShellDev& sh = k.shells[ch.z];
        const int n = ch.x + threadIdx.x;
        if (n >= ch.y || sh.gd) continue;
        const double* V = k.b.v[sh.par];
        const double m = k.nd.mass[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) k.b.F[c * N + n] = damped(k.b.F[c * N + n], sh.eta, m, V[c * N + n], sh.damp_den);
    }
}

__global__ __launch_bounds__(512) void node_clamp_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const Chunk ch = k.chunks[xcd_remap(vi, k.nchunks)];
        const int n = ch.x + threadIdx.x;
        if (n >= ch.y || !k.nd.clamp[n]) continue;
        k.b.F[n] = 0.0;
        k.b.F[N + n] = 0.0;
        k.b.F[2 * N + n] = 0.0;
    }
}

__global__ void fill_ll_kernel(long long* p, int n, long long v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// First offending dof (n_local + c * N_shell) per shell, deterministic min.
__global__ __launch_bounds__(512) void node_scan_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const Chunk ch = k.chunks[xcd_remap(vi, k.nchunks)];
        const ShellDev& sh = k.shells[ch.z];
        const int n = ch.x + threadIdx.x;
        if (n >= ch.y) continue;
        const double lim = sh.lim;
        for (int c = 0; c < 3; ++c) {
            const double f = k.b.F[c * N + n];
            if (!(fabs(f) < lim)) {
                const long long d = (long long)(n - sh.node_begin) + (long long)c * sh.num_nodes;
                atomicMin(&k.b.bad_dof[ch.z], d);
                break;
            }
        }
    }
}

__global__ __launch_bounds__(512) void node_advance_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const Chunk ch = k.chunks[xcd_remap(vi, k.nchunks)];
        const ShellDev& sh = k.shells[ch.z];
        if (k.b.bad_dof[ch.z] != LLONG_MAX) continue;
        const int n = ch.x + threadIdx.x;
        if (n >= ch.y) continue;
        const int par = sh.par;
        const double m = k.nd.mass[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double v = k.b.v[par][c * N + n], x = k.b.x[par][c * N + n];
            update(sh, m, k.b.F[c * N + n], v, x);
            k.b.v[par ^ 1][c * N + n] = v;
            k.b.x[par ^ 1][c * N + n] = x;
        }
    }
}

__global__ __launch_bounds__(512) void node_equil_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const int ci = xcd_remap(vi, k.nchunks);
        const Chunk ch = k.chunks[ci];
        const ShellDev& sh = k.shells[ch.z];
        const int n = ch.x + threadIdx.x;
        double v2[2] = {-1.0, -1.0};
        int i2[2] = {INT_MAX, INT_MAX};
        if (n < ch.y) {
            const double* V = k.b.v[sh.par];
            v2[0] = norm3(k.b.F[n], k.b.F[N + n], k.b.F[2 * N + n]);
            v2[1] = norm3(V[n], V[N + n], V[2 * N + n]);
            i2[0] = i2[1] = n;
        }
        block_argmax<2>(v2, i2);
        if (threadIdx.x == 0) {
            k.b.part[2 * ci] = v2[0];
            k.b.part[2 * ci + 1] = v2[1];
            k.b.part_i[2 * ci] = i2[0];
            k.b.part_i[2 * ci + 1] = i2[1];
        }
    }
}

__global__ __launch_bounds__(512) void node_kinetic_kernel(NodeKArgs k) {
    const int N = k.nd.N;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const int ci = xcd_remap(vi, k.nchunks);
        const Chunk ch = k.chunks[ci];
        const ShellDev& sh = k.shells[ch.z];
        const int n = ch.x + threadIdx.x;
        double hi[1] = {0.0}, lo[1] = {0.0};
        if (n < ch.y) {
            const double* V = k.b.v[sh.par];
            const double vx = V[n], vy = V[N + n], vz = V[2 * N + n];
            hi[0] = (0.5 * k.nd.mass[n]) * (vx * vx + vy * vy + vz * vz);
        }
        block_dd_reduce<1>(hi, lo);
        if (threadIdx.x == 0) {
            k.b.part[4 * ci] = hi[0];
            k.b.part[4 * ci + 1] = lo[0];
        }
    }
}

// Angle deficits (calc_angle_deficits.cpp:40-78): per node 2 pi (pi on the
// boundary) minus the corner angles of its triangles, added in ascending
// triangle order -- the order of the reference's serial loop -- from the
// vertex slots (0-2) of the patch-incidence CSR.  The corner at vertex k is
// acos(-(e_{k-1} . e_k) / (|e_{k-1}| |e_k|)) with e_0 = P_1 - P_0,
// e_1 = P_2 - P_1, e_2 = -e_1 - e_0.
__global__ __launch_bounds__(512) void node_angle_deficit_kernel(NodeKArgs k, const int* pid, int T,
                                                                 const unsigned char* boundary, double* out) {
#pragma clang fp contract(off)
    const int N = k.nd.N;
    const double pi = 3.14159265358979323846;
    for (int vi = blockIdx.x; vi < k.nchunks; vi += gridDim.x) {
        const Chunk ch = k.chunks[xcd_remap(vi, k.nchunks)];
        const ShellDev& sh = k.shells[ch.z];
        const int n = ch.x + threadIdx.x;
        if (n >= ch.y) continue;
        const double* X = k.b.x[sh.par];
        double d = boundary[n] ? pi : 2.0 * pi;
        for (long long e = k.nd.csr_off[n]; e < k.nd.csr_off[n + 1]; ++e) {
            const int ent = k.nd.csr_ent[e];
            const int slot = ent % 6;
            if (slot >= 3) continue;
            const int t = ent / 6;
            double P[3][3];
            for (int v = 0; v < 3; ++v) {
                const int q = pid[(size_t)v * T + t];
                P[v][0] = X[q];
                P[v][1] = X[N + q];
                P[v][2] = X[2 * (size_t)N + q];
            }
            double s[3][3];
            for (int c = 0; c < 3; ++c) {
                s[0][c] = P[1][c] - P[0][c];
                s[1][c] = P[2][c] - P[1][c];
                s[2][c] = -s[1][c] - s[0][c];
            }
            const int km = (slot + 2) % 3;  // the side arriving at this corner
            const double dot = s[km][0] * s[slot][0] + s[km][1] * s[slot][1] + s[km][2] * s[slot][2];
            const double lm = sqrt(s[km][0] * s[km][0] + s[km][1] * s[km][1] + s[km][2] * s[km][2]);
            const double lk = sqrt(s[slot][0] * s[slot][0] + s[slot][1] * s[slot][1] + s[slot][2] * s[slot][2]);
            d += -acos(-dot / (lm * lk));
        }
        out[n] = d;
    }
}

__global__ void reduce_partials_kernel(const double* part, const ShellDev* shells, int num_shells, int tri_chunks,
                                       int ncomp, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= num_shells) return;
    const ShellDev& S = shells[s];
    const int c0 = tri_chunks ? S.tri_chunk_begin : S.node_chunk_begin;
    const int cn = tri_chunks ? S.tri_chunk_count : S.node_chunk_count;
    for (int comp = 0; comp < ncomp; ++comp) {
        double hi = 0.0, lo = 0.0;
        for (int c = c0; c < c0 + cn; ++c) dd_add(hi, lo, part[4 * c + 2 * comp], part[4 * c + 2 * comp + 1]);
        out[s * ncomp + comp] = hi + lo;
    }
}

NodeKArgs make_args(const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks, ShellDev* shells,
                    int num_shells, ClockDev* clock) {
    NodeKArgs k{};
    k.nd = nd;
    k.b = b;
    k.chunks = chunks;
    k.nchunks = nchunks;
    k.shells = shells;
    k.num_shells = num_shells;
    k.clock = clock;
    return k;
}

}  // namespace

void launch_node_step(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                      ShellDev* shells, int num_shells, ClockDev* clock, bool mixed, const StepFlags& fl) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, num_shells, clock);
    k.fl = fl;
    const dim3 g(L.grid), bl(L.block);
    if (mixed) {
        if (fl.check_equil)
            hipLaunchKernelGGL((node_step_kernel<true, true>), g, bl, 0, L.stream, k);
        else
            hipLaunchKernelGGL((node_step_kernel<true, false>), g, bl, 0, L.stream, k);
    } else {
        if (fl.check_equil)
            hipLaunchKernelGGL((node_step_kernel<false, true>), g, bl, 0, L.stream, k);
        else
            hipLaunchKernelGGL((node_step_kernel<false, false>), g, bl, 0, L.stream, k);
    }
    if (fl.check_equil)
        hipLaunchKernelGGL((bookkeeping_kernel<0, true>), dim3(1), dim3(256), 0, L.stream, k);
    else
        hipLaunchKernelGGL((bookkeeping_kernel<0, false>), dim3(1), dim3(256), 0, L.stream, k);
}

void launch_step_bookkeeping(hipStream_t s, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                             ShellDev* shells, int num_shells, ClockDev* clock, const StepFlags& fl) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, num_shells, clock);
    k.fl = fl;
    if (fl.check_equil)
        hipLaunchKernelGGL((bookkeeping_kernel<0, true>), dim3(1), dim3(256), 0, s, k);
    else
        hipLaunchKernelGGL((bookkeeping_kernel<0, false>), dim3(1), dim3(256), 0, s, k);
}

void launch_node_gather(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                        ShellDev* shells, bool mixed) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, 0, nullptr);
    if (mixed)
        hipLaunchKernelGGL(node_gather_kernel<true>, dim3(L.grid), dim3(L.block), 0, L.stream, k);
    else
        hipLaunchKernelGGL(node_gather_kernel<false>, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_node_damping(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, 0, nullptr);
    hipLaunchKernelGGL(node_damping_kernel, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_node_clamp(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, 0, nullptr);
    hipLaunchKernelGGL(node_clamp_kernel, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_node_scan(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                      ShellDev* shells, int num_shells) {
    hipLaunchKernelGGL(fill_ll_kernel, dim3((num_shells + 255) / 256), dim3(256), 0, L.stream, b.bad_dof, num_shells,
                       (long long)LLONG_MAX);
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, num_shells, nullptr);
    hipLaunchKernelGGL(node_scan_kernel, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_node_advance(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, int num_shells, ClockDev* clock) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, num_shells, clock);
    hipLaunchKernelGGL(node_advance_kernel, dim3(L.grid), dim3(L.block), 0, L.stream, k);
    hipLaunchKernelGGL((bookkeeping_kernel<1, false>), dim3(1), dim3(256), 0, L.stream, k);
}

void launch_node_equil(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, int num_shells, ClockDev* clock) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, num_shells, clock);
    k.fl.tol_f = 0.0;
    k.fl.tol_v = 0.0;
    hipLaunchKernelGGL(node_equil_kernel, dim3(L.grid), dim3(L.block), 0, L.stream, k);
    hipLaunchKernelGGL((bookkeeping_kernel<2, true>), dim3(1), dim3(256), 0, L.stream, k);
}

void launch_node_kinetic(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, 0, nullptr);
    hipLaunchKernelGGL(node_kinetic_kernel, dim3(L.grid), dim3(L.block), 0, L.stream, k);
}

void launch_node_angle_deficits(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks,
                                int nchunks, ShellDev* shells, const int* pid, int T, const unsigned char* boundary,
                                double* out) {
    NodeKArgs k = make_args(nd, b, chunks, nchunks, shells, 0, nullptr);
    hipLaunchKernelGGL(node_angle_deficit_kernel, dim3(L.grid), dim3(L.block), 0, L.stream, k, pid, T, boundary, out);
}

void launch_reduce_partials(hipStream_t s, const double* part, const ShellDev* shells, int num_shells, bool tri_chunks,
                            int ncomp, double* out) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((num_shells + 63) / 64), dim3(64), 0, s, part, shells, num_shells,
                       tri_chunks ? 1 : 0, ncomp, out);
}

namespace {
// Each thread keeps four 16-byte loads in flight before storing them (a
// plain one-load grid-stride copy reaches ~12% less on this part).
__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                   long long n) {
    constexpr int U = 4;
    const long long stride = (long long)gridDim.x * blockDim.x * U;
    for (long long b = (long long)blockIdx.x * blockDim.x * U + threadIdx.x; b < n; b += stride) {
        float4 r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = b + (long long)u * blockDim.x;
            if (i < n) r[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = b + (long long)u * blockDim.x;
            if (i < n) dst[i] = r[u];
        }
    }
}
// Read-only stream of the same shape; one partial sum per thread goes to
// out (grid * 256 floats) so the loads are not dropped.
__global__ __launch_bounds__(256) void read_kernel(const float4* __restrict__ src, float* __restrict__ out,
                                                   long long n) {
    constexpr int U = 4;
    const long long stride = (long long)gridDim.x * blockDim.x * U;
    float acc = 0.0f;
    for (long long b = (long long)blockIdx.x * blockDim.x * U + threadIdx.x; b < n; b += stride) {
        float4 r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = b + (long long)u * blockDim.x;
            r[u] = i < n ? src[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += (r[u].x + r[u].y) + (r[u].z + r[u].w);
    }
    out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}
}  // namespace

long long read_out_floats() { return 256LL * 64 * 256; }

void launch_read(hipStream_t s, const void* src, float* out, long long nbytes) {
    const long long n = nbytes / 16;
    const long long blocks = std::min<long long>((n + 1023) / 1024, 256LL * 64);
    hipLaunchKernelGGL(read_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const float4*)src, out, n);
}

void launch_copy(hipStream_t s, const void* src, void* dst, long long nbytes) {
    const long long n = nbytes / 16;
    const long long blocks = std::min<long long>((n + 1023) / 1024, 256LL * 64);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const float4*)src, (float4*)dst, n);
}

}  // namespace ms
