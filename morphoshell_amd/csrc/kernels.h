// Launch interface between the host context (context.cpp) and the device
// kernels (tri_kernels.hip, node_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_types.h"

namespace ms {

// Per-triangle streams.  Index [k * T + t] for k-th component of triangle t
// (SoA over all shells' triangles, T = total triangles of the context).
struct TriData {
    int T;
    int N;                    // total nodes (stride of the position arrays)
    const int* pid;           // 6T   patch node ids (global node index)
    const double* G;          // 6T   [(k*3 + v) * T + t]
    const double* H;          // 18T  [(n*3 + k) * T + t]
    const float* G32;         // 6T   (mixed mode copy, may be null)
    const float* H32;         // 18T
    const double* ainfo;      // acols * T  dial info (blend: abar xx,xy,yy; LCE: c2, cs, s2, lam, -2nu)
    const double* binfo;      // 3T   (blend)
    const double* a0;         // 3T or null (dial from ansatz)
    const double* b0;         // 3T or null
    const double* h;          // T    def_thicknesses
    const double* mu;         // T    def_shear_moduli
    const double* area;       // T    ref_area
    const float* hmuA32;      // 3T   (h, mu, area) in FP32 for mixed mode
    const double* ref;        // 2N   reference x, y of the nodes (SoA, stride N)
};

struct NodeData {
    int N;
    const double* mass;       // N
    const unsigned char* clamp;  // N (1 = constrained)
    const long long* csr_off; // N + 1   patch-incidence CSR offsets
    const int* csr_ent;       // entries t*6 + slot, ascending per node
    const double* ref;        // 2N reference x, y (SoA, stride N)
};

struct Buffers {
    double* x[2];             // 3N positions, two buffers (per-shell parity)
    double* v[2];             // 3N velocities
    double* F;                // 3N forces (ops path)
    double* contrib;          // 18T FP64 per-slot force contributions
    float* contrib32;         // 18T FP32 contributions (mixed mode)
    double* part;             // per-chunk partial sums / maxima scratch
    int* part_i;
    long long* bad_dof;       // per shell, ops-path blow-up scan
    CheckRec* clog;           // per-shell rings of clog_cap protocol events, or null
    int clog_cap;
};

enum TriMode { TRI_FORCES = 0, TRI_ENERGY = 1, TRI_GEOMETRY = 2, TRI_DIAGNOSTICS = 3, TRI_STRESS = 4 };

struct Launch {
    int grid;
    int block;
    hipStream_t stream;
};

// Triangle launches.  skip_halted: ms_step semantics (halted shells skipped).
void launch_tri_forces(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, bool mixed, bool lce, bool uni_data, bool skip_halted);
// Energy densities (dens: 2T, may be null) and per-chunk double-double
// partial sums of A*W_s and A*W_b into b.part[4*chunk .. +3].
void launch_tri_energy(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, double* dens);
// a, b, n of the chunks' triangles; triangle t's component k lands at
// [k * out_stride + t - out_begin] of ga, gb, gn.
void launch_tri_geometry(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, double* ga, double* gb, double* gn, int out_begin, int out_stride);
// Output-time Cauchy stress (calc_stresses.cpp:45-179) into out (8T, SoA):
// eigenvalue 1, eigenvector 1 (3), eigenvalue 2, eigenvector 2 (3), the
// eigenvalues still carrying the reference area (stress_math.h).
void launch_tri_stresses(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, double* out);
// Output-time diagnostics into out (3T): Gauss curvature det(b)/det(a), mean
// curvature tr(a^-1 b)/2 and the strain measure sqrt(tr((a^-1 (a-abar))^2)/8)
// (calc_curvatures.cpp:35-94, calc_strains.cpp:35-78, in invariant form).
void launch_tri_diagnostics(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                            ShellDev* shells, double* out);

// Node launches.
struct StepFlags {
    int check_equil;       // fused equilibrium reduction on this step
    int halt_on_eq;
    double tol_f, tol_v;
    long long lce_gen;     // fused tile step: generation of the tile records whose
                           // LCE metric cache it wrote; -1 otherwise
};
void launch_node_step(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                      ShellDev* shells, int num_shells, ClockDev* clock, bool mixed, const StepFlags& fl);
void launch_node_gather(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                        ShellDev* shells, bool mixed);
void launch_node_damping(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells);
void launch_node_clamp(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells);
// ms_advance_dynamics: blow-up scan into b.bad_dof (per shell, min dof) ...
void launch_node_scan(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                      ShellDev* shells, int num_shells);
// ... then the update of every shell whose bad_dof is none, with bookkeeping.
void launch_node_advance(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, int num_shells, ClockDev* clock);
// Equilibrium maxima from the F buffer into per-shell fields (no halting).
void launch_node_equil(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, int num_shells, ClockDev* clock);
// Kinetic energy partials (double-double) into b.part[4*chunk + 0..1].
void launch_node_kinetic(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells);
// Angle deficits of the chunks' nodes into out[n] (global node index);
// boundary[n] = 1 on boundary nodes (calc_angle_deficits.cpp:40-78).
void launch_node_angle_deficits(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks,
                                int nchunks, ShellDev* shells, const int* pid, int T, const unsigned char* boundary,
                                double* out);
// Per-shell compensated combine of chunk partials: out[shell*ncomp + c] for
// components c < ncomp stored as (hi, lo) pairs at part[4*chunk + 2c].
void launch_reduce_partials(hipStream_t s, const double* part, const ShellDev* shells, int num_shells, bool tri_chunks,
                            int ncomp, double* out);
// Plain 16-byte-vector copy of nbytes (traffic-counter calibration).
void launch_copy(hipStream_t s, const void* src, void* dst, long long nbytes);
// Read-only stream of nbytes; out holds read_out_floats() floats.
void launch_read(hipStream_t s, const void* src, float* out, long long nbytes);
long long read_out_floats();

// Tile-ordered triangle records, blocked by 64 entries: entry j lies in block
// j / 64 at lane j % 64, and a block holds REC_SLOTS component slots of 64
// values each, so the components of 64 consecutive entries are one
// contiguous run and one wave's loads address them from a single base with
// constant offsets.  Slots of the FP64 array rec64 and of the FP32 array
// rec32 (same blocking):
enum RecSlot : int {
    RS_G = 0,       // G (6)          rec64, rec32
    RS_H = 6,       // H (18)         rec64, rec32
    RS_AI = 24,     // dial info (up to 5)    rec64
    RS_BI = 29,     // bbar info (3)          rec64
    RS_THICK = 32,  // h                      rec64
    RS_MU = 33,     // mu                     rec64
    RS_AREA = 34,   // ref area               rec64
    RS_A0 = 35,     // ansatz a (3)           rec64
    RS_B0 = 38,     // ansatz b (3)           rec64
    RS_HMUA = 24,   // (h, mu, area) (3)      rec32
    REC_SLOTS = 41,
};

// Fused owner-computes step.  Tile c = tiles[c] = {first node, end node,
// shell, first chunk}: up to MAX_TILE_CHUNKS consecutive node chunks of one
// shell (B = block size nodes each; node i of the tile lies in chunk
// first + i / B).  Entries j in [tri_off[c], tri_end[c]) are the triangles
// (ascending) whose patches touch the tile's nodes; rec64/rec32 hold a copy
// of their per-triangle records at entry j (halo triangles appear in several
// tiles), so every tile streams its records contiguously.  tri_off[c] is a
// multiple of 64 (the entry lists are padded between tiles), so with passes
// of a multiple of 64 entries each wave's entries form one record block.
// The tile's nodes and the
// halo nodes [halo_off[c], halo_off[c+1]) of its triangles are staged in LDS
// (tile-local ids: own node n -> n - first, halo node k -> own count + k);
// lpid[6j + slot] is the tile-local id of a patch node.  The triangles run
// in passes of pass_split entries; dst[6j + slot] is the pass-local LDS
// position of the patch-incidence entry (j, slot), or 0xFFFF when that node
// lies outside the tile, and poff[(c * max_passes + pass) * poff_stride + i]
// is the first pass-local position of own node i's entries.
constexpr int MAX_TILE_CHUNKS = 2;
struct TileData {
    const int4* tiles;             // ntiles
    int ntiles;
    const int* tri_off;            // ntiles + 1 (first entry of each tile, multiple of 64)
    const int* tri_end;            // ntiles (one past the tile's last entry)
    const int* tris;               // triangle of entry j (-1 for padding)
    const unsigned short* lpid;    // 6 per entry
    const unsigned short* dst;     // 6 per entry
    const int* halo_off;           // ntiles + 1
    const int* halo;               // global ids of halo nodes
    const unsigned short* poff;    // max_passes * poff_stride per tile
    int poff_stride;               // chunk nodes per tile + 2 (even: tables start on 32-bit words)
    int pass_split;                // entries per pass (B or 2B)
    int max_passes;
    int max_nodes;                 // max over tiles of own + halo nodes
    int max_entries;               // max over tiles and passes of pass entries
    const double* rec64;           // REC_SLOTS * 64 doubles per block of 64 entries
    const float* rec32;            // REC_SLOTS * 64 floats per block
    const int* active;             // null: every tile; else the tiles to run (nactive)
    int nactive;
    long long gen;                 // generation of the record copies (bumped at every rebuild)
    int area_ref;                  // 1: the FP64 step forms each reference area from the staged
                                   // reference positions (equal to the area stream, checked at
                                   // set-up) and does not read the area slot
};
// Record gather: component k < ncomp of stream `in` (SoA, [k*T + t]) into
// slot `slot + k` of every entry j < Tp with tris[j] >= 0 (elem_bytes 4 or 8).
void launch_tile_gather(hipStream_t s, const void* in, void* out, int elem_bytes, int ncomp, int T, int Tp, int slot,
                        const int* tris);
// Dynamic LDS of a tile launch: positions of the tile's own and halo nodes
// (24 B each; in FP64 also their reference x, y, 16 B), in mixed mode the own
// nodes' running sums (24 B each), one pass of contributions (3 values per
// entry), and the tile's per-pass entry offsets (16-bit, max_passes rows of
// poff_stride).
inline int tile_lds_bytes(const TileData& tl, bool mixed) {
    return (mixed ? 24 : 40) * tl.max_nodes + (mixed ? 24 * tl.poff_stride : 0) + tl.max_entries * 3 * (mixed ? 4 : 8) +
           ((2 * tl.max_passes * tl.poff_stride + 7) & ~7);
}
// Raise the dynamic-LDS limit of the tile kernels where a tile needs > 64 KiB.
hipError_t prepare_tile_kernels(int lds_bytes_fp64, int lds_bytes_mixed);
// One fused step of every shell not halted (the work of launch_tri_forces +
// launch_node_step's node kernel), followed by launch_step_bookkeeping.
// lce: some shell of the context uses the LCE dial protocol (selects the
// kernel variant that evaluates it).  uni_data: every shell has uniform
// thickness, shear modulus and (blend protocol) bbar_info (ShellDev h_u, mu_u,
// b_u), so the variant that reads them from the shell record runs.
void launch_tile_step(const Launch& L, const NodeData& nd, const Buffers& b, const TileData& tl, ShellDev* shells,
                      bool mixed, bool lce, bool uni_data, const StepFlags& fl);
// The per-step bookkeeping launch of launch_node_step (one block).
void launch_step_bookkeeping(hipStream_t s, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                             ShellDev* shells, int num_shells, ClockDev* clock, const StepFlags& fl);

}  // namespace ms
