// Checked stand-ins for the kernel launches of kernels.h, for the host-only
// sanitizer build of the device context.  Each launch verifies, against the
// live device allocations, every range the real kernel indexes: its chunk
// or tile table, the per-triangle and per-node streams at the indices the
// table names, the partial-result slots it writes, the tile metadata and
// record blocks, and its launch shape (block = chunk size, LDS within the
// limit).  Checks run when the launch runs -- for a captured graph, at every
// replay, so a buffer freed while a graph still names it is caught.  The
// per-step bookkeeping is emulated (buffer parity, step and dial clocks,
// scheduled halts) so the host's tables of running shells are exercised;
// no mechanics is computed.

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

#include "hip/hip_runtime.h"
#include "kernels.h"

namespace {
std::map<int, long long>& halt_schedule() {
    static std::map<int, long long> h;
    return h;
}

[[noreturn]] void bad(const char* what, const char* msg, long long a = 0, long long b = 0) {
    std::fprintf(stderr, "MOCK-KERNEL ERROR in %s: %s (%lld, %lld)\n", what, msg, a, b);
    std::fflush(stderr);
    std::abort();
}

// "<launch>: <expression>" as the name of a checked range
#define WHAT(w, expr) (std::string(w) + ": " + #expr).c_str()

template <class T>
void dev(const T* p, long long n, const char* what, bool allow_null = false) {
    if (n < 0) bad(what, "negative extent", n);
    mock::check_dev(p, (size_t)n * sizeof(T), what, allow_null);
}

void check_launch(const ms::Launch& L, const char* what) {
    if (L.grid < 1) bad(what, "grid < 1", L.grid);
    if (L.block < 64 || L.block > 512 || L.block % 64) bad(what, "block is not a multiple of 64 in [64, 512]", L.block);
}

// The shells table and one chunk table; every chunk lies inside its shell's
// range and fits one block of L.block threads.
void check_chunks(const ms::Launch& L, const ms::Chunk* chunks, int nchunks, const ms::ShellDev* shells,
                  int num_shells, bool tri, const char* what) {
    check_launch(L, what);
    if (nchunks < 0) bad(what, "nchunks < 0", nchunks);
    if (nchunks == 0) return;
    dev(chunks, nchunks, WHAT(what, chunks));
    for (int i = 0; i < nchunks; ++i) {
        const ms::Chunk c = chunks[i];
        if (c.z < 0 || (num_shells > 0 && c.z >= num_shells)) bad(what, "chunk names no shell", i, c.z);
        dev(shells + c.z, 1, WHAT(what, shells + c.z));
        const ms::ShellDev& S = shells[c.z];
        const int b0 = tri ? S.tri_begin : S.node_begin, n0 = tri ? S.num_tris : S.num_nodes;
        if (c.x < b0 || c.y > b0 + n0 || c.x >= c.y) bad(what, "chunk outside its shell", c.x, c.y);
        if (c.y - c.x > L.block) bad(what, "chunk larger than the block", c.y - c.x, L.block);
        const int cb = tri ? S.tri_chunk_begin : S.node_chunk_begin, cn = tri ? S.tri_chunk_count : S.node_chunk_count;
        if (c.w < cb || c.w >= cb + cn) bad(what, "chunk's full-table index outside its shell's chunks", c.w);
    }
}

void check_tri_streams(const ms::TriData& td, bool mixed, bool lce, bool uni, const char* what) {
    const long long T = td.T;
    dev(td.pid, 6 * T, WHAT(what, td.pid));
    for (long long i = 0; i < 6 * T; ++i)
        if (td.pid[i] < 0 || td.pid[i] >= td.N) bad(what, "patch id outside the nodes", i, td.pid[i]);
    if (mixed) {
        dev(td.G32, 6 * T, WHAT(what, td.G32));
        dev(td.H32, 18 * T, WHAT(what, td.H32));
        dev(td.hmuA32, 3 * T, WHAT(what, td.hmuA32));
    } else {
        dev(td.G, 6 * T, WHAT(what, td.G));
        dev(td.H, 18 * T, WHAT(what, td.H));
        dev(td.area, T, WHAT(what, td.area));
        if (!uni) {
            dev(td.h, T, WHAT(what, td.h));
            dev(td.mu, T, WHAT(what, td.mu));
        }
    }
    dev(td.ainfo, (lce ? 5 : 3) * T, WHAT(what, td.ainfo));
    if (!uni) dev(td.binfo, 3 * T, WHAT(what, td.binfo));
    dev(td.a0, 3 * T, WHAT(what, td.a0), true);
    dev(td.b0, 3 * T, WHAT(what, td.b0), true);
}

void check_node_streams(const ms::NodeData& nd, const ms::Buffers& b, const char* what) {
    const long long N = nd.N;
    dev(nd.mass, N, WHAT(what, nd.mass));
    dev(nd.ref, 2LL * N, WHAT(what, nd.ref));
    dev(nd.clamp, N, WHAT(what, nd.clamp));
    dev(nd.csr_off, N + 1, WHAT(what, nd.csr_off));
    dev(nd.csr_ent, nd.csr_off[N], WHAT(what, nd.csr_ent));
    for (int i = 0; i < 2; ++i) {
        dev(b.x[i], 3 * N, WHAT(what, b.x[i]));
        dev(b.v[i], 3 * N, WHAT(what, b.v[i]));
    }
}

// Emulated bookkeeping of one step (MODE 0 of the real kernel): failed shells
// halt; scheduled halts fire; every running shell flips its buffer parity
// and advances its clocks.
void bookkeeping(ms::ShellDev* shells, int num_shells, ms::ClockDev* clock, const char* what) {
    dev(shells, num_shells, WHAT(what, shells));
    dev(clock, 1, WHAT(what, clock));
    const long long step = clock->steps;
    for (int s = 0; s < num_shells; ++s) {
        ms::ShellDev& S = shells[s];
        if (S.halted) continue;
        auto it = halt_schedule().find(s);
        if (S.failed || (it != halt_schedule().end() && step >= it->second)) {
            S.halted = 1;
            S.halt_step = step;
            if (S.failed) S.fail_step = step;
            continue;
        }
        S.par ^= 1;
        S.own_steps += 1;
        S.f_override = -1.0;
        if (S.dialling) S.phase_steps += 1;
    }
    clock->steps = step + 1;
}
}  // namespace

// Halt shell s (emulated bookkeeping) at the first step whose clock is >= step.
extern "C" void mock_schedule_halt(int s, long long step) { halt_schedule()[s] = step; }
extern "C" void mock_clear_halts() { halt_schedule().clear(); }

namespace ms {

void launch_tri_forces(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, bool mixed, bool lce, bool uni_data, bool skip_halted) {
    mock::launch(L.stream, "tri_forces", [=] {
        const char* w = "tri_forces";
        check_chunks(L, chunks, nchunks, shells, 0, true, w);
        check_tri_streams(td, mixed, lce, uni_data, w);
        for (int i = 0; i < 2; ++i) dev(b.x[i], 3LL * td.N, WHAT(w, b.x[i]));
        if (!mixed) dev(td.ref, 2LL * td.N, WHAT(w, td.ref));
        if (mixed)
            dev(b.contrib32, 18LL * td.T, WHAT(w, b.contrib32));
        else
            dev(b.contrib, 18LL * td.T, WHAT(w, b.contrib));
        (void)skip_halted;
    });
}

static void tri_output(const char* w, const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks,
                       int nchunks, ShellDev* shells, bool part) {
    check_chunks(L, chunks, nchunks, shells, 0, true, w);
    check_tri_streams(td, false, td.ainfo && false, false, w);
    for (int i = 0; i < 2; ++i) dev(b.x[i], 3LL * td.N, WHAT(w, b.x[i]));
    if (part)
        for (int i = 0; i < nchunks; ++i) dev(b.part + 4LL * i, 4, WHAT(w, b.part + 4LL * i));
}

void launch_tri_energy(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, double* dens) {
    mock::launch(L.stream, "tri_energy", [=] {
        tri_output("tri_energy", L, td, b, chunks, nchunks, shells, true);
        dev(dens, 2LL * td.T, WHAT("tri_energy dens", dens), true);
    });
}

void launch_tri_geometry(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, double* ga, double* gb, double* gn, int out_begin, int out_stride) {
    mock::launch(L.stream, "tri_geometry", [=] {
        const char* w = "tri_geometry";
        tri_output(w, L, td, b, chunks, nchunks, shells, false);
        // every chunk's triangles land inside [out_begin, out_begin + out_stride)
        for (int i = 0; i < nchunks; ++i)
            if (chunks[i].x < out_begin || chunks[i].y > out_begin + out_stride)
                bad(w, "chunk outside the output range", chunks[i].x, chunks[i].y);
        dev(ga, 3LL * out_stride, WHAT(w, ga));
        dev(gb, 3LL * out_stride, WHAT(w, gb));
        dev(gn, 3LL * out_stride, WHAT(w, gn));
    });
}

void launch_tri_stresses(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, double* out) {
    mock::launch(L.stream, "tri_stresses", [=] {
        tri_output("tri_stresses", L, td, b, chunks, nchunks, shells, false);
        dev(out, 8LL * td.T, "tri_stresses out");
    });
}

void launch_tri_diagnostics(const Launch& L, const TriData& td, const Buffers& b, const Chunk* chunks, int nchunks,
                            ShellDev* shells, double* out) {
    mock::launch(L.stream, "tri_diagnostics", [=] {
        tri_output("tri_diagnostics", L, td, b, chunks, nchunks, shells, false);
        dev(out, 3LL * td.T, WHAT("tri_diagnostics out", out));
    });
}

static void node_common(const char* w, const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks,
                        int nchunks, ShellDev* shells, int num_shells) {
    check_chunks(L, chunks, nchunks, shells, num_shells, false, w);
    check_node_streams(nd, b, w);
}

void launch_node_step(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                      ShellDev* shells, int num_shells, ClockDev* clock, bool mixed, const StepFlags& fl) {
    mock::launch(L.stream, "node_step", [=] {
        const char* w = "node_step";
        node_common(w, L, nd, b, chunks, nchunks, shells, num_shells);
        const long long nent = nd.csr_off[nd.N];
        for (long long e = 0; e < nent; ++e) {
            const int q = nd.csr_ent[e];
            if (q < 0) bad(w, "negative CSR entry", e, q);
            if (mixed)
                dev(b.contrib32 + 3LL * q, 3, WHAT(w, b.contrib32 + 3LL * q));
            else
                dev(b.contrib + 3LL * q, 3, WHAT(w, b.contrib + 3LL * q));
        }
        for (int i = 0; i < nchunks; ++i) {
            dev(b.part + 2LL * chunks[i].w, 2, WHAT(w, b.part + 2LL * chunks[i].w));
            dev(b.part_i + 2LL * chunks[i].w, 2, WHAT(w, b.part_i + 2LL * chunks[i].w));
        }
        (void)fl;
        bookkeeping(shells, num_shells, clock, w);
    });
}

void launch_step_bookkeeping(hipStream_t s, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                             ShellDev* shells, int num_shells, ClockDev* clock, const StepFlags& fl) {
    mock::launch(s, "step_bookkeeping", [=] {
        const char* w = "step_bookkeeping";
        dev(chunks, nchunks, WHAT(w, chunks));
        dev(shells, num_shells, WHAT(w, shells));
        for (int sh = 0; sh < num_shells; ++sh) {
            const ShellDev& S = shells[sh];
            if (S.node_chunk_begin < 0 || S.node_chunk_begin + S.node_chunk_count > nchunks)
                bad(w, "shell's node chunks outside the table", sh);
            dev(b.part + 2LL * S.node_chunk_begin, 2LL * S.node_chunk_count, WHAT(w, b.part + 2LL * S.node_chunk_begin));
            dev(b.part_i + 2LL * S.node_chunk_begin, 2LL * S.node_chunk_count, WHAT(w, b.part_i + 2LL * S.node_chunk_begin));
        }
        (void)nd;
        (void)fl;
        bookkeeping(shells, num_shells, clock, w);
    });
}

void launch_node_gather(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                        ShellDev* shells, bool mixed) {
    mock::launch(L.stream, "node_gather", [=] {
        node_common("node_gather", L, nd, b, chunks, nchunks, shells, 0);
        dev(b.F, 3LL * nd.N, WHAT("node_gather", b.F));
        if (mixed)
            dev(b.contrib32, 3LL * nd.csr_off[nd.N], WHAT("node_gather", b.contrib32), false);
        else
            dev(b.contrib, 3LL * nd.csr_off[nd.N], WHAT("node_gather", b.contrib), false);
    });
}

void launch_node_damping(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells) {
    mock::launch(L.stream, "node_damping", [=] {
        node_common("node_damping", L, nd, b, chunks, nchunks, shells, 0);
        dev(b.F, 3LL * nd.N, WHAT("node_damping", b.F));
    });
}

void launch_node_clamp(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells) {
    mock::launch(L.stream, "node_clamp", [=] {
        node_common("node_clamp", L, nd, b, chunks, nchunks, shells, 0);
        dev(b.F, 3LL * nd.N, WHAT("node_clamp", b.F));
    });
}

void launch_node_scan(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                      ShellDev* shells, int num_shells) {
    mock::launch(L.stream, "node_scan", [=] {
        node_common("node_scan", L, nd, b, chunks, nchunks, shells, num_shells);
        dev(b.F, 3LL * nd.N, WHAT("node_scan", b.F));
        dev(b.bad_dof, num_shells, WHAT("node_scan", b.bad_dof));
        for (int s = 0; s < num_shells; ++s) b.bad_dof[s] = LLONG_MAX;
    });
}

void launch_node_advance(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells, int num_shells, ClockDev* clock) {
    mock::launch(L.stream, "node_advance", [=] {
        node_common("node_advance", L, nd, b, chunks, nchunks, shells, num_shells);
        dev(b.F, 3LL * nd.N, WHAT("node_advance", b.F));
        dev(b.bad_dof, num_shells, WHAT("node_advance", b.bad_dof));
        dev(clock, 1, WHAT("node_advance", clock));
        for (int s = 0; s < num_shells; ++s)
            if (b.bad_dof[s] == LLONG_MAX) {
                shells[s].par ^= 1;
                shells[s].own_steps += 1;
            }
        clock->steps += 1;
    });
}

void launch_node_equil(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                       ShellDev* shells, int num_shells, ClockDev* clock) {
    mock::launch(L.stream, "node_equil", [=] {
        node_common("node_equil", L, nd, b, chunks, nchunks, shells, num_shells);
        dev(b.F, 3LL * nd.N, WHAT("node_equil", b.F));
        dev(clock, 1, WHAT("node_equil", clock));
        for (int i = 0; i < nchunks; ++i) dev(b.part + 2LL * i, 2, WHAT("node_equil", b.part + 2LL * i));
    });
}

void launch_node_kinetic(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks, int nchunks,
                         ShellDev* shells) {
    mock::launch(L.stream, "node_kinetic", [=] {
        node_common("node_kinetic", L, nd, b, chunks, nchunks, shells, 0);
        for (int i = 0; i < nchunks; ++i) dev(b.part + 4LL * i, 2, WHAT("node_kinetic", b.part + 4LL * i));
    });
}

void launch_node_angle_deficits(const Launch& L, const NodeData& nd, const Buffers& b, const Chunk* chunks,
                                int nchunks, ShellDev* shells, const int* pid, int T, const unsigned char* boundary,
                                double* out) {
    mock::launch(L.stream, "node_angle_deficits", [=] {
        const char* w = "node_angle_deficits";
        node_common(w, L, nd, b, chunks, nchunks, shells, 0);
        dev(pid, 6LL * T, WHAT(w, pid));
        for (int i = 0; i < nchunks; ++i) {
            dev(boundary + chunks[i].x, chunks[i].y - chunks[i].x, WHAT(w, boundary));
            dev(out + chunks[i].x, chunks[i].y - chunks[i].x, WHAT(w, out));
        }
    });
}

void launch_reduce_partials(hipStream_t s, const double* part, const ShellDev* shells, int num_shells, bool tri_chunks,
                            int ncomp, double* out) {
    mock::launch(s, "reduce_partials", [=] {
        dev(shells, num_shells, WHAT("reduce_partials", shells));
        dev(out, (long long)num_shells * ncomp, WHAT("reduce_partials", out));
        for (int sh = 0; sh < num_shells; ++sh) {
            const ShellDev& S = shells[sh];
            const int c0 = tri_chunks ? S.tri_chunk_begin : S.node_chunk_begin;
            const int cn = tri_chunks ? S.tri_chunk_count : S.node_chunk_count;
            dev(part + 4LL * c0, 4LL * cn, WHAT("reduce_partials", part + 4LL * c0));
            for (int c = 0; c < ncomp; ++c) out[sh * ncomp + c] = 0.0;
        }
    });
}

long long read_out_floats() { return 256LL * 64 * 256; }

void launch_copy(hipStream_t s, const void* src, void* dst, long long nbytes) {
    mock::launch(s, "copy", [=] {
        mock::check_dev(src, (size_t)nbytes, "copy src");
        mock::check_dev(dst, (size_t)nbytes, "copy dst");
    });
}

void launch_read(hipStream_t s, const void* src, float* out, long long nbytes) {
    mock::launch(s, "read", [=] {
        mock::check_dev(src, (size_t)nbytes, "read src");
        const long long n = nbytes / 16;
        const long long blocks = std::min<long long>((n + 1023) / 1024, 256LL * 64);
        dev(out, blocks * 256, WHAT("read out", out));
    });
}

void launch_tile_gather(hipStream_t s, const void* in, void* out, int elem_bytes, int ncomp, int T, int Tp, int slot,
                        const int* tris) {
    mock::launch(s, "tile_gather", [=] {
        const char* w = "tile_gather";
        if (slot < 0 || slot + ncomp > REC_SLOTS) bad(w, "slots outside the record", slot, ncomp);
        mock::check_dev(in, (size_t)ncomp * T * elem_bytes, w);
        mock::check_dev(out, (size_t)((Tp + 63) / 64) * REC_SLOTS * 64 * elem_bytes, w);
        dev(tris, Tp, WHAT(w, tris));
        for (int j = 0; j < Tp; ++j)
            if (tris[j] >= T) bad(w, "tile entry names no triangle", j, tris[j]);
    });
}

hipError_t prepare_tile_kernels(int lds_bytes_fp64, int lds_bytes_mixed) {
    if (lds_bytes_fp64 > 160 * 1024 || lds_bytes_mixed > 160 * 1024) return hipErrorInvalidValue;
    return hipSuccess;
}

void launch_tile_step(const Launch& L, const NodeData& nd, const Buffers& b, const TileData& tl, ShellDev* shells,
                      bool mixed, bool lce, bool uni_data, const StepFlags& fl) {
    mock::launch(L.stream, "tile_step", [=] {
        const char* w = "tile_step";
        check_launch(L, w);
        check_node_streams(nd, b, w);
        if (tile_lds_bytes(tl, mixed) > 160 * 1024) bad(w, "LDS above 160 KiB", tile_lds_bytes(tl, mixed));
        const int nt = tl.active ? tl.nactive : tl.ntiles;
        if (nt < 0) bad(w, "negative tile count", nt);
        dev(tl.tiles, tl.ntiles, WHAT(w, tl.tiles));
        dev(tl.tri_off, tl.ntiles + 1LL, WHAT(w, tl.tri_off));
        dev(tl.tri_end, tl.ntiles, WHAT(w, tl.tri_end));
        dev(tl.halo_off, tl.ntiles + 1LL, WHAT(w, tl.halo_off));
        if (tl.active) dev(tl.active, nt, WHAT(w, tl.active));
        const long long Tp = tl.tri_off[tl.ntiles];
        dev(tl.tris, Tp, WHAT(w, tl.tris));
        dev(tl.lpid, 6 * Tp, WHAT(w, tl.lpid));
        dev(tl.dst, 6 * Tp, WHAT(w, tl.dst));
        const long long blocks = (Tp + 63) / 64;
        dev(tl.rec64, blocks * REC_SLOTS * 64, WHAT(w, tl.rec64));
        dev(tl.rec32, blocks * REC_SLOTS * 64, WHAT(w, tl.rec32));
        for (int vi = 0; vi < nt; ++vi) {
            const int ci = tl.active ? tl.active[vi] : vi;
            if (ci < 0 || ci >= tl.ntiles) bad(w, "active tile outside the table", vi, ci);
            const int4 ch = tl.tiles[ci];
            dev(shells + ch.z, 1, WHAT(w, shells + ch.z));
            const ShellDev& S = shells[ch.z];
            const int nn = ch.y - ch.x;
            if (nn <= 0 || nn > MAX_TILE_CHUNKS * L.block) bad(w, "tile larger than its chunks", ci, nn);
            if (ch.x < S.node_begin || ch.y > S.node_begin + S.num_nodes) bad(w, "tile outside its shell", ci);
            const int h0 = tl.halo_off[ci], h1 = tl.halo_off[ci + 1];
            dev(tl.halo + h0, h1 - h0, WHAT(w, tl.halo + h0), h1 == h0);
            if (nn + (h1 - h0) > tl.max_nodes) bad(w, "tile reaches more nodes than max_nodes", ci);
            for (int i = h0; i < h1; ++i)
                if (tl.halo[i] < S.node_begin || tl.halo[i] >= S.node_begin + S.num_nodes)
                    bad(w, "halo node outside the tile's shell", ci, tl.halo[i]);
            const int j0 = tl.tri_off[ci], j1 = tl.tri_end[ci];
            if (j0 % 64) bad(w, "tile entries do not start at a multiple of 64", ci, j0);
            if (j1 < j0 || j1 > Tp) bad(w, "tile entry range outside the list", j0, j1);
            const int npass = (j1 - j0 + tl.pass_split - 1) / tl.pass_split;
            if (npass > tl.max_passes) bad(w, "more passes than max_passes", ci, npass);
            dev(tl.poff + (size_t)ci * tl.max_passes * tl.poff_stride, (long long)npass * tl.poff_stride, WHAT(w, tl.poff + (size_t)ci * tl.max_passes * tl.poff_stride),
                npass == 0);
            for (int j = j0; j < j1; ++j) {
                if (tl.tris[j] < S.tri_begin || tl.tris[j] >= S.tri_begin + S.num_tris)
                    bad(w, "tile entry's triangle outside the tile's shell", j, tl.tris[j]);
                for (int s = 0; s < 6; ++s) {
                    if (tl.lpid[6LL * j + s] >= nn + (h1 - h0)) bad(w, "tile-local id outside the LDS nodes", j);
                    const unsigned q = tl.dst[6LL * j + s];
                    if (q != 0xFFFFu && (int)q >= tl.max_entries) bad(w, "LDS position beyond max_entries", j, q);
                }
            }
            for (int r = 0; r < MAX_TILE_CHUNKS && r * L.block < nn; ++r) {
                if (ch.w + r < S.node_chunk_begin || ch.w + r >= S.node_chunk_begin + S.node_chunk_count)
                    bad(w, "tile chunk outside its shell's node chunks", ci, ch.w + r);
                dev(b.part + 2LL * (ch.w + r), 2, WHAT(w, b.part + 2LL * (ch.w + r)));
                dev(b.part_i + 2LL * (ch.w + r), 2, WHAT(w, b.part_i + 2LL * (ch.w + r)));
            }
        }
        (void)lce;
        (void)uni_data;
        (void)fl;
    });
}

}  // namespace ms
