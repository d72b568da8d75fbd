// Device context of the per-step path: owns the device copies of every
// shell's per-triangle and per-node streams, the two position/velocity
// buffers, the chunk tables of the ragged batch, the HIP stream and the
// captured step graphs, and implements the C-ABI entry points that mirror the
// reference's per-step calls.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "ms_common.h"

using namespace ms;

#define MS_VERSION "0.1.0"

struct ms_ctx {
    std::string err;
    int precision = MS_FP64;
    int num_shells = 0;
    int T = 0, N = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;

    std::vector<ShellDev> hs;  // host mirror of the per-shell records
    ShellDev* d_shells = nullptr;
    ClockDev* d_clock = nullptr;
    long long steps = 0;       // host mirror of the step clock

    // device streams
    int* pid = nullptr;
    double *G = nullptr, *H = nullptr, *ainfo = nullptr, *binfo = nullptr, *a0 = nullptr, *b0 = nullptr;
    double *h = nullptr, *mu = nullptr, *area = nullptr;
    float *G32 = nullptr, *H32 = nullptr, *hmuA32 = nullptr;
    double* mass = nullptr;
    double* ref = nullptr;  // 2N reference x, y (SoA)
    unsigned char* clamp = nullptr;
    long long* csr_off = nullptr;
    int* csr_ent = nullptr;
    Buffers b{};
    double* scratch3 = nullptr;  // 9T (geometry outputs) / 2T densities
    double* node_out = nullptr;  // N (angle deficits), allocated on first use
    unsigned char* node_flags = nullptr;  // N boundary flags of the angle deficits
    double* red_out = nullptr;   // per-shell reduction outputs (3 per shell)

    // launch shapes
    ms_tuning tune{};
    Chunk* d_tri_chunks = nullptr;
    Chunk* d_node_chunks = nullptr;
    int n_tri_chunks = 0, n_node_chunks = 0;
    size_t part_len = 0;
    std::vector<Chunk> h_tri_chunks, h_node_chunks;  // host copies of the full tables
    // tiles of the fused step: {first node, end node, shell, first chunk} of
    // each group of tile_L nodes (cut into node chunks by build_chunks)
    std::vector<int4> h_tile_groups;
    int tile_L = 0;

    // Step launches over the running shells only (see refresh_active): the
    // running shells' chunks and tiles, rebuilt when the set of halted shells
    // changes; act_all: every shell runs and the full tables are used.
    std::vector<char> act_live;      // running flags the tables were built for (empty: stale)
    bool act_all = true;
    Chunk* d_act_tri = nullptr;
    Chunk* d_act_node = nullptr;
    int* d_act_tiles = nullptr;
    int n_act_tri = 0, n_act_node = 0, n_act_tiles = 0;

    // fused tile step (built on demand for the current node chunks)
    TileData tile{};                 // device metadata + tile-ordered records
    int tile_T = 0;                  // total tile entries (triangles counted per tile)
    long long tile_gen = 0;          // record generations built so far
    bool tiles_ready = false;
    std::vector<void*> tile_meta;    // metadata allocations
    std::vector<void*> tile_allocs;  // tile-ordered record copies
    int64_t tile_bytes = 0;          // metadata
    int64_t tile_rec_bytes = 0;      // record copies
    int acols = 3;
    bool uni_data = true;            // see uni_data()
    bool area_ref = false;           // see TileData::area_ref

    // graphs
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
    int graph_precision = -1;

    // timing
    bool timing = false;
    double last_ms[3] = {0, 0, 0};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};

    int64_t dev_bytes = 0;
    std::vector<void*> allocs;

    // protocol event log (ms_set_check_log): per-shell rings, read positions
    std::vector<long long> clog_read;

    // this device's address of the static scratch array of the allocation-free
    // entry point (null, with the reason in noalloc_err, if it could not be taken)
    double* noalloc_scratch = nullptr;
    std::string noalloc_err;
};

namespace {

// Scratch of ms_calc_geometry_noalloc: a, b and n (3 values each) of up to
// MS_NOALLOC_MAX_TRIS triangles, 36 MiB in every device's copy of this code
// object, held until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: only the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch_for)
// is passed to kernels and copies.
constexpr size_t NOALLOC_VALS = 9 * (size_t)MS_NOALLOC_MAX_TRIS;
static __device__ __align__(256) double scratch[NOALLOC_VALS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on device dev (the current device), taken
// once per device.
hipError_t noalloc_scratch_for(int dev, double** out) {
    *out = nullptr;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_noalloc_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return hipSuccess;
}

int fail(ms_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    set_last_error(msg);
    return code;
}

#define HIPC(ctx, expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return fail((ctx), MS_ERR_HIP, fmt("%s: %s", #expr, hipGetErrorString(e_))); \
    } while (0)

template <class T>
hipError_t dalloc(ms_ctx* c, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e == hipSuccess) {
        c->allocs.push_back(*p);
        c->dev_bytes += (int64_t)(n * sizeof(T));
    }
    return e;
}

// Chunk tables for blocks of tb triangles and nb nodes, all-or-nothing: the
// new tables and partial buffers are allocated and filled before the old
// ones are freed, so a failed allocation leaves the context on its current
// tables (and its current tiles and graph, which name them).  Each shell's
// nodes are cut into groups of L nodes (the tiles of the fused step) and each
// group into chunks of nb, so chunk r of a tile starts r * nb nodes into it.
int build_chunks(ms_ctx* c, int tb, int nb, int L) {
    std::vector<Chunk> tc, nc;
    std::vector<int4> groups;
    std::vector<ShellDev> hs = c->hs;
    for (int s = 0; s < c->num_shells; ++s) {
        ShellDev& S = hs[s];
        S.tri_chunk_begin = (int)tc.size();
        for (int t = S.tri_begin; t < S.tri_begin + S.num_tris; t += tb)
            tc.push_back(make_int4(t, std::min(t + tb, S.tri_begin + S.num_tris), s, (int)tc.size()));
        S.tri_chunk_count = (int)tc.size() - S.tri_chunk_begin;
        S.node_chunk_begin = (int)nc.size();
        const int nend = S.node_begin + S.num_nodes;
        for (int g = S.node_begin; g < nend; g += L) {
            const int gend = std::min(g + L, nend);
            groups.push_back(make_int4(g, gend, s, (int)nc.size()));
            for (int n = g; n < gend; n += nb) nc.push_back(make_int4(n, std::min(n + nb, gend), s, (int)nc.size()));
        }
        S.node_chunk_count = (int)nc.size() - S.node_chunk_begin;
    }
    const size_t part_len = 4 * std::max(tc.size(), nc.size());
    Chunk *dt = nullptr, *dn = nullptr, *at = nullptr, *an = nullptr;
    double* part = nullptr;
    int* part_i = nullptr;
    hipError_t e = hipMalloc((void**)&dt, tc.size() * sizeof(Chunk));
    if (e == hipSuccess) e = hipMalloc((void**)&dn, nc.size() * sizeof(Chunk));
    if (e == hipSuccess) e = hipMalloc((void**)&at, tc.size() * sizeof(Chunk));
    if (e == hipSuccess) e = hipMalloc((void**)&an, nc.size() * sizeof(Chunk));
    if (e == hipSuccess) e = hipMalloc((void**)&part, part_len * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&part_i, part_len * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(dt, tc.data(), tc.size() * sizeof(Chunk), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, nc.data(), nc.size() * sizeof(Chunk), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        for (void* p : {(void*)dt, (void*)dn, (void*)at, (void*)an, (void*)part, (void*)part_i})
            if (p) hipFree(p);
        return fail(c, MS_ERR_HIP, fmt("chunk tables: %s", hipGetErrorString(e)));
    }
    for (void* p : {(void*)c->d_tri_chunks, (void*)c->d_node_chunks, (void*)c->d_act_tri, (void*)c->d_act_node,
                    (void*)c->b.part, (void*)c->b.part_i})
        if (p) hipFree(p);
    c->d_tri_chunks = dt;
    c->d_node_chunks = dn;
    c->d_act_tri = at;
    c->d_act_node = an;
    c->b.part = part;
    c->b.part_i = part_i;
    c->part_len = part_len;
    c->n_tri_chunks = (int)tc.size();
    c->n_node_chunks = (int)nc.size();
    c->h_tri_chunks = std::move(tc);
    c->h_node_chunks = std::move(nc);
    c->h_tile_groups = std::move(groups);
    c->tile_L = L;
    c->hs = std::move(hs);
    c->act_live.clear();
    c->act_all = true;
    c->tiles_ready = false;  // tiles are built over the node chunks
    return MS_OK;
}

void free_tile_records(ms_ctx* c) {
    for (void* p : c->tile_allocs) hipFree(p);
    c->tile_allocs.clear();
    c->tile.rec64 = nullptr;
    c->tile.rec32 = nullptr;
    c->dev_bytes -= c->tile_rec_bytes;
    c->tile_rec_bytes = 0;
}

void free_tiles(ms_ctx* c) {
    free_tile_records(c);
    for (void* p : c->tile_meta) hipFree(p);
    c->tile_meta.clear();
    c->tile = TileData{};
    c->tile_T = 0;
    c->dev_bytes -= c->tile_bytes;
    c->tile_bytes = 0;
    c->tiles_ready = false;
    c->d_act_tiles = nullptr;  // one of the metadata allocations
    c->act_live.clear();
}

// Tile-ordered copies of the per-triangle streams the mechanics reads, in
// the blocked record layout (see RecSlot; gathered on the device; the patch
// ids are replaced by tile-local ids).
int build_tile_records(ms_ctx* c) {
    const int T = c->T, Tp = c->tile_T;
    const size_t vals = (size_t)((Tp + 63) / 64) * REC_SLOTS * 64;
    void *r64 = nullptr, *r32 = nullptr;
    hipError_t ea = hipMalloc(&r64, vals * sizeof(double));
    if (ea == hipSuccess) ea = hipMalloc(&r32, vals * sizeof(float));
    if (ea != hipSuccess) {  // the current records (if any) stay in place
        if (r64) hipFree(r64);
        return fail(c, MS_ERR_HIP, fmt("tile records: %s", hipGetErrorString(ea)));
    }
    free_tile_records(c);
    c->tile_allocs = {r64, r32};
    c->tile_rec_bytes = (int64_t)(vals * (sizeof(double) + sizeof(float)));
    c->dev_bytes += c->tile_rec_bytes;
    c->tile.rec64 = (const double*)r64;
    c->tile.rec32 = (const float*)r32;
    c->tile.gen = ++c->tile_gen;  // LCE metric caches of older records are void
    auto gather = [&](const void* in, void* out, int elem_bytes, int ncomp, int slot) {
        if (in) launch_tile_gather(c->stream, in, out, elem_bytes, ncomp, T, Tp, slot, c->tile.tris);
    };
    gather(c->G, r64, 8, 6, RS_G);
    gather(c->H, r64, 8, 18, RS_H);
    gather(c->ainfo, r64, 8, c->acols, RS_AI);
    gather(c->binfo, r64, 8, 3, RS_BI);
    gather(c->h, r64, 8, 1, RS_THICK);
    gather(c->mu, r64, 8, 1, RS_MU);
    gather(c->area, r64, 8, 1, RS_AREA);
    gather(c->a0, r64, 8, 3, RS_A0);
    gather(c->b0, r64, 8, 3, RS_B0);
    gather(c->G32, r32, 4, 6, RS_G);
    gather(c->H32, r32, 4, 18, RS_H);
    gather(c->hmuA32, r32, 4, 3, RS_HMUA);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    return MS_OK;
}

bool has_lce(const ms_ctx* c);

// Tile metadata of the fused step for the current node chunks (see TileData):
// per tile the ascending triangles whose patch touches one of its nodes, the
// halo nodes they reach, tile-local patch ids, and per pass the LDS positions
// of the patch-incidence entries and each node's first position.
int build_tiles(ms_ctx* c) {
    free_tiles(c);
    const int T = c->T, N = c->N;
    const int B = c->tune.node_block;
    const bool mixed = c->precision == MS_MIXED;
    std::vector<int> pid((size_t)6 * T), ent;
    std::vector<long long> off((size_t)N + 1);
    HIPC(c, hipMemcpy(pid.data(), c->pid, pid.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIPC(c, hipMemcpy(off.data(), c->csr_off, off.size() * sizeof(long long), hipMemcpyDeviceToHost));
    ent.resize((size_t)off[N]);
    HIPC(c, hipMemcpy(ent.data(), c->csr_ent, ent.size() * sizeof(int), hipMemcpyDeviceToHost));
    // tiles: the node groups the chunks were cut from (at most tile_chunks
    // chunks each)
    const int npt = std::max(1, std::min(MAX_TILE_CHUNKS, c->tune.tile_chunks));
    const std::vector<int4>& tiles = c->h_tile_groups;
    // per pass npt * B + 1 offsets, padded to an even count so that every
    // tile's table starts on a 32-bit word (the kernel stages it by words)
    const int stride = npt * B + 2;
    // Each tile's triangles, and for the two pass lengths 2B and B (a
    // multiple of 64: each wave's entries are one record block) the LDS a
    // tile needs: its nodes and halo nodes, the largest pass's entries and
    // the offset table.
    std::vector<std::vector<int>> tile_tris(tiles.size());
    const int cand[2] = {2 * B, B};
    int cand_entries[2] = {0, 0}, cand_passes[2] = {1, 1}, most_nodes = 0;
    {
        std::vector<int> hl, pc;
        for (size_t ci = 0; ci < tiles.size(); ++ci) {
            std::vector<int>& tl = tile_tris[ci];
            const int n0 = tiles[ci].x, n1 = tiles[ci].y;
            for (long long e = off[n0]; e < off[n1]; ++e) tl.push_back(ent[e] / 6);
            std::sort(tl.begin(), tl.end());
            tl.erase(std::unique(tl.begin(), tl.end()), tl.end());
            hl.clear();
            for (int t : tl)
                for (int sl = 0; sl < 6; ++sl) {
                    const int p = pid[(size_t)sl * T + t];
                    if (p < n0 || p >= n1) hl.push_back(p);
                }
            std::sort(hl.begin(), hl.end());
            most_nodes = std::max(most_nodes, n1 - n0 + (int)(std::unique(hl.begin(), hl.end()) - hl.begin()));
            for (int k = 0; k < 2; ++k) {
                const int sp = cand[k];
                pc.assign((tl.size() + sp - 1) / sp, 0);
                cand_passes[k] = std::max(cand_passes[k], (int)pc.size());
                for (long long e = off[n0]; e < off[n1]; ++e)
                    ++pc[(std::lower_bound(tl.begin(), tl.end(), ent[e] / 6) - tl.begin()) / sp];
                for (int v : pc) cand_entries[k] = std::max(cand_entries[k], v);
            }
        }
    }
    // Passes of 2B triangles (half the barriers of B; measured 3% faster in
    // FP64 at one workgroup per CU) unless their LDS would cost a workgroup
    // of those the kernel's registers allow per CU: 512 threads in FP64 (2
    // waves per SIMD), 1024 in mixed mode (4; 768 with LCE shells).  FP64
    // 256-thread blocks of two-chunk (512-node) tiles thus take B and run two
    // workgroups per CU: 0.153 against 0.163 ms per step of the 1.8M-triangle
    // disc at one 512-thread workgroup of 2B passes; mixed mode at 512-node
    // tiles takes B (2B measured 28% slower).  Chosen when the tiles are
    // built; a later precision switch keeps them (correct either way).
    const int resident = std::max(1, (mixed ? (has_lce(c) ? 768 : 1024) : 512) / B);
    int pick = 1;
    {
        TileData probe{};
        probe.poff_stride = stride;
        probe.max_nodes = most_nodes;
        probe.max_entries = cand_entries[0];
        probe.max_passes = cand_passes[0];
        if (resident * (tile_lds_bytes(probe, mixed) + 1024) <= 160 * 1024) pick = 0;
    }
    const int split = cand[pick];
    const int max_passes = cand_passes[pick];
    std::vector<int> toff(tiles.size() + 1, 0), tend(tiles.size(), 0), tris, hoff(tiles.size() + 1, 0), halo;
    std::vector<unsigned short> lpid, dst, poff(tiles.size() * (size_t)max_passes * stride, 0);
    int max_entries = 0, max_nodes = 0;
    std::vector<int> hl;
    std::vector<std::vector<int>> cnt;
    for (size_t ci = 0; ci < tiles.size(); ++ci) {
        const int n0 = tiles[ci].x, n1 = tiles[ci].y, nn = n1 - n0;
        if (nn > npt * B) return fail(c, MS_ERR_ARG, "fused step: a tile is larger than its chunks");
        // the kernel takes chunk r of a tile to start r * B nodes into it
        for (int r = 0; r * B < nn; ++r) {
            const int k = tiles[ci].w + r;
            if (k >= c->n_node_chunks || c->h_node_chunks[k].x != n0 + r * B ||
                c->h_node_chunks[k].y != std::min(n0 + (r + 1) * B, n1))
                return fail(c, MS_ERR_ARG, "fused step: a tile's chunks do not match its nodes");
        }
        const std::vector<int>& tl = tile_tris[ci];
        const int npass = (int)((tl.size() + split - 1) / split);
        hl.clear();
        for (int t : tl)
            for (int s = 0; s < 6; ++s) {
                const int p = pid[(size_t)s * T + t];
                if (p < n0 || p >= n1) hl.push_back(p);
            }
        std::sort(hl.begin(), hl.end());
        hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
        if (nn + (long long)hl.size() > 0xFFFF)
            return fail(c, MS_ERR_ARG, fmt("fused step: tile %zu reaches too many nodes", ci));
        max_nodes = std::max(max_nodes, nn + (int)hl.size());
        // per node and pass, the number of its entries: an entry's pass is its
        // triangle's position in the tile list over `split`; along the node's
        // CSR range (ascending triangles) the pass never decreases
        cnt.assign(npass, std::vector<int>(nn, 0));
        for (int i = 0; i < nn; ++i)
            for (long long e = off[n0 + i]; e < off[n0 + i + 1]; ++e) {
                const int jj = (int)(std::lower_bound(tl.begin(), tl.end(), ent[e] / 6) - tl.begin());
                cnt[jj / split][i] += 1;
            }
        unsigned short* pbase = &poff[ci * (size_t)max_passes * stride];
        // the tile's entries start at a multiple of 64 (padding entries: no
        // triangle, no own slot)
        while (tris.size() % 64) {
            tris.push_back(-1);
            for (int s = 0; s < 6; ++s) {
                lpid.push_back(0);
                dst.push_back(0xFFFF);
            }
        }
        toff[ci] = (int)tris.size();
        for (int p = 0; p < npass; ++p) {
            unsigned short* po = pbase + (size_t)p * stride;
            for (int i = 0; i < nn; ++i) {
                const long long a = po[i] + (long long)cnt[p][i];
                if (a > 0xFFFE) return fail(c, MS_ERR_ARG, fmt("fused step: tile %zu has too many patch entries", ci));
                po[i + 1] = (unsigned short)a;
            }
            max_entries = std::max(max_entries, (int)po[nn]);
        }
        for (size_t jj = 0; jj < tl.size(); ++jj) {
            const int t = tl[jj];
            const int pass = (int)jj / split;
            tris.push_back(t);
            for (int s = 0; s < 6; ++s) {
                const int p = pid[(size_t)s * T + t];
                unsigned short q = 0xFFFF, l;
                if (p >= n0 && p < n1) {
                    const int i = p - n0;
                    l = (unsigned short)i;
                    const int key = t * 6 + s;
                    const int* b = ent.data() + off[p];
                    const int* e = ent.data() + off[p + 1];
                    const int* it = std::lower_bound(b, e, key);
                    if (it == e || *it != key)
                        return fail(c, MS_ERR_ARG, "fused step: patch-incidence table is inconsistent");
                    int earlier = 0;
                    for (int pp = 0; pp < pass; ++pp) earlier += cnt[pp][i];
                    const int rank = (int)(it - b) - earlier;
                    q = (unsigned short)(pbase[(size_t)pass * stride + i] + rank);
                } else {
                    l = (unsigned short)(nn + (std::lower_bound(hl.begin(), hl.end(), p) - hl.begin()));
                }
                lpid.push_back(l);
                dst.push_back(q);
            }
        }
        tend[ci] = (int)tris.size();
        halo.insert(halo.end(), hl.begin(), hl.end());
        hoff[ci + 1] = (int)halo.size();
    }
    toff[tiles.size()] = (int)tris.size();
    if (tris.size() > (size_t)INT_MAX / 6)
        return fail(c, MS_ERR_ARG, "fused step: too many tile entries for 32-bit indexing");
    TileData& td = c->tile;
    td.poff_stride = stride;
    td.pass_split = split;
    td.max_passes = max_passes;
    td.max_nodes = max_nodes;
    td.max_entries = max_entries;
    td.area_ref = c->area_ref ? 1 : 0;
    const int lds64 = tile_lds_bytes(td, false), lds32 = tile_lds_bytes(td, true);
    if (lds64 + 1024 > 160 * 1024)
        return fail(c, MS_ERR_ARG, fmt("fused step: a tile needs %d bytes of LDS (limit 160 KiB)", lds64));
    HIPC(c, prepare_tile_kernels(lds64, lds32));
    auto upload = [&](const auto& v, auto** dptr) -> hipError_t {
        using E = typename std::decay<decltype(v)>::type::value_type;
        void* p = nullptr;
        const size_t nb = std::max<size_t>(1, v.size()) * sizeof(E);
        hipError_t e = hipMalloc(&p, nb);
        if (e != hipSuccess) return e;
        c->tile_meta.push_back(p);
        c->tile_bytes += (int64_t)nb;
        c->dev_bytes += (int64_t)nb;
        *dptr = static_cast<typename std::remove_pointer<decltype(dptr)>::type>(p);
        return v.empty() ? hipSuccess : hipMemcpy(p, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice);
    };
    int *d_toff, *d_tend, *d_tris, *d_hoff, *d_halo, *d_act;
    int4* d_tiles;
    unsigned short *d_lpid, *d_dst, *d_poff;
    HIPC(c, upload(tiles, &d_tiles));
    HIPC(c, upload(std::vector<int>(tiles.size(), 0), &d_act));  // room for the running tiles' list
    HIPC(c, upload(toff, &d_toff));
    HIPC(c, upload(tend, &d_tend));
    HIPC(c, upload(tris, &d_tris));
    HIPC(c, upload(hoff, &d_hoff));
    HIPC(c, upload(halo, &d_halo));
    HIPC(c, upload(lpid, &d_lpid));
    HIPC(c, upload(dst, &d_dst));
    HIPC(c, upload(poff, &d_poff));
    td.tiles = d_tiles;
    td.ntiles = (int)tiles.size();
    td.tri_off = d_toff;
    td.tri_end = d_tend;
    td.tris = d_tris;
    td.halo_off = d_hoff;
    td.halo = d_halo;
    td.lpid = d_lpid;
    td.dst = d_dst;
    td.poff = d_poff;
    td.active = nullptr;
    td.nactive = 0;
    c->d_act_tiles = d_act;
    c->act_live.clear();
    c->tile_T = (int)tris.size();
    if (int r = build_tile_records(c)) return r;
    c->tiles_ready = true;
    return MS_OK;
}

// Host mirror <-> device records.
int pull_shells(ms_ctx* c) {
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipMemcpy(c->hs.data(), c->d_shells, c->hs.size() * sizeof(ShellDev), hipMemcpyDeviceToHost));
    return MS_OK;
}
int push_shells(ms_ctx* c) {
    // a record the host writes may change what the LCE metric cache holds
    for (ShellDev& S : c->hs) S.lce_gen = -1;
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipMemcpy(c->d_shells, c->hs.data(), c->hs.size() * sizeof(ShellDev), hipMemcpyHostToDevice));
    return MS_OK;
}

TriData tri_data(const ms_ctx* c) {
    TriData td{};
    td.T = c->T;
    td.N = c->N;
    td.pid = c->pid;
    td.G = c->G;
    td.H = c->H;
    td.G32 = c->G32;
    td.H32 = c->H32;
    td.ainfo = c->ainfo;
    td.binfo = c->binfo;
    td.a0 = c->a0;
    td.b0 = c->b0;
    td.h = c->h;
    td.mu = c->mu;
    td.area = c->area;
    td.hmuA32 = c->hmuA32;
    td.ref = c->ref;
    return td;
}

NodeData node_data(const ms_ctx* c) {
    NodeData nd{};
    nd.N = c->N;
    nd.mass = c->mass;
    nd.ref = c->ref;
    nd.clamp = c->clamp;
    nd.csr_off = c->csr_off;
    nd.csr_ent = c->csr_ent;
    return nd;
}

int grid_for(const ms_ctx* c, int nchunks, int per_cu) {
    if (per_cu <= 0) return std::max(1, nchunks);
    int g = c->num_cus * per_cu;
    g = ((g + 7) / 8) * 8;
    return std::max(1, std::min(g, nchunks));
}

// Nodes per tile (node group) under tuning n.  The balanced choice (-2): when
// full tiles are fewer than the launch's workgroups G, some workgroups (and
// CUs) would sit idle, so the tiles are cut to the smallest size (at least a
// wave's worth of nodes, 64) whose count still fits one round of G.  With at
// least G full tiles it keeps them: cutting them to fill whole rounds
// (the 1.8M-triangle disc: 1774 tiles of 512 nodes on 512 workgroups, 3.46
// rounds, against 2046 of 444 in 4) measured 6-10% slower, since a partial
// last round runs faster and smaller tiles recompute more halo; filling one
// round (the 290k-triangle disc: 284 tiles of 512 nodes, against 512 of 285)
// measured 1.3x faster.  G is the persistent grid, or (one workgroup per
// tile) the workgroups the kernel's registers let each CU hold.
int resolve_tile_nodes(const ms_ctx* c, const ms_tuning& n, const std::vector<char>* live = nullptr) {
    const int full = std::max(1, std::min(MAX_TILE_CHUNKS, n.tile_chunks)) * n.node_block;
    if (n.tile_nodes > 0) return std::min(n.tile_nodes, full);
    if (n.tile_nodes != -2 || n.fused != 1) return full;
    auto count = [&](int L) {
        long long k = 0;
        for (int s = 0; s < c->num_shells; ++s)
            if (!live || (*live)[s]) k += (c->hs[s].num_nodes + L - 1) / L;
        return k;
    };
    const bool mixed = c->precision == MS_MIXED;
    const long long G = n.node_grid_per_cu > 0
                            ? (long long)grid_for(c, INT_MAX, n.node_grid_per_cu)
                            : (long long)c->num_cus * std::max(1, (mixed ? 1024 : 512) / n.node_block);
    if (count(full) >= G) return full;
    int lo = std::min(full, 64), hi = full;  // count() does not increase with L; count(full) < G
    const long long cap = G;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (count(mid) <= cap)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}
Launch tri_launch(const ms_ctx* c) {
    return Launch{grid_for(c, c->n_tri_chunks, c->tune.tri_grid_per_cu), c->tune.tri_block, c->stream};
}
Launch node_launch(const ms_ctx* c) {
    return Launch{grid_for(c, c->n_node_chunks, c->tune.node_grid_per_cu), c->tune.node_block, c->stream};
}

void params_to_dev(const ms_shell_params& p, ShellDev& S) {
    S.nu = p.poisson_ratio;
    S.gamma = p.bend_stiffening_scale_factor;
    S.dt = p.timestep;
    S.eta = p.damping_factor;
    S.damp_den = p.ref_density * p.min_ref_thickness * p.min_ref_thickness;
    S.lim = 50000.0 * p.char_force_scale;
    S.fchar = p.char_force_scale;
    S.cchar = p.char_speed_scale;
    S.gd = p.gradient_descent ? 1 : 0;
    S.protocol = p.dial_protocol == MS_DIAL_LCE ? 3 : 1;
    S.from_ansatz = p.dial_from_ansatz ? 1 : 0;
}

void destroy_graph(ms_ctx* c);

TileData tile_data(const ms_ctx* c) {
    TileData t = c->tile;
    if (!c->act_all) {
        t.active = c->d_act_tiles;
        t.nactive = c->n_act_tiles;
    }
    return t;
}

// The chunk and tile tables of the step launches: every shell's, or (after
// some shells halted) the running shells' only.  Halted shells are skipped
// inside the step kernels either way, so this changes which workgroups are
// launched, not the results; per-chunk partials keep their full-table index
// (Chunk .w, tile .w).  Rebuilt when the running set differs from the one
// the tables were built for; the captured graphs hold the old tables and are
// destroyed.
int refresh_active(ms_ctx* c) {
    std::vector<char> live(c->num_shells);
    bool all = true;
    for (int s = 0; s < c->num_shells; ++s) {
        live[s] = c->hs[s].halted ? 0 : 1;
        all = all && live[s];
    }
    if (live == c->act_live) return MS_OK;
    destroy_graph(c);
    const bool any = std::find(live.begin(), live.end(), 1) != live.end();
    if (c->tune.fused == 1 && c->tune.tile_nodes == -2 && any) {
        // Balanced tiles follow the running shells: once they need tiles of
        // at most 0.7x the current size (or, all running again, larger ones)
        // every shell is re-cut.  The tables change which workgroups run, not
        // the results.  Should the new tiles fail to build, the two-launch
        // form (chunks only) carries on with the same results.
        const int L = resolve_tile_nodes(c, c->tune, &live);
        if (10 * (long long)L < 7LL * c->tile_L || (all && L > c->tile_L)) {
            if (int r = build_chunks(c, c->tune.tri_block, c->tune.node_block, L)) return r;
            if (int r = push_shells(c)) return r;
            if (build_tiles(c) != MS_OK) {
                c->tune.fused = 0;
                c->err.clear();
            }
        }
    }
    // the full tables until the new ones are in place (an upload error leaves
    // them in use and the running set unrecorded, so the next call retries)
    c->act_live.clear();
    c->act_all = true;
    if (all) {
        c->act_live = live;
        return MS_OK;
    }
    std::vector<Chunk> tc, nc;
    for (const Chunk& ch : c->h_tri_chunks)
        if (live[ch.z]) tc.push_back(ch);
    for (const Chunk& ch : c->h_node_chunks)
        if (live[ch.z]) nc.push_back(ch);
    if (!tc.empty())
        HIPC(c, hipMemcpy(c->d_act_tri, tc.data(), tc.size() * sizeof(Chunk), hipMemcpyHostToDevice));
    if (!nc.empty())
        HIPC(c, hipMemcpy(c->d_act_node, nc.data(), nc.size() * sizeof(Chunk), hipMemcpyHostToDevice));
    c->n_act_tri = (int)tc.size();
    c->n_act_node = (int)nc.size();
    c->n_act_tiles = 0;
    if (c->tiles_ready) {
        std::vector<int4> tiles(c->tile.ntiles);
        HIPC(c, hipMemcpy(tiles.data(), c->tile.tiles, tiles.size() * sizeof(int4), hipMemcpyDeviceToHost));
        std::vector<int> at;
        for (int i = 0; i < (int)tiles.size(); ++i)
            if (live[tiles[i].z]) at.push_back(i);
        if (!at.empty())
            HIPC(c, hipMemcpy(c->d_act_tiles, at.data(), at.size() * sizeof(int), hipMemcpyHostToDevice));
        c->n_act_tiles = (int)at.size();
    }
    c->act_live = live;
    c->act_all = false;
    return MS_OK;
}

// Some shell uses the LCE dial protocol (its dial info has the 5 LCE columns).
bool has_lce(const ms_ctx* c) { return c->acols == 5; }

// Every shell's thickness, shear modulus and (blend protocol) bbar_info are
// uniform over its triangles (set at creation; these arrays never change).
bool uni_data(const ms_ctx* c) { return c->uni_data; }

// One step; with the fused tuning the tile launch replaces the triangle and
// node launches.  timed: events around the two parts (ev[0..2]).
void step_once(ms_ctx* c, bool check, int halt_on_eq, double tol_f, double tol_v, bool timed = false) {
    const bool mixed = c->precision == MS_MIXED;
    TriData td = tri_data(c);
    NodeData nd = node_data(c);
    StepFlags fl{check ? 1 : 0, halt_on_eq, tol_f, tol_v, c->tune.fused == 1 ? c->tile.gen : -1LL};
    if (timed) hipEventRecord(c->ev[0], c->stream);
    if (c->tune.fused == 1) {
        const TileData tld = tile_data(c);
        const int nt = tld.active ? tld.nactive : tld.ntiles;
        const Launch tl{grid_for(c, nt, c->tune.node_grid_per_cu), c->tune.node_block, c->stream};
        launch_tile_step(tl, nd, c->b, tld, c->d_shells, mixed, has_lce(c), uni_data(c), fl);
        if (timed) hipEventRecord(c->ev[1], c->stream);
        launch_step_bookkeeping(c->stream, nd, c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells, c->num_shells,
                                c->d_clock, fl);
    } else {
        const Chunk* tcs = c->act_all ? c->d_tri_chunks : c->d_act_tri;
        const Chunk* ncs = c->act_all ? c->d_node_chunks : c->d_act_node;
        const int ntc = c->act_all ? c->n_tri_chunks : c->n_act_tri;
        const int nnc = c->act_all ? c->n_node_chunks : c->n_act_node;
        const Launch tl{grid_for(c, ntc, c->tune.tri_grid_per_cu), c->tune.tri_block, c->stream};
        const Launch nl{grid_for(c, nnc, c->tune.node_grid_per_cu), c->tune.node_block, c->stream};
        launch_tri_forces(tl, td, c->b, tcs, ntc, c->d_shells, mixed, has_lce(c), uni_data(c), true);
        if (timed) hipEventRecord(c->ev[1], c->stream);
        launch_node_step(nl, nd, c->b, ncs, nnc, c->d_shells, c->num_shells, c->d_clock, mixed, fl);
    }
    if (timed) hipEventRecord(c->ev[2], c->stream);
}

void destroy_graph(ms_ctx* c) {
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
    c->graph_steps = 0;
}

int ensure_graph(ms_ctx* c) {
    const int gs = std::max(1, c->tune.graph_steps);
    if (c->graph_exec && c->graph_steps == gs && c->graph_precision == c->precision) return MS_OK;
    destroy_graph(c);
    hipGraph_t g;
    HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < gs; ++i) step_once(c, false, 0, 0.0, 0.0);
    HIPC(c, hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, MS_ERR_HIP, fmt("hipGraphInstantiate: %s", hipGetErrorString(e)));
    c->graph_steps = gs;
    c->graph_precision = c->precision;
    return MS_OK;
}

int shell_ok(ms_ctx* c, int32_t s) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    if (s < 0 || s >= c->num_shells) return fail(c, MS_ERR_ARG, fmt("shell %d out of range", s));
    return MS_OK;
}

}  // namespace

extern "C" const char* ms_version(void) { return MS_VERSION; }

extern "C" int ms_create(const ms_shell_desc* shells, int32_t num_shells, int32_t precision, ms_ctx** out) {
    if (!shells || num_shells <= 0 || !out || (precision != MS_FP64 && precision != MS_MIXED))
        return fail(nullptr, MS_ERR_ARG, "ms_create: bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, MS_ERR_NO_DEVICE, "ms_create: no HIP device visible");
    // validate and size
    long long Ttot = 0, Ntot = 0;
    int acols = 3;
    bool any_lce = false, any_a0 = false, any_b0 = false;
    for (int s = 0; s < num_shells; ++s) {
        const ms_shell_desc& d = shells[s];
        if (d.num_nodes <= 0 || d.num_tris <= 0 || !d.patch_ids || !d.G || !d.H || !d.ref_pos || !d.ref_areas ||
            !d.thickness || !d.shear_modulus || !d.abar_info || !d.node_masses)
            return fail(nullptr, MS_ERR_ARG, fmt("ms_create: shell %d has a missing array", s));
        const bool lce = d.params.dial_protocol == MS_DIAL_LCE;
        if (!lce && (d.abar_info_cols != 3 || !d.bbar_info || d.bbar_info_cols != 3))
            return fail(nullptr, MS_ERR_ARG, fmt("ms_create: shell %d: blend protocol needs 3-column abar/bbar info", s));
        if (lce && d.abar_info_cols < 3)
            return fail(nullptr, MS_ERR_ARG, fmt("ms_create: shell %d: LCE protocol needs 3-column abar info", s));
        if (d.params.dial_protocol != MS_DIAL_LCE && d.params.dial_protocol != MS_DIAL_BLEND)
            return fail(nullptr, MS_ERR_ARG, fmt("ms_create: shell %d: unknown dial protocol", s));
        any_lce |= lce;
        any_a0 |= (d.a0 != nullptr);
        any_b0 |= (d.b0 != nullptr);
        for (long long i = 0; i < 6LL * d.num_tris; ++i)
            if (d.patch_ids[i] < 0 || d.patch_ids[i] >= d.num_nodes)
                return fail(nullptr, MS_ERR_ARG, fmt("ms_create: shell %d: patch id out of range", s));
        Ttot += d.num_tris;
        Ntot += d.num_nodes;
    }
    if (Ttot > INT_MAX / 18 || Ntot > INT_MAX / 3)
        return fail(nullptr, MS_ERR_ARG, "ms_create: batch too large for 32-bit stream indexing");
    if (any_lce) acols = 5;
    ms_ctx* c = new ms_ctx();
    c->precision = precision;
    c->num_shells = num_shells;
    c->T = (int)Ttot;
    c->N = (int)Ntot;
    c->tune.tri_block = 256;
    c->tune.node_block = 256;
    c->tune.tri_grid_per_cu = 0;
    c->tune.node_grid_per_cu = 0;
    c->tune.graph_steps = 16;
    c->tune.fused = 0;
    c->tune.tile_chunks = 1;
    c->tune.tile_nodes = -1;
    c->acols = acols;
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, dev);
    // set-up of the allocation-free entry point on this device (the context's
    // device is current); a failure only disables that entry point
    if (const hipError_t e = noalloc_scratch_for(dev, &c->noalloc_scratch); e != hipSuccess)
        c->noalloc_err = fmt("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    auto bail = [&](int code) {
        ms_destroy(c);
        return code;
    };
    // A blocking stream: the context's synchronous hipMemcpy/hipMemset calls
    // (legacy null stream) stay ordered with its kernels.
    if (hipStreamCreateWithFlags(&c->stream, hipStreamDefault) != hipSuccess) {
        set_last_error("ms_create: hipStreamCreate failed");
        return bail(MS_ERR_HIP);
    }
    const int T = c->T, N = c->N;
    // host staging in SoA
    std::vector<int> pid(6 * (size_t)T);
    std::vector<double> G(6 * (size_t)T), H(18 * (size_t)T), ai((size_t)acols * T, 0.0), bi(3 * (size_t)T, 0.0);
    std::vector<double> a0(any_a0 ? 3 * (size_t)T : 0), b0(any_b0 ? 3 * (size_t)T : 0);
    std::vector<double> hh(T), mm(T), ar(T), mass(N), rxy(2 * (size_t)N);
    std::vector<unsigned char> cl(N, 0);
    c->hs.assign(num_shells, ShellDev{});
    int tb = 0, nb = 0;
    for (int s = 0; s < num_shells; ++s) {
        const ms_shell_desc& d = shells[s];
        ShellDev& S = c->hs[s];
        params_to_dev(d.params, S);
        S.tri_begin = tb;
        S.num_tris = d.num_tris;
        S.node_begin = nb;
        S.num_nodes = d.num_nodes;
        S.dial_last_eq = 1.0;
        S.dial_spacing = 0.0;
        S.dial_K = 1.0;
        S.f_override = -1.0;
        S.last_eq_step = -1;
        S.halt_step = -1;
        S.fail_step = -1;
        S.has_a0 = d.a0 ? 1 : 0;
        S.has_b0 = d.b0 ? 1 : 0;
        const bool lce = S.protocol == 3;
        {
            // uniform per-triangle data, compared bit for bit
            auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; };
            bool uh = d.num_tris > 0, ub = d.num_tris > 0 && !lce;
            for (int t = 1; t < d.num_tris && (uh || ub); ++t) {
                uh = uh && same(d.thickness[t], d.thickness[0]) && same(d.shear_modulus[t], d.shear_modulus[0]);
                for (int k = 0; k < 3 && ub; ++k) ub = same(d.bbar_info[(size_t)t * 3 + k], d.bbar_info[k]);
            }
            S.uni_hmu = uh ? 1 : 0;
            S.uni_b = ub ? 1 : 0;
            if (!uh || !(ub || lce)) c->uni_data = false;
            S.h_u = uh ? d.thickness[0] : 0.0;
            S.mu_u = uh ? d.shear_modulus[0] : 0.0;
            for (int k = 0; k < 3; ++k) S.b_u[k] = ub ? d.bbar_info[k] : 0.0;
        }
        for (int t = 0; t < d.num_tris; ++t) {
            const size_t g = (size_t)tb + t;
            for (int k = 0; k < 6; ++k) pid[k * (size_t)T + g] = d.patch_ids[6 * (size_t)t + k] + nb;
            for (int k = 0; k < 6; ++k) G[k * (size_t)T + g] = d.G[6 * (size_t)t + k];
            for (int k = 0; k < 18; ++k) H[k * (size_t)T + g] = d.H[18 * (size_t)t + k];
            if (lce) {
                const double* r = d.abar_info + (size_t)t * d.abar_info_cols;
                const double cth = std::cos(r[0]), sth = std::sin(r[0]);
                ai[0 * (size_t)T + g] = cth * cth;
                ai[1 * (size_t)T + g] = cth * sth;
                ai[2 * (size_t)T + g] = sth * sth;
                ai[3 * (size_t)T + g] = r[1];
                ai[4 * (size_t)T + g] = -2.0 * r[2];
            } else {
                for (int k = 0; k < 3; ++k) {
                    ai[k * (size_t)T + g] = d.abar_info[(size_t)t * 3 + k];
                    bi[k * (size_t)T + g] = d.bbar_info[(size_t)t * 3 + k];
                }
            }
            for (int k = 0; k < 3; ++k) {
                if (any_a0) a0[k * (size_t)T + g] = d.a0 ? d.a0[(size_t)k * d.num_tris + t] : (k == 1 ? 0.0 : 1.0);
                if (any_b0) b0[k * (size_t)T + g] = d.b0 ? d.b0[(size_t)k * d.num_tris + t] : 0.0;
            }
            hh[g] = d.thickness[t];
            mm[g] = d.shear_modulus[t];
            ar[g] = d.ref_areas[t];
        }
        for (int n = 0; n < d.num_nodes; ++n) {
            rxy[(size_t)nb + n] = d.ref_pos[n];
            rxy[(size_t)N + nb + n] = d.ref_pos[(size_t)d.num_nodes + n];
            mass[(size_t)nb + n] = d.node_masses[n];
            cl[(size_t)nb + n] = d.constrained ? (d.constrained[n] ? 1 : 0) : 0;
        }
        tb += d.num_tris;
        nb += d.num_nodes;
    }
    // The FP64 tile step can form a triangle's reference area from the
    // reference x, y it stages for the patch fit: 0.5 |x1 y2 - x2 y1| over the
    // edge vectors from patch node 0 to patch nodes 1 and 2 (the triangle's
    // vertices), each product rounded.  It does so when that reproduces every
    // given area bit for bit, as it does for ms_mesh_info's areas of a planar
    // reference state (whose normal has no x, y component).
    c->area_ref = true;
    for (int t = 0; t < T && c->area_ref; ++t) {
        const int p0 = pid[(size_t)t], p1 = pid[(size_t)T + t], p2 = pid[2 * (size_t)T + t];
        const double x1 = rxy[p1] - rxy[p0], y1 = rxy[(size_t)N + p1] - rxy[(size_t)N + p0];
        const double x2 = rxy[p2] - rxy[p0], y2 = rxy[(size_t)N + p2] - rxy[(size_t)N + p0];
        volatile double q1 = x1 * y2, q2 = x2 * y1;  // rounded apart, as on the device
        const double a = 0.5 * std::fabs(q1 - q2);
        c->area_ref = std::memcmp(&a, &ar[t], sizeof(double)) == 0;
    }
    // patch-incidence CSR: entries t*6+slot grouped by node, ascending
    std::vector<long long> off((size_t)N + 1, 0);
    for (size_t i = 0; i < 6 * (size_t)T; ++i) off[(size_t)pid[i] + 1]++;
    for (int n = 0; n < N; ++n) off[n + 1] += off[n];
    std::vector<int> ent(6 * (size_t)T);
    {
        std::vector<long long> cur(off.begin(), off.end() - 1);
        for (int t = 0; t < T; ++t)
            for (int k = 0; k < 6; ++k) {
                const int n = pid[k * (size_t)T + t];
                ent[(size_t)cur[n]++] = t * 6 + k;
            }
    }
#define UP(dst, vec)                                                                                      \
    do {                                                                                                  \
        if (!(vec).empty()) {                                                                             \
            if (dalloc(c, &(dst), (vec).size()) != hipSuccess ||                                          \
                hipMemcpy((dst), (vec).data(), (vec).size() * sizeof((vec)[0]), hipMemcpyHostToDevice) != \
                    hipSuccess) {                                                                         \
                set_last_error("ms_create: device allocation/upload failed for " #dst);                   \
                return bail(MS_ERR_HIP);                                                                  \
            }                                                                                             \
        }                                                                                                 \
    } while (0)
    UP(c->pid, pid);
    UP(c->G, G);
    UP(c->H, H);
    UP(c->ainfo, ai);
    UP(c->binfo, bi);
    UP(c->a0, a0);
    UP(c->b0, b0);
    UP(c->h, hh);
    UP(c->mu, mm);
    UP(c->area, ar);
    UP(c->mass, mass);
    UP(c->ref, rxy);
    UP(c->clamp, cl);
    UP(c->csr_off, off);
    UP(c->csr_ent, ent);
    {
        std::vector<float> g32(G.begin(), G.end()), h32(H.begin(), H.end()), hma(3 * (size_t)T);
        for (int t = 0; t < T; ++t) {
            hma[t] = (float)hh[t];
            hma[T + (size_t)t] = (float)mm[t];
            hma[2 * (size_t)T + t] = (float)ar[t];
        }
        UP(c->G32, g32);
        UP(c->H32, h32);
        UP(c->hmuA32, hma);
    }
#undef UP
    bool ok = true;
    for (int i = 0; i < 2; ++i) {
        ok &= dalloc(c, &c->b.x[i], 3 * (size_t)N) == hipSuccess;
        ok &= dalloc(c, &c->b.v[i], 3 * (size_t)N) == hipSuccess;
    }
    ok &= dalloc(c, &c->b.F, 3 * (size_t)N) == hipSuccess;
    ok &= dalloc(c, &c->b.contrib, 18 * (size_t)T) == hipSuccess;
    ok &= dalloc(c, &c->b.contrib32, 18 * (size_t)T) == hipSuccess;
    ok &= dalloc(c, &c->b.bad_dof, (size_t)num_shells) == hipSuccess;
    ok &= dalloc(c, &c->red_out, 3 * (size_t)num_shells) == hipSuccess;
    ok &= dalloc(c, &c->d_shells, (size_t)num_shells) == hipSuccess;
    ok &= dalloc(c, &c->d_clock, 1) == hipSuccess;
    if (!ok) {
        set_last_error("ms_create: device allocation failed");
        return bail(MS_ERR_HIP);
    }
    for (int i = 0; i < 2; ++i) {
        hipMemset(c->b.x[i], 0, 3 * sizeof(double) * (size_t)N);
        hipMemset(c->b.v[i], 0, 3 * sizeof(double) * (size_t)N);
    }
    hipMemset(c->b.F, 0, 3 * sizeof(double) * (size_t)N);
    hipMemset(c->d_clock, 0, sizeof(ClockDev));
    if (build_chunks(c, c->tune.tri_block, c->tune.node_block, resolve_tile_nodes(c, c->tune)) != MS_OK) {
        set_last_error(c->err);
        return bail(MS_ERR_HIP);
    }
    if (push_shells(c) != MS_OK) {
        set_last_error(c->err);
        return bail(MS_ERR_HIP);
    }
    for (auto& e : c->ev) hipEventCreate(&e);
    if (hipDeviceSynchronize() != hipSuccess) {
        set_last_error("ms_create: device error after upload");
        return bail(MS_ERR_HIP);
    }
    *out = c;
    return MS_OK;
}

extern "C" void ms_destroy(ms_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    destroy_graph(c);
    free_tiles(c);
    for (void* p : c->allocs) hipFree(p);
    if (c->d_tri_chunks) hipFree(c->d_tri_chunks);
    if (c->d_node_chunks) hipFree(c->d_node_chunks);
    if (c->b.part) hipFree(c->b.part);
    if (c->b.part_i) hipFree(c->b.part_i);
    if (c->d_act_tri) hipFree(c->d_act_tri);
    if (c->d_act_node) hipFree(c->d_act_node);
    if (c->scratch3) hipFree(c->scratch3);
    if (c->node_out) hipFree(c->node_out);
    if (c->node_flags) hipFree(c->node_flags);
    if (c->b.clog) hipFree(c->b.clog);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* ms_ctx_last_error(const ms_ctx* c) { return c ? c->err.c_str() : "null context"; }
extern "C" int32_t ms_num_shells(const ms_ctx* c) { return c ? c->num_shells : 0; }
extern "C" int64_t ms_steps_taken(const ms_ctx* c) { return c ? c->steps : 0; }
extern "C" int64_t ms_device_bytes(const ms_ctx* c) { return c ? c->dev_bytes : 0; }

extern "C" int ms_set_state(ms_ctx* c, int32_t s, const double* dofs, const double* vel) {
    if (int r = shell_ok(c, s)) return r;
    if (!dofs) return fail(c, MS_ERR_ARG, "ms_set_state: dofs is null");
    if (int r = pull_shells(c)) return r;
    ShellDev& S = c->hs[s];
    const int par = S.par;
    for (int comp = 0; comp < 3; ++comp) {
        HIPC(c, hipMemcpy(c->b.x[par] + (size_t)comp * c->N + S.node_begin, dofs + (size_t)comp * S.num_nodes,
                          sizeof(double) * S.num_nodes, hipMemcpyHostToDevice));
        if (vel)
            HIPC(c, hipMemcpy(c->b.v[par] + (size_t)comp * c->N + S.node_begin, vel + (size_t)comp * S.num_nodes,
                              sizeof(double) * S.num_nodes, hipMemcpyHostToDevice));
        else
            HIPC(c, hipMemset(c->b.v[par] + (size_t)comp * c->N + S.node_begin, 0, sizeof(double) * S.num_nodes));
    }
    return MS_OK;
}

extern "C" int ms_get_state(ms_ctx* c, int32_t s, double* dofs, double* vel) {
    if (int r = shell_ok(c, s)) return r;
    if (int r = pull_shells(c)) return r;
    const ShellDev& S = c->hs[s];
    for (int comp = 0; comp < 3; ++comp) {
        if (dofs)
            HIPC(c, hipMemcpy(dofs + (size_t)comp * S.num_nodes, c->b.x[S.par] + (size_t)comp * c->N + S.node_begin,
                              sizeof(double) * S.num_nodes, hipMemcpyDeviceToHost));
        if (vel)
            HIPC(c, hipMemcpy(vel + (size_t)comp * S.num_nodes, c->b.v[S.par] + (size_t)comp * c->N + S.node_begin,
                              sizeof(double) * S.num_nodes, hipMemcpyDeviceToHost));
    }
    return MS_OK;
}

extern "C" int ms_get_forces(ms_ctx* c, int32_t s, double* f) {
    if (int r = shell_ok(c, s)) return r;
    if (!f) return fail(c, MS_ERR_ARG, "ms_get_forces: null output");
    HIPC(c, hipStreamSynchronize(c->stream));
    const ShellDev& S = c->hs[s];
    for (int comp = 0; comp < 3; ++comp)
        HIPC(c, hipMemcpy(f + (size_t)comp * S.num_nodes, c->b.F + (size_t)comp * c->N + S.node_begin,
                          sizeof(double) * S.num_nodes, hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_set_forces(ms_ctx* c, int32_t s, const double* f) {
    if (int r = shell_ok(c, s)) return r;
    if (!f) return fail(c, MS_ERR_ARG, "ms_set_forces: null input");
    HIPC(c, hipStreamSynchronize(c->stream));
    const ShellDev& S = c->hs[s];
    for (int comp = 0; comp < 3; ++comp)
        HIPC(c, hipMemcpy(c->b.F + (size_t)comp * c->N + S.node_begin, f + (size_t)comp * S.num_nodes,
                          sizeof(double) * S.num_nodes, hipMemcpyHostToDevice));
    return MS_OK;
}

extern "C" int ms_set_dial_clock(ms_ctx* c, int32_t s, double last_eq, double spacing, int64_t K, int64_t phase,
                                 int32_t dialling) {
    if (int r = shell_ok(c, s)) return r;
    if (K <= 0) return fail(c, MS_ERR_ARG, "ms_set_dial_clock: steps_per_phase must be > 0");
    if (int r = pull_shells(c)) return r;
    ShellDev& S = c->hs[s];
    S.dial_last_eq = last_eq;
    S.dial_spacing = spacing;
    S.dial_K = (double)K;
    S.phase_steps = phase;
    S.dialling = dialling ? 1 : 0;
    S.f_override = -1.0;
    return push_shells(c);
}

extern "C" int ms_get_dial_factor(ms_ctx* c, int32_t s, double* f, int64_t* phase) {
    if (int r = shell_ok(c, s)) return r;
    if (int r = pull_shells(c)) return r;
    const ShellDev& S = c->hs[s];
    const double d = shell_dial_factor(S);
    if (f) *f = d;
    if (phase) *phase = S.phase_steps;
    return MS_OK;
}

extern "C" int ms_set_shell_params(ms_ctx* c, int32_t s, const ms_shell_params* p) {
    if (int r = shell_ok(c, s)) return r;
    if (!p) return fail(c, MS_ERR_ARG, "ms_set_shell_params: null params");
    if (int r = pull_shells(c)) return r;
    ShellDev& S = c->hs[s];
    const int old_protocol = S.protocol;
    params_to_dev(*p, S);
    if (S.protocol != old_protocol) {
        S.protocol = old_protocol;
        return fail(c, MS_ERR_ARG, "ms_set_shell_params: the dial protocol is fixed at ms_create");
    }
    return push_shells(c);
}

extern "C" int ms_set_zero_dial_geometry(ms_ctx* c, int32_t s, const double* a0, const double* b0) {
    if (int r = shell_ok(c, s)) return r;
    if (int r = pull_shells(c)) return r;
    destroy_graph(c);  // the kernel arguments change (a0, b0, tile records)
    ShellDev& S = c->hs[s];
    const int T = c->T;
    if (a0) {
        if (!c->a0) {
            std::vector<double> init(3 * (size_t)T);
            for (int t = 0; t < T; ++t) {
                init[t] = 1.0;
                init[(size_t)T + t] = 0.0;
                init[2 * (size_t)T + t] = 1.0;
            }
            HIPC(c, dalloc(c, &c->a0, init.size()));
            HIPC(c, hipMemcpy(c->a0, init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        for (int k = 0; k < 3; ++k)
            HIPC(c, hipMemcpy(c->a0 + (size_t)k * T + S.tri_begin, a0 + (size_t)k * S.num_tris,
                              sizeof(double) * S.num_tris, hipMemcpyHostToDevice));
        S.has_a0 = 1;
    }
    if (b0) {
        if (!c->b0) {
            HIPC(c, dalloc(c, &c->b0, 3 * (size_t)T));
            HIPC(c, hipMemset(c->b0, 0, 3 * sizeof(double) * (size_t)T));
        }
        for (int k = 0; k < 3; ++k)
            HIPC(c, hipMemcpy(c->b0 + (size_t)k * T + S.tri_begin, b0 + (size_t)k * S.num_tris,
                              sizeof(double) * S.num_tris, hipMemcpyHostToDevice));
        S.has_b0 = 1;
    }
    if (c->tiles_ready) {  // the tile-ordered copies of a0, b0
        if (int r = build_tile_records(c)) {
            // without current records the fused step cannot run: drop the
            // tiles, and the two-launch form (which needs none) carries on
            free_tiles(c);
            c->tune.fused = 0;
            return r;
        }
    }
    return push_shells(c);
}

extern "C" int ms_set_precision(ms_ctx* c, int32_t precision) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    if (precision != MS_FP64 && precision != MS_MIXED) return fail(c, MS_ERR_ARG, "ms_set_precision: bad mode");
    c->precision = precision;
    return MS_OK;
}

extern "C" int ms_get_precision(const ms_ctx* c) { return c ? c->precision : -1; }

// --------------------------------------------------------------- the ops

extern "C" int ms_calc_deformation_forces(ms_ctx* c) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    const bool mixed = c->precision == MS_MIXED;
    launch_tri_forces(tri_launch(c), tri_data(c), c->b, c->d_tri_chunks, c->n_tri_chunks, c->d_shells, mixed,
                      has_lce(c), uni_data(c), false);
    launch_node_gather(node_launch(c), node_data(c), c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells, mixed);
    HIPC(c, hipGetLastError());
    return MS_OK;
}

extern "C" int ms_calc_non_deformation_forces(ms_ctx* c) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    launch_node_damping(node_launch(c), node_data(c), c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells);
    HIPC(c, hipGetLastError());
    return MS_OK;
}

extern "C" int ms_impose_constraints(ms_ctx* c) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    launch_node_clamp(node_launch(c), node_data(c), c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells);
    HIPC(c, hipGetLastError());
    return MS_OK;
}

namespace {
int scan_bad(ms_ctx* c, std::vector<long long>& bad) {
    launch_node_scan(node_launch(c), node_data(c), c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells,
                     c->num_shells);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    bad.resize(c->num_shells);
    HIPC(c, hipMemcpy(bad.data(), c->b.bad_dof, sizeof(long long) * c->num_shells, hipMemcpyDeviceToHost));
    return MS_OK;
}
int read_force(ms_ctx* c, int s, long long dof, double* out) {
    const ShellDev& S = c->hs[s];
    const long long comp = dof / S.num_nodes, n = dof % S.num_nodes;
    HIPC(c, hipMemcpy(out, c->b.F + comp * c->N + S.node_begin + n, sizeof(double), hipMemcpyDeviceToHost));
    return MS_OK;
}
}  // namespace

extern "C" int ms_advance_dynamics(ms_ctx* c, ms_step_report* reports) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    std::vector<long long> bad;
    if (int r = scan_bad(c, bad)) return r;
    launch_node_advance(node_launch(c), node_data(c), c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells,
                        c->num_shells, c->d_clock);
    HIPC(c, hipGetLastError());
    c->steps += 1;
    if (int r = pull_shells(c)) return r;
    int status = MS_OK;
    for (int s = 0; s < c->num_shells; ++s) {
        ms_step_report rep{};
        rep.fail_step = -1;
        rep.bad_dof = -1;
        if (bad[s] != LLONG_MAX) {
            rep.status = MS_ERR_HIGH_FORCE;
            rep.bad_dof = bad[s];
            rep.fail_step = c->steps - 1;
            if (int r = read_force(c, s, bad[s], &rep.bad_force)) return r;
            status = MS_ERR_HIGH_FORCE;
        }
        rep.halted = c->hs[s].halted;
        if (reports) reports[s] = rep;
    }
    if (status != MS_OK) c->err = "implausibly_high_force";
    return status;
}

extern "C" int ms_is_at_equilibrium(ms_ctx* c, int32_t s, double tol_f, double tol_v, ms_step_report* rep) {
    if (int r = shell_ok(c, s)) return r;
    launch_node_equil(node_launch(c), node_data(c), c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells,
                      c->num_shells, c->d_clock);
    HIPC(c, hipGetLastError());
    if (int r = pull_shells(c)) return r;
    const ShellDev& S = c->hs[s];
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->fail_step = -1;
        rep->bad_dof = -1;
        rep->max_force_ratio = S.max_f;
        rep->max_force_node = S.f_node;
        rep->max_speed_ratio = S.max_v;
        rep->max_speed_node = S.v_node;
        rep->checked = 1;
        rep->at_equilibrium = (S.max_f < tol_f && S.max_v < tol_v) ? 1 : 0;
        rep->halted = S.halted;
    }
    return MS_OK;
}

extern "C" int ms_calc_energies(ms_ctx* c, int32_t s, double out[3]) {
    if (int r = shell_ok(c, s)) return r;
    if (!out) return fail(c, MS_ERR_ARG, "ms_calc_energies: null output");
    launch_tri_energy(tri_launch(c), tri_data(c), c->b, c->d_tri_chunks, c->n_tri_chunks, c->d_shells, nullptr);
    launch_reduce_partials(c->stream, c->b.part, c->d_shells, c->num_shells, true, 2, c->red_out);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    double st[2];
    HIPC(c, hipMemcpy(st, c->red_out + 2 * (size_t)s, sizeof(st), hipMemcpyDeviceToHost));
    launch_node_kinetic(node_launch(c), node_data(c), c->b, c->d_node_chunks, c->n_node_chunks, c->d_shells);
    launch_reduce_partials(c->stream, c->b.part, c->d_shells, c->num_shells, false, 1, c->red_out);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    double ke;
    HIPC(c, hipMemcpy(&ke, c->red_out + s, sizeof(double), hipMemcpyDeviceToHost));
    out[0] = st[0];
    out[1] = st[1];
    out[2] = ke;
    return MS_OK;
}

namespace {
int ensure_scratch(ms_ctx* c) {
    if (!c->scratch3) HIPC(c, hipMalloc((void**)&c->scratch3, 9 * sizeof(double) * (size_t)c->T));
    return MS_OK;
}
}  // namespace

extern "C" int ms_calc_geometry(ms_ctx* c, int32_t s, double* a, double* b, double* n) {
    if (int r = shell_ok(c, s)) return r;
    if (int r = ensure_scratch(c)) return r;
    const size_t T = c->T;
    double *ga = c->scratch3, *gb = c->scratch3 + 3 * T, *gn = c->scratch3 + 6 * T;
    launch_tri_geometry(tri_launch(c), tri_data(c), c->b, c->d_tri_chunks, c->n_tri_chunks, c->d_shells, ga, gb, gn,
                        0, c->T);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    const ShellDev& S = c->hs[s];
    double* outs[3] = {a, b, n};
    double* srcs[3] = {ga, gb, gn};
    for (int q = 0; q < 3; ++q) {
        if (!outs[q]) continue;
        for (int k = 0; k < 3; ++k)
            HIPC(c, hipMemcpy(outs[q] + (size_t)k * S.num_tris, srcs[q] + k * T + S.tri_begin,
                              sizeof(double) * S.num_tris, hipMemcpyDeviceToHost));
    }
    return MS_OK;
}

// ms_calc_geometry of one shell through the static scratch array: no device
// allocation.  Calls on one device share the array, so calls that overlap on
// one device (from several host threads) race.
extern "C" int ms_calc_geometry_noalloc(ms_ctx* c, int32_t s, double* a, double* b, double* n) {
    if (int r = shell_ok(c, s)) return r;
    if (!c->noalloc_scratch) return fail(c, MS_ERR_HIP, "ms_calc_geometry_noalloc: " + c->noalloc_err);
    const ShellDev& S = c->hs[s];  // ranges and chunk tables only (fixed at creation and by re-chunking)
    if (S.num_tris > MS_NOALLOC_MAX_TRIS)
        return fail(c, MS_ERR_ARG, fmt("ms_calc_geometry_noalloc: shell %d has %d triangles, above the scratch "
                                       "limit of %d (MS_NOALLOC_MAX_TRIS); use ms_calc_geometry",
                                       s, S.num_tris, MS_NOALLOC_MAX_TRIS));
    const size_t nt = (size_t)S.num_tris;
    double* ga = c->noalloc_scratch;
    double* gb = ga + 3 * nt;
    double* gn = gb + 3 * nt;
    const Launch L{grid_for(c, S.tri_chunk_count, c->tune.tri_grid_per_cu), c->tune.tri_block, c->stream};
    launch_tri_geometry(L, tri_data(c), c->b, c->d_tri_chunks + S.tri_chunk_begin, S.tri_chunk_count, c->d_shells,
                        ga, gb, gn, S.tri_begin, S.num_tris);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    double* outs[3] = {a, b, n};
    const double* srcs[3] = {ga, gb, gn};
    for (int q = 0; q < 3; ++q)
        if (outs[q]) HIPC(c, hipMemcpy(outs[q], srcs[q], 3 * nt * sizeof(double), hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_calc_diagnostics(ms_ctx* c, int32_t s, double* gauss, double* mean, double* strain) {
    if (int r = shell_ok(c, s)) return r;
    if (int r = ensure_scratch(c)) return r;
    const size_t T = c->T;
    launch_tri_diagnostics(tri_launch(c), tri_data(c), c->b, c->d_tri_chunks, c->n_tri_chunks, c->d_shells,
                           c->scratch3);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    const ShellDev& S = c->hs[s];
    double* outs[3] = {gauss, mean, strain};
    for (int q = 0; q < 3; ++q)
        if (outs[q])
            HIPC(c, hipMemcpy(outs[q], c->scratch3 + q * T + S.tri_begin, sizeof(double) * S.num_tris,
                              hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_calc_stresses(ms_ctx* c, int32_t s, double* stresses) {
    if (int r = shell_ok(c, s)) return r;
    if (!stresses) return fail(c, MS_ERR_ARG, "ms_calc_stresses: null output");
    if (int r = ensure_scratch(c)) return r;  // 9T doubles, 8T used
    const size_t T = c->T;
    launch_tri_stresses(tri_launch(c), tri_data(c), c->b, c->d_tri_chunks, c->n_tri_chunks, c->d_shells,
                        c->scratch3);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    const ShellDev& S = c->hs[s];
    for (int k = 0; k < 8; ++k)
        HIPC(c, hipMemcpy(stresses + (size_t)k * S.num_tris, c->scratch3 + k * T + S.tri_begin,
                          sizeof(double) * S.num_tris, hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_calc_angle_deficits(ms_ctx* c, int32_t s, const uint8_t* boundary, double* deficits) {
    if (int r = shell_ok(c, s)) return r;
    if (!boundary || !deficits) return fail(c, MS_ERR_ARG, "ms_calc_angle_deficits: null argument");
    if (!c->node_out) HIPC(c, hipMalloc((void**)&c->node_out, sizeof(double) * (size_t)c->N));
    if (!c->node_flags) HIPC(c, hipMalloc((void**)&c->node_flags, (size_t)c->N));
    const ShellDev& S = c->hs[s];
    std::vector<unsigned char> fl(S.num_nodes);
    for (int n = 0; n < S.num_nodes; ++n) fl[n] = boundary[n] ? 1 : 0;
    HIPC(c, hipMemcpy(c->node_flags + S.node_begin, fl.data(), fl.size(), hipMemcpyHostToDevice));
    const Launch L{grid_for(c, S.node_chunk_count, c->tune.node_grid_per_cu), c->tune.node_block, c->stream};
    launch_node_angle_deficits(L, node_data(c), c->b, c->d_node_chunks + S.node_chunk_begin, S.node_chunk_count,
                               c->d_shells, c->pid, c->T, c->node_flags, c->node_out);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipMemcpy(deficits, c->node_out + S.node_begin, sizeof(double) * S.num_nodes, hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_calc_energy_densities(ms_ctx* c, int32_t s, double* dens) {
    if (int r = shell_ok(c, s)) return r;
    if (int r = ensure_scratch(c)) return r;
    const size_t T = c->T;
    launch_tri_energy(tri_launch(c), tri_data(c), c->b, c->d_tri_chunks, c->n_tri_chunks, c->d_shells, c->scratch3);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    const ShellDev& S = c->hs[s];
    if (dens)
        for (int k = 0; k < 2; ++k)
            HIPC(c, hipMemcpy(dens + (size_t)k * S.num_tris, c->scratch3 + k * T + S.tri_begin,
                              sizeof(double) * S.num_tris, hipMemcpyDeviceToHost));
    return MS_OK;
}

// ------------------------------------------------------------ fused path

extern "C" int ms_step(ms_ctx* c, int64_t nsteps, int64_t equil_every, double tol_f, double tol_v, int32_t halt_on_eq,
                       ms_step_report* reports) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    if (nsteps < 0) return fail(c, MS_ERR_ARG, "ms_step: nsteps < 0");
    if (int r = pull_shells(c)) return r;
    std::vector<int> was_failed(c->num_shells);
    for (int s = 0; s < c->num_shells; ++s) {
        was_failed[s] = c->hs[s].failed;
        c->hs[s].checked = 0;
    }
    if (int r = push_shells(c)) return r;
    if (c->tune.fused == 1 && !c->tiles_ready)
        return fail(c, MS_ERR_ARG, "ms_step: the fused step has no tiles (ms_set_tuning builds them)");
    if (int r = refresh_active(c)) return r;
    const bool use_graph = !c->timing;
    if (use_graph)
        if (int r = ensure_graph(c)) return r;
    double t_tri = 0.0, t_node = 0.0;
    long long n_timed = 0;
    int64_t left = nsteps;
    while (left > 0) {
        int64_t plain = left;
        bool check_now = false;
        if (equil_every > 0) {
            const int64_t to_check = equil_every - 1 - (c->steps % equil_every);
            if (to_check == 0)
                check_now = true;
            else
                plain = std::min<int64_t>(left, to_check);
        }
        if (check_now) {
            step_once(c, true, halt_on_eq, tol_f, tol_v);
            HIPC(c, hipGetLastError());
            c->steps += 1;
            left -= 1;
            continue;
        }
        if (use_graph) {
            while (plain >= c->graph_steps) {
                HIPC(c, hipGraphLaunch(c->graph_exec, c->stream));
                c->steps += c->graph_steps;
                plain -= c->graph_steps;
                left -= c->graph_steps;
            }
        }
        while (plain > 0) {
            if (c->timing) {
                step_once(c, false, 0, 0.0, 0.0, true);
                HIPC(c, hipEventSynchronize(c->ev[2]));
                float a = 0, b = 0;
                hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
                hipEventElapsedTime(&b, c->ev[1], c->ev[2]);
                t_tri += a;
                t_node += b;
                ++n_timed;
            } else {
                step_once(c, false, 0, 0.0, 0.0);
            }
            HIPC(c, hipGetLastError());
            c->steps += 1;
            plain -= 1;
            left -= 1;
        }
    }
    if (n_timed > 0) {
        c->last_ms[0] = t_tri / n_timed;
        c->last_ms[1] = t_node / n_timed;
        c->last_ms[2] = (t_tri + t_node) / n_timed;
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    if (int r = pull_shells(c)) return r;
    // crash reports for shells that failed in this call: recompute the
    // failing step's forces from the retained pre-step state
    bool new_fail = false;
    for (int s = 0; s < c->num_shells; ++s) new_fail |= (c->hs[s].failed && !was_failed[s]);
    std::vector<long long> bad;
    if (new_fail) {
        if (int r = ms_calc_deformation_forces(c)) return r;
        if (int r = ms_calc_non_deformation_forces(c)) return r;
        if (int r = ms_impose_constraints(c)) return r;
        if (int r = scan_bad(c, bad)) return r;
    }
    int status = MS_OK;
    for (int s = 0; s < c->num_shells; ++s) {
        const ShellDev& S = c->hs[s];
        ms_step_report rep{};
        rep.status = S.failed ? MS_ERR_HIGH_FORCE : MS_OK;
        rep.halted = S.halted;
        rep.fail_step = S.fail_step;
        rep.bad_dof = -1;
        if (S.failed && !was_failed[s]) {
            status = MS_ERR_HIGH_FORCE;
            if (bad[s] != LLONG_MAX) {
                rep.bad_dof = bad[s];
                if (int r = read_force(c, s, bad[s], &rep.bad_force)) return r;
            }
        }
        rep.checked = S.checked;
        rep.max_force_ratio = S.max_f;
        rep.max_force_node = S.f_node;
        rep.max_speed_ratio = S.max_v;
        rep.max_speed_node = S.v_node;
        rep.at_equilibrium = S.at_eq;
        if (reports) reports[s] = rep;
    }
    if (status != MS_OK) c->err = "implausibly_high_force";
    return status;
}

extern "C" int ms_relax_setup(ms_ctx* c, int32_t s, const ms_relax_params* p) {
    if (int r = shell_ok(c, s)) return r;
    if (!p || p->steps_per_dial_phase <= 0 || !(p->dial_spacing > 0) || p->dial_spacing > 1.0 ||
        p->equil_interval < 0 || !(p->tol_f > 0) || !(p->tol_v > 0))
        return fail(c, MS_ERR_ARG, "ms_relax_setup: bad parameters");
    if (int r = pull_shells(c)) return r;
    ShellDev& S = c->hs[s];
    S.auto_protocol = 1;
    S.status = 0;
    S.dialling = 1;
    S.check_due = 0;
    S.pause_on_eq = p->pause_on_eq ? 1 : 0;
    S.num_eq = 0;
    S.since_check = 0;
    S.equil_interval = p->equil_interval;
    S.eta1 = p->eta1;
    S.eta2 = p->eta2;
    S.eta = p->eta1;
    S.tol_f = p->tol_f;
    S.tol_v = p->tol_v;
    S.last_eq_step = -1;
    S.own_steps = 0;
    S.dial_K = (double)p->steps_per_dial_phase;
    S.dial_spacing = p->dial_spacing;
    S.dial_last_eq = p->dial_start;
    // main.cpp:261: the phase counter of a restart resumes where the dial
    // factor left it; the first step uses the start factor itself.
    S.phase_steps = (long long)(std::fmod(p->dial_start, p->dial_spacing) * (double)p->steps_per_dial_phase);
    S.f_override = p->dial_start;
    S.f_targets = p->dial_start;  // step 0 always dials (main.cpp:334)
    if (S.phase_steps >= p->steps_per_dial_phase) {  // main.cpp:293-304 at step 0
        S.status = 1;
        S.dialling = 0;
        S.eta = S.eta2;
    }
    S.halted = 0;
    S.failed = 0;
    S.halt_step = -1;
    S.fail_step = -1;
    return push_shells(c);
}

extern "C" int ms_relax_run(ms_ctx* c, int64_t max_steps, int64_t poll_every, ms_relax_report* reports) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    if (max_steps < 0 || poll_every <= 0) return fail(c, MS_ERR_ARG, "ms_relax_run: bad step counts");
    int64_t done = 0;
    while (done < max_steps) {
        if (int r = pull_shells(c)) return r;
        bool running = false;
        for (const auto& S : c->hs) running |= (S.auto_protocol && !S.halted);
        if (!running) break;
        const int64_t n = std::min<int64_t>(poll_every, max_steps - done);
        const int r = ms_step(c, n, 0, 0.0, 0.0, 0, nullptr);
        if (r != MS_OK && r != MS_ERR_HIGH_FORCE) return r;
        done += n;
    }
    if (int r = pull_shells(c)) return r;
    for (int s = 0; s < c->num_shells && reports; ++s) {
        const ShellDev& S = c->hs[s];
        ms_relax_report& R = reports[s];
        R.status = S.status;
        R.halted = S.halted;
        R.failed = S.failed;
        R.num_eq = S.num_eq;
        R.eq_step = S.last_eq_step;
        R.fail_step = S.fail_step;
        R.dial_factor = shell_dial_factor(S);
        R.max_force_ratio = S.max_f;
        R.max_speed_ratio = S.max_v;
        R.steps = S.own_steps;
    }
    c->err.clear();
    return MS_OK;
}

static_assert(sizeof(CheckRec) == sizeof(ms_check_record), "check record layout");

extern "C" int ms_set_check_log(ms_ctx* c, int32_t capacity) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    if (capacity < 0) return fail(c, MS_ERR_ARG, "ms_set_check_log: capacity < 0");
    if (int r = pull_shells(c)) return r;
    CheckRec* p = nullptr;
    if (capacity > 0)
        HIPC(c, hipMalloc((void**)&p, sizeof(CheckRec) * (size_t)capacity * c->num_shells));
    destroy_graph(c);  // the step launches name the log
    if (c->b.clog) hipFree(c->b.clog);
    c->b.clog = p;
    c->b.clog_cap = capacity;
    c->clog_read.assign(c->num_shells, 0);
    for (int s = 0; s < c->num_shells; ++s) c->clog_read[s] = c->hs[s].num_logged;
    return MS_OK;
}

extern "C" int ms_read_check_log(ms_ctx* c, int32_t s, ms_check_record* out, int32_t max_records, int32_t* count,
                                 int64_t* dropped) {
    if (int r = shell_ok(c, s)) return r;
    if (!count || (max_records > 0 && !out) || max_records < 0)
        return fail(c, MS_ERR_ARG, "ms_read_check_log: bad arguments");
    *count = 0;
    if (dropped) *dropped = 0;
    if (!c->b.clog) return fail(c, MS_ERR_ARG, "ms_read_check_log: no check log (ms_set_check_log)");
    if (int r = pull_shells(c)) return r;
    const long long total = c->hs[s].num_logged, cap = c->b.clog_cap;
    long long from = c->clog_read[s];
    if (total - from > cap) {
        if (dropped) *dropped = total - from - cap;
        from = total - cap;
    }
    const long long n = std::min<long long>(total - from, max_records);
    for (long long i = 0; i < n;) {  // at most two runs around the ring's end
        const long long slot = (from + i) % cap, run = std::min(n - i, cap - slot);
        HIPC(c, hipMemcpy(out + i, c->b.clog + (size_t)s * cap + slot, sizeof(CheckRec) * run,
                          hipMemcpyDeviceToHost));
        i += run;
    }
    c->clog_read[s] = from + n;
    *count = (int32_t)n;
    return MS_OK;
}

extern "C" int ms_resume_shell(ms_ctx* c, int32_t s) {
    if (int r = shell_ok(c, s)) return r;
    if (int r = pull_shells(c)) return r;
    ShellDev& S = c->hs[s];
    if (S.auto_protocol && S.halted && !S.failed) {
        if (S.status == 1 && S.last_eq_step == S.halt_step) {
            // paused at an intermediate equilibrium: keep that step's update
            // (its node launch already wrote it) and start the next dialling
            // phase (main.cpp:483-485, 501-507)
            S.par ^= 1;
            S.own_steps += 1;
            S.f_override = S.dial_last_eq + S.dial_spacing * ((double)S.phase_steps / S.dial_K);
            S.status = 0;
            S.dialling = 1;
            S.phase_steps = 0;
            S.eta = S.eta1;
            S.f_targets = shell_dial_factor(S);
        } else if (S.status == 2) {
            // re-open a final equilibrium (e.g. to finish in another precision)
            S.status = 1;
            S.dialling = 0;
            S.eta = S.eta2;
        }
        S.since_check = 0;
        S.check_due = 0;
    }
    S.halted = 0;
    S.failed = 0;
    S.at_eq = 0;
    S.halt_step = -1;
    S.fail_step = -1;
    return push_shells(c);
}

extern "C" int ms_set_tuning(ms_ctx* c, const ms_tuning* t) {
    if (!c || !t) return fail(c, MS_ERR_ARG, "ms_set_tuning: bad arguments");
    ms_tuning n = c->tune;
    auto okblock = [](int b) { return b == 64 || b == 128 || b == 256; };
    if (t->tri_block) {
        if (!okblock(t->tri_block)) return fail(c, MS_ERR_ARG, "tri_block must be 64, 128 or 256");
        n.tri_block = t->tri_block;
    }
    if (t->node_block) {
        if (!okblock(t->node_block) && t->node_block != 512)
            return fail(c, MS_ERR_ARG, "node_block must be 64, 128, 256 or 512");
        n.node_block = t->node_block;
    }
    if (t->tri_grid_per_cu >= 0) n.tri_grid_per_cu = t->tri_grid_per_cu;
    if (t->node_grid_per_cu >= 0) n.node_grid_per_cu = t->node_grid_per_cu;
    if (t->graph_steps > 0) n.graph_steps = t->graph_steps;
    if (t->fused == 0 || t->fused == 1)
        n.fused = t->fused;
    else if (t->fused != -1)
        return fail(c, MS_ERR_ARG, "fused must be 0, 1 or -1");
    if (t->tile_chunks) {
        if (t->tile_chunks < 1 || t->tile_chunks > MAX_TILE_CHUNKS)
            return fail(c, MS_ERR_ARG, fmt("tile_chunks must be 1 to %d (0 keeps)", MAX_TILE_CHUNKS));
        n.tile_chunks = t->tile_chunks;
    }
    if (t->tile_nodes) {
        if (t->tile_nodes < -2) return fail(c, MS_ERR_ARG, "tile_nodes must be positive, -1, -2 or 0 (keeps)");
        n.tile_nodes = t->tile_nodes;
    }
    if (n.tile_nodes > n.tile_chunks * n.node_block)
        return fail(c, MS_ERR_ARG,
                    fmt("tile_nodes %d exceeds tile_chunks * node_block = %d", n.tile_nodes, n.tile_chunks * n.node_block));
    const int L = resolve_tile_nodes(c, n);
    const bool rechunk = n.tri_block != c->tune.tri_block || n.node_block != c->tune.node_block || L != c->tile_L;
    const bool retile = rechunk || n.tile_chunks != c->tune.tile_chunks;
    if (int r = pull_shells(c)) return r;
    destroy_graph(c);
    if (rechunk) {
        // on failure the context keeps its current tables, tiles and tuning
        if (int r = build_chunks(c, n.tri_block, n.node_block, L)) return r;
        if (int r = push_shells(c)) return r;
    }
    c->tune = n;
    if (retile) c->tiles_ready = false;
    if (n.fused == 1 && !c->tiles_ready) {
        if (int r = build_tiles(c)) {
            // the tiles are gone: the two-launch form, which needs none, keeps
            // the context working (the chunks have the new block size)
            c->tune.fused = 0;
            return r;
        }
    }
    return MS_OK;
}

extern "C" int ms_get_tuning(const ms_ctx* c, ms_tuning* t) {
    if (!c || !t) return MS_ERR_ARG;
    *t = c->tune;
    t->tile_nodes = c->tile_L;  // the size in force
    return MS_OK;
}

extern "C" int ms_enable_timing(ms_ctx* c, int32_t on) {
    if (!c) return MS_ERR_ARG;
    c->timing = on != 0;
    return MS_OK;
}

extern "C" int ms_synchronize(ms_ctx* c) {
    if (!c) return fail(nullptr, MS_ERR_ARG, "null context");
    HIPC(c, hipStreamSynchronize(c->stream));
    return MS_OK;
}

namespace {
// reps timed launches of the copy (read = false) or read-only (read = true)
// stream over nbytes
int calibration_stream(int64_t nbytes, int32_t reps, double* out_ms, bool read, const char* name) {
    if (nbytes <= 0 || nbytes % 16 != 0 || reps <= 0)
        return fail(nullptr, MS_ERR_ARG, fmt("%s: nbytes must be a positive multiple of 16", name));
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t bbytes = read ? (size_t)read_out_floats() * sizeof(float) : (size_t)nbytes;
    int rc = MS_OK;
    if (hipMalloc(&a, (size_t)nbytes) != hipSuccess || hipMalloc(&b, bbytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = fail(nullptr, MS_ERR_HIP, fmt("%s: allocation failed", name));
    } else {
        auto launch = [&] {
            if (read)
                launch_read(s, a, (float*)b, nbytes);
            else
                launch_copy(s, a, b, nbytes);
        };
        hipMemsetAsync(a, 0, (size_t)nbytes, s);
        launch();  // warm
        hipEventRecord(e0, s);
        for (int i = 0; i < reps; ++i) launch();
        hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(nullptr, MS_ERR_HIP, fmt("%s: kernel failed", name));
        } else if (out_ms) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            *out_ms = ms / reps;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    if (a) hipFree(a);
    if (b) hipFree(b);
    return rc;
}
}  // namespace

extern "C" int ms_calibration_copy(int64_t nbytes, int32_t reps, double* out_ms) {
    return calibration_stream(nbytes, reps, out_ms, false, "ms_calibration_copy");
}

extern "C" int ms_calibration_read(int64_t nbytes, int32_t reps, double* out_ms) {
    return calibration_stream(nbytes, reps, out_ms, true, "ms_calibration_read");
}

extern "C" int ms_device_info(char* name, int32_t name_len, int32_t* num_cus, int64_t* hbm_bytes,
                              int32_t* mem_clock_khz, int32_t* mem_bus_bits) {
    int dev = 0;
    hipDeviceProp_t p;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) return fail(nullptr, MS_ERR_NO_DEVICE, fmt("ms_device_info: %s", hipGetErrorString(e)));
    if (name && name_len > 0) {
        std::strncpy(name, p.name[0] ? p.name : p.gcnArchName, (size_t)name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (num_cus) *num_cus = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    if (mem_clock_khz) *mem_clock_khz = p.memoryClockRate;
    if (mem_bus_bits) *mem_bus_bits = p.memoryBusWidth;
    return MS_OK;
}

extern "C" int ms_last_step_timing(const ms_ctx* c, double out[3]) {
    if (!c || !out) return MS_ERR_ARG;
    for (int i = 0; i < 3; ++i) out[i] = c->last_ms[i];
    return MS_OK;
}
