// Mesh message-passing network of the learned warm start: parameters,
// device graph, forward, backward, loss and Adam, all on the GPU through the
// kernels of nn_kernels.hip.  It maps per-node features of the flat
// reference mesh and of the programmed metric to node displacements; the
// plain damped dynamics then finish the relaxation (it replaces
// move_nodes_to_ansatz_positions.cpp:26-39 as the source of the start state).
//
// Architecture (width W, L layers):
//   h0 = silu(xn We + be),   e0 = silu(xe Wf + bf)
//   layer l:  m  = silu([h_dst, h_src, e] E1 + c1) E2 + c2        (edge MLP)
//             a_i = sum_{edges j->i} m                            (deterministic CSR sum)
//             h' = h + silu([h, a] N1 + d1) N2 + d2               (node MLP, residual)
//             e' = e + m
//   out = silu(h_L D1 + g1) D2 + g2
// Weights are stored row-major [in x out] in one flat FP32 buffer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ms_common.h"
#include "nn_kernels.h"

using namespace ms;

namespace {

struct Lin {
    int in, out;
    size_t w, b;  // offsets into the flat parameter buffer
};

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

struct ms_nn {
    std::string err;
    ms_nn_arch arch{};
    hipStream_t stream = nullptr;
    Lin enc_n{}, enc_e{}, dec1{}, dec2{};
    std::vector<Lin> e1, e2, n1, n2;
    size_t nparams = 0;
    DBuf<float> params, grads, am, av, partials;
    int adam_t = 0;
    // graph
    int G = 0, N = 0, E = 0;
    DBuf<int> src, dst, in_off, out_off, out_perm, goff, gid;
    DBuf<float> xn, xe;
    // activations (forward cache) and gradients
    std::vector<DBuf<float>> h, he;          // L+1 each
    std::vector<DBuf<float>> cat_e, z1, a1, m, agg, cat_n, y1, s1;  // per layer
    DBuf<float> pre_n, pre_e, d1pre, d1, out;
    DBuf<float> dH, dE, dtmp_e, dtmp_e2, dtmp_e3, dtmp_n, dtmp_n2, dtmp_n3, dout, tgt, terms;
    bool have_forward = false;
    // inference (ms_nn_predict): its own graph and three state buffers
    DBuf<int> p_src, p_dst, p_inoff;
    DBuf<float> p_xn, p_xe, p_h, p_e, p_m, p_out;
    int num_cus = 0;
};

namespace {

int nfail(ms_nn* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    set_last_error(msg);
    return code;
}

#define NNC(m, expr)                                                                                        \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return nfail((m), MS_ERR_HIP, fmt("%s: %s", #expr, hipGetErrorString(e_))); \
    } while (0)

Lin make_lin(size_t& off, int in, int out) {
    Lin l{in, out, off, off + (size_t)in * out};
    off += (size_t)in * out + out;
    return l;
}

// Y[M x out] = X[M x in] @ W + b, or silu of it
void linear_fwd(ms_nn* m, const Lin& L, const float* X, int M, float* Y, bool silu = false) {
    nn::GemmArgs g{};
    g.A = X;
    g.sa0 = L.in;
    g.sa1 = 1;
    g.B = m->params.p + L.w;
    g.sb0 = L.out;
    g.sb1 = 1;
    g.C = Y;
    g.ldc = L.out;
    g.bias = m->params.p + L.b;
    g.M = M;
    g.N = L.out;
    g.K = L.in;
    g.ksplit = 1;
    g.silu = silu ? 1 : 0;
    nn::gemm(m->stream, g, nullptr);
}

// The checks of a graph description, and its destination CSR.
int check_graph(ms_nn* m, const ms_nn_graph_desc* g, const char* who, std::vector<int>& in_off) {
    if (!g || g->num_graphs <= 0 || g->num_nodes <= 0 || g->num_edges <= 0 || !g->node_offsets || !g->edge_src ||
        !g->edge_dst || !g->node_feat || !g->edge_feat)
        return nfail(m, MS_ERR_ARG, fmt("%s: bad arguments", who));
    const int G = g->num_graphs, N = g->num_nodes, E = g->num_edges;
    if (g->node_offsets[0] != 0 || g->node_offsets[G] != N)
        return nfail(m, MS_ERR_ARG, fmt("%s: node offsets must run from 0 to num_nodes", who));
    for (int q = 0; q < G; ++q)
        if (g->node_offsets[q + 1] < g->node_offsets[q])
            return nfail(m, MS_ERR_ARG, fmt("%s: node offsets must not decrease", who));
    in_off.assign(N + 1, 0);
    for (int e = 0; e < E; ++e) {
        const int s = g->edge_src[e], d = g->edge_dst[e];
        if (s < 0 || s >= N || d < 0 || d >= N) return nfail(m, MS_ERR_ARG, fmt("%s: edge node out of range", who));
        if (e > 0 && d < g->edge_dst[e - 1])
            return nfail(m, MS_ERR_ARG, fmt("%s: edges must be sorted by destination", who));
        in_off[d + 1]++;
    }
    for (int n = 0; n < N; ++n) in_off[n + 1] += in_off[n];
    return MS_OK;
}

// grads of Y = X W + b given dY: dW += X^T dY, db += colsum(dY), dX = dY W^T (if dX)
int linear_bwd(ms_nn* m, const Lin& L, const float* X, int M, const float* dY, float* dX) {
    // dW: reduction over the M rows, split over blocks, fixed-order sum
    const int ksplit = std::max(1, std::min(64, M / 1024));
    NNC(m, m->partials.ensure((size_t)ksplit * L.in * L.out));
    nn::GemmArgs g{};
    g.A = X;
    g.sa0 = 1;
    g.sa1 = L.in;
    g.B = dY;
    g.sb0 = L.out;
    g.sb1 = 1;
    g.C = m->grads.p + L.w;
    g.ldc = L.out;
    g.bias = nullptr;
    g.M = L.in;
    g.N = L.out;
    g.K = M;
    g.accumulate = 1;
    g.ksplit = ksplit;
    nn::gemm(m->stream, g, m->partials.p);
    nn::colsum(m->stream, dY, M, L.out, m->grads.p + L.b, true);
    if (dX) {
        nn::GemmArgs d{};
        d.A = dY;
        d.sa0 = L.out;
        d.sa1 = 1;
        d.B = m->params.p + L.w;
        d.sb0 = 1;
        d.sb1 = L.out;
        d.C = dX;
        d.ldc = L.in;
        d.M = M;
        d.N = L.in;
        d.K = L.out;
        d.ksplit = 1;
        nn::gemm(m->stream, d, nullptr);
    }
    return MS_OK;
}

int alloc_acts(ms_nn* m) {
    const int W = m->arch.width, L = m->arch.layers;
    const size_t N = (size_t)m->N, E = (size_t)m->E;
    m->h.resize(L + 1);
    m->he.resize(L + 1);
    m->cat_e.resize(L);
    m->z1.resize(L);
    m->a1.resize(L);
    m->m.resize(L);
    m->agg.resize(L);
    m->cat_n.resize(L);
    m->y1.resize(L);
    m->s1.resize(L);
    for (int l = 0; l <= L; ++l) {
        NNC(m, m->h[l].ensure(N * W));
        NNC(m, m->he[l].ensure(E * W));
    }
    for (int l = 0; l < L; ++l) {
        NNC(m, m->cat_e[l].ensure(E * 3 * W));
        NNC(m, m->z1[l].ensure(E * W));
        NNC(m, m->a1[l].ensure(E * W));
        NNC(m, m->m[l].ensure(E * W));
        NNC(m, m->agg[l].ensure(N * W));
        NNC(m, m->cat_n[l].ensure(N * 2 * W));
        NNC(m, m->y1[l].ensure(N * W));
        NNC(m, m->s1[l].ensure(N * W));
    }
    NNC(m, m->pre_n.ensure(N * W));
    NNC(m, m->pre_e.ensure(E * W));
    NNC(m, m->d1pre.ensure(N * W));
    NNC(m, m->d1.ensure(N * W));
    NNC(m, m->out.ensure(N * m->arch.out_dim));
    NNC(m, m->dH.ensure(N * W));
    NNC(m, m->dE.ensure(E * W));
    NNC(m, m->dtmp_e.ensure(E * W));
    NNC(m, m->dtmp_e2.ensure(E * W));
    NNC(m, m->dtmp_e3.ensure(E * 3 * W));
    NNC(m, m->dtmp_n.ensure(N * W));
    NNC(m, m->dtmp_n2.ensure(N * 2 * W));
    NNC(m, m->dtmp_n3.ensure(N * W));
    NNC(m, m->dout.ensure(N * m->arch.out_dim));
    NNC(m, m->tgt.ensure(N * 3));
    NNC(m, m->terms.ensure(2 * (size_t)std::max(1, m->G)));
    return MS_OK;
}

}  // namespace

extern "C" int ms_nn_create(const ms_nn_arch* arch, ms_nn** out) {
    if (!arch || !out || arch->node_features <= 0 || arch->edge_features <= 0 || arch->width <= 0 ||
        arch->layers <= 0 || arch->out_dim != 3)
        return nfail(nullptr, MS_ERR_ARG, "ms_nn_create: bad architecture");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return nfail(nullptr, MS_ERR_NO_DEVICE, "ms_nn_create: no HIP device visible");
    ms_nn* m = new ms_nn();
    m->arch = *arch;
    const int W = arch->width;
    size_t off = 0;
    m->enc_n = make_lin(off, arch->node_features, W);
    m->enc_e = make_lin(off, arch->edge_features, W);
    for (int l = 0; l < arch->layers; ++l) {
        m->e1.push_back(make_lin(off, 3 * W, W));
        m->e2.push_back(make_lin(off, W, W));
        m->n1.push_back(make_lin(off, 2 * W, W));
        m->n2.push_back(make_lin(off, W, W));
    }
    m->dec1 = make_lin(off, W, W);
    m->dec2 = make_lin(off, W, arch->out_dim);
    m->nparams = off;
    // Glorot-uniform weights, zero biases; the output layer starts small
    std::vector<float> p(off, 0.f);
    std::mt19937_64 rng(arch->seed);
    auto init = [&](const Lin& L, float gain) {
        const float a = gain * std::sqrt(6.f / (float)(L.in + L.out));
        std::uniform_real_distribution<float> U(-a, a);
        for (size_t i = 0; i < (size_t)L.in * L.out; ++i) p[L.w + i] = U(rng);
    };
    init(m->enc_n, 1.f);
    init(m->enc_e, 1.f);
    for (int l = 0; l < arch->layers; ++l) {
        init(m->e1[l], 1.f);
        init(m->e2[l], 0.5f);
        init(m->n1[l], 1.f);
        init(m->n2[l], 0.5f);
    }
    init(m->dec1, 1.f);
    init(m->dec2, 0.1f);
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamDefault) == hipSuccess;
    ok = ok && m->params.ensure(off) == hipSuccess && m->grads.ensure(off) == hipSuccess &&
         m->am.ensure(off) == hipSuccess && m->av.ensure(off) == hipSuccess;
    ok = ok && hipMemcpy(m->params.p, p.data(), off * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemset(m->am.p, 0, off * sizeof(float)) == hipSuccess &&
         hipMemset(m->av.p, 0, off * sizeof(float)) == hipSuccess &&
         hipMemset(m->grads.p, 0, off * sizeof(float)) == hipSuccess;
    if (!ok) {
        ms_nn_destroy(m);
        return nfail(nullptr, MS_ERR_HIP, "ms_nn_create: device allocation failed");
    }
    *out = m;
    return MS_OK;
}

extern "C" void ms_nn_destroy(ms_nn* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (auto* b : {&m->params, &m->grads, &m->am, &m->av, &m->partials, &m->xn, &m->xe, &m->pre_n, &m->pre_e,
                    &m->d1pre, &m->d1, &m->out, &m->dH, &m->dE, &m->dtmp_e, &m->dtmp_e2, &m->dtmp_e3, &m->dtmp_n,
                    &m->dtmp_n2,
                    &m->dtmp_n3, &m->dout, &m->tgt, &m->terms})
        b->release();
    for (auto* b : {&m->src, &m->dst, &m->in_off, &m->out_off, &m->out_perm, &m->goff, &m->gid, &m->p_src, &m->p_dst,
                    &m->p_inoff})
        b->release();
    for (auto* b : {&m->p_xn, &m->p_xe, &m->p_h, &m->p_e, &m->p_m, &m->p_out}) b->release();
    for (auto* v : {&m->h, &m->he, &m->cat_e, &m->z1, &m->a1, &m->m, &m->agg, &m->cat_n, &m->y1, &m->s1})
        for (auto& b : *v) b.release();
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

extern "C" const char* ms_nn_last_error(const ms_nn* m) { return m ? m->err.c_str() : "null network"; }
extern "C" int64_t ms_nn_num_params(const ms_nn* m) { return m ? (int64_t)m->nparams : 0; }

extern "C" int ms_nn_get_params(ms_nn* m, float* out) {
    if (!m || !out) return nfail(m, MS_ERR_ARG, "ms_nn_get_params: bad arguments");
    NNC(m, hipStreamSynchronize(m->stream));
    NNC(m, hipMemcpy(out, m->params.p, m->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_nn_set_params(ms_nn* m, const float* in) {
    if (!m || !in) return nfail(m, MS_ERR_ARG, "ms_nn_set_params: bad arguments");
    NNC(m, hipStreamSynchronize(m->stream));
    NNC(m, hipMemcpy(m->params.p, in, m->nparams * sizeof(float), hipMemcpyHostToDevice));
    NNC(m, hipMemset(m->am.p, 0, m->nparams * sizeof(float)));
    NNC(m, hipMemset(m->av.p, 0, m->nparams * sizeof(float)));
    m->adam_t = 0;
    return MS_OK;
}

extern "C" int ms_nn_set_graph(ms_nn* m, const ms_nn_graph_desc* g) {
    if (!m) return nfail(nullptr, MS_ERR_ARG, "ms_nn_set_graph: null network");
    std::vector<int> in_off;
    if (int r = check_graph(m, g, "ms_nn_set_graph", in_off)) return r;
    const int G = g->num_graphs, N = g->num_nodes, E = g->num_edges;
    std::vector<int> out_off(N + 1, 0), gid(N);
    for (int e = 0; e < E; ++e) out_off[g->edge_src[e] + 1]++;
    for (int n = 0; n < N; ++n) out_off[n + 1] += out_off[n];
    std::vector<int> perm(E), cur(out_off.begin(), out_off.end() - 1);
    for (int e = 0; e < E; ++e) perm[cur[g->edge_src[e]]++] = e;  // ascending edge id per source
    for (int q = 0; q < G; ++q)
        for (int n = g->node_offsets[q]; n < g->node_offsets[q + 1]; ++n) gid[n] = q;
    NNC(m, hipStreamSynchronize(m->stream));
    m->G = G;
    m->N = N;
    m->E = E;
    const int Fn = m->arch.node_features, Fe = m->arch.edge_features;
    NNC(m, m->src.ensure(E));
    NNC(m, m->dst.ensure(E));
    NNC(m, m->in_off.ensure(N + 1));
    NNC(m, m->out_off.ensure(N + 1));
    NNC(m, m->out_perm.ensure(E));
    NNC(m, m->goff.ensure(G + 1));
    NNC(m, m->gid.ensure(N));
    NNC(m, m->xn.ensure((size_t)N * Fn));
    NNC(m, m->xe.ensure((size_t)E * Fe));
    NNC(m, hipMemcpy(m->src.p, g->edge_src, E * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->dst.p, g->edge_dst, E * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->in_off.p, in_off.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->out_off.p, out_off.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->out_perm.p, perm.data(), E * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->goff.p, g->node_offsets, (G + 1) * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->gid.p, gid.data(), N * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->xn.p, g->node_feat, (size_t)N * Fn * sizeof(float), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->xe.p, g->edge_feat, (size_t)E * Fe * sizeof(float), hipMemcpyHostToDevice));
    if (int r = alloc_acts(m)) return r;
    m->have_forward = false;
    return MS_OK;
}

extern "C" int ms_nn_forward(ms_nn* m, float* out) {
    if (!m || m->N == 0) return nfail(m, MS_ERR_STATE, "ms_nn_forward: no graph set");
    const int W = m->arch.width, L = m->arch.layers, N = m->N, E = m->E;
    hipStream_t s = m->stream;
    linear_fwd(m, m->enc_n, m->xn.p, N, m->pre_n.p);
    nn::silu(s, m->pre_n.p, m->h[0].p, (long long)N * W);
    linear_fwd(m, m->enc_e, m->xe.p, E, m->pre_e.p);
    nn::silu(s, m->pre_e.p, m->he[0].p, (long long)E * W);
    for (int l = 0; l < L; ++l) {
        nn::edge_concat(s, m->h[l].p, m->he[l].p, m->src.p, m->dst.p, E, W, m->cat_e[l].p);
        linear_fwd(m, m->e1[l], m->cat_e[l].p, E, m->z1[l].p);
        nn::silu(s, m->z1[l].p, m->a1[l].p, (long long)E * W);
        linear_fwd(m, m->e2[l], m->a1[l].p, E, m->m[l].p);
        nn::segment_sum(s, m->m[l].p, m->in_off.p, nullptr, N, W, W, 0, m->agg[l].p, W, false);
        nn::node_concat(s, m->h[l].p, m->agg[l].p, N, W, m->cat_n[l].p);
        linear_fwd(m, m->n1[l], m->cat_n[l].p, N, m->y1[l].p);
        nn::silu(s, m->y1[l].p, m->s1[l].p, (long long)N * W);
        linear_fwd(m, m->n2[l], m->s1[l].p, N, m->h[l + 1].p);  // dh, then + h
        nn::add(s, m->h[l + 1].p, m->h[l].p, (long long)N * W);
        NNC(m, hipMemcpyAsync(m->he[l + 1].p, m->he[l].p, sizeof(float) * (size_t)E * W, hipMemcpyDeviceToDevice, s));
        nn::add(s, m->he[l + 1].p, m->m[l].p, (long long)E * W);
    }
    linear_fwd(m, m->dec1, m->h[L].p, N, m->d1pre.p);
    nn::silu(s, m->d1pre.p, m->d1.p, (long long)N * W);
    linear_fwd(m, m->dec2, m->d1.p, N, m->out.p);
    NNC(m, hipGetLastError());
    m->have_forward = true;
    if (out) {
        NNC(m, hipStreamSynchronize(s));
        NNC(m, hipMemcpy(out, m->out.p, sizeof(float) * (size_t)N * m->arch.out_dim, hipMemcpyDeviceToHost));
    }
    return MS_OK;
}

extern "C" int ms_nn_predict(ms_nn* m, const ms_nn_graph_desc* g, float* out) {
    if (!m || !out) return nfail(m, MS_ERR_ARG, "ms_nn_predict: bad arguments");
    const int W = m->arch.width, L = m->arch.layers;
    if (W != 32 && W != 64) return nfail(m, MS_ERR_ARG, "ms_nn_predict: the fused kernels take width 32 or 64");
    std::vector<int> in_off;
    if (int r = check_graph(m, g, "ms_nn_predict", in_off)) return r;
    const int N = g->num_nodes, E = g->num_edges, Fn = m->arch.node_features, Fe = m->arch.edge_features;
    hipStream_t s = m->stream;
    if (!m->num_cus) {
        int dev = 0;
        NNC(m, hipGetDevice(&dev));
        NNC(m, hipDeviceGetAttribute(&m->num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    NNC(m, hipStreamSynchronize(s));
    NNC(m, m->p_src.ensure(E));
    NNC(m, m->p_dst.ensure(E));
    NNC(m, m->p_inoff.ensure(N + 1));
    NNC(m, m->p_xn.ensure((size_t)N * Fn));
    NNC(m, m->p_xe.ensure((size_t)E * Fe));
    NNC(m, m->p_h.ensure((size_t)N * W));
    NNC(m, m->p_e.ensure((size_t)E * W));
    NNC(m, m->p_m.ensure((size_t)std::max(N, E) * W));  // messages, then the decoder's hidden layer
    NNC(m, m->p_out.ensure((size_t)N * m->arch.out_dim));
    NNC(m, hipMemcpy(m->p_src.p, g->edge_src, E * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->p_dst.p, g->edge_dst, E * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->p_inoff.p, in_off.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->p_xn.p, g->node_feat, (size_t)N * Fn * sizeof(float), hipMemcpyHostToDevice));
    NNC(m, hipMemcpy(m->p_xe.p, g->edge_feat, (size_t)E * Fe * sizeof(float), hipMemcpyHostToDevice));
    linear_fwd(m, m->enc_n, m->p_xn.p, N, m->p_h.p, true);
    linear_fwd(m, m->enc_e, m->p_xe.p, E, m->p_e.p, true);
    const float* P = m->params.p;
    for (int l = 0; l < L; ++l) {
        nn::MlpLayerArgs a{};
        a.h = m->p_h.p;
        a.e = m->p_e.p;
        a.m = m->p_m.p;
        a.src = m->p_src.p;
        a.dst = m->p_dst.p;
        a.in_off = m->p_inoff.p;
        a.grid = 2 * m->num_cus;
        a.W1 = P + m->e1[l].w;
        a.b1 = P + m->e1[l].b;
        a.W2 = P + m->e2[l].w;
        a.b2 = P + m->e2[l].b;
        a.M = E;
        NNC(m, nn::mlp_layer(s, W, true, a));
        a.W1 = P + m->n1[l].w;
        a.b1 = P + m->n1[l].b;
        a.W2 = P + m->n2[l].w;
        a.b2 = P + m->n2[l].b;
        a.M = N;
        NNC(m, nn::mlp_layer(s, W, false, a));
    }
    linear_fwd(m, m->dec1, m->p_h.p, N, m->p_m.p, true);
    linear_fwd(m, m->dec2, m->p_m.p, N, m->p_out.p);
    NNC(m, hipGetLastError());
    NNC(m, hipStreamSynchronize(s));
    NNC(m, hipMemcpy(out, m->p_out.p, sizeof(float) * (size_t)N * m->arch.out_dim, hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_nn_loss_backward(ms_nn* m, const float* target, float* loss_out, float* grads_out) {
    if (!m || !target) return nfail(m, MS_ERR_ARG, "ms_nn_loss_backward: bad arguments");
    if (!m->have_forward) return nfail(m, MS_ERR_STATE, "ms_nn_loss_backward: run ms_nn_forward first");
    const int W = m->arch.width, L = m->arch.layers, N = m->N, E = m->E;
    hipStream_t s = m->stream;
    NNC(m, hipMemcpyAsync(m->tgt.p, target, sizeof(float) * 3 * (size_t)N, hipMemcpyHostToDevice, s));
    NNC(m, hipMemsetAsync(m->grads.p, 0, sizeof(float) * m->nparams, s));
    // loss = sum_g min(err+, err-) / (3 N)
    nn::loss_terms(s, m->out.p, m->tgt.p, m->goff.p, m->G, m->terms.p);
    const float scale = 1.f / (3.f * (float)N);
    nn::loss_grad(s, m->out.p, m->tgt.p, m->gid.p, m->terms.p, N, scale, m->dout.p);
    // decoder
    if (int r = linear_bwd(m, m->dec2, m->d1.p, N, m->dout.p, m->dtmp_n.p)) return r;
    nn::silu_bwd(s, m->d1pre.p, m->dtmp_n.p, m->dtmp_n.p, (long long)N * W);
    if (int r = linear_bwd(m, m->dec1, m->h[L].p, N, m->dtmp_n.p, m->dH.p)) return r;
    NNC(m, hipMemsetAsync(m->dE.p, 0, sizeof(float) * (size_t)E * W, s));
    for (int l = L - 1; l >= 0; --l) {
        // node residual: d(dh) = dH; dH carries on to h_l
        if (int r = linear_bwd(m, m->n2[l], m->s1[l].p, N, m->dH.p, m->dtmp_n.p)) return r;
        nn::silu_bwd(s, m->y1[l].p, m->dtmp_n.p, m->dtmp_n.p, (long long)N * W);
        if (int r = linear_bwd(m, m->n1[l], m->cat_n[l].p, N, m->dtmp_n.p, m->dtmp_n2.p)) return r;
        nn::take_cols(s, m->dtmp_n2.p, N, 2 * W, 0, W, m->dH.p, true);          // d h_l (node MLP input)
        nn::take_cols(s, m->dtmp_n2.p, N, 2 * W, W, W, m->dtmp_n3.p, false);    // d agg
        // edge messages: dm = dagg[dst] + dE (edge residual)
        nn::gather_rows(s, m->dtmp_n3.p, m->dst.p, E, W, m->dtmp_e.p);
        nn::add(s, m->dtmp_e.p, m->dE.p, (long long)E * W);
        if (int r = linear_bwd(m, m->e2[l], m->a1[l].p, E, m->dtmp_e.p, m->dtmp_e2.p)) return r;
        nn::silu_bwd(s, m->z1[l].p, m->dtmp_e2.p, m->dtmp_e2.p, (long long)E * W);
        if (int r = linear_bwd(m, m->e1[l], m->cat_e[l].p, E, m->dtmp_e2.p, m->dtmp_e3.p)) return r;
        // scatter the concat gradient back: dst part by destination CSR,
        // src part by the source permutation, edge part onto dE
        nn::segment_sum(s, m->dtmp_e3.p, m->in_off.p, nullptr, N, W, 3 * W, 0, m->dH.p, W, true);
        nn::segment_sum(s, m->dtmp_e3.p, m->out_off.p, m->out_perm.p, N, W, 3 * W, W, m->dH.p, W, true);
        nn::take_cols(s, m->dtmp_e3.p, E, 3 * W, 2 * W, W, m->dE.p, true);
    }
    // encoders
    nn::silu_bwd(s, m->pre_n.p, m->dH.p, m->dH.p, (long long)N * W);
    if (int r = linear_bwd(m, m->enc_n, m->xn.p, N, m->dH.p, nullptr)) return r;
    nn::silu_bwd(s, m->pre_e.p, m->dE.p, m->dE.p, (long long)E * W);
    if (int r = linear_bwd(m, m->enc_e, m->xe.p, E, m->dE.p, nullptr)) return r;
    NNC(m, hipGetLastError());
    NNC(m, hipStreamSynchronize(s));
    if (loss_out) {
        std::vector<float> t(2 * (size_t)m->G);
        NNC(m, hipMemcpy(t.data(), m->terms.p, t.size() * sizeof(float), hipMemcpyDeviceToHost));
        double acc = 0.0;
        for (int q = 0; q < m->G; ++q) acc += std::min(t[2 * q], t[2 * q + 1]);
        *loss_out = (float)(acc * scale);
    }
    if (grads_out) NNC(m, hipMemcpy(grads_out, m->grads.p, m->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_nn_adam_step(ms_nn* m, float lr, float b1, float b2, float eps, float wd) {
    if (!m) return nfail(nullptr, MS_ERR_ARG, "null network");
    m->adam_t += 1;
    nn::adam(m->stream, m->params.p, m->grads.p, m->am.p, m->av.p, (long long)m->nparams, lr, b1, b2, eps, m->adam_t,
             wd);
    NNC(m, hipGetLastError());
    return MS_OK;
}

extern "C" int ms_nn_gemm_test(int32_t M, int32_t N, int32_t K, int32_t transpose_a, int32_t ksplit, const float* A,
                               const float* B, float* C) {
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return nfail(nullptr, MS_ERR_ARG, "ms_nn_gemm_test: bad args");
    float *dA = nullptr, *dB = nullptr, *dC = nullptr, *dP = nullptr;
    const size_t zs = (size_t)std::max(1, ksplit);
    int rc = MS_OK;
    if (hipMalloc((void**)&dA, sizeof(float) * (size_t)M * K) != hipSuccess ||
        hipMalloc((void**)&dB, sizeof(float) * (size_t)K * N) != hipSuccess ||
        hipMalloc((void**)&dC, sizeof(float) * (size_t)M * N) != hipSuccess ||
        hipMalloc((void**)&dP, sizeof(float) * zs * M * N) != hipSuccess) {
        rc = nfail(nullptr, MS_ERR_HIP, "ms_nn_gemm_test: allocation failed");
    } else {
        (void)hipMemcpy(dA, A, sizeof(float) * (size_t)M * K, hipMemcpyHostToDevice);
        (void)hipMemcpy(dB, B, sizeof(float) * (size_t)K * N, hipMemcpyHostToDevice);
        nn::GemmArgs g{};
        g.A = dA;
        // A given row-major M x K, or (transpose_a) row-major K x M read transposed
        g.sa0 = transpose_a ? 1 : K;
        g.sa1 = transpose_a ? M : 1;
        g.B = dB;
        g.sb0 = N;
        g.sb1 = 1;
        g.C = dC;
        g.ldc = N;
        g.M = M;
        g.N = N;
        g.K = K;
        g.ksplit = (int)zs;
        nn::gemm(nullptr, g, dP);
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(C, dC, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost) != hipSuccess)
            rc = nfail(nullptr, MS_ERR_HIP, "ms_nn_gemm_test: kernel failed");
    }
    for (float* p : {dA, dB, dC, dP})
        if (p) (void)hipFree(p);
    return rc;
}
