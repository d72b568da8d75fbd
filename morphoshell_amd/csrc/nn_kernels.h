// Launch interface of the message-passing-network kernels (nn_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

namespace ms {
namespace nn {

// C = A @ B (+ bias) (+ C if accumulate).  A(i,k) = A[i*sa0 + k*sa1],
// B(k,j) = B[k*sb0 + j*sb1], C(i,j) = C[i*ldc + j].  Operands are rounded to
// bfloat16 for the matrix cores; accumulation is FP32.
struct GemmArgs {
    const float* A;
    long long sa0, sa1;
    const float* B;
    long long sb0, sb1;
    float* C;
    int ldc;
    const float* bias;
    int M, N, K;
    int accumulate;
    int ksplit;   // > 1: split K over blocks; needs a partial buffer of ksplit*M*N floats and ldc == N
    int kchunk;   // set by gemm()
    int silu;     // 1: C = silu(A B + bias) (ksplit 1, no accumulation)
};

// One half of a message-passing layer for inference, in one launch over
// tiles of 64 rows (rows = edges or nodes):
//   edge: m = silu([h_dst, h_src, e] W1 + b1) W2 + b2;  e += m
//   node: h += silu([h, sum of m over the node's in-edges] W1 + b1) W2 + b2
// The concatenated rows are gathered straight into the first GEMM's LDS A
// tile (in-edge sums in CSR order), both weight matrices sit in LDS as
// bfloat16, and bias and SiLU run in the GEMMs' epilogues, so no E x 3W or
// N x 2W intermediate exists.  Operand roundings and the MFMA sequence are
// those of gemm(), so the result equals the unfused forward bit for bit.
struct MlpLayerArgs {
    const float* W1;  // K1 x W row-major (K1 = 3W edge, 2W node)
    const float* b1;  // W
    const float* W2;  // W x W
    const float* b2;  // W
    int M;            // rows: edges or nodes
    float* h;         // N x W
    float* e;         // E x W (edge)
    float* m;         // E x W: written by the edge half, read by the node half
    const int* src;   // E (edge)
    const int* dst;   // E (edge)
    const int* in_off;  // N + 1, edges sorted by destination (node)
    int grid;         // persistent blocks
};
// hipErrorInvalidValue for a width other than 32 or 64.
hipError_t mlp_layer(hipStream_t s, int W, bool edge, const MlpLayerArgs& a);

void gemm(hipStream_t s, const GemmArgs& a, float* partials);
void silu(hipStream_t s, const float* x, float* y, long long n);
void silu_bwd(hipStream_t s, const float* x, const float* dy, float* dx, long long n);
void add(hipStream_t s, float* y, const float* x, long long n);
void edge_concat(hipStream_t s, const float* h, const float* ef, const int* src, const int* dst, int E, int W,
                 float* out);
void node_concat(hipStream_t s, const float* h, const float* a, int N, int W, float* out);
void segment_sum(hipStream_t s, const float* m, const int* off, const int* perm, int N, int W, int ld_in, int col0,
                 float* out, int ld_out, bool accumulate);
void gather_rows(hipStream_t s, const float* a, const int* idx, int E, int W, float* out);
void take_cols(hipStream_t s, const float* in, int R, int ld, int col0, int W, float* out, bool accumulate);
void colsum(hipStream_t s, const float* dy, int R, int C, float* db, bool accumulate);
void adam(hipStream_t s, float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
          float eps, int t, float wd);
void loss_terms(hipStream_t s, const float* pred, const float* tgt, const int* goff, int G, float* out);
void loss_grad(hipStream_t s, const float* pred, const float* tgt, const int* gid, const float* terms, int N,
               float scale, float* dpred);

}  // namespace nn
}  // namespace ms
