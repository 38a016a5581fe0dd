// Kernels of the learned subgrid closure (nn.hip): BF16 GEMMs on MFMA with
// FP32 accumulation and the elementwise / reduction / optimizer kernels of
// its training and inference.  Device pointers; BF16 as 16-bit patterns.
#pragma once

#include "common.hpp"

namespace coral {
namespace nn {

// C[b] = alpha op(A[b]) op(B[b]) + beta C[b], b < batch (strides in elements;
// 0 shares an operand): op(A) M x K (A is M x K row-major, or K x M when
// transA), op(B) K x N (B is K x N, or N x K when transB), C FP32 M x N.
void gemm(hipStream_t st, int transA, int transB, int M, int N, int K, float alpha, const unsigned short *A, int lda,
          long long sA, const unsigned short *B, int ldb, long long sB, float beta, float *C, int ldc, long long sC,
          int batch);
void cast_bf16(hipStream_t st, const float *x, unsigned short *y, long long n);
// z += bias (per column of ncol), out = act(z) (act: 0 identity, 1 GELU), out16 its bf16 copy
void bias_act(hipStream_t st, float *z, const float *bias, int ncol, long long n, int act, float *out,
              unsigned short *out16);
void gelu_bwd(hipStream_t st, const float *z, const float *da, float *dz, unsigned short *dz16, long long n);
long long colsum_work(long long nrow, int ncol);
void colsum(hipStream_t st, const float *x, long long nrow, int ncol, float *out, float *work, float beta);
void mode_scale(hipStream_t st, const float *c, const float *d, float *out, unsigned short *out16, int P, int ne);
void mode_scale_grad(hipStream_t st, const float *c, const float *g, float *dd, int P, int ne, float beta);
void axpy(hipStream_t st, float *y, const float *x, float alpha, long long n, unsigned short *y16);
long long mse_partials(long long n);
void mse(hipStream_t st, const float *y, const float *t, float *grad, unsigned short *grad16, float scale, long long n,
         float *partial);
void adam(hipStream_t st, float *w, const float *g, float *m, float *v, unsigned short *w16, long long n, float lr,
          float b1, float b2, float eps, int t);

}  // namespace nn
}  // namespace coral
