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
// Products with few output tiles and a long K are split along K into
// partials in work (gemm_work(M, N, K, batch) floats; throws if smaller).
void gemm(hipStream_t st, int transA, int transB, int M, int N, int K, float alpha, const unsigned short *A, int lda,
          long long sA, const unsigned short *B, int ldb, long long sB, float beta, float *C, int ldc, long long sC,
          int batch, float *work = nullptr, long long nwork = 0);
long long gemm_work(int M, int N, int K, int batch);
// C (M x N) = A (M x K, lda) B (K x N, ldb) + bias, C16 = bf16(C), C16T =
// bf16(C) tiled and transposed (tile_transpose's layout; M % 32 == 0), each
// optional: the arithmetic of gemm, bias_act, tile_transpose
void gemm_bias(hipStream_t st, int M, int N, int K, const unsigned short *A, int lda, const unsigned short *B, int ldb,
               const float *bias, float *C, unsigned short *C16, unsigned short *C16T);
// dst (cols x rows) = src (rows x cols)^T, BF16
void transpose16(hipStream_t st, const unsigned short *src, int rows, int cols, unsigned short *dst);
// fused MLP branch: H += bf16(gelu(Hin16 W1 + b1)) W2 + b2, Hout16 = bf16(H);
// W1t = W1^T (HD x W), W2t = W2^T (W x HD), BF16; Zc (R x HD, the
// pre-activation) and A16c (R x HD, the BF16 activation) are written when given
// With B16 (the Fourier basis, N2 x nbp, modes contiguous and zero past nb)
// and CsT16 (the scaled modes transposed, per plane W x nbp) the kernel also
// adds the spectral branch's B Cs (H <- (B Cs + H) + (U + b2), the order of the
// unfused sequence); requires N2 % MLP_ROWS == 0 and nbp % 32 == 0.
constexpr int MLP_ROWS = 32;
bool mlp_supported(int W, int HD);
// H16T, when given, receives Hout16 in the tiled transposed layout (each
// 32-row tile as W columns of 32 rows; tile_transpose), needs R % 32 == 0.
void tile_transpose(hipStream_t st, const unsigned short *src, long long R, int W, unsigned short *dst);
void mlp(hipStream_t st, long long R, int W, int HD, const unsigned short *Hin16, const unsigned short *W1t,
         const float *b1, const unsigned short *W2t, const float *b2, float *H, unsigned short *Hout16, float *Zc,
         unsigned short *A16c, const unsigned short *B16, const unsigned short *CsT16, int N2, int nbp,
         unsigned short *H16T);
// Cm[p] (nb x W) = B^T H16[p] for P planes from BT (nbp x N2, zero rows past
// nb) and HT (H16 tiled and transposed), chunk partials in work (modes_work floats) summed in
// order; needs nbp and W multiples of 16, N2 of 32, (nbp/16)(W/16) <= 32
// (modes_supported).
bool modes_supported(int nbp, int N2, int W);
long long modes_work(int nbp, int N2, int W, int P);
void modes(hipStream_t st, const unsigned short *BT, int nbp, int nb, int N2, const unsigned short *HT, int W, int P,
           float *Cm, float *work, long long nwork);
// out16T[p][c][k] = bf16(c[p][k][c] d[k][c]) for k < nb, 0 up to nbp
void mode_scale_t(hipStream_t st, const float *c, const float *d, unsigned short *out16T, int P, int nb, int W,
                  int nbp);
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

namespace coral {
namespace nn {

// A trained network on the device, in the flat layout of coral_amd.closure.Layout:
// W0 (cin x W), b0 (W), per layer W1 (W x Hd), b1 (Hd), W2 (Hd x W), b2 (W),
// D (nb x W), then Wp (W x cout), bp (cout).
struct NetParams {
    int cin = 0, cout = 0, width = 0, hidden = 0, layers = 0, nb = 0, N2 = 0;
    const float *w = nullptr;              // FP32 weights (biases and mode scales are read in FP32)
    const unsigned short *w16 = nullptr;   // BF16 copy (GEMM operands)
    const unsigned short *wt16 = nullptr;  // per layer W1^T (Hd x W) then W2^T (W x Hd), BF16
    const unsigned short *basisT16 = nullptr;  // the Fourier basis transposed, nbp x N2 (zero rows past nb), BF16
    const unsigned short *basis16p = nullptr;  // the basis N2 x nbp (modes contiguous, zero past nb), BF16
    int nbp = 0;                               // nb rounded up to a multiple of 32
    long long off_W0 = 0, off_b0 = 0, off_Wp = 0, off_bp = 0;
    long long off_W1[16] = {}, off_b1[16] = {}, off_W2[16] = {}, off_b2[16] = {}, off_D[16] = {};
    long long size = 0;                    // total number of parameters
};
// offsets of the blocks for the given shapes (throws for more than 16 layers)
void net_layout(NetParams &n);
// inference workspace for up to Pmax planes of N2 points: H (Pmax N2 x W),
// H16 (the same, BF16), Cm (Pmax nb x W), Cs16 (Pmax W x nbp, BF16), and the
// work spaces of the mode projection (gemm_work(nb, W, N2, Pmax) for the
// GEMM; for the fused path H16T (W x Pmax N2) and modes_work(nbp, N2, W, Pmax))
struct NetWork {
    float *H = nullptr, *Cm = nullptr, *gwork = nullptr, *mwork = nullptr;
    unsigned short *H16 = nullptr, *Cs16 = nullptr, *H16T = nullptr;   // H16T: W x Pmax N2
    long long ngwork = 0, nmwork = 0;
};
// Y (P N2 x cout, FP32) = network(X16 (P N2 x cin, BF16)), P <= the workspace's planes.
// X16 null: the lift's outputs (H, H16, and H16T when net_tiled) are already
// in the workspace; Y null: the last layer's H16 is left for the caller to
// project (Wp, bp).
void net_forward(hipStream_t st, const NetParams &n, const unsigned short *X16, int P, float *Y, const NetWork &w);
// whether net_forward keeps the tiled transposed copy H16T of each layer's input
bool net_tiled(const NetParams &n, const NetWork &w);
// host rounding FP32 -> BF16 (nearest even), the same as the device kernels'
unsigned short host_bf16(float f);

}  // namespace nn
}  // namespace coral
