// Kernels of the learned subgrid closure: BF16 GEMMs on MFMA with FP32
// accumulation, and the elementwise, reduction and optimizer kernels of its
// training and inference.  BF16 values are stored as their 16-bit patterns
// (unsigned short); FP32 -> BF16 rounds to nearest even.
#include "nn.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

namespace coral {
namespace nn {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(unsigned short h) { return __uint_as_float(unsigned(h) << 16); }
// FP32 -> BF16, round to nearest even (gfx950's conversion instruction)
__device__ inline unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, static_cast<__bf16>(f)); }

// ---------------------------------------------------------------------------
// C[b] = alpha * op(A[b]) op(B[b]) + beta * C[b] for b < batch.  op(A) is
// M x K (A stored M x K row-major with leading dimension lda, or K x M when
// transA), op(B) K x N (B stored K x N, or N x K when transB); C is M x N
// row-major, FP32.  One workgroup of four wavefronts per 64 x 64 tile of C,
// each wavefront a 32 x 32 quarter as 2 x 2 tiles of v_mfma_f32_16x16x32_bf16.
// The K loop stages 64 x 32 slices of op(A) and op(B) in LDS, both with k
// contiguous: a lane's fragment (8 consecutive k of one row or column) is one
// 16-byte LDS read.  Operand layout of the instruction: A fragment of lane l
// = row l % 16, k = 8 (l / 16) .. + 7; B fragment = column l % 16, the same k;
// result element i of lane l = C(4 (l / 16) + i, l % 16).
// Split K: with nsplit > 1, blockIdx.z = b nsplit + s covers k in
// [s kc, (s + 1) kc) of product b and writes its partial product to
// C + blockIdx.z sC (the caller's work space); k_splitk_sum adds the partials
// in the order of s.
// ---------------------------------------------------------------------------
constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;   // LDS row stride in bf16 (80 B)

// Optional epilogue of an unsplit, unbatched product: C += bias (per
// column), C16 = bf16(C) (same layout), C16T = bf16(C) in the tiled
// transposed layout of tile_transpose (M % 32 == 0).
struct GemmEpi {
    const float *bias = nullptr;
    unsigned short *C16 = nullptr, *C16T = nullptr;
};
__global__ __launch_bounds__(256) void k_gemm_bf16(int transA, int transB, int M, int N, int K, float alpha,
                                                   const unsigned short *__restrict__ A, int lda, long long sA,
                                                   const unsigned short *__restrict__ B, int ldb, long long sB,
                                                   float beta, float *__restrict__ C, int ldc, long long sC,
                                                   int kc, int nsplit, GemmEpi epi) {
    __shared__ __align__(16) unsigned short As[TM * LDK];
    __shared__ __align__(16) unsigned short Bs[TN * LDK];
    const int bz = blockIdx.z / nsplit, ks = blockIdx.z - bz * nsplit;
    A += bz * sA;
    B += bz * sB;
    C += blockIdx.z * sC;
    const int kbeg = ks * kc, kend = kbeg + kc < K ? kbeg + kc : K;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    // each operand's 64 x 32 slice is 256 vectors of 8 consecutive elements
    // along its contiguous dimension, one per thread; a vector inside the
    // matrix with a 16-byte aligned address is one load, others element by
    // element with bounds checks (ragged edges, odd leading dimensions)
    const bool alA = (lda % 8 == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
    const bool alB = (ldb % 8 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        {
            // A: rows m of the tile x k; contiguous along k (or along m when transA)
            const int a0 = transA ? (tid >> 3) : (tid >> 2);        // k (transA) or m
            const int a1 = transA ? 8 * (tid & 7) : 8 * (tid & 3);  // first m (transA) or k of the vector
            const int gm = m0 + (transA ? a1 : a0), gk = k0 + (transA ? a0 : a1);
            unsigned short v[8];
            const bool full = transA ? (gk < kend && gm + 7 < M) : (gm < M && gk + 7 < kend);
            if (alA && full) {
                const uint4 q = *reinterpret_cast<const uint4 *>(A + (transA ? (size_t)gk * lda + gm
                                                                             : (size_t)gm * lda + gk));
                const unsigned short *e = reinterpret_cast<const unsigned short *>(&q);
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = e[j];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int mm = transA ? gm + j : gm, kk = transA ? gk : gk + j;
                    v[j] = (mm < M && kk < kend) ? (transA ? A[(size_t)kk * lda + mm] : A[(size_t)mm * lda + kk]) : 0;
                }
            }
            if (transA) {
#pragma unroll
                for (int j = 0; j < 8; ++j) As[(a1 + j) * LDK + a0] = v[j];
            } else {
                uint4 q;
                unsigned short *e = reinterpret_cast<unsigned short *>(&q);
#pragma unroll
                for (int j = 0; j < 8; ++j) e[j] = v[j];
                *reinterpret_cast<uint4 *>(&As[a0 * LDK + a1]) = q;
            }
        }
        {
            // B: k x columns n of the tile; contiguous along n (or along k when transB)
            const int b0 = transB ? (tid >> 2) : (tid >> 3);        // n (transB) or k
            const int b1 = transB ? 8 * (tid & 3) : 8 * (tid & 7);  // first k (transB) or n of the vector
            const int gn = n0 + (transB ? b0 : b1), gk = k0 + (transB ? b1 : b0);
            unsigned short v[8];
            const bool full = transB ? (gn < N && gk + 7 < kend) : (gk < kend && gn + 7 < N);
            if (alB && full) {
                const uint4 q = *reinterpret_cast<const uint4 *>(B + (transB ? (size_t)gn * ldb + gk
                                                                             : (size_t)gk * ldb + gn));
                const unsigned short *e = reinterpret_cast<const unsigned short *>(&q);
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = e[j];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int nn_ = transB ? gn : gn + j, kk = transB ? gk + j : gk;
                    v[j] = (nn_ < N && kk < kend) ? (transB ? B[(size_t)nn_ * ldb + kk] : B[(size_t)kk * ldb + nn_]) : 0;
                }
            }
            if (transB) {
                uint4 q;
                unsigned short *e = reinterpret_cast<unsigned short *>(&q);
#pragma unroll
                for (int j = 0; j < 8; ++j) e[j] = v[j];
                *reinterpret_cast<uint4 *>(&Bs[b0 * LDK + b1]) = q;
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) Bs[(b1 + j) * LDK + b0] = v[j];
            }
        }
        __syncthreads();
        const int fr = lane & 15, fk = 8 * (lane >> 4);
        bf8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf8 *>(&As[(wm + 16 * i + fr) * LDK + fk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf8 *>(&Bs[(wn + 16 * j + fr) * LDK + fk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    const int col = lane & 15, rq = 4 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gm = m0 + wm + 16 * i + rq + q, gn = n0 + wn + 16 * j + col;
                if (gm < M && gn < N) {
                    float *p = C + (size_t)gm * ldc + gn;
                    const float v0 = alpha * acc[i][j][q];
                    float v = beta == 0.f ? v0 : v0 + beta * *p;
                    if (epi.bias) v = v + epi.bias[gn];
                    *p = v;
                    if (epi.C16 || epi.C16T) {
                        const unsigned short h = f2bf(v);
                        if (epi.C16) epi.C16[(size_t)gm * ldc + gn] = h;
                        if (epi.C16T) epi.C16T[((size_t)(gm / 32) * N + gn) * 32 + (gm & 31)] = h;
                    }
                }
            }
}

// C[b][m][n] = alpha sum_s P[b nsplit + s][m][n] + beta C[b][m][n]
__global__ void k_splitk_sum(const float *__restrict__ P, int nsplit, int M, int N, int batch, float alpha,
                             float beta, float *__restrict__ C, int ldc, long long sC) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x, mn = (long long)M * N;
    if (i >= mn * batch) return;
    const long long b = i / mn, e = i - b * mn;
    const int m = int(e / N), n = int(e - (long long)m * N);
    const float *p = P + b * nsplit * mn + e;
    float acc = 0.f;
    for (int s = 0; s < nsplit; ++s) acc += p[s * mn];
    float *c = C + b * sC + (long long)m * ldc + n;
    const float v = alpha * acc;
    *c = beta == 0.f ? v : v + beta * *c;
}

// Split of K for products with few output tiles and a long K (per-plane
// projections onto the Fourier modes, weight gradients over a plane's
// points): about 1024 workgroups, at least 512 of K each; 1 = no split.
static void split_of(int M, int N, int K, int batch, int &nsplit, int &kc) {
    const long long tiles = (long long)((M + TM - 1) / TM) * ((N + TN - 1) / TN) * batch;
    nsplit = 1;
    kc = K;
    if (K < 1024 || tiles >= 512) return;
    long long s = (1024 + tiles - 1) / tiles;
    if (s > K / 512) s = K / 512;
    if (s < 2) return;
    kc = int(((K + s - 1) / s + TK - 1) / TK * TK);
    nsplit = (K + kc - 1) / kc;
}
long long gemm_work(int M, int N, int K, int batch) {
    if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return 0;
    int ns, kc;
    split_of(M, N, K, batch, ns, kc);
    return ns > 1 ? (long long)ns * batch * M * N : 0;
}

void gemm(hipStream_t st, int transA, int transB, int M, int N, int K, float alpha, const unsigned short *A, int lda,
          long long sA, const unsigned short *B, int ldb, long long sB, float beta, float *C, int ldc, long long sC,
          int batch, float *work, long long nwork) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    if (K <= 0) throw std::runtime_error("gemm: K must be positive");
    int ns, kc;
    split_of(M, N, K, batch, ns, kc);
    const dim3 grid(unsigned((N + TN - 1) / TN), unsigned((M + TM - 1) / TM), unsigned(batch * ns));
    if (grid.y > 65535u || (long long)batch * ns > 65535) throw std::runtime_error("gemm: grid too large (M or batch)");
    if (ns == 1) {
        hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, transA, transB, M, N, K, alpha, A, lda, sA, B, ldb,
                           sB, beta, C, ldc, sC, K, 1, GemmEpi{});
        check_launch("bf16 gemm");
        return;
    }
    const long long need = (long long)ns * batch * M * N;
    if (!work || nwork < need)
        throw std::runtime_error("gemm: split-K work space of " + std::to_string(need) + " floats required");
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, transA, transB, M, N, K, 1.f, A, lda, sA, B, ldb, sB, 0.f,
                       work, N, (long long)M * N, kc, ns, GemmEpi{});
    const long long n = (long long)batch * M * N;
    hipLaunchKernelGGL(k_splitk_sum, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, work, ns, M, N, batch, alpha,
                       beta, C, ldc, sC);
    check_launch("bf16 gemm (split K)");
}

// C (M x N, ldc = N) = A B + bias, C16 = bf16(C), C16T = bf16(C) tiled and
// transposed (when given; M % 32 == 0): the sequence gemm, bias_act and
// tile_transpose in one kernel, with their arithmetic
void gemm_bias(hipStream_t st, int M, int N, int K, const unsigned short *A, int lda, const unsigned short *B, int ldb,
               const float *bias, float *C, unsigned short *C16, unsigned short *C16T) {
    if (M <= 0 || N <= 0) return;
    if (K <= 0) throw std::runtime_error("gemm: K must be positive");
    if (C16T && M % 32) throw std::runtime_error("gemm: the tiled copy needs M % 32 == 0");
    const dim3 grid(unsigned((N + TN - 1) / TN), unsigned((M + TM - 1) / TM), 1);
    if (grid.y > 65535u) throw std::runtime_error("gemm: grid too large (M)");
    GemmEpi e;
    e.bias = bias;
    e.C16 = C16;
    e.C16T = C16T;
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, 0, 0, M, N, K, 1.f, A, lda, 0LL, B, ldb, 0LL, 0.f, C, N,
                       0LL, K, 1, e);
    check_launch("bf16 gemm + bias");
}

// dst (cols x rows) = src (rows x cols)^T, BF16 (the MLP kernel's weight
// operands, a column of W1 or W2 contiguous along k; the layer input for the
// mode projections): 64 x 64 tiles through LDS, both sides read and written
// along their rows
__global__ __launch_bounds__(256) void k_transpose16(const unsigned short *__restrict__ src, int rows, int cols,
                                                     unsigned short *__restrict__ dst) {
    __shared__ unsigned short t[64][65];
    const long long r0 = (long long)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        if (r0 + r < rows && c0 + c < cols) t[r][c] = src[(size_t)(r0 + r) * cols + c0 + c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int c = e >> 6, r = e & 63;
        if (r0 + r < rows && c0 + c < cols) dst[(size_t)(c0 + c) * rows + r0 + r] = t[r][c];
    }
}
void transpose16(hipStream_t st, const unsigned short *src, int rows, int cols, unsigned short *dst) {
    if (rows <= 0 || cols <= 0) return;
    const long long gx = (rows + 63) / 64, gy = (cols + 63) / 64;
    if (gx > 2147483647LL || gy > 65535) throw std::runtime_error("transpose: too large");
    hipLaunchKernelGGL(k_transpose16, dim3(unsigned(gx), unsigned(gy)), dim3(256), 0, st, src, rows, cols, dst);
    check_launch("bf16 transpose");
}

// ---------------------------------------------------------------------------
// elementwise, reductions, optimizer
// ---------------------------------------------------------------------------
static inline unsigned blocks(long long n, int t = 256) { return unsigned((n + t - 1) / t); }

__global__ void k_cast(const float *__restrict__ x, unsigned short *__restrict__ y, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}
void cast_bf16(hipStream_t st, const float *x, unsigned short *y, long long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cast, dim3(blocks(n)), dim3(256), 0, st, x, y, n);
    check_launch("cast to bf16");
}

// tanh form of GELU, 0.5 z (1 + tanh u) with u = c (z + 0.044715 z^3),
// c = sqrt(2/pi), evaluated through the identity 0.5 (1 + tanh u) =
// 1 / (1 + exp(-2 u)): one hardware exponential and one hardware reciprocal (1 ulp),
// no branches (exp overflows to inf for very negative z, whose reciprocal 0
// gives z s = -0, and underflows to 0 for large z, giving z).  Derivative:
// s + 2 z s (1 - s) du/dz with s = 1 / (1 + exp(-2 u)).
__device__ inline float gelu_sig(float z) {
    // exp(-2 u) = 2^(z (a + b z^2)), a = -2 c log2(e), b = 0.044715 a: one
    // multiply-add and one multiply feed the hardware exponential (v_exp_f32)
    const float a = -2.f * 0.7978845608028654f * 1.4426950408889634f, b = 0.044715f * a;
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z * __builtin_fmaf(b, z * z, a)));
}
__device__ inline float gelu(float z) { return z * gelu_sig(z); }
__device__ inline float gelu_grad(float z) {
    const float c = 0.7978845608028654f;
    const float s = gelu_sig(z);
    return s + 2.f * z * s * (1.f - s) * c * (1.f + 3.f * 0.044715f * z * z);
}

__global__ void k_bias_act(float *__restrict__ z, const float *__restrict__ bias, int ncol, long long n, int act,
                           float *__restrict__ out, unsigned short *__restrict__ out16) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = z[i] + (bias ? bias[i % ncol] : 0.f);
    z[i] = v;
    const float a = act ? gelu(v) : v;
    if (out) out[i] = a;
    if (out16) out16[i] = f2bf(a);
}
void bias_act(hipStream_t st, float *z, const float *bias, int ncol, long long n, int act, float *out,
              unsigned short *out16) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(n)), dim3(256), 0, st, z, bias, ncol, n, act, out, out16);
    check_launch("bias + activation");
}

// ---------------------------------------------------------------------------
// One MLP branch of a layer, fused: for rows r of H (R x W),
//     Z = Hin16 W1 + b1 (R x HD),  A = bf16(gelu(Z)),  U = A W2 + b2,
//     H <- H + U,  Hout16 = bf16(H)
// with Z and A never leaving the workgroup (unless the caches Zc, A16c are
// given, for the backward pass).  The arithmetic is that of the unfused
// sequence gemm, bias_act (GELU), gemm, bias_act, axpy on the same operands.
// W1t (HD x W) and W2t (W x HD) are the BF16 weights transposed, so that a
// lane's B fragment (8 consecutive k of one column) is one 16-byte load.
// One workgroup of four wavefronts per MLP_ROWS = 32 rows (small register
// and LDS footprints, so that several workgroups share a CU and hide each
// other's memory latency).  Stage 1: wavefront w forms Z for the 32 rows and
// hidden columns [w HD / 4, (w + 1) HD / 4), fragments read straight from
// HBM (Hin16 rows) and L2 (W1t); GELU'd BF16 activations go to an LDS tile.
// Stage 2: wavefront w forms U for rows [16 (w / 2), + 16) and columns
// [W / 2 (w % 2), + W / 2), A fragments from LDS.  Hout16 may alias Hin16:
// every read of a row's Hin16 precedes the barrier between the stages and
// every write follows it.
// ---------------------------------------------------------------------------
// FAST: the inference form, with every tile whole (R % 32 == 0) and no
// caches: no row checks and no cache stores in the instruction stream, and
// addresses as 32-bit offsets from the workgroup's base pointers.
template <int W, int HD, bool FAST>
__global__ __launch_bounds__(256) void k_mlp(long long R, const unsigned short *Hin16,
                                             const unsigned short *__restrict__ W1t, const float *__restrict__ b1,
                                             const unsigned short *__restrict__ W2t, const float *__restrict__ b2,
                                             float *__restrict__ H, unsigned short *Hout16, float *__restrict__ Zc,
                                             unsigned short *__restrict__ A16c, const unsigned short *__restrict__ B16,
                                             const unsigned short *__restrict__ CsT16, int N2, int nbp,
                                             unsigned short *__restrict__ H16T) {
    static_assert(W % 32 == 0 && HD % 64 == 0, "MLP shape");
    constexpr int RB = MLP_ROWS / 16;    // 16-row blocks
    constexpr int LDA = HD + 8;          // LDS row stride of the activation tile (bf16)
    constexpr int KS1 = W / 32, CB1 = HD / 64, KS2 = HD / 32, CB2 = W / 32;
    __shared__ __align__(16) unsigned short As[MLP_ROWS * LDA];
    const long long r0 = (long long)blockIdx.x * MLP_ROWS;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int fr = lane & 15, fk = 8 * (lane >> 4), rq = 4 * (lane >> 4);
    const int sr = 16 * (wv >> 1), sc = (W / 2) * (wv & 1), c0 = wv * (HD / 4);
    const bf8 zero8 = {};
    // Every global read is issued up front, so that a workgroup waits for
    // memory once: the stage-1 fragments (Hin16 rows from HBM, W1t from L2),
    // the old H values of the epilogue's rows, then the synthesis operands.
    const unsigned short *Hin_b = Hin16 + r0 * W;
    float *H_b = H + r0 * W;
    unsigned short *Hout_b = Hout16 + r0 * W;
    bf8 a[RB][KS1], w1[KS1][CB1];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int rl = 16 * rb + fr;
#pragma unroll
        for (int k = 0; k < KS1; ++k)
            a[rb][k] = (FAST || r0 + rl < R) ? *reinterpret_cast<const bf8 *>(Hin_b + rl * W + 32 * k + fk) : zero8;
    }
#pragma unroll
    for (int k = 0; k < KS1; ++k)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb)
            w1[k][cb] = *reinterpret_cast<const bf8 *>(W1t + (size_t)(c0 + 16 * cb + fr) * W + 32 * k + fk);
    float hold[CB2][4], bias1[CB1], bias2[CB2];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rl = sr + rq + q;
            hold[cb][q] = (FAST || r0 + rl < R) ? H_b[rl * W + sc + 16 * cb + fr] : 0.f;
        }
#pragma unroll
    for (int cb = 0; cb < CB1; ++cb) bias1[cb] = b1[c0 + 16 * cb + fr];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) bias2[cb] = b2[sc + 16 * cb + fr];
    // the spectral branch's synthesis S = B Cs for the rows of stage 2 when
    // B16 is given (all MLP_ROWS rows in one plane: N2 % MLP_ROWS == 0): A
    // fragments from the plane's basis rows (N2 x nbp, modes contiguous, zero
    // past nb), B fragments from Cs transposed (W x nbp per plane)
    f4 sacc[CB2];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) sacc[cb] = f4{0.f, 0.f, 0.f, 0.f};
    if (B16) {
        const long long p = r0 / N2;
        const int i = int(r0 - p * N2) + sr + fr;
        const unsigned short *brow = B16 + (size_t)i * nbp + fk;
        const unsigned short *cst = CsT16 + (size_t)p * W * nbp + fk;
        for (int k = 0; k < nbp; k += 32) {
            const bf8 av = *reinterpret_cast<const bf8 *>(brow + k);
#pragma unroll
            for (int cb = 0; cb < CB2; ++cb) {
                const bf8 bv = *reinterpret_cast<const bf8 *>(cst + (size_t)(sc + 16 * cb + fr) * nbp + k);
                sacc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, sacc[cb], 0, 0, 0);
            }
        }
    }
    // stage 1
    f4 acc[RB][CB1];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb) acc[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KS1; ++k)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rb][k], w1[k][cb], acc[rb][cb], 0, 0, 0);
    // stage 2's weight fragments, loaded now: their L2 latency hides behind
    // the GELU epilogue instead of stalling the stage-2 k loop
    bf8 w2[KS2][CB2];
#pragma unroll
    for (int k = 0; k < KS2; ++k)
#pragma unroll
        for (int cb = 0; cb < CB2; ++cb)
            w2[k][cb] = *reinterpret_cast<const bf8 *>(W2t + (size_t)(sc + 16 * cb + fr) * HD + 32 * k + fk);
#pragma unroll
    for (int cb = 0; cb < CB1; ++cb) {
        const int col = c0 + 16 * cb + fr;
        const float bias = bias1[cb];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rl = 16 * rb + rq + q;
                const float v = acc[rb][cb][q] + bias;
                const unsigned short h = f2bf(gelu(v));
                As[rl * LDA + col] = h;
                if constexpr (!FAST) {
                    const long long row = r0 + rl;
                    if (Zc && row < R) Zc[row * HD + col] = v;
                    if (A16c && row < R) A16c[row * HD + col] = h;
                }
            }
    }
    __syncthreads();
    // stage 2
    f4 u[CB2];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) u[cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KS2; ++k) {
        const bf8 av = *reinterpret_cast<const bf8 *>(&As[(sr + fr) * LDA + 32 * k + fk]);
#pragma unroll
        for (int cb = 0; cb < CB2; ++cb) u[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, w2[k][cb], u[cb], 0, 0, 0);
    }
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) {
        const int col = sc + 16 * cb + fr;
        const float bias = bias2[cb];
        unsigned short ht[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rl = sr + rq + q;
            // (S + H) + (U + b2): the order of the unfused GEMM (beta = 1), bias and axpy
            const float hs = B16 ? sacc[cb][q] + hold[cb][q] : hold[cb][q];
            const float h = hs + (u[cb][q] + bias);
            ht[q] = f2bf(h);
            if (FAST || r0 + rl < R) {
                H_b[rl * W + col] = h;
                Hout_b[rl * W + col] = ht[q];
            }
        }
        // the new BF16 values in the tiled transposed layout (tile_index: this
        // workgroup's 32 rows are one W x 32 block), four consecutive rows in
        // one 8-byte store, for the next layer's mode projection
        if (H16T && (FAST || r0 + sr + rq + 3 < R)) {
            uint2 v;
            v.x = unsigned(ht[0]) | (unsigned(ht[1]) << 16);
            v.y = unsigned(ht[2]) | (unsigned(ht[3]) << 16);
            *reinterpret_cast<uint2 *>(H16T + ((size_t)(r0 / MLP_ROWS) * W + col) * MLP_ROWS + sr + rq) = v;
        }
    }
}

// The inference form of k_mlp as a persistent, software-pipelined kernel:
// each workgroup walks row tiles blockIdx.x, + gridDim.x, ... and issues the
// global reads of its next tile (input fragments, old H, basis and mode
// fragments) before computing the current one, so that loads stay in flight
// while the MFMAs, the GELU and the stores run; the weight fragments and
// biases are read once per workgroup.  Two register sets alternate (no copies
// between iterations, which would wait for the loads).  The new H rows go
// through LDS and leave with 16-byte stores.  Arithmetic, order and outputs
// are k_mlp<W, HD, true>'s; requires B16 (the spectral term),
// R % MLP_ROWS == 0, N2 % MLP_ROWS == 0 and nbp == 32 NBK.
template <int W, int HD, int NBK>
struct MlpTile {
    static constexpr int RB = MLP_ROWS / 16, KS1 = W / 32, CB2 = W / 32;
    bf8 a[RB][KS1];
    float hold[CB2][4];
    bf8 bas[NBK];
    bf8 cst[NBK][CB2];
};

template <int W, int HD, int NBK>
__device__ inline void mlp_tile_load(MlpTile<W, HD, NBK> &T, long long t, const unsigned short *Hin16,
                                     const float *H, const unsigned short *B16, const unsigned short *CsT16, int N2,
                                     int fr, int fk, int rq, int sr, int sc) {
    using TT = MlpTile<W, HD, NBK>;
    const long long r0 = t * MLP_ROWS;
    const unsigned short *Hin_b = Hin16 + r0 * W;
#pragma unroll
    for (int rb = 0; rb < TT::RB; ++rb)
#pragma unroll
        for (int k = 0; k < TT::KS1; ++k)
            T.a[rb][k] = *reinterpret_cast<const bf8 *>(Hin_b + (16 * rb + fr) * W + 32 * k + fk);
#pragma unroll
    for (int cb = 0; cb < TT::CB2; ++cb)
#pragma unroll
        for (int q = 0; q < 4; ++q) T.hold[cb][q] = H[(r0 + sr + rq + q) * W + sc + 16 * cb + fr];
    const long long p = r0 / N2;
    const int i = int(r0 - p * N2) + sr + fr;
    const unsigned short *brow = B16 + (size_t)i * (32 * NBK) + fk;
    const unsigned short *cst = CsT16 + (size_t)p * W * (32 * NBK) + fk;
#pragma unroll
    for (int k = 0; k < NBK; ++k) {
        T.bas[k] = *reinterpret_cast<const bf8 *>(brow + 32 * k);
#pragma unroll
        for (int cb = 0; cb < TT::CB2; ++cb)
            T.cst[k][cb] = *reinterpret_cast<const bf8 *>(cst + (size_t)(sc + 16 * cb + fr) * (32 * NBK) + 32 * k);
    }
}

template <int W, int HD, int NBK>
__device__ inline void mlp_tile_compute(const MlpTile<W, HD, NBK> &T, long long t, unsigned short *As,
                                        const bf8 (&w1)[W / 32][HD / 64], const bf8 (&w2)[HD / 32][W / 32],
                                        const float (&bias1)[HD / 64], const float (&bias2)[W / 32], float *H,
                                        unsigned short *Hout16, unsigned short *H16T, int fr, int fk, int rq,
                                        int sr, int sc, int c0) {
    constexpr int RB = MLP_ROWS / 16, LDA = HD + 8, KS1 = W / 32, CB1 = HD / 64, KS2 = HD / 32, CB2 = W / 32;
    const long long r0 = t * MLP_ROWS;
    f4 sacc[CB2];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) sacc[cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < NBK; ++k)
#pragma unroll
        for (int cb = 0; cb < CB2; ++cb)
            sacc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(T.bas[k], T.cst[k][cb], sacc[cb], 0, 0, 0);
    f4 acc[RB][CB1];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb) acc[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KS1; ++k)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(T.a[rb][k], w1[k][cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
    for (int cb = 0; cb < CB1; ++cb) {
        const int col = c0 + 16 * cb + fr;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int q = 0; q < 4; ++q) As[(16 * rb + rq + q) * LDA + col] = f2bf(gelu(acc[rb][cb][q] + bias1[cb]));
    }
    __syncthreads();
    f4 u[CB2];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) u[cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KS2; ++k) {
        const bf8 av = *reinterpret_cast<const bf8 *>(&As[(sr + fr) * LDA + 32 * k + fk]);
#pragma unroll
        for (int cb = 0; cb < CB2; ++cb) u[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, w2[k][cb], u[cb], 0, 0, 0);
    }
    // the activation tile is free once every wavefront has read it: the new
    // H values go through it, so that the tile's H and H16 rows (each one
    // contiguous run) leave with 16-byte stores
    __syncthreads();
    constexpr int LDH = W + 4;
    float *Hs = reinterpret_cast<float *>(As);   // [MLP_ROWS][LDH], within the activation tile
    static_assert(MLP_ROWS * LDH * 4 <= MLP_ROWS * LDA * 2, "H staging fits the activation tile");
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) {
        const int col = sc + 16 * cb + fr;
        unsigned short ht[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float h = (sacc[cb][q] + T.hold[cb][q]) + (u[cb][q] + bias2[cb]);
            ht[q] = f2bf(h);
            Hs[(sr + rq + q) * LDH + col] = h;
        }
        if (H16T) {
            uint2 v;
            v.x = unsigned(ht[0]) | (unsigned(ht[1]) << 16);
            v.y = unsigned(ht[2]) | (unsigned(ht[3]) << 16);
            *reinterpret_cast<uint2 *>(H16T + ((size_t)t * W + col) * MLP_ROWS + sr + rq) = v;
        }
    }
    __syncthreads();
    float4 *Ho = reinterpret_cast<float4 *>(H + r0 * W);
    for (int e = threadIdx.x; e < MLP_ROWS * W / 4; e += 256) {
        const int rl = (4 * e) / W, n = 4 * e - rl * W;
        const float *v = Hs + rl * LDH + n;
        Ho[e] = make_float4(v[0], v[1], v[2], v[3]);
    }
    uint4 *H16o = reinterpret_cast<uint4 *>(Hout16 + r0 * W);
    for (int e = threadIdx.x; e < MLP_ROWS * W / 8; e += 256) {
        const int rl = (8 * e) / W, n = 8 * e - rl * W;
        const float *v = Hs + rl * LDH + n;
        uint4 o;
        o.x = unsigned(f2bf(v[0])) | (unsigned(f2bf(v[1])) << 16);
        o.y = unsigned(f2bf(v[2])) | (unsigned(f2bf(v[3])) << 16);
        o.z = unsigned(f2bf(v[4])) | (unsigned(f2bf(v[5])) << 16);
        o.w = unsigned(f2bf(v[6])) | (unsigned(f2bf(v[7])) << 16);
        H16o[e] = o;
    }
    // the staging area is the next tile's activation tile
    __syncthreads();
}

template <int W, int HD, int NBK>
__global__ __launch_bounds__(256) void k_mlp_pipe(long long ntiles, const unsigned short *Hin16,
                                                  const unsigned short *__restrict__ W1t,
                                                  const float *__restrict__ b1,
                                                  const unsigned short *__restrict__ W2t,
                                                  const float *__restrict__ b2, float *H, unsigned short *Hout16,
                                                  const unsigned short *__restrict__ B16,
                                                  const unsigned short *__restrict__ CsT16, int N2,
                                                  unsigned short *__restrict__ H16T) {
    static_assert(W % 32 == 0 && HD % 64 == 0 && NBK >= 1, "MLP shape");
    constexpr int LDA = HD + 8, KS1 = W / 32, CB1 = HD / 64, KS2 = HD / 32, CB2 = W / 32;
    __shared__ __align__(16) unsigned short As[MLP_ROWS * LDA];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int fr = lane & 15, fk = 8 * (lane >> 4), rq = 4 * (lane >> 4);
    const int sr = 16 * (wv >> 1), sc = (W / 2) * (wv & 1), c0 = wv * (HD / 4);
    const long long G = gridDim.x;
    long long t = blockIdx.x;
    if (t >= ntiles) return;
    MlpTile<W, HD, NBK> A, B;
    mlp_tile_load<W, HD, NBK>(A, t, Hin16, H, B16, CsT16, N2, fr, fk, rq, sr, sc);
    bf8 w1[KS1][CB1], w2[KS2][CB2];
    float bias1[CB1], bias2[CB2];
#pragma unroll
    for (int k = 0; k < KS1; ++k)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb)
            w1[k][cb] = *reinterpret_cast<const bf8 *>(W1t + (size_t)(c0 + 16 * cb + fr) * W + 32 * k + fk);
#pragma unroll
    for (int k = 0; k < KS2; ++k)
#pragma unroll
        for (int cb = 0; cb < CB2; ++cb)
            w2[k][cb] = *reinterpret_cast<const bf8 *>(W2t + (size_t)(sc + 16 * cb + fr) * HD + 32 * k + fk);
#pragma unroll
    for (int cb = 0; cb < CB1; ++cb) bias1[cb] = b1[c0 + 16 * cb + fr];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) bias2[cb] = b2[sc + 16 * cb + fr];
    for (;;) {
        if (t + G < ntiles) mlp_tile_load<W, HD, NBK>(B, t + G, Hin16, H, B16, CsT16, N2, fr, fk, rq, sr, sc);
        mlp_tile_compute<W, HD, NBK>(A, t, As, w1, w2, bias1, bias2, H, Hout16, H16T, fr, fk, rq, sr, sc, c0);
        t += G;
        if (t >= ntiles) break;
        if (t + G < ntiles) mlp_tile_load<W, HD, NBK>(A, t + G, Hin16, H, B16, CsT16, N2, fr, fk, rq, sr, sc);
        mlp_tile_compute<W, HD, NBK>(B, t, As, w1, w2, bias1, bias2, H, Hout16, H16T, fr, fk, rq, sr, sc, c0);
        t += G;
        if (t >= ntiles) break;
    }
}

// Tiled transposed layout of an R x W BF16 array (R % 32 == 0): row r, column
// c at ((r / 32) W + c) 32 + r % 32, i.e. each 32-row tile stored as its W
// columns of 32 consecutive rows (2 KB at W = 32): contiguous along the rows,
// as MFMA B fragments read them, and local to a tile (a layout with whole
// columns of R rows would put every lane of a fragment on another page).
__global__ void k_tile_transpose(const unsigned short *__restrict__ src, long long R, int W,
                                 unsigned short *__restrict__ dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R * W) return;
    const long long t = i / ((long long)W * MLP_ROWS);
    const int e = int(i - t * W * MLP_ROWS), c = e / MLP_ROWS, rr = e - c * MLP_ROWS;
    dst[i] = src[(t * MLP_ROWS + rr) * W + c];
}
void tile_transpose(hipStream_t st, const unsigned short *src, long long R, int W, unsigned short *dst) {
    if (R <= 0 || W <= 0) return;
    if (R % MLP_ROWS) throw std::runtime_error("tile transpose: R % 32 != 0");
    const long long n = R * W;
    hipLaunchKernelGGL(k_tile_transpose, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, src, R, W, dst);
    check_launch("tile transpose");
}

// Mode projections of a layer's input, per plane p: Cm[p] (nb x W) =
// B^T H16[p], from the basis transposed (BT: nbp x N2, zero rows past nb) and
// H16 in the tiled transposed layout (HT, R = P N2 rows), both contiguous
// along the points.  A
// workgroup of four wavefronts takes chunk c of a plane's points (MODE_CHUNK
// of them) and all nbp x W outputs as 16 x 16 MFMA tiles, tile t on wavefront
// t % 4, fragments straight from L2 (B) and HBM (H16); its partial sums go to
// part[p][c]; k_modes_sum adds the chunks in order (deterministic).
constexpr int MODE_CHUNK = 1024, MODE_TILES = 8;   // tiles per wavefront at most
constexpr int MODE_LDP = 40, MODE_AHEAD = 4;   // LDS row stride (bf16); steps loaded ahead
template <int TPW, int NCH>
__global__ __launch_bounds__(256) void k_modes(const unsigned short *__restrict__ BT, int nbp, int N2,
                                               const unsigned short *__restrict__ HT, int W,
                                               float *__restrict__ part) {
    // per step of 32 points the workgroup copies the step's basis columns
    // (nbp x 32) and H columns (W x 32) into LDS once, and every tile reads its
    // fragments there (a global byte is read once per workgroup, not once per
    // tile that uses it).  Double-buffered LDS; the global loads run
    // MODE_AHEAD steps ahead in a ring of registers that are never copied into
    // one another.  One barrier per step.
    extern __shared__ __align__(16) unsigned short sm[];   // [2][nbp + W][MODE_LDP]
    const int c = blockIdx.x, p = blockIdx.y, nchunk = gridDim.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int fr = lane & 15, fk = 8 * (lane >> 4), rq = 4 * (lane >> 4);
    const int NB = W / 16, T = (nbp / 16) * NB, R = nbp + W, bufsz = R * MODE_LDP;
    const int i0 = c * MODE_CHUNK, i1 = min(N2, i0 + MODE_CHUNK), nst = (i1 - i0) / 32;
    // the wavefront's tiles: mode block and channel block of t = wv + 4 j
    int aoff[TPW], boff[TPW];
    bool on[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        const int t = wv + 4 * j;
        on[j] = t < T;
        const int mb = on[j] ? t / NB : 0, nb_ = on[j] ? t - mb * NB : 0;
        aoff[j] = (16 * mb + fr) * MODE_LDP + fk;
        boff[j] = (nbp + 16 * nb_ + fr) * MODE_LDP + fk;
    }
    // this thread's 16-byte pieces of a step: row r = q / 4, points 8 (q % 4) ..
    // (a basis row advances 32 points per step, the tiled H 32 W elements)
    const unsigned short *src[NCH];
    int dst[NCH], sstr[NCH];
    bool has[NCH];
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
        const int q = tid + 256 * i, r = q >> 2, part8 = 8 * (q & 3);
        has[i] = r < R;
        src[i] = !has[i] ? BT
                 : r < nbp ? BT + (size_t)r * N2 + i0 + part8
                           : HT + ((size_t)p * N2 + i0) * W + (size_t)(r - nbp) * 32 + part8;
        sstr[i] = r < nbp ? 32 : 32 * W;
        dst[i] = r * MODE_LDP + part8;
    }
    f4 acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
    u32x4 ring[MODE_AHEAD][NCH];
#pragma unroll
    for (int d = 0; d < MODE_AHEAD; ++d)
#pragma unroll
        for (int i = 0; i < NCH; ++i)
            if (d < nst && has[i]) ring[d][i] = *reinterpret_cast<const u32x4 *>(src[i] + (size_t)d * sstr[i]);
    // one step: store the ring's slot D into LDS buffer D % 2, refill the slot
    // MODE_AHEAD steps ahead, wait for the workgroup, feed the MFMAs (written
    // out per slot so that every ring index is a constant)
#define MODE_STEP(D)                                                                                            \
    if (s0 + D < nst) {                                                                                         \
        unsigned short *buf = sm + (D & 1) * bufsz;                                                             \
        _Pragma("unroll") for (int i = 0; i < NCH; ++i) if (has[i])                                             \
            *reinterpret_cast<u32x4 *>(buf + dst[i]) = ring[D][i];                                              \
        const int sn = s0 + D + MODE_AHEAD;                                                                     \
        _Pragma("unroll") for (int i = 0; i < NCH; ++i) if (sn < nst && has[i])                                 \
            ring[D][i] = *reinterpret_cast<const u32x4 *>(src[i] + (size_t)sn * sstr[i]);                       \
        __syncthreads();                                                                                        \
        _Pragma("unroll") for (int j = 0; j < TPW; ++j) if (on[j]) {                                            \
            const bf8 a = *reinterpret_cast<const bf8 *>(buf + aoff[j]);                                        \
            const bf8 b = *reinterpret_cast<const bf8 *>(buf + boff[j]);                                        \
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);                            \
        }                                                                                                       \
    }
    static_assert(MODE_AHEAD == 4, "MODE_STEP is written out for four slots");
    for (int s0 = 0; s0 < nst; s0 += MODE_AHEAD) {
        MODE_STEP(0)
        MODE_STEP(1)
        MODE_STEP(2)
        MODE_STEP(3)
    }
#undef MODE_STEP
    float *out = part + ((size_t)p * nchunk + c) * nbp * W;
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        const int t = wv + 4 * j;
        if (on[j]) {
            const int mb = t / NB, nb_ = t - mb * NB;
#pragma unroll
            for (int q = 0; q < 4; ++q) out[(size_t)(16 * mb + rq + q) * W + 16 * nb_ + fr] = acc[j][q];
        }
    }
}
__global__ void k_modes_sum(const float *__restrict__ part, int nchunk, int nbp, int nb, int W, int P,
                            float *__restrict__ Cm) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)P * nb * W) return;
    const long long p = i / ((long long)nb * W);
    const int e = int(i - p * nb * W);
    const float *src = part + (size_t)p * nchunk * nbp * W + e;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += src[(size_t)c * nbp * W];
    Cm[i] = s;
}
bool modes_supported(int nbp, int N2, int W) {
    // nbp + W <= 256: a step's rows in LDS (40 KB double-buffered), at most
    // four 16-byte pieces per thread
    return nbp > 0 && W > 0 && nbp % 16 == 0 && W % 16 == 0 && N2 > 0 && N2 % 32 == 0 &&
           (nbp / 16) * (W / 16) <= 4 * MODE_TILES && nbp + W <= 256;
}
long long modes_work(int nbp, int N2, int W, int P) {
    return (long long)P * ((N2 + MODE_CHUNK - 1) / MODE_CHUNK) * nbp * W;
}
void modes(hipStream_t st, const unsigned short *BT, int nbp, int nb, int N2, const unsigned short *HT, int W, int P,
           float *Cm, float *work, long long nwork) {
    if (P <= 0) return;
    if (!modes_supported(nbp, N2, W) || nb > nbp)
        throw std::runtime_error("modes: shapes (nbp, W multiples of 16, N2 of 32, at most 32 tiles)");
    const int nchunk = (N2 + MODE_CHUNK - 1) / MODE_CHUNK;
    if (!work || nwork < modes_work(nbp, N2, W, P)) throw std::runtime_error("modes: work space too small");
    const int tpw = ((nbp / 16) * (W / 16) + 3) / 4, nch = (4 * (nbp + W) + 255) / 256;
    const size_t lds = sizeof(unsigned short) * 2 * size_t(nbp + W) * MODE_LDP;
    const dim3 g(nchunk, P);
    bool done = false;
#define MODES_CASE(tpw_, nch_)                                                                                 \
    if (!done && tpw == tpw_ && nch == nch_) {                                                                 \
        hipLaunchKernelGGL((k_modes<tpw_, nch_>), g, dim3(256), lds, st, BT, nbp, N2, HT, W, work);             \
        done = true;                                                                                           \
    }
#define MODES_TPW(tpw_) MODES_CASE(tpw_, 1) MODES_CASE(tpw_, 2) MODES_CASE(tpw_, 3) MODES_CASE(tpw_, 4)
    MODES_TPW(1) MODES_TPW(2) MODES_TPW(3) MODES_TPW(4) MODES_TPW(5) MODES_TPW(6) MODES_TPW(7) MODES_TPW(8)
#undef MODES_TPW
#undef MODES_CASE
    if (!done) throw std::runtime_error("modes: no kernel for these shapes");
    const long long n = (long long)P * nb * W;
    hipLaunchKernelGGL(k_modes_sum, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, work, nchunk, nbp, nb, W, P,
                       Cm);
    check_launch("mode projection");
}

bool mlp_supported(int W, int HD) { return (W == 64 || W == 32) && (HD == 256 || HD == 128); }
void mlp(hipStream_t st, long long R, int W, int HD, const unsigned short *Hin16, const unsigned short *W1t,
         const float *b1, const unsigned short *W2t, const float *b2, float *H, unsigned short *Hout16, float *Zc,
         unsigned short *A16c, const unsigned short *B16, const unsigned short *CsT16, int N2, int nbp,
         unsigned short *H16T) {
    if (R <= 0) return;
    if (H16T && R % MLP_ROWS != 0) throw std::runtime_error("mlp: the transposed copy needs R % 32 == 0");
    if (B16 && (!CsT16 || N2 <= 0 || N2 % MLP_ROWS != 0 || nbp <= 0 || nbp % 32 != 0))
        throw std::runtime_error("mlp: the spectral term needs N2 % 32 == 0 and nbp a positive multiple of 32");
    const long long nblk = (R + MLP_ROWS - 1) / MLP_ROWS;
    if (nblk > 2147483647LL) throw std::runtime_error("mlp: too many rows");
    const dim3 g{unsigned(nblk)}, b{256};
    const bool fast = !Zc && !A16c && R % MLP_ROWS == 0;
    // the pipelined inference kernel: spectral term present, up to four 32-mode blocks
    const bool pipe = fast && B16 && nbp <= 128 && !std::getenv("CORAL_AMD_MLP_UNPIPELINED");
    if (pipe) {
        // compute units of the current device (read once per device)
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("mlp: no current device");
        static int ncu_of[64] = {};
        if (dev < 0 || dev >= 64) throw std::runtime_error("mlp: device index out of range");
        if (!ncu_of[dev] &&
            (hipDeviceGetAttribute(&ncu_of[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
             ncu_of[dev] <= 0))
            throw std::runtime_error("mlp: cannot read the device's compute-unit count");
        const int ncu = ncu_of[dev];
        // as many workgroups as are resident at once (read once per kernel): a
        // persistent workgroup that had to wait for a slot would start its share late
#define MLP_PIPE(w_, hd_, nbk_)                                                                                   \
        if (W == w_ && HD == hd_ && nbp == 32 * nbk_) {                                                          \
            static int per = 0;                                                                                  \
            if (!per &&                                                                                          \
                (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                                   \
                     &per, reinterpret_cast<const void *>(k_mlp_pipe<w_, hd_, nbk_>), 256, 0) != hipSuccess ||  \
                 per < 1))                                                                                       \
                per = 1;                                                                                         \
            const unsigned gp = unsigned(std::min<long long>(nblk, (long long)per * ncu));                      \
            hipLaunchKernelGGL((k_mlp_pipe<w_, hd_, nbk_>), dim3(gp), b, 0, st, nblk, Hin16, W1t, b1, W2t, b2, H, \
                               Hout16, B16, CsT16, N2, H16T);                                                    \
            check_launch("pipelined mlp");                                                                       \
            return;                                                                                              \
        }
        // (width 64 needs more registers than leave two wavefronts per SIMD: it
        // keeps the one-tile kernel)
        MLP_PIPE(32, 128, 1) MLP_PIPE(32, 128, 2) MLP_PIPE(32, 128, 3) MLP_PIPE(32, 128, 4)
#undef MLP_PIPE
    }
#define MLP_CASE(w_, hd_)                                                                                       \
    if (W == w_ && HD == hd_) {                                                                                 \
        if (fast)                                                                                               \
            hipLaunchKernelGGL((k_mlp<w_, hd_, true>), g, b, 0, st, R, Hin16, W1t, b1, W2t, b2, H, Hout16, Zc,  \
                               A16c, B16, CsT16, N2, nbp, H16T);                                                \
        else                                                                                                    \
            hipLaunchKernelGGL((k_mlp<w_, hd_, false>), g, b, 0, st, R, Hin16, W1t, b1, W2t, b2, H, Hout16, Zc, \
                               A16c, B16, CsT16, N2, nbp, H16T);                                                \
        check_launch("fused mlp");                                                                              \
        return;                                                                                                 \
    }
    MLP_CASE(64, 256)
    MLP_CASE(64, 128)
    MLP_CASE(32, 256)
    MLP_CASE(32, 128)
#undef MLP_CASE
    throw std::runtime_error("mlp: width must be 32 or 64 and hidden 128 or 256");
}

__global__ void k_gelu_bwd(const float *__restrict__ z, const float *__restrict__ da, float *__restrict__ dz,
                           unsigned short *__restrict__ dz16, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = da[i] * gelu_grad(z[i]);
    dz[i] = g;
    if (dz16) dz16[i] = f2bf(g);
}
void gelu_bwd(hipStream_t st, const float *z, const float *da, float *dz, unsigned short *dz16, long long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks(n)), dim3(256), 0, st, z, da, dz, dz16, n);
    check_launch("gelu backward");
}

// column sums of x [nrow][ncol]: per row chunk of 256 rows, partial sums in
// work [nchunk][ncol]; then one thread per column adds the chunks in order
// (deterministic)
__global__ void k_colsum_part(const float *__restrict__ x, long long nrow, int ncol, float *__restrict__ work) {
    const int c = threadIdx.x + blockIdx.y * blockDim.x;
    if (c >= ncol) return;
    const long long r0 = (long long)blockIdx.x * 256, r1 = r0 + 256 < nrow ? r0 + 256 : nrow;
    float s = 0.f;
    for (long long r = r0; r < r1; ++r) s += x[r * ncol + c];
    work[(long long)blockIdx.x * ncol + c] = s;
}
__global__ void k_colsum_fin(const float *__restrict__ work, int nchunk, int ncol, float *__restrict__ out,
                             float beta) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    float s = 0.f;
    for (int k = 0; k < nchunk; ++k) s += work[(long long)k * ncol + c];
    out[c] = beta == 0.f ? s : s + beta * out[c];
}
long long colsum_work(long long nrow, int ncol) { return ((nrow + 255) / 256) * (long long)ncol; }
void colsum(hipStream_t st, const float *x, long long nrow, int ncol, float *out, float *work, float beta) {
    if (nrow <= 0 || ncol <= 0) return;
    const long long nchunk = (nrow + 255) / 256;
    if (nchunk > 2147483647LL) throw std::runtime_error("colsum: too many rows");
    const int tpb = ncol < 256 ? ncol : 256;
    hipLaunchKernelGGL(k_colsum_part, dim3(unsigned(nchunk), unsigned((ncol + tpb - 1) / tpb)), dim3(tpb), 0, st, x,
                       nrow, ncol, work);
    hipLaunchKernelGGL(k_colsum_fin, dim3(blocks(ncol)), dim3(256), 0, st, work, int(nchunk), ncol, out, beta);
    check_launch("column sums");
}

// out[p][e] = c[p][e] d[e] (e < ne, per plane p); out16 the bf16 copy
__global__ void k_mode_scale(const float *__restrict__ c, const float *__restrict__ d, float *__restrict__ out,
                             unsigned short *__restrict__ out16, int P, int ne) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)P * ne) return;
    const float v = c[i] * d[i % ne];
    if (out) out[i] = v;
    if (out16) out16[i] = f2bf(v);
}
void mode_scale(hipStream_t st, const float *c, const float *d, float *out, unsigned short *out16, int P, int ne) {
    if (P <= 0 || ne <= 0) return;
    hipLaunchKernelGGL(k_mode_scale, dim3(blocks((long long)P * ne)), dim3(256), 0, st, c, d, out, out16, P, ne);
    check_launch("mode scale");
}
// out16T[p][c][k] = bf16(c[p][k][c] d[k][c]) for k < nb, 0 for nb <= k < nbp
// (the scaled modes transposed and padded, the fused MLP kernel's operand)
__global__ void k_mode_scale_t(const float *__restrict__ c, const float *__restrict__ d,
                               unsigned short *__restrict__ out16T, int P, int nb, int W, int nbp) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)P * W * nbp) return;
    const int k = int(i % nbp);
    const long long pc = i / nbp;
    const int ch = int(pc % W);
    const long long p = pc / W;
    out16T[i] = k < nb ? f2bf(c[(p * nb + k) * W + ch] * d[(long long)k * W + ch]) : (unsigned short)0;
}
void mode_scale_t(hipStream_t st, const float *c, const float *d, unsigned short *out16T, int P, int nb, int W,
                  int nbp) {
    if (P <= 0 || nb <= 0 || W <= 0) return;
    if (nbp < nb) throw std::runtime_error("mode_scale_t: nbp < nb");
    const long long n = (long long)P * W * nbp;
    hipLaunchKernelGGL(k_mode_scale_t, dim3(blocks(n)), dim3(256), 0, st, c, d, out16T, P, nb, W, nbp);
    check_launch("mode scale (transposed)");
}
// dd[e] (+)= sum_p c[p][e] g[p][e], one thread per e, planes in order
__global__ void k_mode_scale_grad(const float *__restrict__ c, const float *__restrict__ g, float *__restrict__ dd,
                                  int P, int ne, float beta) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    float s = 0.f;
    for (int p = 0; p < P; ++p) s += c[(long long)p * ne + e] * g[(long long)p * ne + e];
    dd[e] = beta == 0.f ? s : s + beta * dd[e];
}
void mode_scale_grad(hipStream_t st, const float *c, const float *g, float *dd, int P, int ne, float beta) {
    if (ne <= 0) return;
    hipLaunchKernelGGL(k_mode_scale_grad, dim3(blocks(ne)), dim3(256), 0, st, c, g, dd, P, ne, beta);
    check_launch("mode scale gradient");
}

__global__ void k_axpy(float *__restrict__ y, const float *__restrict__ x, float alpha, long long n,
                       unsigned short *__restrict__ y16) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = y[i] + alpha * x[i];
    y[i] = v;
    if (y16) y16[i] = f2bf(v);
}
void axpy(hipStream_t st, float *y, const float *x, float alpha, long long n, unsigned short *y16) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpy, dim3(blocks(n)), dim3(256), 0, st, y, x, alpha, n, y16);
    check_launch("axpy");
}

// grad = scale (y - t); partial[b] = sum over the block's elements of (y - t)^2
__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ t, float *__restrict__ grad,
                      unsigned short *__restrict__ grad16, float scale, long long n, float *__restrict__ partial) {
    __shared__ float red[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float d = 0.f;
    if (i < n) {
        d = y[i] - t[i];
        const float g = scale * d;
        if (grad) grad[i] = g;
        if (grad16) grad16[i] = f2bf(g);
    }
    red[threadIdx.x] = d * d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
long long mse_partials(long long n) { return (n + 255) / 256; }
void mse(hipStream_t st, const float *y, const float *t, float *grad, unsigned short *grad16, float scale, long long n,
         float *partial) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mse, dim3(blocks(n)), dim3(256), 0, st, y, t, grad, grad16, scale, n, partial);
    check_launch("mse");
}

// Adam (Kingma & Ba 2015) on FP32 master weights, with bias-corrected moments;
// w16 receives the bf16 copy the GEMMs read
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, unsigned short *__restrict__ w16, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}
void adam(hipStream_t st, float *w, const float *g, float *m, float *v, unsigned short *w16, long long n, float lr,
          float b1, float b2, float eps, int t) {
    if (n <= 0) return;
    const float c1 = 1.f - powf(b1, float(t)), c2 = 1.f - powf(b2, float(t));
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, w, g, m, v, w16, n, lr, b1, b2, eps, c1, c2);
    check_launch("adam");
}


// ---------------------------------------------------------------------------
// inference of a trained network (the same sequence of kernels as the
// trainer's forward pass, without keeping activations)
// ---------------------------------------------------------------------------
void net_layout(NetParams &n) {
    if (n.layers < 0 || n.layers > 16) throw std::runtime_error("closure network: at most 16 layers");
    const long long W = n.width, Hd = n.hidden;
    long long o = 0;
    n.off_W0 = o, o += (long long)n.cin * W;
    n.off_b0 = o, o += W;
    for (int l = 0; l < n.layers; ++l) {
        n.off_W1[l] = o, o += W * Hd;
        n.off_b1[l] = o, o += Hd;
        n.off_W2[l] = o, o += Hd * W;
        n.off_b2[l] = o, o += W;
        n.off_D[l] = o, o += (long long)n.nb * W;
    }
    n.off_Wp = o, o += W * n.cout;
    n.off_bp = o, o += n.cout;
    n.size = o;
}

bool net_tiled(const NetParams &n, const NetWork &w) {
    return n.basis16p && n.N2 % MLP_ROWS == 0 && w.H16T && modes_supported(n.nbp, n.N2, n.width);
}

void net_forward(hipStream_t st, const NetParams &n, const unsigned short *X16, int P, float *Y, const NetWork &w) {
    const int W = n.width, Hd = n.hidden, nb = n.nb, N2 = n.N2;
    const long long R = (long long)P * N2;
    if (R > 2147483647LL) throw std::runtime_error("closure network: too many rows");
    const int Ri = int(R);
    const bool fused = n.basis16p && N2 % MLP_ROWS == 0;
    const bool tmodes = net_tiled(n, w);
    // lift with its bias, BF16 copy and (for the mode kernel) tiled transposed
    // copy; without X16 the caller has filled H, H16 (and H16T when
    // net_tiled) itself
    if (X16)
        gemm_bias(st, Ri, W, n.cin, X16, n.cin, n.w16 + n.off_W0, W, n.w + n.off_b0, w.H, w.H16,
                  tmodes ? w.H16T : nullptr);
    for (int l = 0; l < n.layers; ++l) {
        // spectral branch: Cm = B^T H16 per plane (from H16 transposed when the
        // mode kernel takes the shapes), Cs = Cm D, H += B Cs (inside the MLP
        // kernel when the planes split into its row tiles)
        if (tmodes)
            modes(st, n.basisT16, n.nbp, nb, N2, w.H16T, W, P, w.Cm, w.mwork, w.nmwork);
        else
            gemm(st, 0, 0, nb, W, N2, 1.f, n.basisT16, N2, 0, w.H16, W, (long long)N2 * W, 0.f, w.Cm, W,
                 (long long)nb * W, P, w.gwork, w.ngwork);
        const unsigned short *wt = n.wt16 + (long long)l * 2 * W * Hd;
        if (fused) {
            mode_scale_t(st, w.Cm, n.w + n.off_D[l], w.Cs16, P, nb, W, n.nbp);
            mlp(st, R, W, Hd, w.H16, wt, n.w + n.off_b1[l], wt + (long long)W * Hd, n.w + n.off_b2[l], w.H, w.H16,
                nullptr, nullptr, n.basis16p, w.Cs16, N2, n.nbp, tmodes ? w.H16T : nullptr);
        } else {
            mode_scale(st, w.Cm, n.w + n.off_D[l], nullptr, w.Cs16, P, nb * W);
            gemm(st, 1, 0, N2, W, nb, 1.f, n.basisT16, N2, 0, w.Cs16, W, (long long)nb * W, 1.f, w.H, W,
                 (long long)N2 * W, P);
            mlp(st, R, W, Hd, w.H16, wt, n.w + n.off_b1[l], wt + (long long)W * Hd, n.w + n.off_b2[l], w.H, w.H16,
                nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr);
        }
    }
    if (!Y) return;   // the caller projects H16 itself
    gemm(st, 0, 0, Ri, n.cout, W, 1.f, w.H16, W, 0, n.w16 + n.off_Wp, n.cout, 0, 0.f, Y, n.cout, 0, 1);
    bias_act(st, Y, n.w + n.off_bp, n.cout, R * n.cout, 0, nullptr, nullptr);
}

unsigned short host_bf16(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

}  // namespace nn
}  // namespace coral

// ---------------------------------------------------------------------------
// C ABI (device pointers; stream = hipStream_t, NULL for the null stream)
// ---------------------------------------------------------------------------
namespace coral {
void set_last_error(const std::string &msg);
}

#define NN_ABI(body)                             \
    try {                                        \
        body;                                    \
        return 0;                                \
    } catch (const std::exception &e) {          \
        coral::set_last_error(e.what());         \
        return -1;                               \
    }


extern "C" {

int coral_amd_nn_gemm(int32_t transA, int32_t transB, int32_t M, int32_t N, int32_t K, float alpha, const void *A,
                      int32_t lda, int64_t sA, const void *B, int32_t ldb, int64_t sB, float beta, float *C,
                      int32_t ldc, int64_t sC, int32_t batch, float *work, int64_t nwork, void *stream) {
    NN_ABI(coral::nn::gemm((hipStream_t)stream, transA, transB, M, N, K, alpha, (const unsigned short *)A, lda, sA,
                           (const unsigned short *)B, ldb, sB, beta, C, ldc, sC, batch, work, nwork))
}
int64_t coral_amd_nn_gemm_work(int32_t M, int32_t N, int32_t K, int32_t batch) {
    return coral::nn::gemm_work(M, N, K, batch);
}
int coral_amd_nn_transpose_bf16(const void *src, int32_t rows, int32_t cols, void *dst, void *stream) {
    NN_ABI(coral::nn::transpose16((hipStream_t)stream, (const unsigned short *)src, rows, cols, (unsigned short *)dst))
}
int coral_amd_nn_mlp(int64_t R, int32_t width, int32_t hidden, const void *Hin16, const void *W1t, const float *b1,
                     const void *W2t, const float *b2, float *H, void *Hout16, float *Zc, void *A16c,
                     const void *basis16, const void *CsT16, int32_t N2, int32_t nbp, void *H16T, void *stream) {
    NN_ABI(coral::nn::mlp((hipStream_t)stream, R, width, hidden, (const unsigned short *)Hin16,
                          (const unsigned short *)W1t, b1, (const unsigned short *)W2t, b2, H,
                          (unsigned short *)Hout16, Zc, (unsigned short *)A16c, (const unsigned short *)basis16,
                          (const unsigned short *)CsT16, N2, nbp, (unsigned short *)H16T))
}
int coral_amd_nn_gemm_bias(int32_t M, int32_t N, int32_t K, const void *A, int32_t lda, const void *B, int32_t ldb,
                           const float *bias, float *C, void *C16, void *C16T, void *stream) {
    NN_ABI(coral::nn::gemm_bias((hipStream_t)stream, M, N, K, (const unsigned short *)A, lda,
                                (const unsigned short *)B, ldb, bias, C, (unsigned short *)C16,
                                (unsigned short *)C16T))
}
int coral_amd_nn_tile_transpose(const void *src, int64_t R, int32_t width, void *dst, void *stream) {
    NN_ABI(coral::nn::tile_transpose((hipStream_t)stream, (const unsigned short *)src, R, width, (unsigned short *)dst))
}
int32_t coral_amd_nn_modes_supported(int32_t nbp, int32_t N2, int32_t width) {
    return coral::nn::modes_supported(nbp, N2, width) ? 1 : 0;
}
int64_t coral_amd_nn_modes_work(int32_t nbp, int32_t N2, int32_t width, int32_t P) {
    return coral::nn::modes_work(nbp, N2, width, P);
}
int coral_amd_nn_modes(const void *BT, int32_t nbp, int32_t nb, int32_t N2, const void *HT, int32_t width, int32_t P,
                       float *Cm, float *work, int64_t nwork, void *stream) {
    NN_ABI(coral::nn::modes((hipStream_t)stream, (const unsigned short *)BT, nbp, nb, N2, (const unsigned short *)HT,
                            width, P, Cm, work, nwork))
}
int coral_amd_nn_mode_scale_t(const float *c, const float *d, void *out16T, int32_t P, int32_t nb, int32_t width,
                              int32_t nbp, void *stream) {
    NN_ABI(coral::nn::mode_scale_t((hipStream_t)stream, c, d, (unsigned short *)out16T, P, nb, width, nbp))
}
int coral_amd_nn_cast_bf16(const float *x, void *y, int64_t n, void *stream) {
    NN_ABI(coral::nn::cast_bf16((hipStream_t)stream, x, (unsigned short *)y, n))
}
int coral_amd_nn_bias_act(float *z, const float *bias, int32_t ncol, int64_t n, int32_t act, float *out, void *out16,
                          void *stream) {
    NN_ABI(coral::nn::bias_act((hipStream_t)stream, z, bias, ncol, n, act, out, (unsigned short *)out16))
}
int coral_amd_nn_gelu_bwd(const float *z, const float *da, float *dz, void *dz16, int64_t n, void *stream) {
    NN_ABI(coral::nn::gelu_bwd((hipStream_t)stream, z, da, dz, (unsigned short *)dz16, n))
}
int64_t coral_amd_nn_colsum_work(int64_t nrow, int32_t ncol) { return coral::nn::colsum_work(nrow, ncol); }
int coral_amd_nn_colsum(const float *x, int64_t nrow, int32_t ncol, float *out, float *work, float beta, void *stream) {
    NN_ABI(coral::nn::colsum((hipStream_t)stream, x, nrow, ncol, out, work, beta))
}
int coral_amd_nn_mode_scale(const float *c, const float *d, float *out, void *out16, int32_t P, int32_t ne,
                            void *stream) {
    NN_ABI(coral::nn::mode_scale((hipStream_t)stream, c, d, out, (unsigned short *)out16, P, ne))
}
int coral_amd_nn_mode_scale_grad(const float *c, const float *g, float *dd, int32_t P, int32_t ne, float beta,
                                 void *stream) {
    NN_ABI(coral::nn::mode_scale_grad((hipStream_t)stream, c, g, dd, P, ne, beta))
}
int coral_amd_nn_axpy(float *y, const float *x, float alpha, int64_t n, void *y16, void *stream) {
    NN_ABI(coral::nn::axpy((hipStream_t)stream, y, x, alpha, n, (unsigned short *)y16))
}
int64_t coral_amd_nn_mse_partials(int64_t n) { return coral::nn::mse_partials(n); }
int coral_amd_nn_mse(const float *y, const float *t, float *grad, void *grad16, float scale, int64_t n, float *partial,
                     void *stream) {
    NN_ABI(coral::nn::mse((hipStream_t)stream, y, t, grad, (unsigned short *)grad16, scale, n, partial))
}
int coral_amd_nn_adam(float *w, const float *g, float *m, float *v, void *w16, int64_t n, float lr, float b1, float b2,
                      float eps, int32_t t, void *stream) {
    NN_ABI(coral::nn::adam((hipStream_t)stream, w, g, m, v, (unsigned short *)w16, n, lr, b1, b2, eps, t))
}

}  // extern "C"
