// Kernels of the learned subgrid closure: BF16 GEMMs on MFMA with FP32
// accumulation, and the elementwise, reduction and optimizer kernels of its
// training and inference.  BF16 values are stored as their 16-bit patterns
// (unsigned short); FP32 -> BF16 rounds to nearest even.
#include "nn.hpp"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

namespace coral {
namespace nn {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

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

__global__ __launch_bounds__(256) void k_gemm_bf16(int transA, int transB, int M, int N, int K, float alpha,
                                                   const unsigned short *__restrict__ A, int lda, long long sA,
                                                   const unsigned short *__restrict__ B, int ldb, long long sB,
                                                   float beta, float *__restrict__ C, int ldc, long long sC,
                                                   int kc, int nsplit) {
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
                    const float v = alpha * acc[i][j][q];
                    *p = beta == 0.f ? v : v + beta * *p;
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
                           sB, beta, C, ldc, sC, K, 1);
        check_launch("bf16 gemm");
        return;
    }
    const long long need = (long long)ns * batch * M * N;
    if (!work || nwork < need)
        throw std::runtime_error("gemm: split-K work space of " + std::to_string(need) + " floats required");
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, transA, transB, M, N, K, 1.f, A, lda, sA, B, ldb, sB, 0.f,
                       work, N, (long long)M * N, kc, ns);
    const long long n = (long long)batch * M * N;
    hipLaunchKernelGGL(k_splitk_sum, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, work, ns, M, N, batch, alpha,
                       beta, C, ldc, sC);
    check_launch("bf16 gemm (split K)");
}

// dst (cols x rows) = src (rows x cols)^T, BF16, for the MLP kernel's weight
// operands (a column of W1 or W2 contiguous along k)
__global__ void k_transpose16(const unsigned short *__restrict__ src, int rows, int cols,
                              unsigned short *__restrict__ dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const int c = int(i / rows), r = int(i - (long long)c * rows);
    dst[i] = src[(long long)r * cols + c];
}
void transpose16(hipStream_t st, const unsigned short *src, int rows, int cols, unsigned short *dst) {
    if (rows <= 0 || cols <= 0) return;
    const long long n = (long long)rows * cols;
    hipLaunchKernelGGL(k_transpose16, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, src, rows, cols, dst);
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
// 1 / (1 + exp(-2 u)): one exponential and one hardware reciprocal (1 ulp),
// no branches (exp overflows to inf for very negative z, whose reciprocal 0
// gives z s = -0, and underflows to 0 for large z, giving z).  Derivative:
// s + 2 z s (1 - s) du/dz with s = 1 / (1 + exp(-2 u)).
__device__ inline float gelu_sig(float z) {
    const float c = 0.7978845608028654f;
    const float u = c * (z + 0.044715f * z * z * z);
    return __builtin_amdgcn_rcpf(1.f + __expf(-2.f * u));
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
constexpr int MLP_ROWS = 32;
template <int W, int HD>
__global__ __launch_bounds__(256) void k_mlp(long long R, const unsigned short *Hin16,
                                             const unsigned short *__restrict__ W1t, const float *__restrict__ b1,
                                             const unsigned short *__restrict__ W2t, const float *__restrict__ b2,
                                             float *__restrict__ H, unsigned short *Hout16, float *__restrict__ Zc,
                                             unsigned short *__restrict__ A16c) {
    static_assert(W % 32 == 0 && HD % 64 == 0, "MLP shape");
    constexpr int RB = MLP_ROWS / 16;    // 16-row blocks
    constexpr int LDA = HD + 8;          // LDS row stride of the activation tile (bf16)
    constexpr int KS1 = W / 32, CB1 = HD / 64, KS2 = HD / 32, CB2 = W / 32;
    __shared__ __align__(16) unsigned short As[MLP_ROWS * LDA];
    const long long r0 = (long long)blockIdx.x * MLP_ROWS;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int fr = lane & 15, fk = 8 * (lane >> 4), rq = 4 * (lane >> 4);
    const bf8 zero8 = {};
    // stage 1
    bf8 a[RB][KS1];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const long long row = r0 + 16 * rb + fr;
#pragma unroll
        for (int k = 0; k < KS1; ++k)
            a[rb][k] = row < R ? *reinterpret_cast<const bf8 *>(Hin16 + row * W + 32 * k + fk) : zero8;
    }
    f4 acc[RB][CB1];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb) acc[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
    const int c0 = wv * (HD / 4);
#pragma unroll
    for (int k = 0; k < KS1; ++k)
#pragma unroll
        for (int cb = 0; cb < CB1; ++cb) {
            const bf8 b = *reinterpret_cast<const bf8 *>(W1t + (size_t)(c0 + 16 * cb + fr) * W + 32 * k + fk);
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rb][k], b, acc[rb][cb], 0, 0, 0);
        }
#pragma unroll
    for (int cb = 0; cb < CB1; ++cb) {
        const int col = c0 + 16 * cb + fr;
        const float bias = b1[col];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rl = 16 * rb + rq + q;
                const float v = acc[rb][cb][q] + bias;
                const unsigned short h = f2bf(gelu(v));
                As[rl * LDA + col] = h;
                const long long row = r0 + rl;
                if (Zc && row < R) Zc[row * HD + col] = v;
                if (A16c && row < R) A16c[row * HD + col] = h;
            }
    }
    __syncthreads();
    // stage 2
    const int sr = 16 * (wv >> 1), sc = (W / 2) * (wv & 1);
    f4 u[CB2];
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) u[cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KS2; ++k) {
        const bf8 av = *reinterpret_cast<const bf8 *>(&As[(sr + fr) * LDA + 32 * k + fk]);
#pragma unroll
        for (int cb = 0; cb < CB2; ++cb) {
            const bf8 b = *reinterpret_cast<const bf8 *>(W2t + (size_t)(sc + 16 * cb + fr) * HD + 32 * k + fk);
            u[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b, u[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < CB2; ++cb) {
        const int col = sc + 16 * cb + fr;
        const float bias = b2[col];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long long row = r0 + sr + rq + q;
            if (row < R) {
                const float h = H[row * W + col] + (u[cb][q] + bias);
                H[row * W + col] = h;
                Hout16[row * W + col] = f2bf(h);
            }
        }
    }
}

bool mlp_supported(int W, int HD) { return (W == 64 || W == 32) && (HD == 256 || HD == 128); }
void mlp(hipStream_t st, long long R, int W, int HD, const unsigned short *Hin16, const unsigned short *W1t,
         const float *b1, const unsigned short *W2t, const float *b2, float *H, unsigned short *Hout16, float *Zc,
         unsigned short *A16c) {
    if (R <= 0) return;
    const long long nblk = (R + MLP_ROWS - 1) / MLP_ROWS;
    if (nblk > 2147483647LL) throw std::runtime_error("mlp: too many rows");
    const dim3 g{unsigned(nblk)}, b{256};
#define MLP_CASE(w_, hd_)                                                                                     \
    if (W == w_ && HD == hd_) {                                                                               \
        hipLaunchKernelGGL((k_mlp<w_, hd_>), g, b, 0, st, R, Hin16, W1t, b1, W2t, b2, H, Hout16, Zc, A16c); \
        check_launch("fused mlp");                                                                            \
        return;                                                                                               \
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

void net_forward(hipStream_t st, const NetParams &n, const unsigned short *X16, int P, float *Y, const NetWork &w) {
    const int W = n.width, Hd = n.hidden, nb = n.nb, N2 = n.N2;
    const long long R = (long long)P * N2;
    if (R > 2147483647LL) throw std::runtime_error("closure network: too many rows");
    const int Ri = int(R);
    gemm(st, 0, 0, Ri, W, n.cin, 1.f, X16, n.cin, 0, n.w16 + n.off_W0, W, 0, 0.f, w.H, W, 0, 1);
    bias_act(st, w.H, n.w + n.off_b0, W, R * W, 0, nullptr, w.H16);
    for (int l = 0; l < n.layers; ++l) {
        // spectral branch: Cm = B^T H16 per plane, Cs = Cm D, H += B Cs
        gemm(st, 0, 0, nb, W, N2, 1.f, n.basisT16, N2, 0, w.H16, W, (long long)N2 * W, 0.f, w.Cm, W,
             (long long)nb * W, P, w.gwork, w.ngwork);
        mode_scale(st, w.Cm, n.w + n.off_D[l], nullptr, w.Cs16, P, nb * W);
        gemm(st, 1, 0, N2, W, nb, 1.f, n.basisT16, N2, 0, w.Cs16, W, (long long)nb * W, 1.f, w.H, W,
             (long long)N2 * W, P);
        // MLP branch and the residual sum, H16 updated in place
        const unsigned short *wt = n.wt16 + (long long)l * 2 * W * Hd;
        mlp(st, R, W, Hd, w.H16, wt, n.w + n.off_b1[l], wt + (long long)W * Hd, n.w + n.off_b2[l], w.H, w.H16,
            nullptr, nullptr);
    }
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
                     const void *W2t, const float *b2, float *H, void *Hout16, float *Zc, void *A16c, void *stream) {
    NN_ABI(coral::nn::mlp((hipStream_t)stream, R, width, hidden, (const unsigned short *)Hin16,
                          (const unsigned short *)W1t, b1, (const unsigned short *)W2t, b2, H,
                          (unsigned short *)Hout16, Zc, (unsigned short *)A16c))
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
