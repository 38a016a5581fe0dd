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
__device__ inline unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (unsigned short)(u >> 16);
}

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
// ---------------------------------------------------------------------------
constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;   // LDS row stride in bf16 (80 B)

__global__ __launch_bounds__(256) void k_gemm_bf16(int transA, int transB, int M, int N, int K, float alpha,
                                                   const unsigned short *__restrict__ A, int lda, long long sA,
                                                   const unsigned short *__restrict__ B, int ldb, long long sB,
                                                   float beta, float *__restrict__ C, int ldc, long long sC) {
    __shared__ __align__(16) unsigned short As[TM * LDK];
    __shared__ __align__(16) unsigned short Bs[TN * LDK];
    const int bz = blockIdx.z;
    A += bz * sA;
    B += bz * sB;
    C += bz * sC;
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
    for (int k0 = 0; k0 < K; k0 += TK) {
        {
            // A: rows m of the tile x k; contiguous along k (or along m when transA)
            const int a0 = transA ? (tid >> 3) : (tid >> 2);        // k (transA) or m
            const int a1 = transA ? 8 * (tid & 7) : 8 * (tid & 3);  // first m (transA) or k of the vector
            const int gm = m0 + (transA ? a1 : a0), gk = k0 + (transA ? a0 : a1);
            unsigned short v[8];
            const bool full = transA ? (gk < K && gm + 7 < M) : (gm < M && gk + 7 < K);
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
                    v[j] = (mm < M && kk < K) ? (transA ? A[(size_t)kk * lda + mm] : A[(size_t)mm * lda + kk]) : 0;
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
            const bool full = transB ? (gn < N && gk + 7 < K) : (gk < K && gn + 7 < N);
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
                    v[j] = (nn_ < N && kk < K) ? (transB ? B[(size_t)nn_ * ldb + kk] : B[(size_t)kk * ldb + nn_]) : 0;
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

void gemm(hipStream_t st, int transA, int transB, int M, int N, int K, float alpha, const unsigned short *A, int lda,
          long long sA, const unsigned short *B, int ldb, long long sB, float beta, float *C, int ldc, long long sC,
          int batch) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    if (K <= 0) throw std::runtime_error("gemm: K must be positive");
    const dim3 grid(unsigned((N + TN - 1) / TN), unsigned((M + TM - 1) / TM), unsigned(batch));
    if (grid.y > 65535u || grid.z > 65535u) throw std::runtime_error("gemm: grid too large (M or batch)");
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, transA, transB, M, N, K, alpha, A, lda, sA, B, ldb, sB,
                       beta, C, ldc, sC);
    check_launch("bf16 gemm");
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

// tanh form of GELU: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2/pi)
__device__ inline float gelu(float z) {
    const float c = 0.7978845608028654f;
    return 0.5f * z * (1.f + tanhf(c * (z + 0.044715f * z * z * z)));
}
__device__ inline float gelu_grad(float z) {
    const float c = 0.7978845608028654f;
    const float u = c * (z + 0.044715f * z * z * z), t = tanhf(u);
    return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * z * z);
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
        gemm(st, 0, 0, Ri, Hd, W, 1.f, w.H16, W, 0, n.w16 + n.off_W1[l], Hd, 0, 0.f, w.Z, Hd, 0, 1);
        bias_act(st, w.Z, n.w + n.off_b1[l], Hd, R * Hd, 1, nullptr, w.A16);
        gemm(st, 0, 0, Ri, W, Hd, 1.f, w.A16, Hd, 0, n.w16 + n.off_W2[l], W, 0, 0.f, w.U, W, 0, 1);
        gemm(st, 1, 0, nb, W, N2, 1.f, n.basis16, nb, 0, w.H16, W, (long long)N2 * W, 0.f, w.Cm, W, (long long)nb * W,
             P);
        mode_scale(st, w.Cm, n.w + n.off_D[l], nullptr, w.Cs16, P, nb * W);
        gemm(st, 0, 0, N2, W, nb, 1.f, n.basis16, nb, 0, w.Cs16, W, (long long)nb * W, 1.f, w.H, W, (long long)N2 * W,
             P);
        bias_act(st, w.U, n.w + n.off_b2[l], W, R * W, 0, nullptr, nullptr);
        axpy(st, w.H, w.U, 1.f, R * W, w.H16);
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
                      int32_t ldc, int64_t sC, int32_t batch, void *stream) {
    NN_ABI(coral::nn::gemm((hipStream_t)stream, transA, transB, M, N, K, alpha, (const unsigned short *)A, lda, sA,
                           (const unsigned short *)B, ldb, sB, beta, C, ldc, sC, batch))
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
