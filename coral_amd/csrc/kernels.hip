// Device kernels of the plane-layer IMEX step (CDNA4 / gfx950).
//
// Data layout in HBM (one GPU owns the whole domain):
//   state   [sys][total]              complex, sys = ix*nky + jy over kept modes
//   spec    [full][NZ][sys]           Chebyshev coefficients of full variables
//   Fz      [full][NZAA][sys]         after the z transform (Gauss points)
//   G       [full][NZAA][NYAA][nkx]   after the y transform
//   H       [quad][NZAA][NYAA][nkx]   products after the x transform (kept kx)
//   Hy      [quad][NZAA][sys]         after the y transform (kept ky)
//   Qhat    [quad][NZ][sys]           Chebyshev coefficients of the products
#include "kernels.hpp"

#include <cstdio>
#include <stdexcept>

namespace coral {

#define CHECK_LAUNCH() check_launch(__func__)

static inline unsigned blocks_for(long long n, int threads) { return unsigned((n + threads - 1) / threads); }

// ---------------------------------------------------------------------------
// extraction (padStencilExtractShuffle . aux, PL_IMEX_timestepping.f90:1199-1209,
// PL_IMEX_timestepping_marching.f90:430-470)
// ---------------------------------------------------------------------------
// spec[n][s] (+)= mono(s) * sum_k val[k] * state[s][col[k]].  A workgroup owns
// xt systems: it reads each system's state along n (coalesced within the
// system) and writes rows of xt consecutive systems through LDS.
__global__ __launch_bounds__(256) void k_extract(const double2 *__restrict__ state, int total, int nsys, int NZ,
                                                 const int *__restrict__ rowptr, const int *__restrict__ col,
                                                 const double *__restrict__ val, int kx, int ky, KeptModes km,
                                                 double2 *__restrict__ spec, int accumulate, int xt) {
    extern __shared__ double2 tile[];  // [NZ][xt]
    const int s0 = blockIdx.x * xt;
    const int ns = min(xt, nsys - s0);
    for (int e = threadIdx.x; e < ns * NZ; e += blockDim.x) {
        const int sl = e / NZ, n = e - sl * NZ, s = s0 + sl;
        const double2 *st = state + (size_t)s * total;
        double2 acc = make_double2(0.0, 0.0);
        for (int k = rowptr[n]; k < rowptr[n + 1]; ++k) {
            const double v = val[k];
            const double2 a = st[col[k]];
            acc.x += v * a.x;
            acc.y += v * a.y;
        }
        if (kx || ky) acc = cmul(monomial(kx, ky, km.kx(s), km.ky(s)), acc);
        tile[n * xt + sl] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NZ; e += blockDim.x) {
        const int n = e / ns, sl = e - n * ns;
        double2 *o = spec + (size_t)n * nsys + s0 + sl;
        const double2 v = tile[n * xt + sl];
        *o = accumulate ? cadd(*o, v) : v;
    }
}

void launch_extract(hipStream_t st, const double2 *state, int total, int nsys, int NZ, const int *rowptr,
                    const int *col, const double *val, int kx, int ky, KeptModes km, double2 *spec, int accumulate) {
    int xt = 16;
    while (xt > 1 && size_t(NZ) * xt * sizeof(double2) > 49152) xt /= 2;
    const size_t shm = size_t(NZ) * xt * sizeof(double2);
    hipLaunchKernelGGL(k_extract, dim3(blocks_for(nsys, xt)), dim3(256), shm, st, state, total, nsys, NZ, rowptr, col,
                       val, kx, ky, km, spec, accumulate, xt);
    CHECK_LAUNCH();
}

// One pass over each system's state computes every full variable: for a tile
// of xt systems the workgroup walks the systems one after another (all threads
// on one system's CSR rows, so its state is read from HBM once and then hits
// in cache), keeps the results in LDS as [f][n][sys] and writes rows of xt
// consecutive systems.
__global__ __launch_bounds__(512) void k_extract_all(ExtractAllArgs a, int xt) {
    extern __shared__ double2 tile[];  // [nfull][NZ][xt]
    const int NZ = a.NZ, nsys = a.nsys, nf = a.nfull;
    const int s0 = blockIdx.x * xt;
    const int ns = min(xt, nsys - s0);
    const int per = nf * NZ;  // outputs per system
#pragma unroll 2
    for (int e = threadIdx.x; e < ns * per; e += blockDim.x) {
        const int sl = e / per, r = e - sl * per, f = r / NZ, n = r - f * NZ, s = s0 + sl;
        double2 acc = make_double2(0.0, 0.0);
        for (int t = a.full_first[f]; t < a.full_first[f + 1]; ++t) {
            const ExtractTerm &T = a.term[t];
            const double2 *st = T.state + (size_t)s * T.total;
            double2 part = make_double2(0.0, 0.0);
            for (int k = T.rowptr[n]; k < T.rowptr[n + 1]; ++k) {
                const double v = T.val[k];
                const double2 x = st[T.col[k]];
                part.x += v * x.x;
                part.y += v * x.y;
            }
            if (T.kx || T.ky) part = cmul(monomial(T.kx, T.ky, a.km.kx(s), a.km.ky(s)), part);
            acc = cadd(acc, part);
        }
        tile[(size_t)(f * NZ + n) * xt + sl] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nf * NZ * ns; e += blockDim.x) {
        const int row = e / ns, sl = e - row * ns;  // row = f * NZ + n
        const double2 v = tile[(size_t)row * xt + sl];
        if (a.spec32)
            a.spec32[(size_t)row * nsys + s0 + sl] = make_float2(float(v.x), float(v.y));
        else
            a.spec[(size_t)row * nsys + s0 + sl] = v;
    }
}

// Same outputs as k_extract_all, with the XT systems' state vectors staged in
// LDS first: they are contiguous in HBM (one coalesced copy per state), while
// the sparse rows gather entries about one interleave stride (6 positions)
// apart, which as direct loads touch a cache line per lane.  Output rows of XT
// systems are written straight from registers (XT * 16 contiguous bytes).
template <int XT>
__global__ __launch_bounds__(256) void k_extract_lds(ExtractAllArgs a) {
    extern __shared__ double2 sst[];   // state q: [XT][stotal[q]] from XT * (sum of earlier stotal)
    const int NZ = a.NZ, nsys = a.nsys, nf = a.nfull;
    const int s0 = blockIdx.x * XT;
    const int ns = min(XT, nsys - s0);
    int off = 0;
    for (int q = 0; q < a.nstate; ++q) {
        const int cnt = ns * a.stotal[q];
        const double2 *src = a.state[q] + (size_t)s0 * a.stotal[q];
        double2 *dst = sst + off;
        off += XT * a.stotal[q];
        // up to 16 loads per thread in flight before the first LDS store
        // (12 for four systems of 766 unknowns)
        constexpr int BATCH = 16;
        for (int e0 = 0; e0 < cnt; e0 += BATCH * 256) {
            double2 v[BATCH];
#pragma unroll
            for (int i = 0; i < BATCH; ++i) {
                const int e = e0 + threadIdx.x + 256 * i;
                v[i] = e < cnt ? src[e] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int i = 0; i < BATCH; ++i) {
                const int e = e0 + threadIdx.x + 256 * i;
                if (e < cnt) dst[e] = v[i];
            }
        }
    }
    __syncthreads();
    // one output row (f, n) per thread for all XT systems: the row's sparse
    // entries are read once for the XT systems
    const int per = nf * NZ;
    for (int row = threadIdx.x; row < per; row += 256) {
        const int f = row / NZ, n = row - f * NZ;
        double2 acc[XT];
#pragma unroll
        for (int sl = 0; sl < XT; ++sl) acc[sl] = make_double2(0.0, 0.0);
        for (int t = a.full_first[f]; t < a.full_first[f + 1]; ++t) {
            const ExtractTerm &T = a.term[t];
            const double2 *st = sst + XT * T.soff;
            double2 part[XT];
#pragma unroll
            for (int sl = 0; sl < XT; ++sl) part[sl] = make_double2(0.0, 0.0);
            for (int k = T.rowptr[n]; k < T.rowptr[n + 1]; ++k) {
                const double v = T.val[k];
                const int c = T.col[k];
#pragma unroll
                for (int sl = 0; sl < XT; ++sl) {
                    const double2 x = st[sl * T.total + c];
                    part[sl].x += v * x.x;
                    part[sl].y += v * x.y;
                }
            }
#pragma unroll
            for (int sl = 0; sl < XT; ++sl) {
                if (T.kx || T.ky) {
                    const int s = min(s0 + sl, nsys - 1);
                    part[sl] = cmul(monomial(T.kx, T.ky, a.km.kx(s), a.km.ky(s)), part[sl]);
                }
                acc[sl] = cadd(acc[sl], part[sl]);
            }
        }
        if (a.spec32) {
            float2 *out = a.spec32 + (size_t)row * nsys + s0;
#pragma unroll
            for (int sl = 0; sl < XT; ++sl)
                if (sl < ns) out[sl] = make_float2(float(acc[sl].x), float(acc[sl].y));
        } else {
            double2 *out = a.spec + (size_t)row * nsys + s0;
#pragma unroll
            for (int sl = 0; sl < XT; ++sl)
                if (sl < ns) out[sl] = acc[sl];
        }
    }
}

void launch_extract_all(hipStream_t st, const ExtractAllArgs &a) {
    // staged variant: every term's state known, XT systems of all states in LDS
    {
        bool ok = a.nstate > 0;
        size_t words = 0;
        for (int t = 0; t < a.nterms; ++t) ok = ok && a.term[t].sidx >= 0 && a.term[t].sidx < a.nstate;
        for (int q = 0; q < a.nstate; ++q) words += size_t(a.stotal[q]);
        constexpr int XT = 4;
        const size_t shm = words * XT * sizeof(double2);
        if (ok && shm <= 65536) {
            hipLaunchKernelGGL(k_extract_lds<XT>, dim3(blocks_for(a.nsys, XT)), dim3(256), shm, st, a);
            CHECK_LAUNCH();
            return;
        }
    }
    int lim = 65536;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    // 8 systems (128-byte output rows) leave room for two workgroups per CU
    int xt = 8;
    while (xt > 1 && size_t(a.nfull) * a.NZ * xt * sizeof(double2) > size_t(lim) / 2) xt /= 2;
    const size_t shm = size_t(a.nfull) * a.NZ * xt * sizeof(double2);
    if (shm > 65536) set_dynamic_lds((const void *)k_extract_all, shm, __func__);
    hipLaunchKernelGGL(k_extract_all, dim3(blocks_for(a.nsys, xt)), dim3(512), shm, st, a, xt);
    CHECK_LAUNCH();
}

// mean contributions land in the (0,0) column: spec[n][0] (C = double2, or
// float2 in the FP32 chain, where the FP64 sum is rounded once)
template <class C>
__global__ void k_extract_mean(const double2 *__restrict__ zstate, int NZ, int nsys, const int *__restrict__ rowptr,
                               const int *__restrict__ col, const double *__restrict__ val, C *spec) {
    for (int n = threadIdx.x; n < NZ; n += blockDim.x) {
        double acc = 0.0;
        for (int k = rowptr[n]; k < rowptr[n + 1]; ++k) acc += val[k] * zstate[col[k]].x;
        C &o = spec[(size_t)n * nsys];
        o.x = typename CxT<C>::R(double(o.x) + acc);
    }
}

void launch_extract_mean(hipStream_t st, const double2 *zstate, int NZ, int nsys, const int *rowptr, const int *col,
                         const double *val, double2 *spec) {
    hipLaunchKernelGGL(k_extract_mean<double2>, dim3(1), dim3(256), 0, st, zstate, NZ, nsys, rowptr, col, val, spec);
    CHECK_LAUNCH();
}

void launch_extract_mean32(hipStream_t st, const double2 *zstate, int NZ, int nsys, const int *rowptr, const int *col,
                           const double *val, float2 *spec) {
    hipLaunchKernelGGL(k_extract_mean<float2>, dim3(1), dim3(256), 0, st, zstate, NZ, nsys, rowptr, col, val, spec);
    CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// z transforms, plain FP64 GEMM (used when no MFMA variant covers the shape):
// C[b](m, c) = sum_k A(m, k) B[b](k, c), B and C row-major with Ncol columns
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm(const double *__restrict__ A, int lda, const double *__restrict__ Bp,
                                              double *__restrict__ Cp, int M, int Ncol, int K, size_t strideB,
                                              size_t strideC) {
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    Bp += blockIdx.z * strideB;
    Cp += blockIdx.z * strideC;
    const int m0 = blockIdx.y * BM, c0 = blockIdx.x * BN;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < K; k0 += BK) {
        for (int e = tid; e < BM * BK; e += 256) {
            int mm = e / BK, kk = e % BK;
            int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? A[(size_t)gm * lda + gk] : 0.0;
        }
        for (int e = tid; e < BN * BK; e += 256) {
            int cc = e % BN, kk = e / BN;
            int c = c0 + cc, gk = k0 + kk;
            Bs[kk][cc] = (c < Ncol && gk < K) ? Bp[(size_t)gk * Ncol + c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int m = m0 + ty * 4 + i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int c = c0 + tx + 16 * j;
            if (c < Ncol) Cp[(size_t)m * Ncol + c] = acc[i][j];
        }
    }
}

void launch_zbwd(hipStream_t st, const double *Tz, const double2 *spec, double2 *Fz, int NZAA, int NZ, int nsys,
                 int nbatch) {
    int Ncol = 2 * nsys;
    dim3 grid(blocks_for(Ncol, 64), blocks_for(NZAA, 64), nbatch);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, Tz, NZ, (const double *)spec, (double *)Fz, NZAA, Ncol, NZ,
                       (size_t)NZ * Ncol, (size_t)NZAA * Ncol);
    CHECK_LAUNCH();
}

void launch_zfwd(hipStream_t st, const double *Cz, const double2 *Hy, double2 *Qhat, int NZ, int NZAA, int nsys,
                 int nbatch) {
    int Ncol = 2 * nsys;
    dim3 grid(blocks_for(Ncol, 64), blocks_for(NZ, 64), nbatch);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, Cz, NZAA, (const double *)Hy, (double *)Qhat, NZ, Ncol, NZAA,
                       (size_t)NZAA * Ncol, (size_t)NZ * Ncol);
    CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Stockham FFT of `lines` lines held in LDS (line stride LS), radix 2/3/4.
// Stage: in index j + r*N/R, out index (j/Ns)*Ns*R + j%Ns + r*Ns, twiddle
// exp(sign*2*pi*i*(j%Ns)*r/(Ns*R)).  sign = -1 forward, +1 inverse.
// ---------------------------------------------------------------------------
struct FFTPlanDev {
    int N, nrad;
    int radix[16];
    const double2 *tw;
};

// R-point DFT in registers, exp(sign 2 pi i r k / R)
template <int R>
__device__ inline void dft_small(double2 *y, double sign) {
    if (R == 2) {
        double2 a = y[0], c = y[1];
        y[0] = cadd(a, c);
        y[1] = csub(a, c);
    } else if (R == 3) {
        const double c1 = -0.5, s1 = sign * 0.86602540378443864676;
        double2 a = y[0], b1 = y[1], b2 = y[2];
        double2 t = cadd(b1, b2), d = csub(b1, b2);
        y[0] = cadd(a, t);
        double2 m = make_double2(a.x + c1 * t.x, a.y + c1 * t.y);
        y[1] = make_double2(m.x - s1 * d.y, m.y + s1 * d.x);
        y[2] = make_double2(m.x + s1 * d.y, m.y - s1 * d.x);
    } else {  // R == 4
        double2 a = y[0], b = y[1], c = y[2], d = y[3];
        double2 t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = csub(b, d);
        double2 wt3 = make_double2(-sign * t3.y, sign * t3.x);
        y[0] = cadd(t0, t2);
        y[2] = csub(t0, t2);
        y[1] = cadd(t1, wt3);
        y[3] = csub(t1, wt3);
    }
}

// one Stockham stage over `lines` lines; every thread owns at most MAXB butterflies
// (slots are unrolled so the butterfly registers never spill)
template <int R, int MAXB>
__device__ inline void fft_stage(double2 *buf, int LS, int lines, int N, int Ns, const double2 *tw, double sign) {
    const int nb = N / R;
    const int total = nb * lines;
    const int step = N / (Ns * R);
    double2 v[MAXB][R];
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
        const int b = threadIdx.x + c * blockDim.x;
        if (b < total) {
            const int line = b / nb, j = b - line * nb;
            const double2 *L = buf + line * LS;
            const int jm = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double2 x = L[j + r * nb];
                if (r > 0) {
                    double2 w = tw[jm * r * step];  // exp(-2 pi i k / N)
                    if (sign > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[c][r] = x;
            }
            dft_small<R>(v[c], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
        const int b = threadIdx.x + c * blockDim.x;
        if (b < total) {
            const int line = b / nb, j = b - line * nb;
            double2 *L = buf + line * LS;
            const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; ++r) L[base + r * Ns] = v[c][r];
        }
    }
    __syncthreads();
}

// Stockham FFT of `lines` lines held in LDS (line stride LS), radix 2/3/4.
// Stage: in index j + r*N/R, out index (j/Ns)*Ns*R + j%Ns + r*Ns, twiddle
// exp(sign*2*pi*i*(j%Ns)*r/(Ns*R)).  sign = -1 forward, +1 inverse.  With
// blockDim = 256 and lines*N <= 1536, MAXB = 3 slots hold every butterfly.
template <int MAXB>
__device__ void fft_lds(double2 *buf, int LS, int lines, const FFTPlanDev &p, double sign) {
    int Ns = 1;
    for (int st = 0; st < p.nrad; ++st) {
        const int R = p.radix[st];
        if (R == 4)
            fft_stage<4, MAXB>(buf, LS, lines, p.N, Ns, p.tw, sign);
        else if (R == 2)
            fft_stage<2, MAXB>(buf, LS, lines, p.N, Ns, p.tw, sign);
        else
            fft_stage<3, MAXB>(buf, LS, lines, p.N, Ns, p.tw, sign);
        Ns *= R;
    }
}

static FFTPlanDev to_dev(const FFTPlan &p) {
    FFTPlanDev d;
    d.N = p.N;
    d.nrad = p.nrad;
    for (int i = 0; i < 16; ++i) d.radix[i] = p.radix[i];
    d.tw = p.tw;
    return d;
}

// lines per workgroup of the y passes: lines*N <= 1536 keeps the butterflies of
// every stage (lines*N/2 at most) within fft_lds<3> capacity (3 per thread x 256)
static int lines_for(int N) {
    int lines = 1;
    while (lines * 2 * N <= 1536) lines *= 2;
    return lines;
}

// y inverse FFT: Fz[b][z][ix*nky + jy] -> G[b][z][y][ix]
__global__ __launch_bounds__(256) void k_ybwd(FFTPlanDev p, const double2 *__restrict__ Fz, double2 *__restrict__ G,
                                              int nkx, int nky, int NY, int lines) {
    extern __shared__ double2 sbuf[];
    const int N = p.N, LS = N + 1;
    const int ntile = (nkx + lines - 1) / lines;
    const int tile = blockIdx.x % ntile;
    const size_t bz = blockIdx.x / ntile;
    const int ix0 = tile * lines, nl = min(lines, nkx - ix0);
    for (int e = threadIdx.x; e < lines * LS; e += blockDim.x) sbuf[e] = make_double2(0.0, 0.0);
    __syncthreads();
    const double2 *src = Fz + bz * nkx * nky + (size_t)ix0 * nky;
    for (int e = threadIdx.x; e < nl * nky; e += blockDim.x) {
        int l = e / nky, jy = e - l * nky;
        int slot = jy < NY / 2 ? jy : N - (NY - 1) + jy;
        sbuf[l * LS + slot] = src[e];
    }
    __syncthreads();
    fft_lds<3>(sbuf, LS, lines, p, +1.0);
    double2 *dst = G + bz * N * nkx;
    for (int e = threadIdx.x; e < N * nl; e += blockDim.x) {
        int y = e / nl, l = e - y * nl;
        dst[(size_t)y * nkx + ix0 + l] = sbuf[l * LS + y];
    }
}

static bool fft_fits(int lines, int N) {
    if (lines * N <= 1536) return true;
    throw std::runtime_error("FFT length " + std::to_string(N) + " exceeds the LDS FFT capacity");
}

void launch_ybwd(hipStream_t st, const FFTPlan &py, const double2 *Fz, double2 *G, int nbatch, int NZAA, int nkx,
                 int nky, int NY) {
    if (launch_ypass_wave(st, false, py, Fz, G, nbatch, NZAA, nkx, nky, NY)) return;
    int lines = lines_for(py.N);
    if (!fft_fits(lines, py.N)) return;
    int ntile = (nkx + lines - 1) / lines;
    size_t shmem = size_t(lines) * (py.N + 1) * sizeof(double2);
    hipLaunchKernelGGL(k_ybwd, dim3(ntile * nbatch * NZAA), dim3(256), shmem, st, to_dev(py), Fz, G, nkx, nky, NY,
                       lines);
    CHECK_LAUNCH();
}

// y forward FFT: H[b][z][y][ix] -> Hy[b][z][ix*nky + jy]
__global__ __launch_bounds__(256) void k_yfwd(FFTPlanDev p, const double2 *__restrict__ H, double2 *__restrict__ Hy,
                                              int nkx, int nky, int NY, int lines) {
    extern __shared__ double2 sbuf[];
    const int N = p.N, LS = N + 1;
    const int ntile = (nkx + lines - 1) / lines;
    const int tile = blockIdx.x % ntile;
    const size_t bz = blockIdx.x / ntile;
    const int ix0 = tile * lines, nl = min(lines, nkx - ix0);
    const double2 *src = H + bz * N * nkx;
    for (int e = threadIdx.x; e < N * lines; e += blockDim.x) {
        int y = e / lines, l = e - y * lines;
        sbuf[l * LS + y] = (l < nl) ? src[(size_t)y * nkx + ix0 + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<3>(sbuf, LS, lines, p, -1.0);
    double2 *dst = Hy + bz * nkx * nky + (size_t)ix0 * nky;
    for (int e = threadIdx.x; e < nl * nky; e += blockDim.x) {
        int l = e / nky, jy = e - l * nky;
        int slot = jy < NY / 2 ? jy : N - (NY - 1) + jy;
        dst[e] = sbuf[l * LS + slot];
    }
}

void launch_yfwd(hipStream_t st, const FFTPlan &py, const double2 *H, double2 *Hy, int nbatch, int NZAA, int nkx,
                 int nky, int NY) {
    if (launch_ypass_wave(st, true, py, H, Hy, nbatch, NZAA, nkx, nky, NY)) return;
    int lines = lines_for(py.N);
    if (!fft_fits(lines, py.N)) return;
    int ntile = (nkx + lines - 1) / lines;
    size_t shmem = size_t(lines) * (py.N + 1) * sizeof(double2);
    hipLaunchKernelGGL(k_yfwd, dim3(ntile * nbatch * NZAA), dim3(256), shmem, st, to_dev(py), H, Hy, nkx, nky, NY,
                       lines);
    CHECK_LAUNCH();
}

// x pass: per (z, y) line: c2r of all full variables, products, r2c (kept kx)
__device__ inline void atomic_max_nonneg(unsigned long long *addr, double v) {
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

__global__ __launch_bounds__(256) void k_xpass(FFTPlanDev p, XPassArgs a, int LX) {
    extern __shared__ double2 sbuf[];
    const int N = p.N, LS = N + 1;
    double2 *fb = sbuf;
    double *phys = (double *)(sbuf + LX * LS);
    const long long nlines = (long long)a.NZAA * a.NYAA;
    const long long l0 = (long long)blockIdx.x * LX;
    const int nl = (int)min((long long)LX, nlines - l0);
    const int nkx = a.nkx;
    for (int f = 0; f < a.nfull; ++f) {
        for (int e = threadIdx.x; e < LX * LS; e += blockDim.x) fb[e] = make_double2(0.0, 0.0);
        __syncthreads();
        const double2 *g = a.G + ((size_t)f * nlines + l0) * nkx;
        for (int e = threadIdx.x; e < nl * nkx; e += blockDim.x) {
            int l = e / nkx, k = e - l * nkx;
            double2 v = g[e];
            if (k == 0) {
                fb[l * LS] = make_double2(v.x, 0.0);
            } else {
                fb[l * LS + k] = v;
                fb[l * LS + N - k] = cconj(v);
            }
        }
        __syncthreads();
        fft_lds<3>(fb, LS, LX, p, +1.0);
        for (int e = threadIdx.x; e < LX * N; e += blockDim.x) {
            int l = e / N, x = e - l * N;
            phys[(f * LX + l) * N + x] = fb[l * LS + x].x;
        }
        __syncthreads();
    }
    // CFL maxima (compute_cfl_based_timestep, marching.f90:200-254)
    if (a.cflmax && a.nfull >= 3) {
        double mu = 0.0, mv = 0.0, mw = 0.0;
        for (int e = threadIdx.x; e < nl * N; e += blockDim.x) {
            int l = e / N, x = e - l * N;
            int z = int((l0 + l) / a.NYAA);
            mu = fmax(mu, fabs(phys[(0 * LX + l) * N + x]));
            mv = fmax(mv, fabs(phys[(1 * LX + l) * N + x]));
            mw = fmax(mw, fabs(phys[(2 * LX + l) * N + x] / a.zweight[z]));
        }
        for (int off = 32; off > 0; off >>= 1) {
            mu = fmax(mu, __shfl_xor(mu, off));
            mv = fmax(mv, __shfl_xor(mv, off));
            mw = fmax(mw, __shfl_xor(mw, off));
        }
        if ((threadIdx.x & 63) == 0) {
            atomic_max_nonneg(&a.cflmax[0], mu);
            atomic_max_nonneg(&a.cflmax[1], mv);
            atomic_max_nonneg(&a.cflmax[2], mw);
        }
    }
    for (int q = 0; q < a.nquad; ++q) {
        const int fa = a.qa[q], fbi = a.qb[q];
        for (int e = threadIdx.x; e < LX * N; e += blockDim.x) {
            int l = e / N, x = e - l * N;
            double va = fa < 0 ? 1.0 : phys[(fa * LX + l) * N + x];
            double vb = fbi < 0 ? 1.0 : phys[(fbi * LX + l) * N + x];
            fb[l * LS + x] = make_double2(va * vb, 0.0);
        }
        __syncthreads();
        fft_lds<3>(fb, LS, LX, p, -1.0);
        double2 *h = a.H + ((size_t)q * nlines + l0) * nkx;
        for (int e = threadIdx.x; e < nl * nkx; e += blockDim.x) {
            int l = e / nkx, k = e - l * nkx;
            h[e] = fb[l * LS + k];
        }
        __syncthreads();
    }
}

static size_t xpass_lds(int LX, int N, int nfull) {
    return size_t(LX) * (N + 1) * sizeof(double2) + size_t(nfull) * LX * N * sizeof(double);
}

void launch_xpass(hipStream_t st, const FFTPlan &px, const XPassArgs &a) {
    if (launch_xpass_wave(st, px, a)) return;
    if (a.fp32) throw std::runtime_error("the FP32 x pass needs a wavefront plan (N = 96, 144, 192, 384, 768)");
    // lines per workgroup: FFT butterflies must fit fft_lds<3> (LX*N/2 <= 3*256)
    // and the workgroup's LDS must stay within 64 KiB
    int LX = 4;
    while (LX > 1 && (LX * px.N > 1536 || xpass_lds(LX, px.N, a.nfull) > 65536)) LX /= 2;
    if (LX * px.N > 1536 || xpass_lds(LX, px.N, a.nfull) > 65536) {
        throw std::runtime_error("x pass does not fit (N=" + std::to_string(px.N) + ", nfull=" +
                                 std::to_string(a.nfull) + ")");
    }
    long long nlines = (long long)a.NZAA * a.NYAA;
    size_t shmem = xpass_lds(LX, px.N, a.nfull);
    hipLaunchKernelGGL(k_xpass, dim3(blocks_for(nlines, LX)), dim3(256), shmem, st, to_dev(px), a, LX);
    CHECK_LAUNCH();
}

// c2r only (diagnostics and the transform seam) for lengths without a
// wavefront plan: G[b][line][kx] -> phys[b][line][x]
__global__ __launch_bounds__(256) void k_xbwd(FFTPlanDev p, const double2 *__restrict__ G, double *__restrict__ phys,
                                              long long nlines, int nkx, int LX) {
    extern __shared__ double2 sbuf[];
    const int N = p.N, LS = N + 1;
    const int b = blockIdx.y;
    const long long l0 = (long long)blockIdx.x * LX;
    const int nl = (int)min((long long)LX, nlines - l0);
    for (int e = threadIdx.x; e < LX * LS; e += blockDim.x) sbuf[e] = make_double2(0.0, 0.0);
    __syncthreads();
    const double2 *g = G + ((size_t)b * nlines + l0) * nkx;
    for (int e = threadIdx.x; e < nl * nkx; e += blockDim.x) {
        int l = e / nkx, k = e - l * nkx;
        double2 v = g[e];
        if (k == 0) {
            sbuf[l * LS] = make_double2(v.x, 0.0);
        } else {
            sbuf[l * LS + k] = v;
            sbuf[l * LS + N - k] = cconj(v);
        }
    }
    __syncthreads();
    fft_lds<3>(sbuf, LS, LX, p, +1.0);
    double *o = phys + ((size_t)b * nlines + l0) * N;
    for (int e = threadIdx.x; e < nl * N; e += blockDim.x) {
        int l = e / N, x = e - l * N;
        o[e] = sbuf[l * LS + x].x;
    }
}

// r2c only, for lengths without a wavefront plan: phys[b][line][x] ->
// H[b][line][kx] (kept kx)
__global__ __launch_bounds__(256) void k_xfwd(FFTPlanDev p, const double *__restrict__ phys, double2 *__restrict__ H,
                                              long long nlines, int nkx, int LX) {
    extern __shared__ double2 sbuf[];
    const int N = p.N, LS = N + 1;
    const int b = blockIdx.y;
    const long long l0 = (long long)blockIdx.x * LX;
    const int nl = (int)min((long long)LX, nlines - l0);
    const double *src = phys + ((size_t)b * nlines + l0) * N;
    for (int e = threadIdx.x; e < LX * N; e += blockDim.x) {
        int l = e / N, x = e - l * N;
        sbuf[l * LS + x] = make_double2(l < nl ? src[e] : 0.0, 0.0);
    }
    __syncthreads();
    fft_lds<3>(sbuf, LS, LX, p, -1.0);
    double2 *h = H + ((size_t)b * nlines + l0) * nkx;
    for (int e = threadIdx.x; e < nl * nkx; e += blockDim.x) {
        int l = e / nkx, k = e - l * nkx;
        h[e] = sbuf[l * LS + k];
    }
}

static int xlines_for(int N) {
    int LX = 1;
    while (LX * 2 * N <= 1536) LX *= 2;
    return LX;
}

void launch_xbwd(hipStream_t st, const FFTPlan &px, const double2 *G, double *phys, int nbatch, long long nlines,
                 int nkx) {
    if (launch_xbwd_wave(st, px, G, phys, nbatch, nlines, nkx)) return;
    int LX = xlines_for(px.N);
    if (!fft_fits(LX, px.N)) return;
    size_t shmem = size_t(LX) * (px.N + 1) * sizeof(double2);
    hipLaunchKernelGGL(k_xbwd, dim3(blocks_for(nlines, LX), nbatch), dim3(256), shmem, st, to_dev(px), G, phys, nlines,
                       nkx, LX);
    CHECK_LAUNCH();
}

void launch_xfwd(hipStream_t st, const FFTPlan &px, const double *phys, double2 *H, int nbatch, long long nlines,
                 int nkx) {
    if (launch_xfwd_wave(st, px, phys, H, nbatch, nlines, nkx)) return;
    int LX = xlines_for(px.N);
    if (!fft_fits(LX, px.N)) return;
    size_t shmem = size_t(LX) * (px.N + 1) * sizeof(double2);
    hipLaunchKernelGGL(k_xfwd, dim3(blocks_for(nlines, LX), nbatch), dim3(256), shmem, st, to_dev(px), phys, H, nlines,
                       nkx, LX);
    CHECK_LAUNCH();
}

// remove_vertical_mean_of_quadratic_var (PL_IMEX_timestepping.f90:223-235)
template <class C>
__global__ void k_remove_zmean(C *Qhat, int NZ, int nsys) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 1; k <= NZ / 2 - 1; ++k) {
        double w = 0.5 / (2.0 * k + 1.0) - 0.5 / (2.0 * k - 1.0);
        const C q = Qhat[(size_t)(2 * k) * nsys + s];
        acc.x -= double(q.x) * w;
        acc.y -= double(q.y) * w;
    }
    Qhat[s] = mkc<C>(typename CxT<C>::R(acc.x), typename CxT<C>::R(acc.y));
}

void launch_remove_zmean(hipStream_t st, double2 *Qhat, int NZ, int nsys) {
    hipLaunchKernelGGL(k_remove_zmean<double2>, dim3(blocks_for(nsys, 256)), dim3(256), 0, st, Qhat, NZ, nsys);
    CHECK_LAUNCH();
}

void launch_remove_zmean32(hipStream_t st, float2 *Qhat, int NZ, int nsys) {
    hipLaunchKernelGGL(k_remove_zmean<float2>, dim3(blocks_for(nsys, 256)), dim3(256), 0, st, Qhat, NZ, nsys);
    CHECK_LAUNCH();
}

// Boundary-value extraction and vertical decomposition of one full variable's
// Chebyshev coefficients spec[n][s] (n < NZ), after they are formed and
// before the z synthesis (compute_full_variables_in_physical_space,
// PL_IMEX_timestepping_marching.f90:544-577):
//   extract 1 (BotSurf) / 2 (TopSurf): row 0 <- the series' value at the
//     bottom / top wall (sum of (-1)^n c_n / of c_n, in the order of n), rows
//     1 .. NZ-1 zeroed;
//   zdecomp 1 (barotropic, PL_IMEX_timestepping.f90:252-274): row 0 <- row 0
//     plus the even rows 2q times their integral weights, rows 2q-1 and 2q
//     zeroed, q = 1 .. NZ/2-1;
//   zdecomp 2 (baroclinic, :238-250): row 0 <- minus those weighted sums (the
//     series' z-integral removed).
// One thread per system.
template <class C>
__global__ void k_full_zfilter(C *spec, int NZ, int nsys, int extract, int zdecomp) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    using R = typename CxT<C>::R;
    const C zero = mkc<C>(R(0), R(0));
    if (extract) {
        C acc = spec[s];
        for (int n = 1; n < NZ; ++n) {
            C &v = spec[(size_t)n * nsys + s];
            if (extract == 2 || (n & 1) == 0) {
                acc.x += v.x;
                acc.y += v.y;
            } else {
                acc.x -= v.x;
                acc.y -= v.y;
            }
            v = zero;
        }
        spec[s] = acc;
    }
    if (zdecomp) {
        double2 acc = zdecomp == 1 ? make_double2(double(spec[s].x), double(spec[s].y)) : make_double2(0.0, 0.0);
        for (int k = 1; k <= NZ / 2 - 1; ++k) {
            const double w = 0.5 / (2.0 * k + 1.0) - 0.5 / (2.0 * k - 1.0);
            C &e = spec[(size_t)(2 * k) * nsys + s];
            if (zdecomp == 1) {
                acc.x += double(e.x) * w;
                acc.y += double(e.y) * w;
                spec[(size_t)(2 * k - 1) * nsys + s] = zero;
                e = zero;
            } else {
                acc.x -= double(e.x) * w;
                acc.y -= double(e.y) * w;
            }
        }
        spec[s] = mkc<C>(R(acc.x), R(acc.y));
    }
}

void launch_full_zfilter(hipStream_t st, double2 *spec, int NZ, int nsys, int extract, int zdecomp) {
    hipLaunchKernelGGL(k_full_zfilter<double2>, dim3(blocks_for(nsys, 256)), dim3(256), 0, st, spec, NZ, nsys, extract,
                       zdecomp);
    CHECK_LAUNCH();
}

void launch_full_zfilter32(hipStream_t st, float2 *spec, int NZ, int nsys, int extract, int zdecomp) {
    hipLaunchKernelGGL(k_full_zfilter<float2>, dim3(blocks_for(nsys, 256)), dim3(256), 0, st, spec, NZ, nsys, extract,
                       zdecomp);
    CHECK_LAUNCH();
}

// Galerkin coefficients of one coupled variable from the Chebyshev
// coefficients of its physical values (import_quicksave,
// PL_IMEX_timestepping_output.f90:340-356): g[j] = sum_{n < ncol} Sinv[j][n] Q[n][s],
// stored at position pos[j] of every kxky system; the (0,0) column stays zero.
__global__ void k_insert_var(const double2 *__restrict__ Q, int nsys, int ncol, const double *__restrict__ Sinv,
                             const int *__restrict__ pos, double2 *__restrict__ field, int total) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)ncol * nsys) return;
    const int j = int(idx / nsys), s = int(idx % nsys);
    double2 g = make_double2(0.0, 0.0);
    for (int n = 0; n < ncol; ++n) {
        const double w = Sinv[(size_t)j * ncol + n];
        const double2 q = Q[(size_t)n * nsys + s];
        g.x += w * q.x;
        g.y += w * q.y;
    }
    field[(size_t)s * total + pos[j]] = s == 0 ? make_double2(0.0, 0.0) : g;
}

void launch_insert_var(hipStream_t st, const double2 *Q, int nsys, int ncol, const double *Sinv, const int *pos,
                       double2 *field, int total) {
    hipLaunchKernelGGL(k_insert_var, dim3(blocks_for((long long)ncol * nsys, 256)), dim3(256), 0, st, Q, nsys, ncol,
                       Sinv, pos, field, total);
    CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// NL assembly (PL_IMEX_timestepping_marching.f90:306-367 kxky, :372-397 mean)
// ---------------------------------------------------------------------------
// kxky systems: a workgroup owns 16 systems x 64 positions.  Lanes run over
// systems first, so every read of Qhat[q][m][s] covers 16 consecutive
// systems; the result is transposed through LDS and written as 64-position
// rows of each system.
__global__ __launch_bounds__(256) void k_nl(NLArgs a) {
    __shared__ double2 tile[64][17];
    const int nsys = a.nsys, total = a.total;
    const int ntp = (total + 63) / 64;
    const int s0 = (blockIdx.x / ntp) * 16, p0 = (blockIdx.x % ntp) * 64;
    const int sl = threadIdx.x & 15, pr = threadIdx.x >> 4;
    const int s = s0 + sl;
    const bool sv = s < nsys;
    const double kx = a.km.kx(sv ? s : 0), ky = a.km.ky(sv ? s : 0);
    const size_t qstride = (size_t)a.NZ * nsys;
    for (int i = 0; i < 4; ++i) {
        const int pl = pr + 16 * i, pos = p0 + pl;
        double2 acc = make_double2(0.0, 0.0);
        if (pos < total && sv && s != 0) {
            const int e = a.pos_eqn[pos], r = a.pos_row[pos];
            for (int t = a.eqn_begin[e]; t < a.eqn_begin[e + 1]; ++t) {
                const int w = a.t_w[t], m0 = r + a.t_order[t];
                const double2 *Q = a.Qhat + a.t_qvar[t] * qstride + s;
                const double *W = a.wv + a.t_woff[t] + r * (2 * w + 1);
                double2 sum = make_double2(0.0, 0.0);
                for (int d = -w; d <= w; ++d) {
                    const int m = m0 + d;
                    if (m < 0 || m >= a.NZ) continue;
                    const double wv = W[d + w];
                    const double2 qv = Q[(size_t)m * nsys];
                    sum.x += wv * qv.x;
                    sum.y += wv * qv.y;
                }
                const double2 c = monomial(a.t_kx[t], a.t_ky[t], kx, ky);
                acc = cadd(acc, cscale(cmul(c, sum), a.t_dsca[t]));
            }
        }
        tile[pl][sl] = acc;
    }
    __syncthreads();
    for (int i = 0; i < 4; ++i) {
        const int e = threadIdx.x + 256 * i, sl2 = e >> 6, pl = e & 63;
        const int s2 = s0 + sl2, pos = p0 + pl;
        if (s2 < nsys && pos < total) a.Khat[(size_t)s2 * total + pos] = tile[pl][sl2];
    }
}

// One workgroup per 64 consecutive systems (one per lane) and NLP_POS
// positions (NLP_POS / NLP_WAVES per wavefront).  A position's program is the
// same for every system, so the loops are wave-uniform: the program tables
// are read with scalar loads and every Qhat access is one contiguous row of
// 64 systems; the results go through an LDS tile so that each system's run of
// positions is stored contiguously.  Neighbouring positions read overlapping
// Qhat rows; 64 positions over 16 wavefronts keep those re-reads inside one
// CU (4 positions per wavefront, as few sequential programs as possible).
constexpr int NLP_POS = 64, NLP_WAVES = 16, NLP_THREADS = 64 * NLP_WAVES;

__device__ inline double2 to_d2(double2 v) { return v; }
__device__ inline double2 to_d2(float2 v) { return make_double2(double(v.x), double(v.y)); }

// QT: element type of Qhat (double2, or float2 in the FP32 chain; the sums are FP64)
template <class QT>
__global__ __launch_bounds__(NLP_THREADS) void k_nl_prog(NLProgArgs a, const QT *__restrict__ Qsrc) {
    __shared__ double2 tile[NLP_POS][65];
    const int nsys = a.nsys, total = a.total;
    const int npb = (total + NLP_POS - 1) / NLP_POS;
    const int pb = blockIdx.x % npb, sb = blockIdx.x / npb;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = sb * 64 + lane;
    const bool sv = s < nsys;
    const int sc = sv ? s : 0;
    const QT *__restrict__ Q = Qsrc + sc;
    const int *__restrict__ pos_ptr = a.pos_ptr;
    const int *__restrict__ blk_ptr = a.blk_ptr;
    const int *__restrict__ blk_term = a.blk_term;
    const int *__restrict__ ent_qm = a.ent_qm;
    const double *__restrict__ ent_w = a.ent_w;
#pragma unroll 1
    for (int k = 0; k < NLP_POS / NLP_WAVES; ++k) {
        const int pl = wv * (NLP_POS / NLP_WAVES) + k;
        const int pos = pb * NLP_POS + pl;
        double2 acc = make_double2(0.0, 0.0);
        if (pos < total) {
            const int b0 = pos_ptr[pos], b1 = pos_ptr[pos + 1];
            for (int b = b0; b < b1; ++b) {
                const int e0 = blk_ptr[b], e1 = blk_ptr[b + 1];
                double2 sum = make_double2(0.0, 0.0);
                int e = e0;
                for (; e + 4 <= e1; e += 4) {
                    double2 q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) q[u] = to_d2(Q[(size_t)ent_qm[e + u] * nsys]);
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        sum.x = fma(ent_w[e + u], q[u].x, sum.x);
                        sum.y = fma(ent_w[e + u], q[u].y, sum.y);
                    }
                }
                for (; e < e1; ++e) {
                    const double2 q = to_d2(Q[(size_t)ent_qm[e] * nsys]);
                    sum.x = fma(ent_w[e], q.x, sum.x);
                    sum.y = fma(ent_w[e], q.y, sum.y);
                }
                acc = cadd(acc, cmul(a.cmono[(size_t)blk_term[b] * nsys + sc], sum));
            }
        }
        // the kxky (0,0) mode stays zero
        tile[pl][lane] = (sv && s != 0) ? acc : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int i = 0; i < 64 * NLP_POS / NLP_THREADS; ++i) {
        const int e = threadIdx.x + NLP_THREADS * i, sl = e / NLP_POS, pl = e % NLP_POS;
        const int s2 = sb * 64 + sl, pos = pb * NLP_POS + pl;
        if (s2 < nsys && pos < total) a.Khat[(size_t)s2 * total + pos] = tile[pl][sl];
    }
}

void launch_nl_prog(hipStream_t st, const NLProgArgs &a) {
    const int npb = (a.total + NLP_POS - 1) / NLP_POS;
    if (a.Qhat32)
        hipLaunchKernelGGL(k_nl_prog<float2>, dim3(blocks_for(a.nsys, 64) * npb), dim3(NLP_THREADS), 0, st, a,
                           a.Qhat32);
    else
        hipLaunchKernelGGL(k_nl_prog<double2>, dim3(blocks_for(a.nsys, 64) * npb), dim3(NLP_THREADS), 0, st, a,
                           a.Qhat);
    CHECK_LAUNCH();
}

__global__ void k_term_monomials(const int *t_kx, const int *t_ky, int nterms, KeptModes km, int nsys, double2 *out) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nsys * nterms) return;
    const int t = int(idx / nsys), s = int(idx % nsys);
    out[idx] = monomial(t_kx[t], t_ky[t], km.kx(s), km.ky(s));
}

void launch_term_monomials(hipStream_t st, const int *t_kx, const int *t_ky, int nterms, KeptModes km, int nsys,
                           double2 *cmono) {
    if (nterms == 0) return;
    hipLaunchKernelGGL(k_term_monomials, dim3(blocks_for((long long)nsys * nterms, 256)), dim3(256), 0, st, t_kx,
                       t_ky, nterms, km, nsys, cmono);
    CHECK_LAUNCH();
}

// mean systems: real part of the (0,0) column, then the sources' constant part
// (PL_IMEX_timestepping_marching.f90:369-417: the NL terms, then the sources,
// accumulate into K_hat)
__global__ void k_nl_mean(NLArgs a) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= a.total) return;
    const int e = a.pos_eqn[pos], r = a.pos_row[pos];
    double acc = 0.0;
    for (int t = a.eqn_begin[e]; t < a.eqn_begin[e + 1]; ++t) {
        const int w = a.t_w[t], m0 = r + a.t_order[t];
        const size_t qoff = a.t_qvar[t] * (size_t)a.NZ * a.nsys;
        const double *W = a.wv + a.t_woff[t] + r * (2 * w + 1);
        double sum = 0.0;
        for (int d = -w; d <= w; ++d) {
            const int m = m0 + d;
            if (m < 0 || m >= a.NZ) continue;
            const size_t i = qoff + (size_t)m * a.nsys;
            sum += W[d + w] * (a.Qhat32 ? double(a.Qhat32[i].x) : a.Qhat[i].x);
        }
        acc += a.t_dsca[t] * sum;
    }
    if (a.src) acc += a.src[pos];
    a.Khat[pos] = make_double2(acc, 0.0);
}

void launch_nl(hipStream_t st, const NLArgs &a) {
    if (a.Qhat32 && !a.mean) throw std::runtime_error("FP32 Qhat: the kxky NL terms run on their per-position program");
    if (a.mean) {
        hipLaunchKernelGGL(k_nl_mean, dim3(blocks_for(a.total, 256)), dim3(256), 0, st, a);
    } else {
        const int ntp = (a.total + 63) / 64;
        hipLaunchKernelGGL(k_nl, dim3(blocks_for(a.nsys, 16) * ntp), dim3(256), 0, st, a);
    }
    CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// right-hand sides (mass_field_to_rhs, add_K_std_to_rhs, add_K_hat_to_rhs,
// PL_IMEX_timestepping_marching.f90:93-171): one thread per (system, row),
// operators as ELL rows of the merged groups
// ---------------------------------------------------------------------------
// A workgroup covers RHS_ROWS consecutive rows of one system: the combined
// vectors X = sum_k ma_k mv_k and V = sum_k xa_k xv_k are staged once in LDS
// over the rows and the operator band around them, so the ELL products read
// LDS instead of issuing one global load per entry and vector.
constexpr int RHS_ROWS = 256;

__global__ __launch_bounds__(RHS_ROWS) void k_rhs(RhsArgs a) {
    extern __shared__ double2 xs[];   // [2][RHS_ROWS + bl + bu]
    const int nchunk = (a.total + RHS_ROWS - 1) / RHS_ROWS;
    const int s = blockIdx.x / nchunk, i0 = (blockIdx.x - s * nchunk) * RHS_ROWS;
    const int lo = max(i0 - a.bl, 0), hi = min(i0 + RHS_ROWS + a.bu, a.total), span = RHS_ROWS + a.bl + a.bu;
    const size_t base = (size_t)s * a.total;
    double2 *X = xs, *V = xs + span;
    for (int c = lo + threadIdx.x; c < hi; c += blockDim.x) {
        double2 u = make_double2(0.0, 0.0), v = make_double2(0.0, 0.0);
        for (int k = 0; k < a.nm; ++k) {
            const double2 t = a.mv[k][base + c];
            u.x = fma(a.ma[k], t.x, u.x);
            u.y = fma(a.ma[k], t.y, u.y);
        }
        for (int k = 0; k < a.nx; ++k) {
            const double2 t = a.xv[k][base + c];
            v.x = fma(a.xa[k], t.x, v.x);
            v.y = fma(a.xa[k], t.y, v.y);
        }
        X[c - lo] = u;
        V[c - lo] = v;
    }
    __syncthreads();
    const int i = i0 + threadIdx.x;
    if (i >= a.total) return;
    double2 y = make_double2(0.0, 0.0);
    for (int g = 0; g < a.mass.ngroups; ++g) {
        const int W = a.mass.ell_w[g], off = a.mass.ell_off[g] + i * W;
        double2 sum = make_double2(0.0, 0.0);
        for (int e = 0; e < W; ++e) {
            const double bv = a.mass.ell_val[off + e];
            const double2 u = X[a.mass.ell_col[off + e] - lo];
            sum.x = fma(bv, u.x, sum.x);
            sum.y = fma(bv, u.y, sum.y);
        }
        y = cadd(y, cmul(a.cmass[(size_t)s * a.mass.ngroups + g], sum));
    }
    if (a.nx > 0) {
        for (int g = 0; g < a.stif.ngroups; ++g) {
            const int W = a.stif.ell_w[g], off = a.stif.ell_off[g] + i * W;
            double2 sum = make_double2(0.0, 0.0);
            for (int e = 0; e < W; ++e) {
                const double bv = a.stif.ell_val[off + e];
                const double2 v = V[a.stif.ell_col[off + e] - lo];
                sum.x = fma(bv, v.x, sum.x);
                sum.y = fma(bv, v.y, sum.y);
            }
            y = cadd(y, cmul(a.cstif[(size_t)s * a.stif.ngroups + g], sum));
        }
    }
    const size_t idx = base + i;
    for (int k = 0; k < a.nk; ++k) {
        const double2 t = a.kv[k][idx];
        y.x = fma(a.kb[k], t.x, y.x);
        y.y = fma(a.kb[k], t.y, y.y);
    }
    if (a.zero00 && s == 0) y = make_double2(0.0, 0.0);
    a.y[idx] = y;
}

// Whole-system variant for total <= 256 * RPT: one workgroup per system, RPT
// rows per thread.  No halo rows are re-read, and the loads of the thread's
// RPT rows of the input vectors are issued together.
template <int RPT>
__global__ __launch_bounds__(256) void k_rhs_sys(RhsArgs a) {
    extern __shared__ double2 xs[];   // X [total], V [total]
    const int total = a.total, s = blockIdx.x, tid = threadIdx.x;
    const size_t base = (size_t)s * total;
    double2 *X = xs, *V = xs + total;
    // rows tid + 256 r of vector src (zero past the end)
    auto load_rows = [&](const double2 *src, double2 (&t)[RPT]) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = tid + 256 * r;
            t[r] = i < total ? src[base + i] : make_double2(0.0, 0.0);
        }
    };
    auto axpy_rows = [&](double c, const double2 (&t)[RPT], double2 (&acc)[RPT]) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            acc[r].x = fma(c, t[r].x, acc[r].x);
            acc[r].y = fma(c, t[r].y, acc[r].y);
        }
    };
    // the loads of the first RM mass-side, RX stiffness-side and RK K vectors
    // are all issued before any is used (one HBM round trip, not one per
    // vector); further vectors are loaded one at a time
    constexpr int RM = 2, RX = 1, RK = 3;
    double2 mr[RM][RPT], xr[RX][RPT], kr[RK][RPT];
#pragma unroll
    for (int k = 0; k < RM; ++k)
        if (k < a.nm) load_rows(a.mv[k], mr[k]);
#pragma unroll
    for (int k = 0; k < RX; ++k)
        if (k < a.nx) load_rows(a.xv[k], xr[k]);
#pragma unroll
    for (int k = 0; k < RK; ++k)
        if (k < a.nk) load_rows(a.kv[k], kr[k]);
    {
        double2 u[RPT], v[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) u[r] = v[r] = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < RM; ++k)
            if (k < a.nm) axpy_rows(a.ma[k], mr[k], u);
        for (int k = RM; k < a.nm; ++k) {
            double2 t[RPT];
            load_rows(a.mv[k], t);
            axpy_rows(a.ma[k], t, u);
        }
#pragma unroll
        for (int k = 0; k < RX; ++k)
            if (k < a.nx) axpy_rows(a.xa[k], xr[k], v);
        for (int k = RX; k < a.nx; ++k) {
            double2 t[RPT];
            load_rows(a.xv[k], t);
            axpy_rows(a.xa[k], t, v);
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = tid + 256 * r;
            if (i < total) {
                X[i] = u[r];
                V[i] = v[r];
            }
        }
    }
    __syncthreads();
    double2 y[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int i = min(tid + 256 * r, total - 1);
        double2 acc = make_double2(0.0, 0.0);
        for (int g = 0; g < a.mass.ngroups; ++g) {
            const int W = a.mass.ell_w[g], off = a.mass.ell_off[g] + i * W;
            double2 sum = make_double2(0.0, 0.0);
            for (int e = 0; e < W; ++e) {
                const double bv = a.mass.ell_val[off + e];
                const double2 u = X[a.mass.ell_col[off + e]];
                sum.x = fma(bv, u.x, sum.x);
                sum.y = fma(bv, u.y, sum.y);
            }
            acc = cadd(acc, cmul(a.cmass[(size_t)s * a.mass.ngroups + g], sum));
        }
        if (a.nx > 0) {
            for (int g = 0; g < a.stif.ngroups; ++g) {
                const int W = a.stif.ell_w[g], off = a.stif.ell_off[g] + i * W;
                double2 sum = make_double2(0.0, 0.0);
                for (int e = 0; e < W; ++e) {
                    const double bv = a.stif.ell_val[off + e];
                    const double2 v = V[a.stif.ell_col[off + e]];
                    sum.x = fma(bv, v.x, sum.x);
                    sum.y = fma(bv, v.y, sum.y);
                }
                acc = cadd(acc, cmul(a.cstif[(size_t)s * a.stif.ngroups + g], sum));
            }
        }
        y[r] = acc;
    }
#pragma unroll
    for (int k = 0; k < RK; ++k)
        if (k < a.nk) axpy_rows(a.kb[k], kr[k], y);
    for (int k = RK; k < a.nk; ++k) {
        double2 t[RPT];
        load_rows(a.kv[k], t);
        axpy_rows(a.kb[k], t, y);
    }
    const bool zero = a.zero00 && s == 0;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int i = tid + 256 * r;
        if (i < total) a.y[base + i] = zero ? make_double2(0.0, 0.0) : y[r];
    }
}

void launch_rhs(hipStream_t st, const RhsArgs &a) {
    {
        const int rpt = (a.total + 255) / 256;
        const size_t shm = 2 * size_t(a.total) * sizeof(double2);
        const dim3 grid(a.nsys), block(256);
        switch (rpt) {
        case 1: hipLaunchKernelGGL(k_rhs_sys<1>, grid, block, shm, st, a); CHECK_LAUNCH(); return;
        case 2: hipLaunchKernelGGL(k_rhs_sys<2>, grid, block, shm, st, a); CHECK_LAUNCH(); return;
        case 3: hipLaunchKernelGGL(k_rhs_sys<3>, grid, block, shm, st, a); CHECK_LAUNCH(); return;
        case 4: hipLaunchKernelGGL(k_rhs_sys<4>, grid, block, shm, st, a); CHECK_LAUNCH(); return;
        case 5: hipLaunchKernelGGL(k_rhs_sys<5>, grid, block, shm, st, a); CHECK_LAUNCH(); return;
        case 6: hipLaunchKernelGGL(k_rhs_sys<6>, grid, block, shm, st, a); CHECK_LAUNCH(); return;
        default: break;   // longer systems: row chunks with halos
        }
    }
    const int nchunk = (a.total + RHS_ROWS - 1) / RHS_ROWS;
    const size_t shm = 2 * size_t(RHS_ROWS + a.bl + a.bu) * sizeof(double2);
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((long long)a.nsys * nchunk)), dim3(RHS_ROWS), shm, st, a);
    CHECK_LAUNCH();
}

__global__ void k_group_coefs(OpDev op, KeptModes km, int nsys, double2 *out) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nsys * op.ngroups) return;
    const int s = int(idx / op.ngroups), g = int(idx % op.ngroups);
    out[idx] = group_coef(op, g, km.kx(s), km.ky(s));
}

void launch_group_coefs(hipStream_t st, const OpDev &op, KeptModes km, int nsys, double2 *out) {
    if (op.ngroups == 0) return;
    hipLaunchKernelGGL(k_group_coefs, dim3(blocks_for((long long)nsys * op.ngroups, 256)), dim3(256), 0, st, op, km,
                       nsys, out);
    CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// banded LU, one workgroup (one wavefront) per wavenumber: restates LAPACK
// xGBTF2 as called per (kx,ky) in PL_algebra_z_3D.f90:245-276.  The working
// band lives in `work` (LAPACK layout), the result is split into column
// records Lf (multipliers) and Uf (inverse pivot + column of U above it).
// ---------------------------------------------------------------------------
// Scalar traits of the factor entries: complex (double2), or real (double)
// for systems in real form.  The real reciprocal is 1/x, which is what the
// complex reciprocal gives for a zero imaginary part.
template <typename T>
struct Fx;
template <>
struct Fx<double2> {
    __device__ static double abs1(double2 z) { return fabs(z.x) + fabs(z.y); }
    __device__ static bool is_zero(double2 z) { return z.x == 0.0 && z.y == 0.0; }
    __device__ static double2 zero() { return make_double2(0.0, 0.0); }
    __device__ static double2 recip(double2 z) { return crecip(z); }
    __device__ static double2 mul(double2 a, double2 b) { return cmul(a, b); }
    __device__ static double2 sub(double2 a, double2 b) { return csub(a, b); }
};
template <>
struct Fx<double> {
    __device__ static double abs1(double z) { return fabs(z); }
    __device__ static bool is_zero(double z) { return z == 0.0; }
    __device__ static double zero() { return 0.0; }
    __device__ static double recip(double z) { return 1.0 / z; }
    __device__ static double mul(double a, double b) { return a * b; }
    __device__ static double sub(double a, double b) { return a - b; }
};

// Partial pivoting picks the largest |Re| + |Im| (LAPACK's cabs1); rotating
// rows and columns by powers of i leaves that measure unchanged, so the real
// form pivots on the same rows as the complex operator.
template <typename T>
__device__ void gbtf2_wave(T *AB, int n, int kl, int ku, int ldab, unsigned char *piv, int *info) {
    using F = Fx<T>;
    const int kv = kl + ku;
    const int tid = threadIdx.x;
    int ju = 0;
    for (int j = 0; j < n; ++j) {
        const int km = min(kl, n - 1 - j);
        T *colj = AB + (size_t)j * ldab;
        double v = (tid <= km) ? F::abs1(colj[kv + tid]) : -1.0;
        int ix = tid;
        for (int off = 32; off > 0; off >>= 1) {
            double ov = __shfl_xor(v, off);
            int oi = __shfl_xor(ix, off);
            if (ov > v || (ov == v && oi < ix)) {
                v = ov;
                ix = oi;
            }
        }
        const int jp = ix;
        if (tid == 0) piv[j] = (unsigned char)jp;
        T pv = colj[kv + jp];
        if (F::is_zero(pv)) {
            if (tid == 0 && *info == 0) *info = j + 1;
            __syncthreads();
            continue;
        }
        ju = max(ju, min(j + ku + jp, n - 1));
        if (jp != 0) {
            for (int c = j + tid; c <= ju; c += blockDim.x) {
                T *cc = AB + (size_t)c * ldab;
                T t = cc[kv + jp + j - c];
                cc[kv + jp + j - c] = cc[kv + j - c];
                cc[kv + j - c] = t;
            }
        }
        __syncthreads();
        if (km > 0) {
            T rp = F::recip(colj[kv]);
            if (tid >= 1 && tid <= km) colj[kv + tid] = F::mul(colj[kv + tid], rp);
            __syncthreads();
            const int ncols = ju - j;
            for (int e = tid; e < ncols * km; e += blockDim.x) {
                int c = j + 1 + e / km, i = 1 + e % km;
                T *cc = AB + (size_t)c * ldab;
                T u = cc[kv + j - c];
                T l = colj[kv + i];
                T t = cc[kv + j + i - c];
                cc[kv + j + i - c] = F::sub(t, F::mul(l, u));
            }
        }
        __syncthreads();
    }
}

template <typename T>
__device__ void split_factors(const T *AB, int n, int kl, int ku, int ldab, T *Lf, T *Uf) {
    using F = Fx<T>;
    const int kv = kl + ku;
    for (int e = threadIdx.x; e < n * kl; e += blockDim.x) {
        int j = e / kl, k = e % kl;
        Lf[e] = (k < min(kl, n - 1 - j)) ? AB[(size_t)j * ldab + kv + 1 + k] : F::zero();
    }
    for (int e = threadIdx.x; e < n * (kv + 1); e += blockDim.x) {
        int j = e / (kv + 1), k = e % (kv + 1);
        T v;
        if (k == 0)
            v = F::recip(AB[(size_t)j * ldab + kv]);
        else
            v = (j - k >= 0) ? AB[(size_t)j * ldab + kv - k] : F::zero();
        Uf[e] = v;
    }
}

// Factor records for the 16-lane solve.  The four matrices one wavefront
// solves are interleaved column by column, so that the wavefront's loads of
// one step fall on adjacent records.  For matrix M (group g = M/4, slot
// m = M%4), with KLP = solve16_klp(kl):
//   L:   column j at ((g*ncol16 + j)*4 + m)*KLP: entry k < kl is
//        L(j + 1 + k, j) (zero for k >= n-1-j and for the padding columns
//        j >= n);
//   U:   column j at ((g*nb + j)*4 + m)*(kv+1), nb = 16*ceil(n/16) columns per
//        matrix (those beyond n stay zero): entry 0 is 1/U(j,j), entry k the
//        superdiagonal U(j-k, j).
template <typename T>
__device__ void split_factors16(const T *AB, int n, int kl, int ku, int ldab, int ncol16, int s, T *L16, T *Ur) {
    using F = Fx<T>;
    const int kv = kl + ku, nb = 16 * ((n + 15) >> 4), g = s >> 2, m = s & 3, klp = solve16_klp(kl);
    for (int e = threadIdx.x; e < ncol16 * klp; e += blockDim.x) {
        const int j = e / klp, k = e - j * klp;
        L16[(((size_t)g * ncol16 + j) * 4 + m) * klp + k] =
            (j < n && k < min(kl, n - 1 - j)) ? AB[(size_t)j * ldab + kv + 1 + k] : F::zero();
    }
    for (int e = threadIdx.x; e < n * (kv + 1); e += blockDim.x) {
        int j = e / (kv + 1), k = e % (kv + 1);
        T v;
        if (k == 0)
            v = F::recip(AB[(size_t)j * ldab + kv]);
        else
            v = (j - k >= 0) ? AB[(size_t)j * ldab + kv - k] : F::zero();
        Ur[(((size_t)g * nb + j) * 4 + m) * (kv + 1) + k] = v;
    }
}

// Re(z * i^e) for e in {0, 1, 3}
__device__ inline double real_rot(double2 z, int e) { return e == 0 ? z.x : (e == 1 ? -z.y : z.y); }

template <typename T>
__global__ __launch_bounds__(64) void k_factor(FactorArgs a) {
    constexpr bool REAL = sizeof(T) == sizeof(double);
    // one workgroup per factored matrix: a structural block (16-lane layout) or
    // a whole system
    const bool l16 = a.ncol16 > 0;
    const int nb = l16 ? a.nblk : 1;
    const int Sl = blockIdx.x, S = a.s0 + Sl, s = S / nb, blk = S - s * nb;
    const int N = a.total, w = a.nl + a.nu + 1;
    const int n = l16 ? a.blk_n : N, kl = l16 ? a.blk_nl : a.nl, ku = l16 ? a.blk_nu : a.nu;
    const int kv = kl + ku, ldab = 2 * kl + ku + 1;
    const int *rmap = l16 && a.blk_row ? a.blk_row + (size_t)blk * n : nullptr;
    const int *cmap = l16 && a.blk_col ? a.blk_col + (size_t)blk * n : nullptr;
    T *AB = reinterpret_cast<T *>(a.work) + (size_t)Sl * n * ldab;
    const bool ident = a.identity00 && s == 0;
    const double kx = a.km.kx(s), ky = a.km.ky(s);
    for (int e = threadIdx.x; e < n * ldab; e += blockDim.x) {
        const int j = e / ldab, rr = e - j * ldab;
        const int i = rr - kv + j;
        double2 val = make_double2(0.0, 0.0);
        int rot = 0;
        if (rr >= kl && i >= 0 && i < n) {
            const int I = rmap ? rmap[i] : i, J = cmap ? cmap[j] : j;
            if (ident || I < 0 || J < 0) {
                // identity: the (0,0) system and the padding of a block (the
                // identity system stays unrotated; its solve skips the rotation)
                val.x = (i == j) ? 1.0 : 0.0;
            } else {
                const int d = J - I;
                if (d >= -a.nl && d <= a.nu) {
                    for (int g = 0; g < a.mass.ngroups; ++g) {
                        double b = a.mass.band[((size_t)g * N + I) * w + d + a.nl];
                        if (b != 0.0) val = cadd(val, cscale(group_coef(a.mass, g, kx, ky), b));
                    }
                    for (int g = 0; g < a.stif.ngroups; ++g) {
                        double b = a.stif.band[((size_t)g * N + I) * w + d + a.nl];
                        if (b != 0.0) val = cadd(val, cscale(group_coef(a.stif, g, kx, ky), b * a.sdt));
                    }
                }
                if constexpr (REAL) rot = ((a.rot[J] >> 1) - (a.rot[I] & 1)) & 3;
            }
        }
        if constexpr (REAL)
            AB[e] = real_rot(val, rot);
        else
            AB[e] = val;
    }
    __syncthreads();
    int *info = a.info + S;
    if (threadIdx.x == 0) *info = 0;
    __syncthreads();
    // (16-lane layout: pivot records padded to whole blocks of 16 columns)
    gbtf2_wave<T>(AB, n, kl, ku, ldab, a.piv + (size_t)S * (l16 ? 16 * ((n + 15) >> 4) : n), info);
    __syncthreads();
    T *Lout = REAL ? reinterpret_cast<T *>(a.Lr) : reinterpret_cast<T *>(a.Lf);
    T *Uout = REAL ? reinterpret_cast<T *>(a.Ur) : reinterpret_cast<T *>(a.Uf);
    if (l16)
        split_factors16<T>(AB, n, kl, ku, ldab, a.ncol16, S, Lout, Uout);
    else
        split_factors<T>(AB, n, kl, ku, ldab, Lout + (size_t)S * n * kl, Uout + (size_t)S * n * (kv + 1));
}

void launch_factor(hipStream_t st, const FactorArgs &a) {
    if (a.rot)
        hipLaunchKernelGGL(k_factor<double>, dim3(a.ns), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL(k_factor<double2>, dim3(a.ns), dim3(64), 0, st, a);
    CHECK_LAUNCH();
}

__global__ __launch_bounds__(64) void k_factor_lapack(double2 *work, int n, int kl, int ku, double2 *Lf, double2 *Uf,
                                                      unsigned char *piv, int *info) {
    const int s = blockIdx.x, kv = kl + ku, ldab = 2 * kl + ku + 1;
    double2 *AB = work + (size_t)s * n * ldab;
    if (threadIdx.x == 0) info[s] = 0;
    __syncthreads();
    gbtf2_wave<double2>(AB, n, kl, ku, ldab, piv + (size_t)s * n, info + s);
    __syncthreads();
    split_factors<double2>(AB, n, kl, ku, ldab, Lf + (size_t)s * n * kl, Uf + (size_t)s * n * (kv + 1));
}

void launch_factor_lapack(hipStream_t st, double2 *work, int n, int kl, int ku, int nsys, double2 *Lf, double2 *Uf,
                          unsigned char *piv, int *info) {
    hipLaunchKernelGGL(k_factor_lapack, dim3(nsys), dim3(64), 0, st, work, n, kl, ku, Lf, Uf, piv, info);
    CHECK_LAUNCH();
}

// LAPACK band input factored into the records of the 16-lane solve (the
// layout k_factor writes for ncol16 > 0, one block per system)
__global__ __launch_bounds__(64) void k_factor_lapack16(double2 *work, int n, int kl, int ku, int ncol16, double2 *L16,
                                                        double2 *Ufull, unsigned char *piv, int *info) {
    const int s = blockIdx.x, ldab = 2 * kl + ku + 1;
    double2 *AB = work + (size_t)s * n * ldab;
    if (threadIdx.x == 0) info[s] = 0;
    __syncthreads();
    gbtf2_wave<double2>(AB, n, kl, ku, ldab, piv + (size_t)s * 16 * ((n + 15) >> 4), info + s);
    __syncthreads();
    split_factors16<double2>(AB, n, kl, ku, ldab, ncol16, s, L16, Ufull);
}

void launch_factor_lapack16(hipStream_t st, double2 *work, int n, int kl, int ku, int nsys, int ncol16, double2 *L16,
                            double2 *Ufull, unsigned char *piv, int *info) {
    hipLaunchKernelGGL(k_factor_lapack16, dim3(nsys), dim3(64), 0, st, work, n, kl, ku, ncol16, L16, Ufull, piv, info);
    CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Helpers shared by the solves.  Factor entries are T = double2 (complex) or
// double (real form); the right-hand side is always complex.
// ---------------------------------------------------------------------------
typedef double nd2 __attribute__((ext_vector_type(2)));
// streaming (non-temporal) loads of factor entries
__device__ inline double2 ld_nt(const double2 *p) {
    nd2 v = __builtin_nontemporal_load(reinterpret_cast<const nd2 *>(p));
    return make_double2(v.x, v.y);
}
__device__ inline double ld_nt(const double *p) { return __builtin_nontemporal_load(p); }

// component-wise select: selects on double2 aggregates are lowered through the
// stack, selects on the two doubles stay in registers
__device__ inline double2 csel(bool c, double2 a, double2 b) {
    return make_double2(c ? a.x : b.x, c ? a.y : b.y);
}
__device__ inline double tsel(bool c, double a, double b) { return c ? a : b; }
__device__ inline double2 tsel(bool c, double2 a, double2 b) { return csel(c, a, b); }

// y - c * b
__device__ inline double2 submul(double2 y, double2 c, double2 b) {
    y.x = fma(-c.x, b.x, y.x);
    y.x = fma(c.y, b.y, y.x);
    y.y = fma(-c.x, b.y, y.y);
    y.y = fma(-c.y, b.x, y.y);
    return y;
}
__device__ inline double2 submul(double2 y, double c, double2 b) {
    y.x = fma(-c, b.x, y.x);
    y.y = fma(-c, b.y, y.y);
    return y;
}
__device__ inline double2 scale(double2 z, double2 c) { return cmul(z, c); }
__device__ inline double2 scale(double2 z, double c) { return make_double2(z.x * c, z.y * c); }

// real-form rotations of the right-hand side (z * i^-psi) and of the solution (z * i^phi)
__device__ inline double2 rot_in(double2 z, int code) { return csel(code & 1, make_double2(z.y, -z.x), z); }
__device__ inline double2 rot_out(double2 z, int code) { return csel(code & 2, make_double2(-z.y, z.x), z); }

// ---------------------------------------------------------------------------
// banded solve (xGBTRS 'N', PL_algebra_z_3D.f90:189-243), one wavefront per
// wavenumber, right-hand side staged in LDS (small systems)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_solve(const T *__restrict__ Lf, const T *__restrict__ Uf,
                                              const unsigned char *__restrict__ piv,
                                              const unsigned char *__restrict__ rot, int ident0, int n, int kl,
                                              int kv, const double2 *__restrict__ rhs, double2 *__restrict__ x) {
    extern __shared__ double2 b[];
    const int s = blockIdx.x, tid = threadIdx.x;
    const bool dorot = rot && !(ident0 && s == 0);
    const T *L = Lf + (size_t)s * n * kl;
    const T *U = Uf + (size_t)s * n * (kv + 1);
    const unsigned char *P = piv + (size_t)s * n;
    for (int i = tid; i < n; i += blockDim.x) {
        const double2 v = rhs[(size_t)s * n + i];
        b[i] = dorot ? rot_in(v, rot[i]) : v;
    }
    __syncthreads();
    for (int j = 0; j < n - 1; ++j) {
        const int p = P[j];
        if (p != 0) {
            if (tid == 0) {
                double2 t = b[j];
                b[j] = b[j + p];
                b[j + p] = t;
            }
            __syncthreads();
        }
        const double2 bj = b[j];
        const int lm = min(kl, n - 1 - j);
        for (int k = tid; k < lm; k += blockDim.x) b[j + 1 + k] = submul(b[j + 1 + k], L[(size_t)j * kl + k], bj);
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {
        if (tid == 0) b[j] = scale(b[j], U[(size_t)j * (kv + 1)]);
        __syncthreads();
        const double2 xj = b[j];
        const int km = min(kv, j);
        for (int k = tid; k < km; k += blockDim.x)
            b[j - 1 - k] = submul(b[j - 1 - k], U[(size_t)j * (kv + 1) + 1 + k], xj);
        __syncthreads();
    }
    for (int i = tid; i < n; i += blockDim.x) x[(size_t)s * n + i] = dorot ? rot_out(b[i], rot[i]) : b[i];
}

// ---------------------------------------------------------------------------
// Register-resident banded solve: one wavefront per wavenumber.  Position i of
// the right-hand side lives in lane i%64 of a 64-entry block; the wave keeps the
// active block and its neighbour in registers and stages finished blocks of the
// forward result through the output array.  Pivots and the current unknown are
// broadcast with v_readlane (uniform scalar).  Factor records are fetched per
// lane, PF steps ahead, into a ring of registers indexed statically by the
// unrolled step, so each step's coefficients are already in flight.
// Requirements: kl < 64 and kl + ku < 64.
// ---------------------------------------------------------------------------
__device__ inline double readlane_d(double v, int l) {
    long long i = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(i & 0xffffffffLL), l);
    int hi = __builtin_amdgcn_readlane((int)(i >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double2 readlane_c(double2 v, int l) { return make_double2(readlane_d(v.x, l), readlane_d(v.y, l)); }

// p[i] if cond else 0, as an unconditional load of an in-bounds element (index
// clamped to 0) followed by a value select, so no pointer select is generated
__device__ inline double2 ld_masked(const double2 *p, int i, bool cond) {
    const double2 v = p[cond ? i : 0];
    return csel(cond, v, make_double2(0.0, 0.0));
}

// forward coefficient of lane `lane` for step j: L(j+1+k, j), k = (lane - j%64 - 1) mod 64.
// The load address is always inside the system's factor block (clamped); invalid
// lanes and steps get zero, so the update is a no-op there.
template <typename T>
__device__ inline T fwd_coef(const T *L, int j, int lane, int n, int kl) {
    const int t = j & 63;
    const int k = (lane - t - 1) & 63;
    const int lm = min(kl, n - 1 - j);
    const int jc = min(j, n - 2);
    const T v = ld_nt(L + (size_t)jc * kl + (k < kl ? k : 0));
    return tsel(k < lm, v, Fx<T>::zero());
}

// backward coefficient of lane `lane` for step j: lane j%64 gets 1/U(j,j), the
// others U(j-1-k, j) with k = (j%64 - 1 - lane) mod 64; zero outside the band or for j >= n.
template <typename T>
__device__ inline T bwd_coef(const T *U, int j, int lane, int n, int kv) {
    const int t = j & 63;
    const int jc = max(0, min(j, n - 1));
    const int k = (t - 1 - lane) & 63;
    const bool diag = lane == t;
    const int off = diag ? 0 : 1 + (k < kv ? k : 0);
    const T v = ld_nt(U + (size_t)jc * (kv + 1) + off);
    const bool valid = (j >= 0) && (j < n) && (diag || k < min(kv, j));
    return tsel(valid, v, Fx<T>::zero());
}

template <int PF, typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_solve_reg(
    const T *__restrict__ Lf, const T *__restrict__ Uf, const unsigned char *__restrict__ piv,
    const unsigned char *__restrict__ rot, int ident0, int n, int kl, int kv, const double2 *__restrict__ rhs,
    double2 *__restrict__ x) {
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    const bool dorot = rot && !(ident0 && s == 0);
    const T *L = Lf + (size_t)s * n * kl;
    const T *U = Uf + (size_t)s * n * (kv + 1);
    const unsigned char *P = piv + (size_t)s * n;
    const double2 *b = rhs + (size_t)s * n;
    double2 *xo = x + (size_t)s * n;
    const int R = (n + 63) >> 6;
    const T zero = Fx<T>::zero();
    auto code = [&](int i) -> int { return dorot ? (int)rot[min(i, n - 1)] : 0; };

    // ---- forward: L with interleaved row interchanges (xGBTRS, first loop) ----
    double2 cur = rot_in(ld_masked(b, lane, lane < n), code(lane));
    double2 nxt = rot_in(ld_masked(b, 64 + lane, 64 + lane < n), code(64 + lane));
    int pv = P[min(lane, n - 1)];
    pv = (lane < n) ? pv : 0;
    T ring[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) ring[u] = fwd_coef(L, u, lane, n, kl);
    for (int r = 0; r < R; ++r) {
        for (int t0 = 0; t0 < 64; t0 += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int t = t0 + u;
                const int j = (r << 6) + t;
                const int p = __builtin_amdgcn_readlane(pv, t);
                if (p != 0) {
                    // interchange positions j and j+p (lane tp of this block or lane tp-64 of the next)
                    const int tp = t + p;
                    const double2 a = readlane_c(cur, t);
                    const double2 ca = readlane_c(cur, min(tp, 63));
                    const double2 cb = readlane_c(nxt, max(tp - 64, 0));
                    const double2 c2 = csel(tp < 64, ca, cb);
                    cur = csel(lane == t, c2, cur);
                    cur = csel(lane == tp, a, cur);
                    nxt = csel(lane == tp - 64, a, nxt);
                }
                const double2 bj = readlane_c(cur, t);
                const T c = ring[u];
                const bool here = lane > t;
                cur = submul(cur, tsel(here, c, zero), bj);
                if (t >= 64 - kl) nxt = submul(nxt, tsel(here, zero, c), bj);
                ring[u] = fwd_coef(L, j + PF, lane, n, kl);
            }
        }
        const int i0 = (r << 6) + lane;
        if (i0 < n) xo[i0] = cur;
        if (r < R - 1) {
            cur = nxt;
            const int i2 = ((r + 2) << 6) + lane;
            nxt = rot_in(ld_masked(b, i2, i2 < n), code(i2));
            const int i1 = ((r + 1) << 6) + lane;
            pv = P[min(i1, n - 1)];
            pv = (i1 < n) ? pv : 0;
        }
    }
    __threadfence_block();

    // ---- backward: U with kv super-diagonals, column oriented (xTBSV) ----
    // cur holds block R-1 of the forward result
    double2 prv = ld_masked(xo, ((R - 2) << 6) + lane, R >= 2);
#pragma unroll
    for (int u = 0; u < PF; ++u) ring[u] = bwd_coef(U, (R << 6) - 1 - u, lane, n, kv);
    for (int r = R - 1; r >= 0; --r) {
        for (int t0 = 63; t0 >= 0; t0 -= PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int t = t0 - u;
                const int j = (r << 6) + t;
                const T c = ring[u];
                const double2 xj = readlane_c(scale(cur, c), t);
                cur = csel(lane == t, xj, cur);
                const bool here = lane < t;
                const T cu = tsel(lane == t, zero, c);
                cur = submul(cur, tsel(here, cu, zero), xj);
                if (t < kv) prv = submul(prv, tsel(here, zero, cu), xj);
                ring[u] = bwd_coef(U, j - PF, lane, n, kv);
            }
        }
        const int i0 = (r << 6) + lane;
        if (i0 < n) xo[i0] = rot_out(cur, code(i0));
        if (r > 0) {
            cur = prv;
            prv = ld_masked(xo, ((r - 2) << 6) + lane, r >= 2);
        }
    }
}

// ---------------------------------------------------------------------------
// Banded solve with four systems per wavefront, sixteen lanes each (kl <= 16).
// Position 16b + l of a system lives in lane l of its 16-lane row, block b in
// one register.  Forward (xGBTRS first loop): at step j = 16b + t the row
// interchange j <-> j + p gathers position j + p with ds_bpermute, the value
// of position j is broadcast with DPP row_newbcast, and rows j+1 .. j+16 are
// exactly the 16 lanes (lanes > t in block b, lanes <= t in block b+1), so
// every lane does useful work.  Backward (xTBSV, column oriented): x_j is
// broadcast and NQ registers (blocks b .. b-NQ+1) take the column update.
// Factor records: the interleaved layout of split_factors16, with padding
// around the U array so that out-of-record reads stay in bounds (their values
// are masked).
// ---------------------------------------------------------------------------
// row_newbcast writes every lane of the row, so no old value is needed
template <int T_>
__device__ inline int bcast16_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + T_, 0xF, 0xF, false);
}
template <int T_>
__device__ inline double bcast16_d(double v) {
    const long long i = __double_as_longlong(v);
    const int lo = bcast16_i<T_>((int)(i & 0xffffffffLL)), hi = bcast16_i<T_>((int)(i >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int T_>
__device__ inline double2 bcast16_c(double2 v) { return make_double2(bcast16_d<T_>(v.x), bcast16_d<T_>(v.y)); }
template <int T_>
__device__ inline double bcast16(double v) { return bcast16_d<T_>(v); }
template <int T_>
__device__ inline double2 bcast16(double2 v) { return bcast16_c<T_>(v); }

__device__ inline double bperm_d(int addr, double v) {
    const long long i = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(i & 0xffffffffLL));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(i >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double2 bperm_c(int addr, double2 v) { return make_double2(bperm_d(addr, v.x), bperm_d(addr, v.y)); }

// Position words of a block (see Solve16Args::words): row of the system at
// bits 0-14, column at 15-29 (SOLVE16_NONE: padding), rotation of the
// right-hand side (psi) at bit 30, of the solution (phi) at bit 31.
__device__ inline int w_row(unsigned w) { return int(w & 0x7FFFu); }
__device__ inline int w_col(unsigned w) { return int((w >> 15) & 0x7FFFu); }

// Loads: multipliers are read once and streamed (non-temporal); the U
// windows of consecutive backward steps overlap in cache lines and use
// ordinary loads, which keep those lines in the cache for the next step.
//
// Scheduling note for the 16-lane solve.  Every value a block starts with
// (coefficients, right-hand sides, pivots, intermediate results) is loaded
// one block or PFB steps ahead into a register that the same slot's previous
// value occupied.  Each reload's address passes through an empty asm that
// consumes the values the slot's old contents fed, so the load is issued
// after the old value's last use and can take over its register; otherwise
// the compiler hoists the load, gives it a fresh register and copies it into
// place at the end of the loop, and that copy waits for every load issued
// since.
template <typename T, int NQ, int PFB>
struct Solve16 {
    // step j's U entries sit in ring slot (j mod 16) mod PFB and are reloaded
    // for step j - PFB, whose slot must be the same
    static_assert(PFB > 0 && 16 % PFB == 0, "the backward prefetch distance must divide 16");
    // forward step t of a block: swap, eliminate, prefetch the next block's
    // coefficient.  lo is the lane's position in its row, opaque to the
    // compiler so that the lane masks are formed where they are used.
    template <int t>
    __device__ static inline void fwd_step(double2 &cur, double2 &nxt, const uint4 &pv, T (&ring)[16], const T *Lg,
                                           unsigned &lcol, int lo, int rowbase, int klp) {
        const T zero = Fx<T>::zero();
        const unsigned pw = t < 4 ? pv.x : t < 8 ? pv.y : t < 12 ? pv.z : pv.w;
        const int p = int((pw >> (8 * (t & 3))) & 0xffu);
        const int tp = t + p;
        const double2 w = csel(tp < 16, cur, nxt);
        const double2 bj = bperm_c((rowbase + (tp & 15)) << 2, w);
        const double2 a = bcast16_c<t>(cur);
        cur = csel(lo == tp, a, cur);
        nxt = csel(lo + 16 == tp, a, nxt);
        cur = csel(lo == t, bj, cur);
        // lanes whose row is past the band (entry (lo - t - 1) mod 16 >= kl)
        // loaded the record's last entry and take no update
        const T c = tsel(unsigned((lo + 15 - t) & 15) < unsigned(klp), ring[t], zero);
        const bool below = lo > t;
        cur = submul(cur, tsel(below, c, zero), bj);
        nxt = submul(nxt, tsel(below, zero, c), bj);
        asm volatile("" : "+v"(lcol) : "v"(cur.x), "v"(cur.y), "v"(nxt.x), "v"(nxt.y));
        // entry k = (lo - t - 1) mod 16 of column t of the next block; k >= kl
        // reads entry kl - 1 (masked at its step)
        const unsigned kk = min(unsigned((lo + 15 - t) & 15), unsigned(klp - 1));
        ring[t] = ld_nt(Lg + (size_t)(lcol + kk) + t * 4 * klp);
        __builtin_amdgcn_sched_barrier(0);   // and not later: the prefetch distance stays 16 steps
    }

    // Lg: the wavefront's multiplier records (uniform); lcol: the offset of
    // the lane's record of the next block's first column
    template <int t>
    __device__ static inline void fwd_block(double2 &cur, double2 &nxt, const uint4 &pv, T (&ring)[16], const T *Lg,
                                            unsigned &lcol, int lo, int rowbase, int klp) {
        fwd_step<t>(cur, nxt, pv, ring, Lg, lcol, lo, rowbase, klp);
        if constexpr (t < 15) fwd_block<t + 1>(cur, nxt, pv, ring, Lg, lcol, lo, rowbase, klp);
    }

    // backward step t (descending) of a block.  Lane t's q = 0 entry of the
    // ring is record offset 0 of column j, i.e. 1/U(j, j); it is broadcast to
    // the row instead of being loaded separately.  x_j goes to xo.  Entries
    // outside a column's record were loaded from the group's zeros, so the
    // updates need no masks; block 0's lanes >= t are finished (their values
    // live in xo) and may take any update.
    // Ug: the wavefront's trimmed U records (uniform); ucol: per column (offset
    // of its four records in the group, record length); m: the lane's matrix
    // slot.  Entry e = t - lo + 16q of column j is U(j-1-k, j), k = e - 1.
    template <int t>
    __device__ static inline void bwd_step(double2 (&r)[NQ], double2 &xo, T (&cring)[PFB][NQ], const T *Ug,
                                           const int2 *ucol, int m, int lo, int jnext) {
        constexpr int slot = t % PFB;
        const double2 xj = scale(bcast16_c<t>(r[0]), bcast16<t>(cring[slot][0]));
        xo = csel(lo == t, xj, xo);
#pragma unroll
        for (int q = 0; q < NQ; ++q) r[q] = submul(r[q], cring[slot][q], xj);
        // prefetch step jnext = j - PFB (its position in the block: (t - PFB) mod
        // 16); steps before column 0 are never used and load column 0
        constexpr int tn = (t - PFB) & 15;
        const int2 cl = ucol[max(jnext, 0)];   // uniform
        int lb = cl.x + m * cl.y;
#pragma unroll
        for (int q = 0; q < NQ; ++q) asm volatile("" : "+v"(lb) : "v"(r[q].x), "v"(r[q].y));
        asm volatile("" : "+v"(lb) : "v"(xo.x), "v"(xo.y));
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tn + 16 * q - lo;
            const unsigned off = unsigned(e) < unsigned(cl.y) ? unsigned(lb + e) : 0u;
            cring[slot][q] = Ug[off];
        }
        __builtin_amdgcn_sched_barrier(0);   // and not later: the prefetch distance stays PFB steps
    }

    template <int t>
    __device__ static inline void bwd_block(double2 (&r)[NQ], double2 &xo, T (&cring)[PFB][NQ], const T *Ug,
                                            const int2 *ucol, int m, int lo, int j0) {
        bwd_step<t>(r, xo, cring, Ug, ucol, m, lo, j0 + t - PFB);
        if constexpr (t > 0) bwd_block<t - 1>(r, xo, cring, Ug, ucol, m, lo, j0);
    }
};

// the solve of one 64-lane workgroup wg of a launch (four block matrices)
template <typename T, int NQ, int PFB>
__device__ __forceinline__ void solve16_body(const Solve16Args &a, int wg) {
    using S = Solve16<T, NQ, PFB>;
    extern __shared__ unsigned sW[];   // position words of every block
    const T *__restrict__ L16 = reinterpret_cast<const T *>(a.L16);
    const T *__restrict__ Ur = reinterpret_cast<const T *>(a.U);
    const int n = a.n, kv = a.kv, ncol16 = a.ncol16;
    const int lane = threadIdx.x, l = lane & 15, rowbase = lane & 48;
    const int wstride = n + 2 * SOLVE16_WPAD;
    for (int e = lane; e < a.nblk * wstride; e += 64) sW[e] = a.words[e];
    // U column table after the words (read with uniform addresses in the
    // backward sweep; in LDS, so that those reads never wait for global loads)
    const int nbc = 16 * ((n + 15) >> 4);
    int2 *sU = reinterpret_cast<int2 *>(sW + ((a.nblk * wstride + 1) & ~1));
    for (int e = lane; e < nbc; e += 64) sU[e] = a.ucol[e];
    // factored matrix (structural block) of this 16-lane row
    const int nmat = a.nsys * a.nblk;
    const int sg = wg * 4 + (lane >> 4);
    const bool live = sg < nmat;
    const int M = live ? sg : nmat - 1;
    const int sys = M / a.nblk, blk = M - sys * a.nblk;
    // the kxky (0,0) system is factored as the identity in block coordinates,
    // which maps block position i to row w_row and column w_col; it solves
    // x = rhs when its right-hand side is read at the columns too
    const bool idsys = a.ident0 && sys == 0;
    const bool dorot = a.rotate && !idsys;
    const int g = M >> 2, m = M & 3;
    const int klp = solve16_klp(a.kl);
    const T *Lg = L16 + (size_t)g * ncol16 * 4 * klp;   // wavefront's records; matrix m's column j at (4j + m)*klp
    const int B = (n + 15) >> 4;
    // pivot offsets of block b: 16 bytes at P[b] (zero past n)
    const uint4 *P = reinterpret_cast<const uint4 *>(a.piv + (size_t)M * B * 16);
    const double2 *b = a.rhs + (size_t)sys * a.total;
    double2 *y = a.x + (size_t)sys * a.total;
    double2 *sink = a.sink + lane;   // stores of padding positions and idle lanes
    __syncthreads();
    const unsigned *W = sW + blk * wstride + SOLVE16_WPAD;   // W[i]: position i (padding outside [0, n))
    auto code = [&](unsigned w) -> int { return dorot ? int(w >> 30) : 0; };
    auto src_row = [&](unsigned w) -> int { return idsys ? w_col(w) : w_row(w); };
    auto row_or0 = [&](unsigned w) -> int { const int r = src_row(w); return r != SOLVE16_NONE ? r : 0; };
    auto col_or0 = [&](unsigned w) -> int { const int c = w_col(w); return c != SOLVE16_NONE ? c : 0; };
    auto fix_rhs = [&](double2 v, unsigned w) -> double2 {
        return rot_in(csel(src_row(w) != SOLVE16_NONE, v, make_double2(0.0, 0.0)), code(w));
    };
    auto fix_y = [&](double2 v, unsigned w) -> double2 {
        return csel(w_col(w) != SOLVE16_NONE, v, make_double2(0.0, 0.0));
    };
    auto out_ptr = [&](unsigned w) -> double2 * {
        const int c = w_col(w);
        return (live && c != SOLVE16_NONE) ? y + c : sink;
    };

    // ---- forward ----
    {
        const unsigned w0 = W[l], w1 = W[16 + l];
        double2 cur = fix_rhs(b[row_or0(w0)], w0), nxt = fix_rhs(b[row_or0(w1)], w1);
        uint4 pv = P[0];
        T ring[16];
        const unsigned lbase = unsigned(m * klp);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const unsigned kk = min(unsigned((l + 15 - t) & 15), unsigned(klp - 1));
            ring[t] = ld_nt(Lg + lbase + kk + t * 4 * klp);
            __builtin_amdgcn_sched_barrier(0);   // in the order of use
        }
        for (int bk = 0; bk < B; ++bk) {
            int lo = l;
            asm volatile("" : "+v"(lo));
            // loaded here, consumed at the end of this block: the right-hand
            // side of block bk+2 and the pivots of block bk+1
            const unsigned w2 = W[bk * 16 + 32 + l];
            const double2 nx2 = b[row_or0(w2)];
            const uint4 pvn = P[min(bk + 1, B - 1)];
            unsigned lcol = lbase + unsigned(bk * 16 + 16) * unsigned(4 * klp);
            S::template fwd_block<0>(cur, nxt, pv, ring, Lg, lcol, lo, rowbase, klp);
            *out_ptr(W[bk * 16 + l]) = cur;   // intermediate, kept at the unknown's slot
            cur = nxt;
            nxt = fix_rhs(nx2, w2);
            pv = pvn;
        }
    }
    __threadfence_block();

    // ---- backward ----
    // trimmed U records of the wavefront's group (see Solve16Args::U)
    const T *Ug = Ur + (size_t)g * a.ustride;
    const int2 *ucol = sU;
    double2 r[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const unsigned wq = W[(B - 1 - q) * 16 + l];
        r[q] = fix_y(y[col_or0(wq)], wq);
    }
    T cring[PFB][NQ];
#pragma unroll
    for (int u = 0; u < PFB; ++u) {
        // step j = 16B - 1 - u lives in slot (15 - u) % PFB
        const int j = B * 16 - 1 - u;
        const int t = j & 15;
        const int2 cl = ucol[j];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = t + 16 * q - l;
            cring[(15 - u) % PFB][q] = Ug[unsigned(e) < unsigned(cl.y) ? unsigned(cl.x + m * cl.y + e) : 0u];
        }
        // issued in the order of use: the loop's first wait then counts the
        // same way whether it is entered from here or from its own end
        __builtin_amdgcn_sched_barrier(0);
    }
    double2 xo = make_double2(0.0, 0.0);
    for (int bk = B - 1; bk >= 0; --bk) {
        int lo = l;
        asm volatile("" : "+v"(lo));
        // loaded here, consumed at the end of this block: the forward result of
        // block bk-NQ, which enters the window of blocks being updated
        const unsigned wy = W[(bk - NQ) * 16 + l];
        const double2 ynx = y[col_or0(wy)];
        S::template bwd_block<15>(r, xo, cring, Ug, ucol, m, lo, bk * 16);
        const unsigned w0 = W[bk * 16 + l];
        *out_ptr(w0) = rot_out(xo, code(w0));
#pragma unroll
        for (int q = 0; q < NQ - 1; ++q) r[q] = r[q + 1];
        r[NQ - 1] = fix_y(ynx, wy);
    }
}

template <typename T, int NQ, int PFB>
__global__ __launch_bounds__(64) void k_solve16(Solve16Args a) {
    solve16_body<T, NQ, PFB>(a, blockIdx.x);
}

// Solves of several launches in one: workgroup b belongs to the launch r with
// start[r] <= b < start[r + 1] (nrun launches, their arguments in args) and
// runs as that launch's workgroup b - start[r]
template <typename T, int NQ, int PFB>
__global__ __launch_bounds__(64) void k_solve16_batch(const Solve16Args *__restrict__ args,
                                                      const int *__restrict__ start, int nrun) {
    const int b = blockIdx.x;
    int r = 0;
    while (r + 1 < nrun && start[r + 1] <= b) ++r;
    const Solve16Args a = args[r];
    solve16_body<T, NQ, PFB>(a, b - start[r]);
}

template <typename T>
static bool launch_solve16(hipStream_t st, const Solve16Args &a) {
    if (a.kl > 16 || a.n < 32 || a.ncol16 < 16 * ((a.n + 15) / 16 + 1) || !a.ucol) return false;
    const int nq = 1 + (a.kv + 15) / 16;
    const dim3 grid((a.nsys * a.nblk + 3) / 4);
    const size_t lds = (size_t(a.nblk) * (a.n + 2 * SOLVE16_WPAD) + 1) / 2 * 2 * sizeof(unsigned) +
                       size_t(16 * ((a.n + 15) / 16)) * sizeof(int2);
    if (lds > 32768 || !a.sink) return false;
    switch (nq) {
    case 2: hipLaunchKernelGGL((k_solve16<T, 2, 8>), grid, dim3(64), lds, st, a); break;
    case 3: hipLaunchKernelGGL((k_solve16<T, 3, 8>), grid, dim3(64), lds, st, a); break;
    case 4: hipLaunchKernelGGL((k_solve16<T, 4, 8>), grid, dim3(64), lds, st, a); break;
    default: return false;
    }
    CHECK_LAUNCH();
    return true;
}

std::vector<unsigned> solve16_words(int total, int nblk, int blk_n, const int *blk_row, const int *blk_col,
                                    const unsigned char *rot) {
    const size_t stride = size_t(blk_n) + 2 * SOLVE16_WPAD;
    std::vector<unsigned> w(size_t(nblk) * stride, SOLVE16_NOWORD);
    for (int b = 0; b < nblk; ++b)
        for (int i = 0; i < blk_n; ++i) {
            const size_t e = size_t(b) * blk_n + i;
            const int r = blk_row ? blk_row[e] : (i < total ? i : -1);
            const int c = blk_col ? blk_col[e] : (i < total ? i : -1);
            unsigned v = unsigned(r < 0 ? SOLVE16_NONE : r) | (unsigned(c < 0 ? SOLVE16_NONE : c) << 15);
            if (rot && r >= 0) v |= unsigned(rot[r] & 1) << 30;
            if (rot && c >= 0) v |= unsigned((rot[c] >> 1) & 1) << 31;
            w[size_t(b) * stride + SOLVE16_WPAD + i] = v;
        }
    return w;
}

// record length of column j: one past the last nonzero entry over all matrices
template <typename T>
__global__ void k_ucol_len(const T *__restrict__ U, int n, int kv, int nmat, int *len) {
    const int nb = 16 * ((n + 15) >> 4);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nmat * nb) return;
    const int M = int(idx / nb), j = int(idx % nb);
    const T *rec = U + ((size_t)((M >> 2) * nb + j) * 4 + (M & 3)) * (kv + 1);
    int last = 0;
    for (int e = kv; e > 0; --e)
        if (!Fx<T>::is_zero(rec[e])) {
            last = e;
            break;
        }
    if (last > 0) atomicMax(len + j, last + 1);
}

template <typename T>
__global__ void k_u_trim(const T *__restrict__ U, int n, int kv, int nmat, const int2 *__restrict__ ucol,
                         long long ustride, T *__restrict__ Ut) {
    const int nb = 16 * ((n + 15) >> 4);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nmat * nb) return;
    const int M = int(idx / nb), j = int(idx % nb), g = M >> 2, m = M & 3;
    const T *rec = U + ((size_t)(g * nb + j) * 4 + m) * (kv + 1);
    const int2 cl = ucol[j];
    T *out = Ut + (size_t)g * ustride + cl.x + m * cl.y;
    for (int e = 0; e < cl.y; ++e) out[e] = rec[e];
}

void solve16_ucol(hipStream_t st, const void *Ufull, int elem, int n, int kv, int nmat, int *scratch_len,
                  std::vector<int2> &ucol, long long *ustride) {
    const int nb = 16 * ((n + 15) >> 4);
    const unsigned grid = unsigned(((long long)nmat * nb + 255) / 256);
    std::vector<int> len(nb, 1);   // the diagonal entry at least
    if (hipMemcpyAsync(scratch_len, len.data(), nb * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
        throw std::runtime_error("upload of U column lengths failed");
    if (elem == 8)
        hipLaunchKernelGGL(k_ucol_len<double>, dim3(grid), dim3(256), 0, st, (const double *)Ufull, n, kv, nmat,
                           scratch_len);
    else
        hipLaunchKernelGGL(k_ucol_len<double2>, dim3(grid), dim3(256), 0, st, (const double2 *)Ufull, n, kv, nmat,
                           scratch_len);
    CHECK_LAUNCH();
    if (hipMemcpyAsync(len.data(), scratch_len, nb * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        throw std::runtime_error("download of U column lengths failed");
    ucol.assign(nb, make_int2(0, 0));
    long long off = 16;   // [0, 16) of every group stays zero
    for (int j = 0; j < nb; ++j) {
        ucol[j] = make_int2(int(off), len[j]);
        off += 4LL * len[j];
    }
    *ustride = off;
}

void solve16_trim_u_copy(hipStream_t st, const void *Ufull, int elem, int n, int kv, int nmat, const int2 *ucol_dev,
                         long long ustride, void *Utrim) {
    const int nb = 16 * ((n + 15) >> 4);
    const unsigned grid = unsigned(((long long)nmat * nb + 255) / 256);
    if (elem == 8)
        hipLaunchKernelGGL(k_u_trim<double>, dim3(grid), dim3(256), 0, st, (const double *)Ufull, n, kv, nmat, ucol_dev,
                           ustride, (double *)Utrim);
    else
        hipLaunchKernelGGL(k_u_trim<double2>, dim3(grid), dim3(256), 0, st, (const double2 *)Ufull, n, kv, nmat,
                           ucol_dev, ustride, (double2 *)Utrim);
    CHECK_LAUNCH();
}

int solve16_ncol(int n, int kl, int kv) {
    // (positions are packed in 15 bits of the position words)
    if (kl > 16 || n < 32 || n >= SOLVE16_NONE || 1 + (kv + 15) / 16 > 4) return 0;
    return 16 * ((n + 15) / 16 + 1);
}

bool solve16_shape(const Solve16Args &a, int &nwg, size_t &lds, int &nq) {
    if (a.kl > 16 || a.n < 32 || a.ncol16 < 16 * ((a.n + 15) / 16 + 1) || !a.ucol || !a.sink) return false;
    nq = 1 + (a.kv + 15) / 16;
    if (nq < 2 || nq > 4) return false;
    nwg = (a.nsys * a.nblk + 3) / 4;
    lds = (size_t(a.nblk) * (a.n + 2 * SOLVE16_WPAD) + 1) / 2 * 2 * sizeof(unsigned) +
          size_t(16 * ((a.n + 15) / 16)) * sizeof(int2);
    return lds <= 32768;
}

bool launch_solve16_batch(hipStream_t st, bool real, int nq, int nwg, size_t lds, const Solve16Args *args,
                          const int *start, int nrun) {
    if (nwg <= 0 || nrun <= 0 || lds > 32768) return false;
    const dim3 grid{unsigned(nwg)};
#define SOLVE16_BATCH(T_)                                                                                         \
    switch (nq) {                                                                                                 \
    case 2: hipLaunchKernelGGL((k_solve16_batch<T_, 2, 8>), grid, dim3(64), lds, st, args, start, nrun); break;  \
    case 3: hipLaunchKernelGGL((k_solve16_batch<T_, 3, 8>), grid, dim3(64), lds, st, args, start, nrun); break;  \
    case 4: hipLaunchKernelGGL((k_solve16_batch<T_, 4, 8>), grid, dim3(64), lds, st, args, start, nrun); break;  \
    default: return false;                                                                                        \
    }
    if (real) {
        SOLVE16_BATCH(double)
    } else {
        SOLVE16_BATCH(double2)
    }
#undef SOLVE16_BATCH
    CHECK_LAUNCH();
    return true;
}

bool launch_solve16_real(hipStream_t st, const Solve16Args &a) { return launch_solve16<double>(st, a); }

bool launch_solve16_complex(hipStream_t st, const Solve16Args &a) {
    Solve16Args c = a;
    c.rotate = 0;
    return launch_solve16<double2>(st, c);
}

template <typename T>
static void launch_solve_t(hipStream_t st, const T *Lf, const T *Uf, const unsigned char *piv,
                           const unsigned char *rot, int ident0, int n, int kl, int kv, int nsys, const double2 *rhs,
                           double2 *x) {
    if (kl >= 1 && kl < 64 && kv < 64 && n >= 64) {
        hipLaunchKernelGGL((k_solve_reg<16, T>), dim3(nsys), dim3(64), 0, st, Lf, Uf, piv, rot, ident0, n, kl, kv,
                           rhs, x);
    } else {
        hipLaunchKernelGGL(k_solve<T>, dim3(nsys), dim3(64), size_t(n) * sizeof(double2), st, Lf, Uf, piv, rot,
                           ident0, n, kl, kv, rhs, x);
    }
    CHECK_LAUNCH();
}

void launch_solve(hipStream_t st, const double2 *Lf, const double2 *Uf, const unsigned char *piv, int n, int kl,
                  int kv, int nsys, const double2 *rhs, double2 *x) {
    launch_solve_t<double2>(st, Lf, Uf, piv, nullptr, 0, n, kl, kv, nsys, rhs, x);
}

void launch_solve_real(hipStream_t st, const double *Lr, const double *Ur, const unsigned char *piv,
                       const unsigned char *rot, int ident0, int n, int kl, int kv, int nsys, const double2 *rhs,
                       double2 *x) {
    launch_solve_t<double>(st, Lr, Ur, piv, rot, ident0, n, kl, kv, nsys, rhs, x);
}

}  // namespace coral
