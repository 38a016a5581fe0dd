// y passes of the horizontal transforms on wavefront-local FFTs.
//
// A workgroup owns YL consecutive kx columns of one (variable, z) plane, so
// every global access covers YL consecutive kx (YL complex = 128 B) or a whole
// kept-ky line; each wavefront transforms its own YL/4 lines in LDS without
// workgroup barriers.  These are the y-direction parts of c2r_transform
// (backward c2c in y) and r2c_transform (forward c2c in y, truncated to the
// kept ky), slabs.fftw3mpi/transforms.f90:155-301.
#include "kernels.hpp"
#include "wavefft.hpp"

#include <cstdio>

namespace coral {

using wfft::StagesLSZ;
using wfft::swz;

constexpr int YL = 8;        // kx columns per workgroup
constexpr int YWAVES = 4;    // wavefronts per workgroup (YL / YWAVES lines each)
constexpr int YT = 64 * YWAVES;   // threads per workgroup

template <int N>
struct YPlan;
template <>
struct YPlan<96> {
    template <int SIGN, int S, int LS> using Run = StagesLSZ<96, SIGN, S, LS, 1, 4, 4, 6>;
};
template <>
struct YPlan<144> {
    template <int SIGN, int S, int LS> using Run = StagesLSZ<144, SIGN, S, LS, 1, 4, 4, 3, 3>;
};
template <>
struct YPlan<192> {
    template <int SIGN, int S, int LS> using Run = StagesLSZ<192, SIGN, S, LS, 1, 4, 4, 4, 3>;
};
template <>
struct YPlan<384> {
    template <int SIGN, int S, int LS> using Run = StagesLSZ<384, SIGN, S, LS, 1, 4, 4, 4, 6>;
};
template <>
struct YPlan<768> {
    template <int SIGN, int S, int LS> using Run = StagesLSZ<768, SIGN, S, LS, 1, 8, 4, 4, 6>;
};

// kept ky index jy -> slot of the NYAA-long line
__device__ inline int ky_slot(int jy, int NY, int N) { return jy < NY / 2 ? jy : N - (NY - 1) + jy; }

// G[b][z][y][ix] = sum_jy Fz[b][z][ix*nky + jy] exp(+i ky y)
template <int N>
__global__ __launch_bounds__(64 * YWAVES) void k_ybwd_w(const double2 *__restrict__ Fz, double2 *__restrict__ G,
                                                       const double2 *__restrict__ tw, int nkx, int nky, int NY) {
    constexpr int LS = N + 1, LPW = YL / YWAVES;
    extern __shared__ double2 buf[];   // YL * LS (beyond 64 KiB for N = 768)
    const int ntile = (nkx + YL - 1) / YL;
    const int tile = blockIdx.x % ntile;
    const size_t bz = blockIdx.x / ntile;
    const int ix0 = tile * YL, nl = min(YL, nkx - ix0);
    const int nsys = nkx * nky;
    // dealiased slots are zero
    for (int e = threadIdx.x; e < YL * (N - nky); e += blockDim.x) {
        const int l = e / (N - nky), k = e - l * (N - nky);
        buf[l * LS + NY / 2 + k] = make_double2(0.0, 0.0);
    }
    // all of a thread's loads are issued before the first LDS store (a loop
    // of load, wait, store would make them one HBM round trip each)
    const double2 *src = Fz + bz * nsys + (size_t)ix0 * nky;
    {
        constexpr int PER = (YL * N + YT - 1) / YT;   // nky <= N
        const int cnt = nl * nky;
        double2 v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            v[i] = e < cnt ? src[e] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            if (e < YL * nky) {
                const int l = e / nky, jy = e - l * nky;
                buf[l * LS + ky_slot(jy, NY, N)] = v[i];
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    YPlan<N>::template Run<+1, LPW, LS>::run(buf + w * LPW * LS, tw, lane);
    __syncthreads();
    double2 *dst = G + bz * N * nkx + ix0;
    {
        constexpr int PER = (N * YL + YT - 1) / YT;   // the last round partial when YT does not divide N*YL
        double2 v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int y = e / YL, l = e - y * YL;
            v[i] = (N * YL % YT == 0 || e < N * YL)
                       ? buf[l * LS + swz<YPlan<N>::template Run<+1, LPW, LS>::OUT_SWZ>(y)]
                       : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int y = e / YL, l = e - y * YL;
            if ((N * YL % YT == 0 || e < N * YL) && l < nl) dst[(size_t)y * nkx + l] = v[i];
        }
    }
}

// Hy[b][z][ix*nky + jy] = sum_y H[b][z][y][ix] exp(-i ky y)
template <int N>
__global__ __launch_bounds__(64 * YWAVES) void k_yfwd_w(const double2 *__restrict__ H, double2 *__restrict__ Hy,
                                                       const double2 *__restrict__ tw, int nkx, int nky, int NY) {
    constexpr int LS = N + 1, LPW = YL / YWAVES;
    extern __shared__ double2 buf[];   // YL * LS (beyond 64 KiB for N = 768)
    const int ntile = (nkx + YL - 1) / YL;
    const int tile = blockIdx.x % ntile;
    const size_t bz = blockIdx.x / ntile;
    const int ix0 = tile * YL, nl = min(YL, nkx - ix0);
    const int nsys = nkx * nky;
    // all of a thread's loads are issued before the first LDS store
    const double2 *src = H + bz * N * nkx + ix0;
    {
        constexpr int PER = (N * YL + YT - 1) / YT;   // the last round partial when YT does not divide N*YL
        double2 v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int y = e / YL, l = e - y * YL;
            v[i] = ((N * YL % YT == 0 || e < N * YL) && l < nl) ? src[(size_t)y * nkx + l] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int y = e / YL, l = e - y * YL;
            if (N * YL % YT == 0 || e < N * YL) buf[l * LS + y] = v[i];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    YPlan<N>::template Run<-1, LPW, LS>::run(buf + w * LPW * LS, tw, lane);
    __syncthreads();
    double2 *dst = Hy + bz * nsys + (size_t)ix0 * nky;
    {
        constexpr int PER = (YL * N + YT - 1) / YT;   // nky <= N
        const int cnt = nl * nky;
        double2 v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int l = e / nky, jy = e - l * nky;
            v[i] = e < cnt ? buf[l * LS + swz<YPlan<N>::template Run<-1, LPW, LS>::OUT_SWZ>(ky_slot(jy, NY, N))]
                           : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            if (e < cnt) dst[e] = v[i];
        }
    }
}

// LDS of a workgroup of the wavefront y pass at length N
static size_t ywave_lds(int N) { return size_t(YL) * (N + 1) * sizeof(double2); }

template <int N>
static void go(hipStream_t st, bool fwd, const FFTPlan &py, const double2 *in, double2 *out, int nbatch, int NZAA,
               int nkx, int nky, int NY) {
    const int ntile = (nkx + YL - 1) / YL;
    const dim3 grid(ntile * nbatch * NZAA);
    const size_t shm = ywave_lds(N);
    if (fwd) {
        if (shm > 65536) set_dynamic_lds((const void *)k_yfwd_w<N>, shm, "y pass (forward)");
        hipLaunchKernelGGL(k_yfwd_w<N>, grid, dim3(64 * YWAVES), shm, st, in, out, py.tw, nkx, nky, NY);
    } else {
        if (shm > 65536) set_dynamic_lds((const void *)k_ybwd_w<N>, shm, "y pass (inverse)");
        hipLaunchKernelGGL(k_ybwd_w<N>, grid, dim3(64 * YWAVES), shm, st, in, out, py.tw, nkx, nky, NY);
    }
    check_launch("y pass");
}

bool launch_ypass_wave(hipStream_t st, bool fwd, const FFTPlan &py, const double2 *in, double2 *out, int nbatch,
                       int NZAA, int nkx, int nky, int NY) {
    if (nky > py.N || NY / 2 + (py.N - nky) > py.N) return false;
    // the tile of a workgroup must fit the device's LDS (98 KB at N = 768);
    // otherwise the workgroup-level LDS FFT takes the lines
    if (ywave_lds(py.N) > lds_limit_current()) return false;
    switch (py.N) {
    case 96: go<96>(st, fwd, py, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 144: go<144>(st, fwd, py, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 192: go<192>(st, fwd, py, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 384: go<384>(st, fwd, py, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 768: go<768>(st, fwd, py, in, out, nbatch, NZAA, nkx, nky, NY); return true;   // 98 KB of LDS
    default: return false;
    }
}

}  // namespace coral
