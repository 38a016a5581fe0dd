// y passes of the horizontal transforms on wavefront-local FFTs.
//
// A workgroup owns YL consecutive kx columns of one (variable, z) plane, so
// every global access covers YL consecutive kx (YL complex = 128 B, 64 B at
// 768 points) or a whole kept-ky line; each wavefront transforms its own
// lines in LDS without workgroup barriers (YL/4 of them, one at 768 points
// and at 384 in FP64).  These are the y-direction parts of c2r_transform
// (backward c2c in y) and r2c_transform (forward c2c in y, truncated to the
// kept ky), slabs.fftw3mpi/transforms.f90:155-301.
#include "kernels.hpp"
#include "wavefft.hpp"

#include <cstdio>

namespace coral {

using wfft::StagesLSZ;
using wfft::swz;


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

// kx columns per workgroup for element type C: one 128-byte segment, 64
// bytes at 768 points, where a 128-byte tile of lines fills 98 KB of LDS and
// leaves one workgroup per CU; the 49 KB tile fits three (512^2 x 257: FP64
// yfwd 5.34 -> 4.61 ms, ybwd 2.30 -> 2.06 ms per launch)
constexpr int ycol_bytes(int N) { return N == 768 ? 64 : 128; }
template <int N, class C>
constexpr int ycols() { return int(ycol_bytes(N) / sizeof(C)); }

// wavefronts per workgroup.  At 768 points (and 384 in FP64) one line per
// wavefront rather than four wavefronts of two or four lines each: the
// tile's LDS bounds the workgroups per CU, so four wavefronts per workgroup
// left one to three per SIMD, each with two or four lines' butterflies in
// its registers (768 points, 128-byte tiles: FP64 yfwd 5.90 -> 5.34 ms, FP32
// 3.89 -> 3.00 ms per launch).  The FP32 384-point pass measured slower with
// 16 wavefronts and keeps four.
template <int N, class C>
constexpr int ywaves() { return N == 768 || (N == 384 && sizeof(C) == 16) ? ycols<N, C>() : 4; }

// G[b][z][y][ix] = sum_jy Fz[b][z][ix*nky + jy] exp(+i ky y)
template <int N, class C>
__global__ __launch_bounds__((64 * ywaves<N, C>())) void k_ybwd_w(const C *__restrict__ Fz, C *__restrict__ G,
                                                               const C *__restrict__ tw, int nkx, int nky, int NY) {
    constexpr int YL = ycols<N, C>(), YT = 64 * ywaves<N, C>();
    constexpr int LS = N + 1, LPW = YL / ywaves<N, C>();
    extern __shared__ unsigned char buf_raw[];   // YL * LS elements (beyond 64 KiB for N = 768)
    C *buf = reinterpret_cast<C *>(buf_raw);
    const int ntile = (nkx + YL - 1) / YL;
    const int tile = blockIdx.x % ntile;
    const size_t bz = blockIdx.x / ntile;
    const int ix0 = tile * YL, nl = min(YL, nkx - ix0);
    const int nsys = nkx * nky;
    // dealiased slots are zero
    for (int e = threadIdx.x; e < YL * (N - nky); e += blockDim.x) {
        const int l = e / (N - nky), k = e - l * (N - nky);
        buf[l * LS + NY / 2 + k] = mkc<C>(0, 0);
    }
    // all of a thread's loads are issued before the first LDS store (a loop
    // of load, wait, store would make them one HBM round trip each)
    const C *src = Fz + bz * nsys + (size_t)ix0 * nky;
    {
        constexpr int PER = (YL * N + YT - 1) / YT;   // nky <= N
        const int cnt = nl * nky;
        C v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            v[i] = e < cnt ? src[e] : mkc<C>(0, 0);
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
    C *dst = G + bz * N * nkx + ix0;
    {
        constexpr int PER = (N * YL + YT - 1) / YT;   // the last round partial when YT does not divide N*YL
        C v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int y = e / YL, l = e - y * YL;
            v[i] = (N * YL % YT == 0 || e < N * YL)
                       ? buf[l * LS + swz<YPlan<N>::template Run<+1, LPW, LS>::OUT_SWZ>(y)]
                       : mkc<C>(0, 0);
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
template <int N, class C>
__global__ __launch_bounds__((64 * ywaves<N, C>())) void k_yfwd_w(const C *__restrict__ H, C *__restrict__ Hy,
                                                               const C *__restrict__ tw, int nkx, int nky, int NY) {
    constexpr int YL = ycols<N, C>(), YT = 64 * ywaves<N, C>();
    constexpr int LS = N + 1, LPW = YL / ywaves<N, C>();
    extern __shared__ unsigned char buf_raw[];   // YL * LS elements (beyond 64 KiB for N = 768)
    C *buf = reinterpret_cast<C *>(buf_raw);
    const int ntile = (nkx + YL - 1) / YL;
    const int tile = blockIdx.x % ntile;
    const size_t bz = blockIdx.x / ntile;
    const int ix0 = tile * YL, nl = min(YL, nkx - ix0);
    const int nsys = nkx * nky;
    // all of a thread's loads are issued before the first LDS store
    const C *src = H + bz * N * nkx + ix0;
    {
        constexpr int PER = (N * YL + YT - 1) / YT;   // the last round partial when YT does not divide N*YL
        C v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int y = e / YL, l = e - y * YL;
            v[i] = ((N * YL % YT == 0 || e < N * YL) && l < nl) ? src[(size_t)y * nkx + l] : mkc<C>(0, 0);
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
    C *dst = Hy + bz * nsys + (size_t)ix0 * nky;
    {
        constexpr int PER = (YL * N + YT - 1) / YT;   // nky <= N
        const int cnt = nl * nky;
        C v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            const int l = e / nky, jy = e - l * nky;
            v[i] = e < cnt ? buf[l * LS + swz<YPlan<N>::template Run<-1, LPW, LS>::OUT_SWZ>(ky_slot(jy, NY, N))]
                           : mkc<C>(0, 0);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * YT;
            if (e < cnt) dst[e] = v[i];
        }
    }
}

// LDS of a workgroup of the wavefront y pass at length N
template <class C>
static size_t ywave_lds(int N) { return size_t(ycol_bytes(N) / sizeof(C)) * (N + 1) * sizeof(C); }

template <int N, class C>
static void go(hipStream_t st, bool fwd, const C *tw, const C *in, C *out, int nbatch, int NZAA, int nkx, int nky,
               int NY) {
    constexpr int YL = ycols<N, C>();
    const int ntile = (nkx + YL - 1) / YL;
    const dim3 grid(ntile * nbatch * NZAA);
    const size_t shm = ywave_lds<C>(N);
    if (fwd) {
        if (shm > 65536) set_dynamic_lds((const void *)k_yfwd_w<N, C>, shm, "y pass (forward)");
        hipLaunchKernelGGL((k_yfwd_w<N, C>), grid, dim3(64 * ywaves<N, C>()), shm, st, in, out, tw, nkx, nky, NY);
    } else {
        if (shm > 65536) set_dynamic_lds((const void *)k_ybwd_w<N, C>, shm, "y pass (inverse)");
        hipLaunchKernelGGL((k_ybwd_w<N, C>), grid, dim3(64 * ywaves<N, C>()), shm, st, in, out, tw, nkx, nky, NY);
    }
    check_launch("y pass");
}

template <class C>
static bool ypass_wave(hipStream_t st, bool fwd, int N, const C *tw, const C *in, C *out, int nbatch, int NZAA,
                       int nkx, int nky, int NY) {
    if (!tw || nky > N || NY / 2 + (N - nky) > N) return false;
    // the tile of a workgroup must fit the device's LDS (98 KB at N = 768);
    // otherwise the workgroup-level LDS FFT takes the lines
    if (ywave_lds<C>(N) > lds_limit_current()) return false;
    switch (N) {
    case 96: go<96>(st, fwd, tw, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 144: go<144>(st, fwd, tw, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 192: go<192>(st, fwd, tw, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 384: go<384>(st, fwd, tw, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    case 768: go<768>(st, fwd, tw, in, out, nbatch, NZAA, nkx, nky, NY); return true;
    default: return false;
    }
}

bool launch_ypass_wave(hipStream_t st, bool fwd, const FFTPlan &py, const double2 *in, double2 *out, int nbatch,
                       int NZAA, int nkx, int nky, int NY) {
    return ypass_wave<double2>(st, fwd, py.N, py.tw, in, out, nbatch, NZAA, nkx, nky, NY);
}

void launch_ypass32(hipStream_t st, bool fwd, const FFTPlan &py, const float2 *in, float2 *out, int nbatch, int NZAA,
                    int nkx, int nky, int NY) {
    if (!ypass_wave<float2>(st, fwd, py.N, py.twf, in, out, nbatch, NZAA, nkx, nky, NY))
        throw std::runtime_error("the FP32 y pass needs a wavefront plan (N = 96, 144, 192, 384, 768) that fits LDS");
}

}  // namespace coral
