// Wavefront-local Stockham FFTs for the horizontal transforms.
//
// One wavefront (64 lanes) owns S complex sequences of length N held in its
// own slice of LDS; stages exchange data through that slice with wavefront
// ordering only (no workgroup barrier).  A group of LN = 128, 192, ... lanes
// (consecutive wavefronts of a workgroup) can share the sequences instead:
// its stages are ordered by the workgroup barrier, so every wavefront of the
// workgroup runs the same stages.  Lengths, radices and stage strides
// are template constants, so every index is resolved at compile time.
//
// Stage with radix R after Ns points have been combined (Ns = product of the
// earlier radices), T = N / R butterflies per sequence:
//   butterfly j reads   j + r*T                           (r = 0..R-1)
//   twiddle             exp(sign * 2 pi i * (j % Ns) * r / (Ns * R))
//   writes              (j / Ns) * Ns * R + j % Ns + r * Ns
// The last stage (Ns * R = N) writes j + r*T: the element set a butterfly
// owns at the end of one transform is the set a first-stage butterfly of the
// same radix reads, which lets the x pass form products in registers between
// the inverse and the forward transform.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

#include <type_traits>

namespace coral {
namespace wfft {

// LDS traffic of one wavefront is ordered; the fences keep the compiler from
// moving LDS accesses across a stage boundary.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// orders the LDS traffic of a group of LN lanes between stages
template <int LN>
__device__ inline void group_sync() {
    static_assert(LN % 64 == 0, "lane groups are whole wavefronts");
    if constexpr (LN > 64)
        __syncthreads();
    else
        wave_sync();
}

// z * (SIGN i)
template <int SIGN, class C>
__device__ inline C mul_si(C z) {
    return SIGN > 0 ? mkc<C>(-z.y, z.x) : mkc<C>(z.y, -z.x);
}

// R-point DFT in registers: X[k] = sum_n x[n] exp(SIGN 2 pi i n k / R), on
// double2 or float2 elements (constants in the element's precision)
template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<2, SIGN> {
    template <class C>
    __device__ static inline void run(C *v) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int SIGN>
struct Dft<3, SIGN> {
    template <class C>
    __device__ static inline void run(C *v) {
        using T = typename CxT<C>::R;
        const T s1 = T(SIGN * 0.86602540378443864676), h = T(0.5);  // sin(2 pi / 3)
        const C a = v[0], t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const C m = mkc<C>(a.x - h * t.x, a.y - h * t.y);
        v[0] = cadd(a, t);
        v[1] = mkc<C>(m.x - s1 * d.y, m.y + s1 * d.x);
        v[2] = mkc<C>(m.x + s1 * d.y, m.y - s1 * d.x);
    }
};

template <int SIGN>
struct Dft<4, SIGN> {
    template <class C>
    __device__ static inline void run(C *v) {
        const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const C t2 = cadd(v[1], v[3]), t3 = mul_si<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

// 6 = 2 x 3: X[k] = E[k%3] + w^k O[k%3], X[k+3] = E[k] - w^k O[k], w = exp(SIGN pi i / 3)
template <int SIGN>
struct Dft<6, SIGN> {
    template <class C>
    __device__ static inline void run(C *v) {
        using T = typename CxT<C>::R;
        C e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
        Dft<3, SIGN>::run(e);
        Dft<3, SIGN>::run(o);
        const T s = T(SIGN * 0.86602540378443864676), h = T(0.5);
        const C o1 = mkc<C>(h * o[1].x - s * o[1].y, h * o[1].y + s * o[1].x);
        const C o2 = mkc<C>(-h * o[2].x - s * o[2].y, -h * o[2].y + s * o[2].x);
        v[0] = cadd(e[0], o[0]);
        v[3] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[4] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[5] = csub(e[2], o2);
    }
};

// 8 = 2 x 4: X[k] = E[k] + w^k O[k], X[k+4] = E[k] - w^k O[k], w = exp(SIGN pi i / 4)
template <int SIGN>
struct Dft<8, SIGN> {
    template <class C>
    __device__ static inline void run(C *v) {
        using T = typename CxT<C>::R;
        C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, SIGN>::run(e);
        Dft<4, SIGN>::run(o);
        const T h = T(0.70710678118654752440), sg = T(SIGN);
        const C o1 = mkc<C>(h * (o[1].x - sg * o[1].y), h * (o[1].y + sg * o[1].x));
        const C o2 = mul_si<SIGN>(o[2]);
        const C o3 = mkc<C>(-h * (o[3].x + sg * o[3].y), h * (sg * o[3].x - o[3].y));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// exp(SIGN 2 pi i m / N) from the table tw[k] = exp(-2 pi i k / N)
template <int SIGN, class C>
__device__ inline C twiddle(const C *__restrict__ tw, int m) {
    const C w = tw[m];
    return SIGN > 0 ? mkc<C>(w.x, -w.y) : w;
}

// p[r] = exp(SIGN 2 pi i r m / N) for r = 1..R-1 (r m < N).  FP64: one table
// load, the other powers by a chain of products (each adds an ulp or so,
// invisible at FP64); loads of more powers were measured slower, since the
// twiddle loads sit at the head of every stage.  FP32: every power from the
// table, each rounded once, since the chain's ulps would double the FFT's
// rounding error at that precision.
template <int R, int SIGN, class C>
__device__ inline void twiddle_powers(const C *__restrict__ tw, int m, C (&p)[R]) {
    if constexpr (std::is_same<C, float2>::value) {
#pragma unroll
        for (int r = 1; r < R; ++r) p[r] = twiddle<SIGN>(tw, r * m);
    } else {
        p[1] = twiddle<SIGN>(tw, m);
#pragma unroll
        for (int r = 2; r < R; ++r) p[r] = cmul(p[r - 1], p[1]);
    }
}

// LDS bank swizzles (used by runs that opt in with SW).  A wavefront's
// 16-byte accesses are served eight lanes
// at a time, conflict-free when the eight elements fall in distinct 16-byte
// slots of a 128-byte row (element index mod 8).  Reads of a stage are runs
// of eight consecutive elements (conflict-free under any permutation inside
// rows of eight); the writes of the early stages are strided.  Each stage
// therefore stores element e of a sequence at swz<W>(e), W chosen for its
// write pattern, and the next stage reads through the same map:
//   W = 1: stride-4 writes (radix 4, NS = 1):  slot ^= row & 3
//   W = 2: radix 4 after NS = 4:               slot ^= 4 * ((row >> 1) & 1)
//   W = 3: stride-6 writes (radix 6, NS = 1):  slot ^= row & 1
//   W = 4: stride-8 writes (radix 8, NS = 1):  slot ^= row & 7
// (row = e / 8); other stages write whole rows and keep W = 0.
template <int W>
__device__ inline int swz(int e) {
    if constexpr (W == 1) return e ^ ((e >> 3) & 3);
    else if constexpr (W == 2) return e ^ (((e >> 4) & 1) << 2);
    else if constexpr (W == 3) return e ^ ((e >> 3) & 1);
    else if constexpr (W == 4) return e ^ ((e >> 3) & 7);
    else return e;
}
// swizzle a stage with radix R after NS points writes with
template <int R, int NS>
constexpr int stage_swz() { return (NS == 1 && R == 4) ? 1 : (NS == 4 && R == 4) ? 2 : (NS == 1 && R == 6) ? 3 : 0; }

// One LDS -> LDS stage over S sequences (sequence s at buf + s*LS); the
// S*T butterflies are spread over the LN lanes of the group round by round.  The input was
// written with swizzle WI, the output is written with stage_swz<R, NS>() when
// SW is set, unswizzled otherwise.
template <int N, int R, int NS, int SIGN, int S, int LS, int WI = 0, bool SW = false, int LN = 64, class C>
__device__ inline void stage_lds(C *buf, const C *__restrict__ tw, int lane) {
    constexpr int WO = SW ? stage_swz<R, NS>() : 0;
    constexpr int T = N / R;
    constexpr int TOT = S * T;
    constexpr int ROUNDS = (TOT + LN - 1) / LN;
    static_assert(N % R == 0, "radix must divide the length");
    C v[ROUNDS][R];
#pragma unroll
    for (int c = 0; c < ROUNDS; ++c) {
        const int b = lane + LN * c;
        if (TOT % LN == 0 || b < TOT) {
            const int s = b / T, j = b - s * T;
            const C *L = buf + s * LS;
            const int jm = j % NS;
            C wp[R];
            if constexpr (NS > 1) twiddle_powers<R, SIGN>(tw, jm * (N / (NS * R)), wp);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                C x = L[swz<WI>(j + r * T)];
                if (NS > 1 && r > 0) x = cmul(x, wp[r]);
                v[c][r] = x;
            }
            Dft<R, SIGN>::run(v[c]);
        }
    }
    group_sync<LN>();
#pragma unroll
    for (int c = 0; c < ROUNDS; ++c) {
        const int b = lane + LN * c;
        if (TOT % LN == 0 || b < TOT) {
            const int s = b / T, j = b - s * T;
            C *L = buf + s * LS;
            const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) L[swz<WO>(base + r * NS)] = v[c][r];
        }
    }
    group_sync<LN>();
}

// A run of LDS stages with radices Rs..., starting after NS points on input
// written with swizzle WI; sequences LS elements apart; SW: the stages write
// swizzled; LN lanes share the work.  OUT_SWZ is the swizzle of the run's
// output (what its reader must apply).
template <int LN, int N, int SIGN, int S, int LS, int WI, bool SW, int NS, int... Rs>
struct GStagesLSW;

template <int LN, int N, int SIGN, int S, int LS, int WI, bool SW, int NS>
struct GStagesLSW<LN, N, SIGN, S, LS, WI, SW, NS> {
    static constexpr int OUT_SWZ = WI;
    template <class C>
    __device__ static inline void run(C *, const C *__restrict__, int) {}
};

template <int LN, int N, int SIGN, int S, int LS, int WI, bool SW, int NS, int R, int... Rs>
struct GStagesLSW<LN, N, SIGN, S, LS, WI, SW, NS, R, Rs...> {
    using Next = GStagesLSW<LN, N, SIGN, S, LS, (SW ? stage_swz<R, NS>() : 0), SW, NS * R, Rs...>;
    static constexpr int OUT_SWZ = Next::OUT_SWZ;
    template <class C>
    __device__ static inline void run(C *buf, const C *__restrict__ tw, int lane) {
        stage_lds<N, R, NS, SIGN, S, LS, WI, SW, LN>(buf, tw, lane);
        Next::run(buf, tw, lane);
    }
};

// one wavefront per group of sequences
template <int N, int SIGN, int S, int LS, int WI, bool SW, int NS, int... Rs>
using StagesLSW = GStagesLSW<64, N, SIGN, S, LS, WI, SW, NS, Rs...>;
// contiguous sequences shared by LN lanes, unswizzled
template <int LN, int N, int SIGN, int S, int NS, int... Rs>
using GStages = GStagesLSW<LN, N, SIGN, S, N, 0, false, NS, Rs...>;

// runs on unswizzled input, without (StagesLS) or with (StagesLSZ) swizzled writes
template <int N, int SIGN, int S, int LS, int NS, int... Rs>
using StagesLS = StagesLSW<N, SIGN, S, LS, 0, false, NS, Rs...>;
template <int N, int SIGN, int S, int LS, int NS, int... Rs>
using StagesLSZ = StagesLSW<N, SIGN, S, LS, 0, true, NS, Rs...>;
// contiguous sequences (LS = N)
template <int N, int SIGN, int S, int NS, int... Rs>
using Stages = StagesLS<N, SIGN, S, N, NS, Rs...>;

}  // namespace wfft
}  // namespace coral
