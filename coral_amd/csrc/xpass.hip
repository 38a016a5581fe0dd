// x pass of the nonlinear step on one wavefront (two for 768-point lines)
// per (z, y) line.
//
// For each line: inverse real FFTs of the full variables (two real lines are
// packed into one complex sequence, f0 + i f1), the quadratic products and
// the CFL maxima in registers, then forward real FFTs of the products (again
// two per complex sequence, separated with Hermitian symmetry) truncated to
// the kept kx.  This is the x-direction part of c2r_transform / r2c_transform
// (slabs.fftw3mpi/transforms.f90:155-301) fused with the products of
// PL_IMEX_timestepping_marching.f90:279-300 and compute_cfl_based_timestep
// (PL_IMEX_timestepping_marching.f90:200-254).
#include "kernels.hpp"
#include "wavefft.hpp"

#include <cstdio>

namespace coral {

using wfft::Dft;
using wfft::group_sync;
using wfft::Stages;

// Radix plans.  Inverse: first stage R0 reads the packed spectrum from HBM,
// InvMid runs in LDS, the last stage RL leaves each lane holding the points
// j + r*N/RL of its lines.  Forward: first stage RL starts from those
// registers, FwdRest finishes in LDS in natural order.  LN lanes (one or
// more wavefronts) share LPW lines, LPW * N / RL a multiple of LN.  W0 / WL:
// the swizzles (wfft::swz) of the first inverse and first forward stages' LDS
// writes, which InvMid / FwdRest read through.
template <int N>
struct XPlan;
template <>
struct XPlan<96> {
    static constexpr int R0 = 4, RL = 6, LPW = 4, W0 = 0, WL = 0, LN = 64;
    template <int S> using InvMid = Stages<96, +1, S, 4, 4>;
    template <int S> using FwdRest = Stages<96, -1, S, 6, 4, 4>;
};
template <>
struct XPlan<144> {
    static constexpr int R0 = 4, RL = 3, LPW = 4, W0 = 0, WL = 0, LN = 64;
    template <int S> using InvMid = Stages<144, +1, S, 4, 4, 3>;
    template <int S> using FwdRest = Stages<144, -1, S, 3, 3, 4, 4>;
};
template <>
struct XPlan<192> {
    static constexpr int R0 = 4, RL = 3, LPW = 1, W0 = 0, WL = 0, LN = 64;
    template <int S> using InvMid = Stages<192, +1, S, 4, 4, 4>;
    template <int S> using FwdRest = Stages<192, -1, S, 3, 4, 4, 4>;
};
template <>
struct XPlan<384> {
    // the stride-8 writes of the first inverse stage put all 8 lanes of a
    // row of eight on the same banks: swizzled (0.613 -> 0.599 ms per
    // launch); the 2-way conflicts of the first forward stage's stride-6
    // writes cost less than their swizzle (0.624 ms with it)
    static constexpr int R0 = 8, RL = 6, LPW = 1, W0 = 4, WL = 0, LN = 64;
    template <int S> using InvMid = wfft::StagesLSW<384, +1, S, 384, W0, false, 8, 8>;
    template <int S> using FwdRest = wfft::StagesLSW<384, -1, S, 384, WL, false, 6, 8, 8>;
};
template <>
struct XPlan<768> {
    // two wavefronts per line: one lane holds 6 points of each full variable
    // pair (96 VGPRs of them per lane in FP64 on one wavefront, which left
    // two wavefronts per SIMD)
    static constexpr int R0 = 8, RL = 6, LPW = 1, W0 = 0, WL = 0, LN = 128;
    template <int S> using InvMid = wfft::GStages<LN, 768, +1, S, 8, 4, 4>;
    template <int S> using FwdRest = wfft::GStages<LN, 768, -1, S, 6, 4, 4, 8>;
};

// XGP: product pairs per forward batch
template <int N, int NPI, int XGP>
struct XShape {
    using P = XPlan<N>;
    static constexpr int LPW = P::LPW, R0 = P::R0, RL = P::RL, LN = P::LN;
    static constexpr int T0 = N / R0, TL = N / RL;
    static constexpr int NT = LPW * TL / LN;  // (line, j) tasks per lane at the register stages
    static constexpr int SIN = LPW;           // inverse sequences in LDS at once (one pair per line)
    static constexpr int SF = LPW * XGP;      // forward sequences per batch
    static constexpr int SMAX = SIN > SF ? SIN : SF;
    static_assert((LPW * TL) % LN == 0, "register stage must fill the lane group");
};

// value of full variable A (1 for A < 0) at point r of a task
template <int A, int NPI, int RL, class C>
__device__ inline typename CxT<C>::R fget(const C (&f)[NPI][RL], int r) {
    if constexpr (A < 0)
        return 1;
    else
        return (A & 1) ? f[A >> 1][r].y : f[A >> 1][r].x;
}

// o[r] = F_a at the RL points of a task.  a is the same for the whole
// wavefront: a scalar branch picks the copy (the empty asm keeps the branches
// from being turned into per-element selects, one v_cndmask per candidate).
template <int NPI, int RL, int K = 0, class C>
__device__ inline void pick(const C (&f)[NPI][RL], int a, typename CxT<C>::R (&o)[RL]) {
    if constexpr (K < 2 * NPI) {
        if (a == K) {
            asm volatile("");
#pragma unroll
            for (int r = 0; r < RL; ++r) o[r] = fget<K, NPI, RL>(f, r);
            return;
        }
        pick<NPI, RL, K + 1>(f, a, o);
    } else {
#pragma unroll
        for (int r = 0; r < RL; ++r) o[r] = 1;   // a < 0: the constant ONE
    }
}

template <int NPI, int RL, class C>
__device__ inline void product(const C (&f)[NPI][RL], int a, int b, typename CxT<C>::R (&p)[RL]) {
    typename CxT<C>::R fa[RL], fb[RL];
    pick<NPI, RL>(f, a, fa);
    pick<NPI, RL>(f, b, fb);
#pragma unroll
    for (int r = 0; r < RL; ++r) p[r] = fa[r] * fb[r];
}

// Per-line balancing of packed pairs (the FP32 chain).  Two real lines packed
// into one complex FFT share its rounding: each comes out with an error of
// about eps times the larger of the two, so a line much smaller than its
// partner (w beside t, a temperature product beside a velocity product)
// would lose most of its digits.  In FP32 each line of a pair is therefore
// scaled by the power of two that brings its largest component into
// [0.5, 1) before the FFT, and scaled back after it; both scalings are
// exact.  The maxima are reduced per line through LDS (bit patterns of
// non-negative floats order as unsigned integers).
template <class C>
struct Balanced {
    static constexpr bool value = false;
};
template <>
struct Balanced<float2> {
    static constexpr bool value = true;
};
constexpr int XBAL_SLOTS = 16;   // per lane group: 2 lines x (lines per group x pairs per batch)

__device__ inline unsigned absmax_bits(float2 v) { return __float_as_uint(fmaxf(fabsf(v.x), fabsf(v.y))); }
__device__ inline unsigned absmax_bits(float v) { return __float_as_uint(fabsf(v)); }

// power of two bringing a positive maximum into [0.5, 1) (1 for zero), and
// its inverse; exponents are clamped to the normal range
__device__ inline float bal_scale(unsigned mbits) {
    const float m = __uint_as_float(mbits);
    if (!(m > 0.f)) return 1.f;
    int e;
    frexpf(m, &e);
    e = e < -125 ? -125 : (e > 125 ? 125 : e);
    return ldexpf(1.f, -e);
}
__device__ inline float bal_unscale(unsigned mbits) {
    const float m = __uint_as_float(mbits);
    if (!(m > 0.f)) return 1.f;
    int e;
    frexpf(m, &e);
    e = e < -125 ? -125 : (e > 125 ? 125 : e);
    return ldexpf(1.f, e);
}

// forward transforms of product pairs [pq0, pq0 + G) and the store of the kept kx
template <int N, int NPI, int XGP, int G, class C>
__device__ inline void xfwd_batch(const XPassArgs &a, const C (&fv)[XShape<N, NPI, XGP>::NT][NPI][XPlan<N>::RL],
                                  C *buf, const C *__restrict__ tw, int lane, long long line0, long long nlines,
                                  int pq0, unsigned *xmax) {
    using X = XShape<N, NPI, XGP>;
    using T = typename CxT<C>::R;
    C *H = reinterpret_cast<C *>(a.H);
    constexpr int RL = X::RL, TL = X::TL, LPW = X::LPW, LN = X::LN;
    if constexpr (Balanced<C>::value) {
        static_assert(2 * LPW * G <= XBAL_SLOTS, "balancing slots");
        // products of every task first, their per-line maxima, then the
        // scaled FFTs (slot 2 (l G + gp) + {0, 1}: the pair's two lines)
        T p0[X::NT][G][RL], p1[X::NT][G][RL];
        if (lane < 2 * LPW * G) xmax[lane] = 0u;
        group_sync<LN>();
#pragma unroll
        for (int c = 0; c < X::NT; ++c) {
            const int t = lane + LN * c, l = t / TL;
#pragma unroll
            for (int gp = 0; gp < G; ++gp) {
                const int q0 = 2 * (pq0 + gp), q1 = q0 + 1;
                const bool has1 = q1 < a.nquad;
                const int a0 = __builtin_amdgcn_readfirstlane(a.qa[q0]), b0 = __builtin_amdgcn_readfirstlane(a.qb[q0]);
                const int a1 = has1 ? __builtin_amdgcn_readfirstlane(a.qa[q1]) : -1;
                const int b1 = has1 ? __builtin_amdgcn_readfirstlane(a.qb[q1]) : -1;
                product<NPI, RL>(fv[c], a0, b0, p0[c][gp]);
                if (has1)
                    product<NPI, RL>(fv[c], a1, b1, p1[c][gp]);
                else
#pragma unroll
                    for (int r = 0; r < RL; ++r) p1[c][gp][r] = 0;
                unsigned m0 = 0u, m1 = 0u;
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    m0 = max(m0, absmax_bits(p0[c][gp][r]));
                    m1 = max(m1, absmax_bits(p1[c][gp][r]));
                }
                atomicMax(&xmax[2 * (l * G + gp)], m0);
                atomicMax(&xmax[2 * (l * G + gp) + 1], m1);
            }
        }
        group_sync<LN>();
#pragma unroll
        for (int c = 0; c < X::NT; ++c) {
            const int t = lane + LN * c, l = t / TL, j = t - l * TL;
#pragma unroll
            for (int gp = 0; gp < G; ++gp) {
                const T s0 = bal_scale(xmax[2 * (l * G + gp)]), s1 = bal_scale(xmax[2 * (l * G + gp) + 1]);
                C y[RL];
#pragma unroll
                for (int r = 0; r < RL; ++r) y[r] = mkc<C>(p0[c][gp][r] * s0, p1[c][gp][r] * s1);
                Dft<RL, -1>::run(y);
                C *L = buf + (l * G + gp) * N;
#pragma unroll
                for (int r = 0; r < RL; ++r) L[wfft::swz<XPlan<N>::WL>(j * RL + r)] = y[r];
            }
        }
    } else {
#pragma unroll
    for (int c = 0; c < X::NT; ++c) {
        const int t = lane + LN * c, l = t / TL, j = t - l * TL;
#pragma unroll
        for (int gp = 0; gp < G; ++gp) {
            const int q0 = 2 * (pq0 + gp), q1 = q0 + 1;
            const bool has1 = q1 < a.nquad;
            const int a0 = __builtin_amdgcn_readfirstlane(a.qa[q0]), b0 = __builtin_amdgcn_readfirstlane(a.qb[q0]);
            const int a1 = has1 ? __builtin_amdgcn_readfirstlane(a.qa[q1]) : -1;
            const int b1 = has1 ? __builtin_amdgcn_readfirstlane(a.qb[q1]) : -1;
            T p0[RL], p1[RL];
            product<NPI, RL>(fv[c], a0, b0, p0);
            if (has1)
                product<NPI, RL>(fv[c], a1, b1, p1);
            else
#pragma unroll
                for (int r = 0; r < RL; ++r) p1[r] = 0;
            C y[RL];
#pragma unroll
            for (int r = 0; r < RL; ++r) y[r] = mkc<C>(p0[r], p1[r]);
            Dft<RL, -1>::run(y);
            C *L = buf + (l * G + gp) * N;
#pragma unroll
            for (int r = 0; r < RL; ++r) L[wfft::swz<XPlan<N>::WL>(j * RL + r)] = y[r];
        }
    }
    }
    group_sync<LN>();
    XPlan<N>::template FwdRest<LPW * G>::run(buf, tw, lane);
    const int nkx = a.nkx;
#pragma unroll
    for (int l = 0; l < LPW; ++l) {
        const long long line = line0 + l;
        if (line < nlines) {
#pragma unroll
            for (int gp = 0; gp < G; ++gp) {
                const int q0 = 2 * (pq0 + gp), q1 = q0 + 1;
                const C *L = buf + (l * G + gp) * N;
                C *h0 = H + ((size_t)q0 * nlines + line) * nkx;
                C *h1 = H + ((size_t)q1 * nlines + line) * nkx;
                T hf0 = T(0.5), hf1 = T(0.5);
                if constexpr (Balanced<C>::value) {
                    hf0 *= bal_unscale(xmax[2 * (l * G + gp)]);
                    hf1 *= bal_unscale(xmax[2 * (l * G + gp) + 1]);
                }
                for (int k = lane; k < nkx; k += LN) {
                    const C y = L[k], yc = cconj(L[k == 0 ? 0 : N - k]);
                    // p0 = (Y + conj Y(-k)) / 2,  p1 = -i (Y - conj Y(-k)) / 2
                    h0[k] = mkc<C>(hf0 * (y.x + yc.x), hf0 * (y.y + yc.y));
                    if (q1 < a.nquad) h1[k] = mkc<C>(hf1 * (y.y - yc.y), -hf1 * (y.x - yc.x));
                }
            }
        }
    }
    group_sync<LN>();
}

// NK: the kept kx count when it is known at compile time (N/3, the 3/2 rule),
// 0 when it is read from the arguments
// C: element type of G, H and the FFTs (double2, or float2 for the FP32 chain)
template <int N, int NPI, int XGP, int NK, class C>
__global__ __launch_bounds__(256) void k_xpass_wave(XPassArgs a, const C *__restrict__ tw, long long nlines) {
    using X = XShape<N, NPI, XGP>;
    using T = typename CxT<C>::R;
    constexpr int R0 = X::R0, RL = X::RL, T0 = X::T0, TL = X::TL, LPW = X::LPW, SIN = X::SIN, LN = X::LN;
    extern __shared__ unsigned char lds_raw[];
    C *lds = reinterpret_cast<C *>(lds_raw);
    __shared__ double cfl_part[4][3];
    __shared__ unsigned xmax_all[4][XBAL_SLOTS];   // per-line maxima of packed pairs (FP32 only)
    // lane in the group of LN lanes that shares LPW lines, group in the workgroup
    const int lane = threadIdx.x % LN, grp = threadIdx.x / LN, wid = threadIdx.x >> 6;
    unsigned *xmax = xmax_all[grp];
    C *buf = lds + grp * X::SMAX * N;
    const C *G = reinterpret_cast<const C *>(a.G);
    const long long line0 = ((long long)blockIdx.x * (blockDim.x / LN) + grp) * LPW;
    const int nkx = NK > 0 ? NK : a.nkx;

    // ---- inverse transforms, one packed pair of full variables at a time
    // (half the LDS of both pairs at once, which buys a wavefront per SIMD);
    // the last stage leaves fv[task][pair][r] at x = j + r*TL in registers ----
    C fv[X::NT][NPI][RL];
#pragma unroll
    for (int p = 0; p < NPI; ++p) {
        // first stage: packed Hermitian spectra straight from HBM
        if constexpr (Balanced<C>::value) {
            static_assert(2 * LPW <= XBAL_SLOTS, "balancing slots");
            constexpr int TOT = SIN * T0, ROUNDS = (TOT + LN - 1) / LN;
            // the two lines' values first (kept slots only, kx = 0 imaginary
            // parts dropped), their per-line maxima, then the scaled packing
            C va[ROUNDS][R0], vb[ROUNDS][R0];
            if (lane < 2 * LPW) xmax[lane] = 0u;
            group_sync<LN>();
#pragma unroll
            for (int c = 0; c < ROUNDS; ++c) {
                const int b = lane + LN * c;
                if (TOT % LN == 0 || b < TOT) {
                    const int l = b / T0, j = b - l * T0;
                    const long long line = line0 + l;
                    const bool valid = line < nlines;
                    const int f0 = 2 * p, f1 = 2 * p + 1;
                    const bool has1 = f1 < a.nfull;
                    const size_t lo = valid ? (size_t)line : 0;
                    const C *g0 = G + ((size_t)f0 * nlines + lo) * nkx;
                    const C *g1 = G + ((size_t)(has1 ? f1 : f0) * nlines + lo) * nkx;
                    unsigned ma = 0u, mb = 0u;
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        const int e = j + r * T0;
                        const int lo = r * T0, hi = lo + T0 - 1;
                        const bool all_pos = NK > 0 && hi < NK;
                        const bool all_neg = NK > 0 && lo > N - NK;
                        const bool all_zero = NK > 0 && lo >= NK && hi <= N - NK;
                        if (all_zero) {
                            va[c][r] = vb[c][r] = mkc<C>(0, 0);
                            continue;
                        }
                        const bool pos = all_pos || (!all_neg && e < nkx);
                        const bool neg = all_neg || (!all_pos && e > N - nkx);
                        const int k = pos ? e : (neg ? N - e : 0);
                        const C A = g0[k], B = g1[k];
                        const bool e0 = r == 0 && j == 0;
                        const bool keep = valid && (all_pos || all_neg || pos || neg);
                        va[c][r] = mkc<C>(keep ? A.x : T(0), (keep && !e0) ? A.y : T(0));
                        vb[c][r] = mkc<C>((keep && has1) ? B.x : T(0), (keep && has1 && !e0) ? B.y : T(0));
                        ma = max(ma, absmax_bits(va[c][r]));
                        mb = max(mb, absmax_bits(vb[c][r]));
                    }
                    atomicMax(&xmax[2 * l], ma);
                    atomicMax(&xmax[2 * l + 1], mb);
                }
            }
            group_sync<LN>();
#pragma unroll
            for (int c = 0; c < ROUNDS; ++c) {
                const int b = lane + LN * c;
                if (TOT % LN == 0 || b < TOT) {
                    const int l = b / T0, j = b - l * T0;
                    const T sa = bal_scale(xmax[2 * l]), sb = bal_scale(xmax[2 * l + 1]);
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        const int e = j + r * T0;
                        const int lo = r * T0;
                        const bool all_neg = NK > 0 && lo > N - NK;
                        const bool all_pos = NK > 0 && lo + T0 - 1 < NK;
                        const bool pos = all_pos || (!all_neg && e < nkx);
                        const T ax = va[c][r].x * sa, ay = va[c][r].y * sa;
                        const T bx = vb[c][r].x * sb, by = vb[c][r].y * sb;
                        // pos: A + iB;  neg: conj(A) + i conj(B)
                        va[c][r] = mkc<C>(pos ? ax - by : ax + by, pos ? ay + bx : bx - ay);
                    }
                    Dft<R0, +1>::run(va[c]);
#pragma unroll
                    for (int r = 0; r < R0; ++r) buf[l * N + wfft::swz<XPlan<N>::W0>(j * R0 + r)] = va[c][r];
                }
            }
            group_sync<LN>();
        } else {
            constexpr int TOT = SIN * T0, ROUNDS = (TOT + LN - 1) / LN;
            C v[ROUNDS][R0];
#pragma unroll
            for (int c = 0; c < ROUNDS; ++c) {
                const int b = lane + LN * c;
                if (TOT % LN == 0 || b < TOT) {
                    const int l = b / T0, j = b - l * T0;
                    const long long line = line0 + l;
                    const bool valid = line < nlines;
                    const int f0 = 2 * p, f1 = 2 * p + 1;
                    const bool has1 = f1 < a.nfull;
                    const size_t lo = valid ? (size_t)line : 0;
                    const C *g0 = G + ((size_t)f0 * nlines + lo) * nkx;
                    const C *g1 = G + ((size_t)(has1 ? f1 : f0) * nlines + lo) * nkx;
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        const int e = j + r * T0;
                        // with NK known, the slots [r T0, (r+1) T0) that hold
                        // only positive, only negative or only dealiased kx
                        // are known once the loop is unrolled
                        const int lo = r * T0, hi = lo + T0 - 1;
                        const bool all_pos = NK > 0 && hi < NK;
                        const bool all_neg = NK > 0 && lo > N - NK;
                        const bool all_zero = NK > 0 && lo >= NK && hi <= N - NK;
                        if (all_zero) {
                            v[c][r] = mkc<C>(0, 0);
                            continue;
                        }
                        const bool pos = all_pos || (!all_neg && e < nkx);
                        const bool neg = all_neg || (!all_pos && e > N - nkx);
                        const int k = pos ? e : (neg ? N - e : 0);
                        C A = g0[k], B = g1[k];
                        const bool e0 = r == 0 && j == 0;   // kx = 0: imaginary parts dropped
                        const T bx = has1 ? B.x : T(0), by = (has1 && !e0) ? B.y : T(0);
                        const T ay = !e0 ? A.y : T(0);
                        // pos: A + iB;  neg: conj(A) + i conj(B)
                        T zx = pos ? A.x - by : A.x + by;
                        T zy = pos ? ay + bx : bx - ay;
                        const bool keep = valid && (all_pos || all_neg || pos || neg);
                        v[c][r] = mkc<C>(keep ? zx : T(0), keep ? zy : T(0));
                    }
                    Dft<R0, +1>::run(v[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < ROUNDS; ++c) {
                const int b = lane + LN * c;
                if (TOT % LN == 0 || b < TOT) {
                    const int l = b / T0, j = b - l * T0;
#pragma unroll
                    for (int r = 0; r < R0; ++r) buf[l * N + wfft::swz<XPlan<N>::W0>(j * R0 + r)] = v[c][r];
                }
            }
            group_sync<LN>();
        }
        XPlan<N>::template InvMid<SIN>::run(buf, tw, lane);
        // last stage into registers
#pragma unroll
        for (int c = 0; c < X::NT; ++c) {
            const int t = lane + LN * c, l = t / TL, j = t - l * TL;
            const C *L = buf + l * N;
            C wp[RL];
            wfft::twiddle_powers<RL, +1>(tw, j, wp);
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                C x = L[j + r * TL];
                if (r > 0) x = cmul(x, wp[r]);
                fv[c][p][r] = x;
            }
            Dft<RL, +1>::run(fv[c][p]);
            if constexpr (Balanced<C>::value) {
                const T ua = bal_unscale(xmax[2 * l]), ub = bal_unscale(xmax[2 * l + 1]);
#pragma unroll
                for (int r = 0; r < RL; ++r) fv[c][p][r] = mkc<C>(fv[c][p][r].x * ua, fv[c][p][r].y * ub);
            }
        }
        group_sync<LN>();
    }

    // ---- CFL maxima over u, v, w / weight_z (full variables 0, 1, 2) ----
    if constexpr (NPI >= 2) {
      if (a.cflmax && a.nfull >= 3) {
        double mu = 0.0, mv = 0.0, mw = 0.0;
#pragma unroll
        for (int c = 0; c < X::NT; ++c) {
            const int t = lane + LN * c, l = t / TL;
            const long long line = line0 + l;
            if (line < nlines) {
                // max |w| / zw = (max |w|) / zw: rounded division by a
                // positive number keeps the order, so one division per task
                double wm = 0.0;
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    mu = fmax(mu, fabs(double(fv[c][0][r].x)));
                    mv = fmax(mv, fabs(double(fv[c][0][r].y)));
                    wm = fmax(wm, fabs(double(fv[c][1][r].x)));
                }
                mw = fmax(mw, wm / a.zweight[line / a.NYAA]);
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            mu = fmax(mu, __shfl_xor(mu, off));
            mv = fmax(mv, __shfl_xor(mv, off));
            mw = fmax(mw, __shfl_xor(mw, off));
        }
        if ((threadIdx.x & 63) == 0) {   // per wavefront, not per lane group
            cfl_part[wid][0] = mu;
            cfl_part[wid][1] = mv;
            cfl_part[wid][2] = mw;
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            double m = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, cfl_part[w][threadIdx.x]);
            atomicMax(&a.cflmax[threadIdx.x], (unsigned long long)__double_as_longlong(m));
        }
      }
    }

    // ---- products and forward transforms, XGP pairs at a time ----
    const int npairs = (a.nquad + 1) / 2;
    for (int pq0 = 0; pq0 < npairs; pq0 += XGP) {
        if (npairs - pq0 >= XGP)
            xfwd_batch<N, NPI, XGP, XGP>(a, fv, buf, tw, lane, line0, nlines, pq0, xmax);
        else
            xfwd_batch<N, NPI, XGP, 1>(a, fv, buf, tw, lane, line0, nlines, pq0, xmax);
    }
}

template <int N, int NPI, int XGP, int NK, class C>
static bool launch_wave(hipStream_t st, const C *tw, const XPassArgs &a) {
    using X = XShape<N, NPI, XGP>;
    constexpr int WPG = X::LN / 64;   // wavefronts per lane group
    static_assert(4 % WPG == 0, "a workgroup holds whole lane groups");
    const long long nlines = (long long)a.NZAA * a.NYAA;
    const size_t per_group = size_t(X::SMAX) * N * sizeof(C);
    const int wpb = a.max_wpb >= 1 && a.max_wpb <= 4 ? a.max_wpb : 4;
    int gpb = wpb / WPG > 0 ? wpb / WPG : 1;   // lane groups per workgroup
    while (gpb > 1 && gpb * per_group > 65536) gpb /= 2;
    if (gpb * per_group > 65536) return false;
    const long long per_block = (long long)gpb * X::LPW;
    const unsigned blocks = unsigned((nlines + per_block - 1) / per_block);
    hipLaunchKernelGGL((k_xpass_wave<N, NPI, XGP, NK, C>), dim3(blocks), dim3(X::LN * gpb), gpb * per_group, st, a,
                       tw, nlines);
    check_launch("x pass");
    return true;
}

template <int N, class C>
static bool launch_wave_c(hipStream_t st, const C *tw, const XPassArgs &a) {
    // one product pair per forward batch: two pairs would fill the lanes of
    // the forward stages better, but their registers cost a wave per SIMD
    const bool third = 3 * a.nkx == N;
    switch ((a.nfull + 1) / 2) {
    case 1: return third ? launch_wave<N, 1, 1, N / 3>(st, tw, a) : launch_wave<N, 1, 1, 0>(st, tw, a);
    case 2: return third ? launch_wave<N, 2, 1, N / 3>(st, tw, a) : launch_wave<N, 2, 1, 0>(st, tw, a);
    case 3: return third ? launch_wave<N, 3, 1, N / 3>(st, tw, a) : launch_wave<N, 3, 1, 0>(st, tw, a);
    default: return false;
    }
}

template <int N>
static bool launch_wave_n(hipStream_t st, const FFTPlan &px, const XPassArgs &a) {
    if (a.fp32) {
        if (!px.twf) return false;
        return launch_wave_c<N>(st, px.twf, a);
    }
    return launch_wave_c<N>(st, px.tw, a);
}

// ---- the x transforms alone (FP64) ----
// c2r of nbatch variables, G[b][line][kx] -> phys[b][line][x], and r2c,
// phys -> H[b][line][kx] (kept kx), on the x pass's plans: two variables of a
// line (b = 2p, 2p + 1, blockIdx.y = p) share one complex sequence, the inverse
// from the packed Hermitian spectra as in the x pass, the forward separated
// with Hermitian symmetry as its products are.  The c2r / r2c of
// slabs.fftw3mpi/transforms.f90:155-301 in x, for the transform seam, the
// physical outputs, the side products and the closure's inputs and outputs.
template <int N, int NK>
__global__ __launch_bounds__(256) void k_xbwd_wave(const double2 *__restrict__ G, double *__restrict__ phys,
                                                   const double2 *__restrict__ tw, long long nlines, int nkx_arg,
                                                   int nbatch) {
    using X = XShape<N, 1, 1>;
    using C = double2;
    constexpr int R0 = X::R0, RL = X::RL, T0 = X::T0, TL = X::TL, LPW = X::LPW, SIN = X::SIN, LN = X::LN;
    extern __shared__ unsigned char lds_raw[];
    C *buf = reinterpret_cast<C *>(lds_raw) + (threadIdx.x / LN) * X::SMAX * N;
    const int lane = threadIdx.x % LN;
    const long long line0 = ((long long)blockIdx.x * (blockDim.x / LN) + threadIdx.x / LN) * LPW;
    const int nkx = NK > 0 ? NK : nkx_arg;
    const int f0 = 2 * blockIdx.y, f1 = f0 + 1;
    const bool has1 = f1 < nbatch;
    {
        constexpr int TOT = SIN * T0, ROUNDS = (TOT + LN - 1) / LN;
        C v[ROUNDS][R0];
#pragma unroll
        for (int c = 0; c < ROUNDS; ++c) {
            const int b = lane + LN * c;
            if (TOT % LN == 0 || b < TOT) {
                const int l = b / T0, j = b - l * T0;
                const long long line = line0 + l;
                const bool valid = line < nlines;
                const size_t lo = valid ? (size_t)line : 0;
                const C *g0 = G + ((size_t)f0 * nlines + lo) * nkx;
                const C *g1 = G + ((size_t)(has1 ? f1 : f0) * nlines + lo) * nkx;
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const int e = j + r * T0;
                    const int lo_e = r * T0, hi_e = lo_e + T0 - 1;
                    const bool all_pos = NK > 0 && hi_e < NK;
                    const bool all_neg = NK > 0 && lo_e > N - NK;
                    const bool all_zero = NK > 0 && lo_e >= NK && hi_e <= N - NK;
                    if (all_zero) {
                        v[c][r] = mkc<C>(0, 0);
                        continue;
                    }
                    const bool pos = all_pos || (!all_neg && e < nkx);
                    const bool neg = all_neg || (!all_pos && e > N - nkx);
                    const int k = pos ? e : (neg ? N - e : 0);
                    const C A = g0[k], B = g1[k];
                    const bool e0 = r == 0 && j == 0;   // kx = 0: imaginary parts dropped
                    const double bx = has1 ? B.x : 0.0, by = (has1 && !e0) ? B.y : 0.0;
                    const double ay = !e0 ? A.y : 0.0;
                    // pos: A + iB;  neg: conj(A) + i conj(B)
                    const double zx = pos ? A.x - by : A.x + by;
                    const double zy = pos ? ay + bx : bx - ay;
                    const bool keep = valid && (all_pos || all_neg || pos || neg);
                    v[c][r] = mkc<C>(keep ? zx : 0.0, keep ? zy : 0.0);
                }
                Dft<R0, +1>::run(v[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < ROUNDS; ++c) {
            const int b = lane + LN * c;
            if (TOT % LN == 0 || b < TOT) {
                const int l = b / T0, j = b - l * T0;
#pragma unroll
                for (int r = 0; r < R0; ++r) buf[l * N + wfft::swz<XPlan<N>::W0>(j * R0 + r)] = v[c][r];
            }
        }
        group_sync<LN>();
    }
    XPlan<N>::template InvMid<SIN>::run(buf, tw, lane);
    // last stage in registers: x = j + r TL of both lines, stored to consecutive x
#pragma unroll
    for (int c = 0; c < X::NT; ++c) {
        const int t = lane + LN * c, l = t / TL, j = t - l * TL;
        const long long line = line0 + l;
        const C *L = buf + l * N;
        C wp[RL], x[RL];
        wfft::twiddle_powers<RL, +1>(tw, j, wp);
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            x[r] = L[j + r * TL];
            if (r > 0) x[r] = cmul(x[r], wp[r]);
        }
        Dft<RL, +1>::run(x);
        if (line < nlines) {
            double *o0 = phys + ((size_t)f0 * nlines + line) * N + j;
            double *o1 = phys + ((size_t)f1 * nlines + line) * N + j;
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                o0[r * TL] = x[r].x;
                if (has1) o1[r * TL] = x[r].y;
            }
        }
    }
}

template <int N>
__global__ __launch_bounds__(256) void k_xfwd_wave(const double *__restrict__ phys, double2 *__restrict__ H,
                                                   const double2 *__restrict__ tw, long long nlines, int nkx,
                                                   int nbatch) {
    using X = XShape<N, 1, 1>;
    using C = double2;
    constexpr int RL = X::RL, TL = X::TL, LPW = X::LPW, LN = X::LN;
    extern __shared__ unsigned char lds_raw[];
    C *buf = reinterpret_cast<C *>(lds_raw) + (threadIdx.x / LN) * X::SMAX * N;
    const int lane = threadIdx.x % LN;
    const long long line0 = ((long long)blockIdx.x * (blockDim.x / LN) + threadIdx.x / LN) * LPW;
    const int f0 = 2 * blockIdx.y, f1 = f0 + 1;
    const bool has1 = f1 < nbatch;
    // first forward stage from the two lines' values at x = j + r TL
#pragma unroll
    for (int c = 0; c < X::NT; ++c) {
        const int t = lane + LN * c, l = t / TL, j = t - l * TL;
        const long long line = line0 + l;
        const bool valid = line < nlines;
        const double *i0 = phys + ((size_t)f0 * nlines + (valid ? line : 0)) * N + j;
        const double *i1 = phys + ((size_t)(has1 ? f1 : f0) * nlines + (valid ? line : 0)) * N + j;
        C y[RL];
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            const double a = i0[r * TL], b = i1[r * TL];
            y[r] = mkc<C>(valid ? a : 0.0, (valid && has1) ? b : 0.0);
        }
        Dft<RL, -1>::run(y);
#pragma unroll
        for (int r = 0; r < RL; ++r) buf[l * N + wfft::swz<XPlan<N>::WL>(j * RL + r)] = y[r];
    }
    group_sync<LN>();
    XPlan<N>::template FwdRest<LPW>::run(buf, tw, lane);
#pragma unroll
    for (int l = 0; l < LPW; ++l) {
        const long long line = line0 + l;
        if (line < nlines) {
            const C *L = buf + l * N;
            C *h0 = H + ((size_t)f0 * nlines + line) * nkx;
            C *h1 = H + ((size_t)f1 * nlines + line) * nkx;
            for (int k = lane; k < nkx; k += LN) {
                const C y = L[k], yc = cconj(L[k == 0 ? 0 : N - k]);
                // f0 = (Y + conj Y(-k)) / 2,  f1 = -i (Y - conj Y(-k)) / 2
                h0[k] = mkc<C>(0.5 * (y.x + yc.x), 0.5 * (y.y + yc.y));
                if (has1) h1[k] = mkc<C>(0.5 * (y.y - yc.y), -0.5 * (y.x - yc.x));
            }
        }
    }
}

// grid of the x transforms: lane groups per workgroup within 64 KiB of LDS,
// blockIdx.y over the variable pairs
template <int N>
static bool xt_shape(long long nlines, int nbatch, dim3 &grid, dim3 &block, size_t &shm) {
    using X = XShape<N, 1, 1>;
    const size_t per_group = size_t(X::SMAX) * N * sizeof(double2);
    int gpb = 4 / (X::LN / 64);
    while (gpb > 1 && gpb * per_group > 65536) gpb /= 2;
    if (gpb * per_group > 65536) return false;
    const long long per_block = (long long)gpb * X::LPW;
    grid = dim3(unsigned((nlines + per_block - 1) / per_block), unsigned((nbatch + 1) / 2));
    block = dim3(X::LN * gpb);
    shm = gpb * per_group;
    return true;
}

template <int N>
static bool xbwd_n(hipStream_t st, const double2 *tw, const double2 *G, double *phys, int nbatch, long long nlines,
                   int nkx) {
    dim3 grid, block;
    size_t shm;
    if (!xt_shape<N>(nlines, nbatch, grid, block, shm)) return false;
    if (3 * nkx == N)
        hipLaunchKernelGGL((k_xbwd_wave<N, N / 3>), grid, block, shm, st, G, phys, tw, nlines, nkx, nbatch);
    else
        hipLaunchKernelGGL((k_xbwd_wave<N, 0>), grid, block, shm, st, G, phys, tw, nlines, nkx, nbatch);
    check_launch("x c2r");
    return true;
}

template <int N>
static bool xfwd_n(hipStream_t st, const double2 *tw, const double *phys, double2 *H, int nbatch, long long nlines,
                   int nkx) {
    dim3 grid, block;
    size_t shm;
    if (!xt_shape<N>(nlines, nbatch, grid, block, shm)) return false;
    hipLaunchKernelGGL((k_xfwd_wave<N>), grid, block, shm, st, phys, H, tw, nlines, nkx, nbatch);
    check_launch("x r2c");
    return true;
}

bool launch_xbwd_wave(hipStream_t st, const FFTPlan &px, const double2 *G, double *phys, int nbatch, long long nlines,
                      int nkx) {
    if (nbatch < 1 || nlines < 1 || 2 * nkx > px.N || !px.tw) return false;
    switch (px.N) {
    case 96: return xbwd_n<96>(st, px.tw, G, phys, nbatch, nlines, nkx);
    case 144: return xbwd_n<144>(st, px.tw, G, phys, nbatch, nlines, nkx);
    case 192: return xbwd_n<192>(st, px.tw, G, phys, nbatch, nlines, nkx);
    case 384: return xbwd_n<384>(st, px.tw, G, phys, nbatch, nlines, nkx);
    case 768: return xbwd_n<768>(st, px.tw, G, phys, nbatch, nlines, nkx);
    default: return false;
    }
}

bool launch_xfwd_wave(hipStream_t st, const FFTPlan &px, const double *phys, double2 *H, int nbatch, long long nlines,
                      int nkx) {
    if (nbatch < 1 || nlines < 1 || 2 * nkx > px.N || !px.tw) return false;
    switch (px.N) {
    case 96: return xfwd_n<96>(st, px.tw, phys, H, nbatch, nlines, nkx);
    case 144: return xfwd_n<144>(st, px.tw, phys, H, nbatch, nlines, nkx);
    case 192: return xfwd_n<192>(st, px.tw, phys, H, nbatch, nlines, nkx);
    case 384: return xfwd_n<384>(st, px.tw, phys, H, nbatch, nlines, nkx);
    case 768: return xfwd_n<768>(st, px.tw, phys, H, nbatch, nlines, nkx);
    default: return false;
    }
}

bool launch_xpass_wave(hipStream_t st, const FFTPlan &px, const XPassArgs &a) {
    if (a.nfull < 1 || 2 * a.nkx > px.N) return false;
    switch (px.N) {
    case 96: return launch_wave_n<96>(st, px, a);
    case 144: return launch_wave_n<144>(st, px, a);
    case 192: return launch_wave_n<192>(st, px, a);
    case 384: return launch_wave_n<384>(st, px, a);
    case 768: return launch_wave_n<768>(st, px, a);
    default: return false;
    }
}

}  // namespace coral
