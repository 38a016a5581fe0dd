// Chebyshev transforms along z in FP32 on MFMA (v_mfma_f32_16x16x4f32) with
// compensated accumulation: the FP32 chain of the mixed-precision step.
//
// The folding into even and odd halves, the fragment layout of the matrix
// operand in LDS and the column strips per wavefront are those of the FP64
// kernels (ztransform.hip).  The products are accumulated in chunks: each
// unrolled group of k-steps goes into a fresh FP32 accumulator on the matrix
// core, and the chunk's sum is added to a running (hi, lo) pair with Knuth's
// TwoSum (s = hi + c, b = s - hi, err = (hi - (s - b)) + (c - b)), so the
// rounding of the running sum is carried in lo instead of lost; the output is
// hi + lo rounded once.
#include "kernels.hpp"

#include <functional>

namespace coral {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int Z32_PF_BWD = 4, Z32_PF_FWD = 8, Z32_THREADS = 512;

__device__ inline f4 mfma32(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// D(i, j) of a 16x16 tile: lane j + 16*(i/4), element i%4 (16x16x4f32, one block)
__device__ inline int z32_row(int kq, int it) { return 4 * kq + it; }

// (hi, lo) += c, elementwise TwoSum
__device__ inline void two_sum_acc(f4 &hi, f4 &lo, f4 c) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float s = hi[i] + c[i];
        const float b = s - hi[i];
        const float e = (hi[i] - (s - b)) + (c[i] - b);
        hi[i] = s;
        lo[i] += e;
    }
}

template <int TE, int TO>
__global__ __launch_bounds__(Z32_THREADS) void k_zbwd32(const float *__restrict__ frag, const float *__restrict__ fragO,
                                                        int NEK, int NOK, const float *__restrict__ X,
                                                        float *__restrict__ F, int NZ, int NZAA, int ncol, int nb,
                                                        int r0) {
    extern __shared__ float lds32[];
    const int nfe = TE * NEK * 64, nfrag = nfe + TO * NOK * 64;
    for (int e = threadIdx.x; e < nfrag; e += blockDim.x) lds32[e] = e < nfe ? frag[e] : fragO[e - nfe];
    __syncthreads();
    const float *GE = lds32, *GO = lds32 + TE * NEK * 64;
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int j = lane & 15, kq = lane >> 4;
    const int ntc = (ncol + 15) >> 4, ntiles = nb * ntc;
    const int Ke = (NZAA + 1) / 2, Ko = NZAA / 2, NE = (NZ + 1) / 2, NO = NZ / 2;
    for (int tile = blockIdx.x * nw + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * nw) {
        const int b = tile / ntc, col = (tile - b * ntc) * 16 + j;
        const bool cv = col < ncol;
        const float *Xb = X + (size_t)b * NZ * ncol + (cv ? col : 0);
        f4 he[TE], le[TE], ho[TO], lo[TO];
#pragma unroll
        for (int r = 0; r < TE; ++r) he[r] = le[r] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < TO; ++r) ho[r] = lo[r] = f4{0.f, 0.f, 0.f, 0.f};
        auto ld_e = [&](int kk) -> float {
            const int m = 4 * kk + kq;
            return Xb[(size_t)(cv && m < NE ? 2 * m : 0) * ncol];
        };
        auto ld_o = [&](int kk) -> float {
            const int m = 4 * kk + kq;
            return Xb[(size_t)(cv && m < NO ? 2 * m + 1 : 0) * ncol];
        };
        float re[Z32_PF_BWD], ro[Z32_PF_BWD];
#pragma unroll
        for (int u = 0; u < Z32_PF_BWD; ++u) {
            re[u] = ld_e(u);
            ro[u] = ld_o(u);
        }
        for (int k0 = 0; k0 < NEK; k0 += Z32_PF_BWD) {
            // one chunk: the group's k-steps into fresh accumulators
            f4 ce[TE], co[TO];
#pragma unroll
            for (int r = 0; r < TE; ++r) ce[r] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < TO; ++r) co[r] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u = 0; u < Z32_PF_BWD; ++u) {
                const int kk = k0 + u;
                if (kk < NEK) {
                    const int m = 4 * kk + kq;
                    const float be = (cv && m < NE) ? re[u] : 0.f;
                    const float bo = (cv && m < NO) ? ro[u] : 0.f;
#pragma unroll
                    for (int r = 0; r < TE; ++r) ce[r] = mfma32(GE[(r * NEK + kk) * 64 + lane], be, ce[r]);
                    if (kk < NOK) {
#pragma unroll
                        for (int r = 0; r < TO; ++r) co[r] = mfma32(GO[(r * NOK + kk) * 64 + lane], bo, co[r]);
                    }
                }
                re[u] = ld_e(kk + Z32_PF_BWD);
                ro[u] = ld_o(kk + Z32_PF_BWD);
            }
#pragma unroll
            for (int r = 0; r < TE; ++r) two_sum_acc(he[r], le[r], ce[r]);
#pragma unroll
            for (int r = 0; r < TO; ++r) two_sum_acc(ho[r], lo[r], co[r]);
        }
        if (cv) {
            float *Fb = F + (size_t)b * NZAA * ncol + col;
#pragma unroll
            for (int r = 0; r < TE; ++r)
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int z = 16 * (r0 + r) + z32_row(kq, it);
                    const int ro_ = r < TO ? r : 0;
                    // E and O with their low parts, each rounded once
                    const float E = he[r][it], Elo = le[r][it];
                    const float O = (r < TO) ? ho[ro_][it] : 0.f, Olo = (r < TO) ? lo[ro_][it] : 0.f;
                    if (z < Ke) Fb[(size_t)z * ncol] = (E + O) + (Elo + Olo);
                    if (z < Ko) Fb[(size_t)(NZAA - 1 - z) * ncol] = (E - O) + (Elo - Olo);
                }
        }
    }
}

template <int RE, int RO>
__global__ __launch_bounds__(Z32_THREADS) void k_zfwd32(const float *__restrict__ frag, const float *__restrict__ fragO,
                                                        int KEK, int KOK, const float *__restrict__ Y,
                                                        float *__restrict__ Q, int NZ, int NZAA, int ncol, int nb,
                                                        int re0, int ro0) {
    extern __shared__ float lds32[];
    const int nfe = RE * KEK * 64, nfrag = nfe + RO * KOK * 64;
    for (int e = threadIdx.x; e < nfrag; e += blockDim.x) lds32[e] = e < nfe ? frag[e] : fragO[e - nfe];
    __syncthreads();
    const float *FE = lds32, *FO = lds32 + RE * KEK * 64;
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int j = lane & 15, kq = lane >> 4;
    const int ntc = (ncol + 15) >> 4, ntiles = nb * ntc;
    const int Ke = (NZAA + 1) / 2, Ko = NZAA / 2;
    for (int tile = blockIdx.x * nw + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * nw) {
        const int b = tile / ntc, col = (tile - b * ntc) * 16 + j;
        const bool cv = col < ncol;
        const float *Yb = Y + (size_t)b * NZAA * ncol + (cv ? col : 0);
        f4 he[RE], le[RE], ho[RO], lo[RO];
#pragma unroll
        for (int r = 0; r < RE; ++r) he[r] = le[r] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < RO; ++r) ho[r] = lo[r] = f4{0.f, 0.f, 0.f, 0.f};
        auto ld0 = [&](int kk) -> float {
            const int z = 4 * kk + kq;
            return Yb[(size_t)(cv && z < Ke ? z : 0) * ncol];
        };
        auto ld1 = [&](int kk) -> float {
            const int z = 4 * kk + kq;
            return Yb[(size_t)(cv && z < Ko ? NZAA - 1 - z : 0) * ncol];
        };
        float r0[Z32_PF_FWD], r1[Z32_PF_FWD];
#pragma unroll
        for (int u = 0; u < Z32_PF_FWD; ++u) {
            r0[u] = ld0(u);
            r1[u] = ld1(u);
        }
        for (int k0 = 0; k0 < KEK; k0 += Z32_PF_FWD) {
            f4 ce[RE], co[RO];
#pragma unroll
            for (int r = 0; r < RE; ++r) ce[r] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < RO; ++r) co[r] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u = 0; u < Z32_PF_FWD; ++u) {
                const int kk = k0 + u;
                if (kk < KEK) {
                    const int z = 4 * kk + kq;
                    const bool oh = cv && z < Ke, om = cv && z < Ko;
                    const float h = oh ? r0[u] : 0.f, hm = om ? r1[u] : 0.f;
                    const float S = h + hm, D = om ? h - hm : 0.f;
#pragma unroll
                    for (int r = 0; r < RE; ++r) ce[r] = mfma32(FE[(r * KEK + kk) * 64 + lane], S, ce[r]);
                    if (kk < KOK) {
#pragma unroll
                        for (int r = 0; r < RO; ++r) co[r] = mfma32(FO[(r * KOK + kk) * 64 + lane], D, co[r]);
                    }
                }
                r0[u] = ld0(kk + Z32_PF_FWD);
                r1[u] = ld1(kk + Z32_PF_FWD);
            }
#pragma unroll
            for (int r = 0; r < RE; ++r) two_sum_acc(he[r], le[r], ce[r]);
#pragma unroll
            for (int r = 0; r < RO; ++r) two_sum_acc(ho[r], lo[r], co[r]);
        }
        if (cv) {
            float *Qb = Q + (size_t)b * NZ * ncol + col;
#pragma unroll
            for (int r = 0; r < RE; ++r)
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int n = 2 * (16 * (re0 + r) + z32_row(kq, it));
                    if (n < NZ) Qb[(size_t)n * ncol] = he[r][it] + le[r][it];
                }
#pragma unroll
            for (int r = 0; r < RO; ++r)
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int n = 2 * (16 * (ro0 + r) + z32_row(kq, it)) + 1;
                    if (n < NZ) Qb[(size_t)n * ncol] = ho[r][it] + lo[r][it];
                }
        }
    }
}

// ---- host side ----

static void fill_frags32(std::vector<float> &out, int ntile, int nk, const std::function<double(int, int)> &A, int M,
                         int K) {
    for (int t = 0; t < ntile; ++t)
        for (int kk = 0; kk < nk; ++kk)
            for (int lane = 0; lane < 64; ++lane) {
                const int m = 16 * t + (lane & 15), k = 4 * kk + (lane >> 4);
                out.push_back((m < M && k < K) ? float(A(m, k)) : 0.f);
            }
}

#define Z32BWD_SHAPES(X) X(1, 1) X(2, 1) X(2, 2) X(3, 2) X(3, 3) X(4, 3) X(4, 4) X(5, 5) X(6, 6) X(7, 6) X(7, 7)
#define Z32FWD_SHAPES(X) X(1, 1) X(2, 1) X(2, 2) X(3, 2) X(3, 3) X(4, 3) X(4, 4) X(5, 4) X(5, 5)

static bool z32bwd_ok(int ne, int no) {
#define ZS(a, b) if (ne == a && no == b) return true;
    Z32BWD_SHAPES(ZS)
#undef ZS
    return false;
}
static bool z32fwd_ok(int ne, int no) {
#define ZS(a, b) if (ne == a && no == b) return true;
    Z32FWD_SHAPES(ZS)
#undef ZS
    return false;
}

bool make_zfrags32(int NZ, int NZAA, const std::vector<double> &Tz, const std::vector<double> &Cz, size_t lds_limit,
                   ZFrags32 &zf) {
    const int Ke = (NZAA + 1) / 2, Ko = NZAA / 2, NE = (NZ + 1) / 2, NO = NZ / 2;
    zf = ZFrags32{};
    zf.TE = (Ke + 15) / 16;
    zf.TO = (Ko + 15) / 16;
    zf.NEK = (NE + 3) / 4;
    zf.NOK = (NO + 3) / 4;
    fill_frags32(zf.bwd, zf.TE, zf.NEK, [&](int z, int m) { return Tz[(size_t)z * NZ + 2 * m]; }, Ke, NE);
    fill_frags32(zf.bwd, zf.TO, zf.NOK, [&](int z, int m) { return Tz[(size_t)z * NZ + 2 * m + 1]; }, Ko, NO);
    zf.RE = (NE + 15) / 16;
    zf.RO = (NO + 15) / 16;
    zf.KEK = (Ke + 3) / 4;
    zf.KOK = (Ko + 3) / 4;
    fill_frags32(zf.fwd, zf.RE, zf.KEK, [&](int m, int z) { return Cz[(size_t)(2 * m) * NZAA + z]; }, NE, Ke);
    fill_frags32(zf.fwd, zf.RO, zf.KOK, [&](int m, int z) { return Cz[(size_t)(2 * m + 1) * NZAA + z]; }, NO, Ko);
    for (int P = 1; P <= ZMAXPARTS && !zf.bwd_ok; ++P) {
        bool ok = true;
        for (int p = 0; p < P && ok; ++p) {
            const int r0 = p * zf.TE / P, r1 = (p + 1) * zf.TE / P;
            const int ne = r1 - r0, no = std::max(0, std::min(r1, zf.TO) - r0);
            ok = ne > 0 && no > 0 && z32bwd_ok(ne, no) && size_t(ne * zf.NEK + no * zf.NOK) * 64 * sizeof(float) <= lds_limit;
            zf.bp[p] = ZPart{r0, ne, r0, no};
        }
        if (ok) {
            zf.bwd_ok = true;
            zf.nbp = P;
        }
    }
    for (int P = 1; P <= ZMAXPARTS && !zf.fwd_ok; ++P) {
        bool ok = true;
        for (int p = 0; p < P && ok; ++p) {
            const int e0 = p * zf.RE / P, e1 = (p + 1) * zf.RE / P, o0 = p * zf.RO / P, o1 = (p + 1) * zf.RO / P;
            ok = e1 > e0 && o1 > o0 && z32fwd_ok(e1 - e0, o1 - o0) &&
                 size_t((e1 - e0) * zf.KEK + (o1 - o0) * zf.KOK) * 64 * sizeof(float) <= lds_limit;
            zf.fp[p] = ZPart{e0, e1 - e0, o0, o1 - o0};
        }
        if (ok) {
            zf.fwd_ok = true;
            zf.nfp = P;
        }
    }
    return zf.bwd_ok && zf.fwd_ok;
}

static int grid32(int ntiles, int ncu) {
    const int want = (ntiles + Z32_THREADS / 64 - 1) / (Z32_THREADS / 64);
    return want < ncu ? want : ncu;
}

void launch_zbwd32(hipStream_t st, const ZFrags32Dev &z, const float2 *spec, float2 *Fz, int NZ, int NZAA, int nsys,
                   int nb) {
    if (!z.bwd_ok) throw std::runtime_error("no FP32 MFMA plan for the z synthesis at this NZ");
    const int ncol = 2 * nsys, g = grid32(nb * ((ncol + 15) / 16), z.ncu);
    const float *X = (const float *)spec;
    float *F = (float *)Fz;
    for (int p = 0; p < z.nbp; ++p) {
        const ZPart &pt = z.bp[p];
        const size_t shm = size_t(pt.ne * z.NEK + pt.no * z.NOK) * 64 * sizeof(float);
        const float *fe = z.bwd + size_t(pt.e0) * z.NEK * 64, *fo = z.bwd + size_t(z.TE * z.NEK + pt.o0 * z.NOK) * 64;
        bool done = false;
#define ZS(a, b)                                                                                             \
        if (!done && pt.ne == a && pt.no == b) {                                                             \
            if (shm > 65536) set_dynamic_lds((const void *)k_zbwd32<a, b>, shm, "FP32 z synthesis");         \
            hipLaunchKernelGGL((k_zbwd32<a, b>), dim3(g), dim3(Z32_THREADS), shm, st, fe, fo, z.NEK, z.NOK, X, F, NZ, \
                               NZAA, ncol, nb, pt.e0);                                                       \
            done = true;                                                                                     \
        }
        Z32BWD_SHAPES(ZS)
#undef ZS
        if (!done) throw std::runtime_error("FP32 z synthesis: part shape not compiled");
    }
    check_launch("FP32 z synthesis");
}

void launch_zfwd32(hipStream_t st, const ZFrags32Dev &z, const float2 *Hy, float2 *Qhat, int NZ, int NZAA, int nsys,
                   int nb) {
    if (!z.fwd_ok) throw std::runtime_error("no FP32 MFMA plan for the z analysis at this NZ");
    const int ncol = 2 * nsys, g = grid32(nb * ((ncol + 15) / 16), z.ncu);
    const float *Y = (const float *)Hy;
    float *Q = (float *)Qhat;
    for (int p = 0; p < z.nfp; ++p) {
        const ZPart &pt = z.fp[p];
        const size_t shm = size_t(pt.ne * z.KEK + pt.no * z.KOK) * 64 * sizeof(float);
        const float *fe = z.fwd + size_t(pt.e0) * z.KEK * 64, *fo = z.fwd + size_t(z.RE * z.KEK + pt.o0 * z.KOK) * 64;
        bool done = false;
#define ZS(a, b)                                                                                             \
        if (!done && pt.ne == a && pt.no == b) {                                                             \
            if (shm > 65536) set_dynamic_lds((const void *)k_zfwd32<a, b>, shm, "FP32 z analysis");          \
            hipLaunchKernelGGL((k_zfwd32<a, b>), dim3(g), dim3(Z32_THREADS), shm, st, fe, fo, z.KEK, z.KOK, Y, Q, NZ, \
                               NZAA, ncol, nb, pt.e0, pt.o0);                                                \
            done = true;                                                                                     \
        }
        Z32FWD_SHAPES(ZS)
#undef ZS
        if (!done) throw std::runtime_error("FP32 z analysis: part shape not compiled");
    }
    check_launch("FP32 z analysis");
}

}  // namespace coral
