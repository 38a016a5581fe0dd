// Chebyshev transforms along z as FP64 MFMA products (v_mfma_f64_16x16x4f64).
//
// The Gauss-Chebyshev points are symmetric, x_{NZAA-1-z} = -x_z, and
// T_n(-x) = (-1)^n T_n(x).  Both transforms are therefore folded into an even
// and an odd half of about half the size:
//   synthesis (z bwd):  E[z] = sum_{n even} T_n(x_z) c_n,  O[z] = sum_{n odd} ...
//                       F[z] = E + O,  F[NZAA-1-z] = E - O          (z < NZAA/2)
//   analysis  (z fwd):  S[z] = H[z] + H[NZAA-1-z],  D[z] = H[z] - H[NZAA-1-z]
//                       Q[n even] = sum_z Cz[n][z] S[z],  Q[n odd] = sum_z Cz[n][z] D[z]
// (the middle point of an odd NZAA enters S once and D not at all).  This is
// the z-direction part of c2r_transform / r2c_transform (REDFT01 / REDFT10,
// slabs.fftw3mpi/transforms.f90:155-301) as dense real matrices acting on the
// real and imaginary parts of every kept (kx, ky) column.
//
// Layout: X[b][k][col] and Y[b][m][col], col = 2*s + (re, im), rows contiguous.
// The matrix operand lives in LDS as MFMA A fragments, frag[(tile*nk + kstep)*64
// + lane] = A[16*tile + lane%16][4*kstep + lane/16]; a wavefront owns a 16-column
// strip (B operand: lane = column + 16*k) and every output row of it.
#include "kernels.hpp"

#include <cstdio>

namespace coral {

typedef double d4 __attribute__((ext_vector_type(4)));

// k-steps of B operands in flight (synthesis / analysis), threads per workgroup
constexpr int ZPF_BWD = 4, ZPF_FWD = 8, ZTHREADS = 512;

__device__ inline d4 mfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// D(i, j) of a 16x16 tile sits in lane j + 16*(i%4), element i/4.
// fragE / fragO: the fragments of the TE even and TO odd row tiles this launch
// computes, output row tiles r0 .. r0 + TE - 1 (a part of all tiles when they
// do not fit one workgroup's LDS together)
template <int TE, int TO>
__global__ __launch_bounds__(ZTHREADS) void k_zbwd_mfma(const double *__restrict__ frag,
                                                    const double *__restrict__ fragO, int NEK, int NOK,
                                                    const double *__restrict__ X, double *__restrict__ F, int NZ,
                                                    int NZAA, int ncol, int nb, int r0) {
    extern __shared__ double lds[];
    const int nfe = TE * NEK * 64, nfrag = nfe + TO * NOK * 64;
    for (int e = threadIdx.x; e < nfrag; e += blockDim.x) lds[e] = e < nfe ? frag[e] : fragO[e - nfe];
    __syncthreads();
    const double *GE = lds, *GO = lds + TE * NEK * 64;
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int j = lane & 15, kq = lane >> 4;
    const int ntc = (ncol + 15) >> 4, ntiles = nb * ntc;
    const int Ke = (NZAA + 1) / 2, Ko = NZAA / 2, NE = (NZ + 1) / 2, NO = NZ / 2;
    for (int tile = blockIdx.x * nw + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * nw) {
        const int b = tile / ntc, col = (tile - b * ntc) * 16 + j;
        const bool cv = col < ncol;
        const double *Xb = X + (size_t)b * NZ * ncol + (cv ? col : 0);
        d4 ae[TE], ao[TO];
#pragma unroll
        for (int r = 0; r < TE; ++r) ae[r] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < TO; ++r) ao[r] = d4{0.0, 0.0, 0.0, 0.0};
        // B operands ZPF_BWD k-steps ahead, in a ring whose slot is the step's
        // position in an unrolled group of ZPF_BWD steps; loaded values stay raw
        // until their step
        auto ld_e = [&](int kk) -> double {
            const int m = 4 * kk + kq;
            return Xb[(size_t)(cv && m < NE ? 2 * m : 0) * ncol];
        };
        auto ld_o = [&](int kk) -> double {
            const int m = 4 * kk + kq;
            return Xb[(size_t)(cv && m < NO ? 2 * m + 1 : 0) * ncol];
        };
        double re[ZPF_BWD], ro[ZPF_BWD];
#pragma unroll
        for (int u = 0; u < ZPF_BWD; ++u) {
            re[u] = ld_e(u);
            ro[u] = ld_o(u);
        }
        for (int k0 = 0; k0 < NEK; k0 += ZPF_BWD) {
#pragma unroll
            for (int u = 0; u < ZPF_BWD; ++u) {
                const int kk = k0 + u;
                if (kk < NEK) {
                    const int m = 4 * kk + kq;
                    const double be = (cv && m < NE) ? re[u] : 0.0;
                    const double bo = (cv && m < NO) ? ro[u] : 0.0;
#pragma unroll
                    for (int r = 0; r < TE; ++r) ae[r] = mfma(GE[(r * NEK + kk) * 64 + lane], be, ae[r]);
                    if (kk < NOK) {
#pragma unroll
                        for (int r = 0; r < TO; ++r) ao[r] = mfma(GO[(r * NOK + kk) * 64 + lane], bo, ao[r]);
                    }
                }
                re[u] = ld_e(kk + ZPF_BWD);
                ro[u] = ld_o(kk + ZPF_BWD);
            }
        }
        if (cv) {
            double *Fb = F + (size_t)b * NZAA * ncol + col;
#pragma unroll
            for (int r = 0; r < TE; ++r)
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int z = 16 * (r0 + r) + 4 * it + kq;
                    const double E = ae[r][it];
                    const double O = (r < TO) ? ao[r < TO ? r : 0][it] : 0.0;
                    if (z < Ke) Fb[(size_t)z * ncol] = E + O;
                    if (z < Ko) Fb[(size_t)(NZAA - 1 - z) * ncol] = E - O;
                }
        }
    }
}

// fragE / fragO: the fragments of the RE even and RO odd output row tiles this
// launch computes, starting at tiles re0 / ro0
template <int RE, int RO>
__global__ __launch_bounds__(ZTHREADS) void k_zfwd_mfma(const double *__restrict__ frag,
                                                    const double *__restrict__ fragO, int KEK, int KOK,
                                                    const double *__restrict__ Y, double *__restrict__ Q, int NZ,
                                                    int NZAA, int ncol, int nb, int re0, int ro0) {
    extern __shared__ double lds[];
    const int nfe = RE * KEK * 64, nfrag = nfe + RO * KOK * 64;
    for (int e = threadIdx.x; e < nfrag; e += blockDim.x) lds[e] = e < nfe ? frag[e] : fragO[e - nfe];
    __syncthreads();
    const double *FE = lds, *FO = lds + RE * KEK * 64;
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int j = lane & 15, kq = lane >> 4;
    const int ntc = (ncol + 15) >> 4, ntiles = nb * ntc;
    const int Ke = (NZAA + 1) / 2, Ko = NZAA / 2;
    for (int tile = blockIdx.x * nw + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * nw) {
        const int b = tile / ntc, col = (tile - b * ntc) * 16 + j;
        const bool cv = col < ncol;
        const double *Yb = Y + (size_t)b * NZAA * ncol + (cv ? col : 0);
        d4 ae[RE], ao[RO];
#pragma unroll
        for (int r = 0; r < RE; ++r) ae[r] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < RO; ++r) ao[r] = d4{0.0, 0.0, 0.0, 0.0};
        // rows z and NZAA-1-z of the B operand ZPF_FWD k-steps ahead (ring as in
        // the synthesis), folded into S and D at their step
        auto ld0 = [&](int kk) -> double {
            const int z = 4 * kk + kq;
            return Yb[(size_t)(cv && z < Ke ? z : 0) * ncol];
        };
        auto ld1 = [&](int kk) -> double {
            const int z = 4 * kk + kq;
            return Yb[(size_t)(cv && z < Ko ? NZAA - 1 - z : 0) * ncol];
        };
        double r0[ZPF_FWD], r1[ZPF_FWD];
#pragma unroll
        for (int u = 0; u < ZPF_FWD; ++u) {
            r0[u] = ld0(u);
            r1[u] = ld1(u);
        }
        for (int k0 = 0; k0 < KEK; k0 += ZPF_FWD) {
#pragma unroll
            for (int u = 0; u < ZPF_FWD; ++u) {
                const int kk = k0 + u;
                if (kk < KEK) {
                    const int z = 4 * kk + kq;
                    const bool oh = cv && z < Ke, om = cv && z < Ko;
                    const double h = oh ? r0[u] : 0.0, hm = om ? r1[u] : 0.0;
                    const double S = h + hm, D = om ? h - hm : 0.0;
#pragma unroll
                    for (int r = 0; r < RE; ++r) ae[r] = mfma(FE[(r * KEK + kk) * 64 + lane], S, ae[r]);
                    if (kk < KOK) {
#pragma unroll
                        for (int r = 0; r < RO; ++r) ao[r] = mfma(FO[(r * KOK + kk) * 64 + lane], D, ao[r]);
                    }
                }
                r0[u] = ld0(kk + ZPF_FWD);
                r1[u] = ld1(kk + ZPF_FWD);
            }
        }
        if (cv) {
            double *Qb = Q + (size_t)b * NZ * ncol + col;
#pragma unroll
            for (int r = 0; r < RE; ++r)
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int n = 2 * (16 * (re0 + r) + 4 * it + kq);
                    if (n < NZ) Qb[(size_t)n * ncol] = ae[r][it];
                }
#pragma unroll
            for (int r = 0; r < RO; ++r)
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int n = 2 * (16 * (ro0 + r) + 4 * it + kq) + 1;
                    if (n < NZ) Qb[(size_t)n * ncol] = ao[r][it];
                }
        }
    }
}

// ---- host side ----

static int lds_limit(int device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) v = 65536;
    return v;
}

static void fill_frags(std::vector<double> &out, int ntile, int nk, const std::function<double(int, int)> &A, int M,
                       int K) {
    for (int t = 0; t < ntile; ++t)
        for (int kk = 0; kk < nk; ++kk)
            for (int lane = 0; lane < 64; ++lane) {
                const int m = 16 * t + (lane & 15), k = 4 * kk + (lane >> 4);
                out.push_back((m < M && k < K) ? A(m, k) : 0.0);
            }
}

// the (even, odd) row-tile counts a launch can be compiled for
#define ZBWD_SHAPES(X) X(1, 1) X(2, 1) X(2, 2) X(3, 2) X(3, 3) X(4, 3) X(4, 4) X(5, 5) X(7, 6)
#define ZFWD_SHAPES(X) X(1, 1) X(2, 1) X(2, 2) X(3, 2) X(3, 3) X(4, 3) X(5, 4)

static bool zbwd_shape_ok(int ne, int no) {
#define ZS(a, b) if (ne == a && no == b) return true;
    ZBWD_SHAPES(ZS)
#undef ZS
    return false;
}
static bool zfwd_shape_ok(int ne, int no) {
#define ZS(a, b) if (ne == a && no == b) return true;
    ZFWD_SHAPES(ZS)
#undef ZS
    return false;
}

bool make_zfrags(int NZ, int NZAA, const std::vector<double> &Tz, const std::vector<double> &Cz, int device,
                 ZFrags &zf) {
    const int Ke = (NZAA + 1) / 2, Ko = NZAA / 2, NE = (NZ + 1) / 2, NO = NZ / 2;
    zf = ZFrags{};
    // synthesis: rows z < Ke (even part) / z < Ko (odd part), k over even / odd n
    zf.TE = (Ke + 15) / 16;
    zf.TO = (Ko + 15) / 16;
    zf.NEK = (NE + 3) / 4;
    zf.NOK = (NO + 3) / 4;
    fill_frags(zf.bwd, zf.TE, zf.NEK, [&](int z, int m) { return Tz[(size_t)z * NZ + 2 * m]; }, Ke, NE);
    fill_frags(zf.bwd, zf.TO, zf.NOK, [&](int z, int m) { return Tz[(size_t)z * NZ + 2 * m + 1]; }, Ko, NO);
    // analysis: rows over even / odd n, k over z < Ke / z < Ko
    zf.RE = (NE + 15) / 16;
    zf.RO = (NO + 15) / 16;
    zf.KEK = (Ke + 3) / 4;
    zf.KOK = (Ko + 3) / 4;
    fill_frags(zf.fwd, zf.RE, zf.KEK, [&](int m, int z) { return Cz[(size_t)(2 * m) * NZAA + z]; }, NE, Ke);
    fill_frags(zf.fwd, zf.RO, zf.KOK, [&](int m, int z) { return Cz[(size_t)(2 * m + 1) * NZAA + z]; }, NO, Ko);
    const size_t lim = (size_t)lds_limit(device);
    // synthesis: part p takes row tiles [r0, r1) of both halves (an even and
    // an odd tile of the same r feed the same mirrored output rows)
    for (int P = 1; P <= ZMAXPARTS && !zf.bwd_ok; ++P) {
        bool ok = true;
        for (int p = 0; p < P && ok; ++p) {
            const int r0 = p * zf.TE / P, r1 = (p + 1) * zf.TE / P;
            const int ne = r1 - r0, no = std::max(0, std::min(r1, zf.TO) - r0);
            ok = ne > 0 && no > 0 && zbwd_shape_ok(ne, no) &&
                 size_t(ne * zf.NEK + no * zf.NOK) * 64 * sizeof(double) <= lim;
            zf.bp[p] = ZPart{r0, ne, r0, no};
        }
        if (ok) {
            zf.bwd_ok = true;
            zf.nbp = P;
        }
    }
    // analysis: even and odd output rows are independent; part p takes its
    // share of each
    for (int P = 1; P <= ZMAXPARTS && !zf.fwd_ok; ++P) {
        bool ok = true;
        for (int p = 0; p < P && ok; ++p) {
            const int e0 = p * zf.RE / P, e1 = (p + 1) * zf.RE / P, o0 = p * zf.RO / P, o1 = (p + 1) * zf.RO / P;
            ok = e1 > e0 && o1 > o0 && zfwd_shape_ok(e1 - e0, o1 - o0) &&
                 size_t((e1 - e0) * zf.KEK + (o1 - o0) * zf.KOK) * 64 * sizeof(double) <= lim;
            zf.fp[p] = ZPart{e0, e1 - e0, o0, o1 - o0};
        }
        if (ok) {
            zf.fwd_ok = true;
            zf.nfp = P;
        }
    }
    return zf.bwd_ok || zf.fwd_ok;
}

template <int TE, int TO>
static void go_bwd(hipStream_t st, const ZFragsDev &z, const ZPart &pt, const double *X, double *F, int NZ, int NZAA,
                   int ncol, int nb, int nblocks) {
    const size_t shm = size_t(TE * z.NEK + TO * z.NOK) * 64 * sizeof(double);
    if (shm > 65536) set_dynamic_lds((const void *)k_zbwd_mfma<TE, TO>, shm, "z synthesis");
    const double *fe = z.bwd + size_t(pt.e0) * z.NEK * 64, *fo = z.bwd + size_t(z.TE * z.NEK + pt.o0 * z.NOK) * 64;
    hipLaunchKernelGGL((k_zbwd_mfma<TE, TO>), dim3(nblocks), dim3(ZTHREADS), shm, st, fe, fo, z.NEK, z.NOK, X, F, NZ,
                       NZAA, ncol, nb, pt.e0);
}

template <int RE, int RO>
static void go_fwd(hipStream_t st, const ZFragsDev &z, const ZPart &pt, const double *Y, double *Q, int NZ, int NZAA,
                   int ncol, int nb, int nblocks) {
    const size_t shm = size_t(RE * z.KEK + RO * z.KOK) * 64 * sizeof(double);
    if (shm > 65536) set_dynamic_lds((const void *)k_zfwd_mfma<RE, RO>, shm, "z analysis");
    const double *fe = z.fwd + size_t(pt.e0) * z.KEK * 64, *fo = z.fwd + size_t(z.RE * z.KEK + pt.o0 * z.KOK) * 64;
    hipLaunchKernelGGL((k_zfwd_mfma<RE, RO>), dim3(nblocks), dim3(ZTHREADS), shm, st, fe, fo, z.KEK, z.KOK, Y, Q, NZ,
                       NZAA, ncol, nb, pt.e0, pt.o0);
}

static int grid_for(int ntiles, int ncu) {
    const int want = (ntiles + ZTHREADS / 64 - 1) / (ZTHREADS / 64);
    return want < ncu ? want : ncu;
}

bool launch_zbwd_mfma(hipStream_t st, const ZFragsDev &z, const double2 *spec, double2 *Fz, int NZ, int NZAA, int nsys,
                      int nb) {
    if (!z.bwd_ok) return false;
    const int ncol = 2 * nsys, ntiles = nb * ((ncol + 15) / 16), g = grid_for(ntiles, z.ncu);
    const double *X = (const double *)spec;
    double *F = (double *)Fz;
    for (int p = 0; p < z.nbp; ++p) {
        const ZPart &pt = z.bp[p];
#define ZS(a, b) \
        if (pt.ne == a && pt.no == b) { go_bwd<a, b>(st, z, pt, X, F, NZ, NZAA, ncol, nb, g); continue; }
        ZBWD_SHAPES(ZS)
#undef ZS
        return false;   // not reached: the plan holds compiled shapes only
    }
    check_launch("z transform");
    return true;
}

bool launch_zfwd_mfma(hipStream_t st, const ZFragsDev &z, const double2 *Hy, double2 *Qhat, int NZ, int NZAA,
                      int nsys, int nb) {
    if (!z.fwd_ok) return false;
    const int ncol = 2 * nsys, ntiles = nb * ((ncol + 15) / 16), g = grid_for(ntiles, z.ncu);
    const double *Y = (const double *)Hy;
    double *Q = (double *)Qhat;
    for (int p = 0; p < z.nfp; ++p) {
        const ZPart &pt = z.fp[p];
#define ZS(a, b) \
        if (pt.ne == a && pt.no == b) { go_fwd<a, b>(st, z, pt, Y, Q, NZ, NZAA, ncol, nb, g); continue; }
        ZFWD_SHAPES(ZS)
#undef ZS
        return false;   // not reached: the plan holds compiled shapes only
    }
    check_launch("z transform");
    return true;
}

}  // namespace coral
