/*
 * ORACLE -- test infrastructure only.  Nothing in the product links or calls
 * this file; tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
 * it as the checker.
 *
 * CPU restatement of the per-(kx,ky) banded LU factor and solve that Coral
 * performs with LAPACK:
 *   - factor: zgbtrf called per wavenumber in
 *     src/plane_layer/PL_algebra_z_3D.f90:245-276 (band copied into rows
 *     kl+1..2kl+ku+1 of LDAB = 2kl+ku+1 at :258);
 *   - solve:  zgbtrs('N', n, kl, ku, nrhs=1, ...) per wavenumber in
 *     src/plane_layer/PL_algebra_z_3D.f90:189-243;
 *   - the real mean-field systems use dgbtrf/dgbtrs the same way
 *     (src/plane_layer/PL_algebra_d_1D.f90:127-215).
 * LAPACK is a third-party dependency (Intel MKL 2024.0 in the reference's
 * environment, etc/set_env_examples/set_env_template.sh:29).  This file restates
 * the published unblocked algorithm (LAPACK 3.x xGBTF2 / xGBTRS): partial
 * pivoting on max(|Re|+|Im|) (IZAMAX/DCABS1), reciprocal-pivot column scaling,
 * rank-1 update limited to the fill extent JU.  Band storage is LAPACK's:
 * element A(i,j) (0-based) lives at ab[j*ldab + kl+ku+i-j].
 * Batched over systems with OpenMP; every system uses the same n, kl, ku.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef double _Complex zc;

static inline double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* xGBTF2 on one system.  Returns LAPACK info (0 ok, j+1 if U(j,j) == 0). */
static int zgbtf2_one(int n, int kl, int ku, zc *ab, int ldab, int32_t *ipiv)
{
    const int kv = ku + kl;
    int info = 0;
#define AB(i, j) ab[(size_t)(j) * ldab + (i)]
    /* zero the fill-in part of the first columns (xGBTF2: columns ku+2..kv) */
    for (int j = ku + 1; j < (kv < n ? kv : n); ++j)
        for (int i = kv - j; i < kl; ++i)
            AB(i, j) = 0.0;
    int ju = 0;
    for (int j = 0; j < n; ++j) {
        /* zero the fill-in elements of column j+kv */
        if (j + kv < n)
            for (int i = 0; i < kl; ++i) AB(i, j + kv) = 0.0;
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        /* pivot: first index of max |Re|+|Im| among rows j..j+km of column j */
        int jp = 0;
        double best = cabs1(AB(kv, j));
        for (int i = 1; i <= km; ++i) {
            double v = cabs1(AB(kv + i, j));
            if (v > best) { best = v; jp = i; }
        }
        ipiv[j] = j + jp + 1; /* 1-based, as LAPACK */
        if (AB(kv + jp, j) != 0.0) {
            int cand = j + ku + jp;
            if (cand > n - 1) cand = n - 1;
            if (cand > ju) ju = cand;
            if (jp != 0) {
                /* swap rows j and j+jp over columns j..ju */
                for (int c = j; c <= ju; ++c) {
                    zc t = AB(kv + jp + j - c, c);
                    AB(kv + jp + j - c, c) = AB(kv + j - c, c);
                    AB(kv + j - c, c) = t;
                }
            }
            if (km > 0) {
                zc rp = 1.0 / AB(kv, j);
                for (int i = 1; i <= km; ++i) AB(kv + i, j) *= rp;
                if (ju > j) {
                    /* rank-1 update of rows j+1..j+km, columns j+1..ju */
                    for (int c = j + 1; c <= ju; ++c) {
                        zc u = AB(kv + j - c, c);
                        if (u != 0.0)
                            for (int i = 1; i <= km; ++i)
                                AB(kv + j + i - c, c) -= AB(kv + i, j) * u;
                    }
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
    }
#undef AB
    return info;
}

/* xGBTRS('N') with one right-hand side. */
static void zgbtrs_one(int n, int kl, int ku, const zc *ab, int ldab, const int32_t *ipiv, zc *b)
{
    const int kd = ku + kl;
#define AB(i, j) ab[(size_t)(j) * ldab + (i)]
    if (kl > 0) {
        for (int j = 0; j < n - 1; ++j) {
            int lm = kl < n - 1 - j ? kl : n - 1 - j;
            int l = ipiv[j] - 1;
            if (l != j) { zc t = b[l]; b[l] = b[j]; b[j] = t; }
            zc bj = b[j];
            for (int i = 1; i <= lm; ++i) b[j + i] -= AB(kd + i, j) * bj;
        }
    }
    /* back substitution with U (upper band of width kl+ku), column oriented (ZTBSV) */
    for (int j = n - 1; j >= 0; --j) {
        if (b[j] != 0.0) {
            b[j] /= AB(kd, j);
            zc t = b[j];
            int i0 = j - kd > 0 ? j - kd : 0;
            for (int i = j - 1; i >= i0; --i) b[i] -= t * AB(kd + i - j, j);
        }
    }
#undef AB
}

/* Batched complex factorisation: nsys systems stored back to back
 * (each ldab*n entries).  info[s] receives each system's LAPACK info. */
int oracle_zgbtrf_batch(int n, int kl, int ku, int nsys, double *ab_, int ldab,
                        int32_t *ipiv, int32_t *info)
{
    zc *ab = (zc *)ab_;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : bad)
    for (int s = 0; s < nsys; ++s) {
        info[s] = zgbtf2_one(n, kl, ku, ab + (size_t)s * ldab * n, ldab, ipiv + (size_t)s * n);
        bad += info[s] != 0;
    }
    return bad;
}

int oracle_zgbtrs_batch(int n, int kl, int ku, int nsys, const double *ab_, int ldab,
                        const int32_t *ipiv, double *b_, int ldb)
{
    const zc *ab = (const zc *)ab_;
    zc *b = (zc *)b_;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < nsys; ++s)
        zgbtrs_one(n, kl, ku, ab + (size_t)s * ldab * n, ldab, ipiv + (size_t)s * n,
                   b + (size_t)s * ldb);
    return 0;
}

/* Banded complex mat-vec y = A x for nsys systems whose band matrices are
 * sum_p coef[s,p] * B_p with real, shared pieces B_p stored densely by diagonal:
 * B_p(i, i+d) = band[p][(d+nl) * n + i] for d in [-nl, nu].  This restates the
 * per-wavenumber CSR products of src/plane_layer/PL_algebra_z_3D.f90:127-157
 * (mkl_zcsrmv), whose matrices are the same sums of pieces
 * (PL_IMEX_timestepping.f90:1500-1574). */
int oracle_band_matvec_batch(int n, int nl, int nu, int npieces, const double *band,
                             int nsys, const double *coef_, const double *x_, int ldx,
                             double *y_, int ldy)
{
    const zc *coef = (const zc *)coef_;
    const zc *x = (const zc *)x_;
    zc *y = (zc *)y_;
    const int w = nl + nu + 1;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < nsys; ++s) {
        const zc *xs = x + (size_t)s * ldx;
        zc *ys = y + (size_t)s * ldy;
        for (int i = 0; i < n; ++i) ys[i] = 0.0;
        for (int p = 0; p < npieces; ++p) {
            zc c = coef[(size_t)s * npieces + p];
            if (c == 0.0) continue;
            const double *bp = band + (size_t)p * w * n;
            for (int i = 0; i < n; ++i) {
                zc acc = 0.0;
                int dlo = -nl, dhi = nu;
                if (i + dlo < 0) dlo = -i;
                if (i + dhi > n - 1) dhi = n - 1 - i;
                for (int d = dlo; d <= dhi; ++d) acc += bp[(size_t)(d + nl) * n + i] * xs[i + d];
                ys[i] += c * acc;
            }
        }
    }
    return 0;
}
