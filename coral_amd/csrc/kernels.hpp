// Kernel launch wrappers (device code in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

#include <functional>
#include <vector>

namespace coral {

// ---- extraction: spec[n][s] (+)= mono(s) * sum val[k] * state[s][col[k]] ----
void launch_extract(hipStream_t st, const double2 *state, int total, int nsys, int NZ, const int *rowptr,
                    const int *col, const double *val, int kx, int ky, KeptModes km, double2 *spec, int accumulate);
// all kxky terms of all full variables in one pass over the state: a
// workgroup owns xt systems and produces every full variable for them
struct ExtractTerm {
    const int *rowptr, *col;
    const double *val;
    const double2 *state;
    int total, kx, ky;
    int sidx;                               // index of state among ExtractAllArgs::state
    int soff;                               // sum of stotal over the states before it
};
constexpr int MAX_EXTRACT_STATES = 4;
constexpr int MAX_EXTRACT_TERMS = 24;
constexpr int MAX_FULL = 16;
struct ExtractAllArgs {
    int nterms, nfull;
    ExtractTerm term[MAX_EXTRACT_TERMS];    // sorted by full variable
    int full_first[MAX_FULL + 1];           // terms of full f: [full_first[f], full_first[f+1])
    int nsys, NZ;
    KeptModes km;
    double2 *spec;                          // [nfull][NZ][nsys]
    float2 *spec32;                         // when set, the output goes here in FP32 instead
    // distinct state arrays read by the terms ((total, NY, NX/2) each)
    int nstate;
    const double2 *state[MAX_EXTRACT_STATES];
    int stotal[MAX_EXTRACT_STATES];
};
void launch_extract_all(hipStream_t st, const ExtractAllArgs &a);
// mean contributions into the (0,0) column: spec[n][0] += sum val[k] * Re(zstate[col[k]])
void launch_extract_mean(hipStream_t st, const double2 *zstate, int NZ, int nsys, const int *rowptr, const int *col,
                         const double *val, double2 *spec);
void launch_extract_mean32(hipStream_t st, const double2 *zstate, int NZ, int nsys, const int *rowptr, const int *col,
                           const double *val, float2 *spec);

// ---- z transforms ----
// Layouts (z-major): spec[b][n][s], Fz[b][z][s], Hy[b][z][s], Qhat[b][n][s] (complex s)
// Fz[b][z][s] = sum_n Tz[z][n] spec[b][n][s]          (plain FP64 GEMM fallback)
void launch_zbwd(hipStream_t st, const double *Tz, const double2 *spec, double2 *Fz, int NZAA, int NZ, int nsys,
                 int nbatch);
// Qhat[b][n][s] = sum_z Cz[n][z] Hy[b][z][s]
void launch_zfwd(hipStream_t st, const double *Cz, const double2 *Hy, double2 *Qhat, int NZ, int NZAA, int nsys,
                 int nbatch);
// MFMA fragments of the folded (even/odd) transform matrices
// A launch of a z transform over a part of the output row tiles: even tiles
// [e0, e0 + ne), odd tiles [o0, o0 + no).  All tiles go in one part when their
// fragments fit one workgroup's LDS, else in parts that do.
struct ZPart {
    int e0, ne, o0, no;
};
constexpr int ZMAXPARTS = 8;
struct ZFrags {
    int TE = 0, TO = 0, NEK = 0, NOK = 0;  // synthesis: row tiles (even, odd), k steps (even, odd)
    int RE = 0, RO = 0, KEK = 0, KOK = 0;  // analysis
    bool bwd_ok = false, fwd_ok = false;    // an MFMA plan exists (one part or several)
    int nbp = 0, nfp = 0;                   // parts of the synthesis / analysis
    ZPart bp[ZMAXPARTS], fp[ZMAXPARTS];
    std::vector<double> bwd, fwd;
};
struct ZFragsDev {
    int TE, TO, NEK, NOK, RE, RO, KEK, KOK;
    bool bwd_ok, fwd_ok;
    int ncu;
    const double *bwd, *fwd;
    int nbp, nfp;
    ZPart bp[ZMAXPARTS], fp[ZMAXPARTS];
};
bool make_zfrags(int NZ, int NZAA, const std::vector<double> &Tz, const std::vector<double> &Cz, int device,
                 ZFrags &zf);
// return false (nothing launched) when no MFMA variant covers the shape
bool launch_zbwd_mfma(hipStream_t st, const ZFragsDev &z, const double2 *spec, double2 *Fz, int NZ, int NZAA, int nsys,
                      int nb);
bool launch_zfwd_mfma(hipStream_t st, const ZFragsDev &z, const double2 *Hy, double2 *Qhat, int NZ, int NZAA,
                      int nsys, int nb);

// FP32 z transforms on MFMA with compensated accumulation (ztransform32.hip):
// the same folded products in single precision; fragments are float
struct ZFrags32 {
    int TE = 0, TO = 0, NEK = 0, NOK = 0, RE = 0, RO = 0, KEK = 0, KOK = 0;
    bool bwd_ok = false, fwd_ok = false;
    int nbp = 0, nfp = 0;
    ZPart bp[ZMAXPARTS], fp[ZMAXPARTS];
    std::vector<float> bwd, fwd;
};
struct ZFrags32Dev {
    int TE, TO, NEK, NOK, RE, RO, KEK, KOK;
    bool bwd_ok, fwd_ok;
    int ncu;
    const float *bwd, *fwd;
    int nbp, nfp;
    ZPart bp[ZMAXPARTS], fp[ZMAXPARTS];
};
// false when no compiled shape covers NZ (both directions needed)
bool make_zfrags32(int NZ, int NZAA, const std::vector<double> &Tz, const std::vector<double> &Cz, size_t lds_limit,
                   ZFrags32 &zf);
void launch_zbwd32(hipStream_t st, const ZFrags32Dev &z, const float2 *spec, float2 *Fz, int NZ, int NZAA, int nsys,
                   int nb);
void launch_zfwd32(hipStream_t st, const ZFrags32Dev &z, const float2 *Hy, float2 *Qhat, int NZ, int NZAA, int nsys,
                   int nb);

// ---- FFT passes ----
struct FFTPlan {
    int N = 0, nrad = 0;
    int radix[16];
    const double2 *tw = nullptr;  // device, N entries exp(-2 pi i k / N)
    const float2 *twf = nullptr;  // the same rounded to FP32 (FP32 transform chain)
};
// y inverse: G[b][z][y][ix] = sum_jy Fz[b][z][ix*nky+jy] e^{+i ky y}
void launch_ybwd(hipStream_t st, const FFTPlan &py, const double2 *Fz, double2 *G, int nbatch, int NZAA,
                 int nkx, int nky, int NY);
// x pass: c2r of every full variable, products, CFL maxima, r2c of every quadratic variable
struct XPassArgs {
    const double2 *G;    // [nfull][NZAA][NYAA][nkx]
    double2 *H;          // [nquad][NZAA][NYAA][nkx]
    int fp32;            // G and H hold float2 (the FP32 chain; wavefront plans only)
    int nfull, nquad, NZAA, NYAA, nkx;
    const int *qa, *qb;  // quadratic factors (-1 = ONE)
    const double *zweight;
    unsigned long long *cflmax;  // [3] bit patterns of non-negative doubles
    int max_wpb;                 // wavefront kernels: at most this many wavefronts per workgroup (0 = 4)
};
void launch_xpass(hipStream_t st, const FFTPlan &px, const XPassArgs &a);
// wavefront-per-line x pass for N in {96, 144, 192, 384, 768} and up to 6 full
// variables; returns false (nothing launched) for other shapes
bool launch_xpass_wave(hipStream_t st, const FFTPlan &px, const XPassArgs &a);
// the x transforms alone on the x pass's wavefront plans (FP64, N = 96, 144,
// 192, 384, 768); false (nothing launched) for other lengths
bool launch_xbwd_wave(hipStream_t st, const FFTPlan &px, const double2 *G, double *phys, int nbatch, long long nlines,
                      int nkx);
bool launch_xfwd_wave(hipStream_t st, const FFTPlan &px, const double *phys, double2 *H, int nbatch, long long nlines,
                      int nkx);
// wavefront-FFT y passes for NYAA in {96, 144, 192, 384, 768}; false (nothing launched) otherwise
bool launch_ypass_wave(hipStream_t st, bool fwd, const FFTPlan &py, const double2 *in, double2 *out, int nbatch,
                       int NZAA, int nkx, int nky, int NY);
// the same passes on float2 arrays (the FP32 chain; twiddles py.twf), forward
// or inverse; throws when no wavefront plan covers NYAA
void launch_ypass32(hipStream_t st, bool fwd, const FFTPlan &py, const float2 *in, float2 *out, int nbatch, int NZAA,
                    int nkx, int nky, int NY);
// y forward: Hy[b][z][ix*nky+jy] = sum_y H[b][z][y][ix] e^{-i ky y}
void launch_yfwd(hipStream_t st, const FFTPlan &py, const double2 *H, double2 *Hy, int nbatch, int NZAA,
                 int nkx, int nky, int NY);
// c2r / r2c passes alone (diagnostics, transform seam); lines = (z, y) pairs
void launch_xbwd(hipStream_t st, const FFTPlan &px, const double2 *G, double *phys, int nbatch, long long nlines,
                 int nkx);
void launch_xfwd(hipStream_t st, const FFTPlan &px, const double *phys, double2 *H, int nbatch, long long nlines,
                 int nkx);
// field[s][pos[j]] = sum_{n<ncol} Sinv[j][n] Q[n][s] (state import of one variable)
void launch_insert_var(hipStream_t st, const double2 *Q, int nsys, int ncol, const double *Sinv, const int *pos,
                       double2 *field, int total);
// remove the z-integral of Qhat[s][:] (PL_IMEX_timestepping.f90:223-235)
void launch_remove_zmean(hipStream_t st, double2 *Qhat, int NZ, int nsys);
void launch_remove_zmean32(hipStream_t st, float2 *Qhat, int NZ, int nsys);
// boundary-value extraction (1 bottom, 2 top) and vertical decomposition (1
// barotropic, 2 baroclinic) of one full variable's Chebyshev coefficients [NZ][nsys]
void launch_full_zfilter(hipStream_t st, double2 *spec, int NZ, int nsys, int extract, int zdecomp);
void launch_full_zfilter32(hipStream_t st, float2 *spec, int NZ, int nsys, int extract, int zdecomp);

// ---- NL assembly into K_hat ----
struct NLArgs {
    const double2 *Qhat;      // [nquad][NZ][nsys]
    const float2 *Qhat32;     // when set, Qhat is read from this FP32 copy (mean systems only)
    int nsys, NZ, total, mean;
    const int *pos_eqn, *pos_row, *eqn_begin;
    const int *t_qvar, *t_kx, *t_ky, *t_w, *t_order, *t_woff;
    const double *t_dsca, *wv;
    const double *src;        // mean systems: added to each position's sum (the sources' part), or null
    KeptModes km;
    double2 *Khat;            // [nsys][total]
};
void launch_nl(hipStream_t st, const NLArgs &a);

// kxky NL assembly from a per-position program: position p runs term blocks
// [pos_ptr[p], pos_ptr[p+1]); block b adds cmono[blk_term[b]][s] * sum over
// entries [blk_ptr[b], blk_ptr[b+1]) of ent_w * Qhat[ent_qm][s]
// (ent_qm = q*NZ + m, ent_w = dsca * trunc(I^c M^d)[row][m]).
struct NLProgArgs {
    const int *pos_ptr, *blk_term, *blk_ptr, *ent_qm;
    const double *ent_w;
    const double2 *cmono;  // [nterms][nsys]
    int nterms;
    const double2 *Qhat;   // [nquad][NZ][nsys]
    const float2 *Qhat32;  // when set, Qhat is read from this FP32 copy (sums stay FP64)
    int nsys, total;
    double2 *Khat;         // [nsys][total]
};
void launch_nl_prog(hipStream_t st, const NLProgArgs &a);
// cmono[t*nsys + s] = (i kx)^kx_t (i ky)^ky_t
void launch_term_monomials(hipStream_t st, const int *t_kx, const int *t_ky, int nterms, KeptModes km, int nsys,
                           double2 *cmono);

// ---- operators: groups of real matrices times sums of horizontal monomials ----
struct OpDev {
    int ngroups;
    const int *mono_off, *mono_a, *mono_b;
    const double *band;                 // [g][total][nl+nu+1]
    const int *ell_w, *ell_off, *ell_col;
    const double *ell_val;
};

__device__ inline double2 group_coef(const OpDev &op, int g, double kx, double ky) {
    double2 c = make_double2(0.0, 0.0);
    for (int m = op.mono_off[g]; m < op.mono_off[g + 1]; ++m) c = cadd(c, monomial(op.mono_a[m], op.mono_b[m], kx, ky));
    return c;
}

// ---- right-hand sides ----
// y[s][i] = sum_g cM_g(s) (M_g u[s])[i] + sum_g cL_g(s) (L_g v[s])[i] + sum_k kb_k K_k[s][i],
// u = sum_k ma_k U_k, v = sum_k xa_k X_k  (mass_field_to_rhs + add_K_std_to_rhs +
// add_K_hat_to_rhs, PL_IMEX_timestepping_marching.f90:93-171).  With a
// diagonally implicit scheme, dt L x_j = (M x_j - rhs_j) / a_jj, so the stage
// right-hand sides need M only: rhs_i = M (x^n + sum_j c_ij x_j) - sum_j c_ij rhs_j
// + sum_j dt ah_{i+1,j} K_hat_j with c_ij = a_ij / a_jj.  zero00: y = 0 at s = 0.
struct RhsArgs {
    OpDev mass, stif;
    const double2 *cmass, *cstif;  // [nsys][ngroups] group coefficients
    int total, nsys;
    int bl, bu;                    // band of the operators: ELL columns of row i lie in [i - bl, i + bu]
    int nm;
    const double2 *mv[8];
    double ma[8];
    int nx;
    const double2 *xv[8];
    double xa[8];
    int nk;
    const double2 *kv[8];
    double kb[8];
    int zero00;
    double2 *y;
};
void launch_rhs(hipStream_t st, const RhsArgs &a);
// out[s*ngroups + g] = sum over the group's monomials of (i kx)^a (i ky)^b
void launch_group_coefs(hipStream_t st, const OpDev &op, KeptModes km, int nsys, double2 *out);

// ---- banded LU factorization and solve ----
struct FactorArgs {
    OpDev mass, stif;
    int nl, nu, total;
    double sdt;               // -dt * a11
    int s0, ns;               // systems [s0, s0+ns)
    int identity00;           // kxky (0,0) -> identity
    KeptModes km;
    double2 *work;            // [ns][total][ldab]
    double2 *Lf;              // [nsys][total][kl]
    double2 *Uf;              // [nsys][total][kv+1]  (Uf[.][j][0] = 1/U(j,j), Uf[.][j][1+k] = U(j-1-k, j))
    unsigned char *piv;       // [nsys][total] pivot offsets ([nmat][16*ceil(blk_n/16)] in the 16-lane layout)
    int *info;                // [nsys]
    // real form (rot != nullptr): entries rotated by i^(phi_col - psi_row) are
    // real; factors go to Lr/Ur (same layout, real) and work is used as double
    const unsigned char *rot; // [total] psi | phi << 1
    double *Lr, *Ur;
    int ncol16;               // > 0: multipliers in the 16-lane layout with this many columns per system
    // Structural blocks (16-lane layout only): each system is factored as nblk
    // independent blocks of blk_n unknowns, band (blk_nl, blk_nu); block b takes
    // rows blk_row[b*blk_n + i] and columns blk_col[b*blk_n + j] of the operator
    // (-1: padding, identity).  s0/ns, piv, info, Lr/Ur, Lf/Uf then count blocks
    // (block index = s*nblk + b).  nblk = 1 with null maps: the whole system.
    int nblk, blk_n, blk_nl, blk_nu;
    const int *blk_row, *blk_col;
};
void launch_factor(hipStream_t st, const FactorArgs &a);
// generic batched factorization of LAPACK band input (for the unit seam)
void launch_factor_lapack(hipStream_t st, double2 *work, int n, int kl, int ku, int nsys, double2 *Lf, double2 *Uf,
                          unsigned char *piv, int *info);
// the same factorization into the records of the 16-lane solve (ncol16 =
// solve16_ncol(n, kl, kl + ku) > 0): L16 [ceil(nsys/4)*4][ncol16][kl],
// Ufull with SOLVE16_PAD elements before it, pivots [nsys][16*ceil(n/16)];
// all three cleared by the caller
void launch_factor_lapack16(hipStream_t st, double2 *work, int n, int kl, int ku, int nsys, int ncol16, double2 *L16,
                            double2 *Ufull, unsigned char *piv, int *info);
void launch_solve(hipStream_t st, const double2 *Lf, const double2 *Uf, const unsigned char *piv, int n, int kl,
                  int kv, int nsys, const double2 *rhs, double2 *x);
// real-form solve: x = i^phi . (LU)^-1 . P . i^-psi rhs, with real factors
// 16-lane solve (four systems per wavefront): columns per system of its
// multiplier records, or 0 when the band does not fit (kl > 16, kv > 47)
int solve16_ncol(int n, int kl, int kv);
// entries per multiplier record of the 16-lane solve: the kl multipliers of a
// column (at least one entry)
__host__ __device__ inline int solve16_klp(int kl) { return kl < 1 ? 1 : kl; }
// U arrays of the solves carry this many elements of padding on each side (the
// 16-lane solve reads up to 16 elements before the first record and 64 past
// the last)
constexpr int SOLVE16_PAD = 2048;
// 16-lane solve of nsys systems of `total` unknowns, each made of nblk
// structural blocks of n unknowns and band (kl, kv - kl) (see FactorArgs);
// factors in the layout k_factor writes for ncol16 > 0.
struct Solve16Args {
    const void *L16;             // double (real form) or double2
    // U records trimmed per column (solve16_trim_u): for wavefront group g,
    // column j, matrix slot m and entry e < len_j at U + g*ustride + cb_j +
    // m*len_j + e, with ucol[j] = (cb_j, len_j); U + g*ustride + [0, 16) is zero
    const void *U;
    const int2 *ucol;
    long long ustride;
    int ncol16;
    const unsigned char *piv;    // [nsys*nblk][16*ceil(n/16)], zero past n
    // position words (solve16_words): row and column of the system at each
    // position of a block, and their real-form rotations
    const unsigned *words;
    int rotate;                  // apply the rotations (real form)
    int ident0;                  // system 0 is the identity and is not rotated
    int total, n, kl, kv, nsys, nblk;
    const double2 *rhs;
    double2 *x;
    double2 *sink;               // 64 elements of scratch that absorb stores of padding positions
};
// Trimmed U records of the 16-lane layout: the full records of kv+1 entries
// at Ufull (as k_factor writes them, nmat matrices of n unknowns, elem bytes
// per entry) are cut to the longest nonzero prefix each column has in any
// matrix.  solve16_ucol fills ucol (16*ceil(n/16) columns: offset of the
// column's four records in a group, record length) and the group stride
// (scratch_len: that many ints of device memory); solve16_trim_u_copy then
// repacks into Utrim (ngroups * ustride elements, cleared by the caller).
void solve16_ucol(hipStream_t st, const void *Ufull, int elem, int n, int kv, int nmat, int *scratch_len,
                  std::vector<int2> &ucol, long long *ustride);
void solve16_trim_u_copy(hipStream_t st, const void *Ufull, int elem, int n, int kv, int nmat, const int2 *ucol_dev,
                         long long ustride, void *Utrim);
// "no row / column" in a position word (padding of a block), and the word of
// a position outside the block
constexpr int SOLVE16_NONE = 0x7FFF;
constexpr unsigned SOLVE16_NOWORD = 0x3FFFFFFFu;
// words of padding on each side of a block's words (the solve reads up to 96
// positions before and 79 after the block)
constexpr int SOLVE16_WPAD = 128;
// Position words of the nblk blocks of blk_n positions (blk_row/blk_col as in
// FactorArgs, null for one block of the whole system; rot as in SystemTables,
// or null): row | col << 15 | psi(row) << 30 | phi(col) << 31, block b at
// b*(blk_n + 2*SOLVE16_WPAD) + SOLVE16_WPAD, padding SOLVE16_NOWORD.
std::vector<unsigned> solve16_words(int total, int nblk, int blk_n, const int *blk_row, const int *blk_col,
                                    const unsigned char *rot);
// workgroups, LDS bytes and NQ of the 16-lane solve of a (false: it does not take a)
bool solve16_shape(const Solve16Args &a, int &nwg, size_t &lds, int &nq);
// the 16-lane solves of nrun launches of one element type (real: the real
// form; complex entries must carry rotate = 0) and NQ in one launch: device
// copies of their arguments and of start[0..nrun] (cumulative workgroups),
// nwg = start[nrun] workgroups, lds the largest of theirs; results equal each
// launch's own
bool launch_solve16_batch(hipStream_t st, bool real, int nq, int nwg, size_t lds, const Solve16Args *args,
                          const int *start, int nrun);
bool launch_solve16_real(hipStream_t st, const Solve16Args &a);
bool launch_solve16_complex(hipStream_t st, const Solve16Args &a);
// (ident0: system 0 is the identity and is not rotated)
void launch_solve_real(hipStream_t st, const double *Lr, const double *Ur, const unsigned char *piv,
                       const unsigned char *rot, int ident0, int n, int kl, int kv, int nsys, const double2 *rhs,
                       double2 *x);

}  // namespace coral
