// Host-side operator assembly: turns the parsed recipe into the tables the
// device kernels consume.  Setup-time only (not on the timed path).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/coral_amd.h"

namespace coral {

struct Dense {
    int rows = 0, cols = 0;
    std::vector<double> a;
    Dense() = default;
    Dense(int r, int c) : rows(r), cols(c), a(size_t(r) * c, 0.0) {}
    double &operator()(int i, int j) { return a[size_t(i) * cols + j]; }
    double operator()(int i, int j) const { return a[size_t(i) * cols + j]; }
};

Dense matmul(const Dense &A, const Dense &B);
Dense identity(int n);

// Chebyshev integration / multiplication matrices of size N+padding
// (chebyshev_elementary.f90:28-151).
void chebyshev_elementary(int N, double center, double gap, int padding, Dense &I, Dense &M);
// trunc_{NZ x NZ}(I^iz M^mz) with padding 15 (PL_IMEX_timestepping.f90:726-791).
Dense cheb_IM(int NZ, double center, double gap, int iz, int mz);
// Galerkin stencil N x (N - bc/10) (chebyshev_galerkin_2.f90:7-440).
Dense galerkin_stencil(int N, int bc);
// d/dz on NZ Chebyshev coefficients (PL_IMEX_timestepping.f90:276-319, :488-506).
Dense chebyshev_derivative(int NZ, double center, double gap);

// One NL term, in the form the assembly kernel reads: for row r of equation
// `eqn`, value = dsca * (i kx)^kx (i ky)^ky * sum_{d=-w..w} W[r][d+w] * Qhat[qvar][r+order+d].
struct NLTable {
    int eqn, qvar, kx, ky, w, order;
    double dsca;
    int woff;  // offset into SystemTables::nl_w
};

// A linear operator as groups of real matrices, each multiplied by a sum of
// horizontal monomials: Op(kx,ky) = sum_g (sum_{m in g} (i kx)^a_m (i ky)^b_m) B_g.
// Pieces whose matrices are identical share one group (e.g. the kx^2 and ky^2
// parts of a Laplacian).
struct OpTables {
    int ngroups = 0;
    std::vector<int> mono_off;         // [ngroups+1] into mono_a/mono_b
    std::vector<int> mono_a, mono_b;
    std::vector<double> band;          // [g][total][w]: B_g(i, i+d) at (g*total + i)*w + d + nl
    std::vector<int> ell_w, ell_off;   // per group: entries per row, offset of the group's block
    std::vector<int> ell_col;          // [off + i*ell_w + e] column, padded with i
    std::vector<double> ell_val;       // padded with 0
};

struct SystemTables {
    bool kxky = true;
    int nvars = 0, total = 0, nl = 0, nu = 0;
    std::vector<int> var_size, eqn_size, var_first, eqn_first, eqn_orders, bc_codes;
    std::vector<int> posP, posQ;          // var-order / eqn-order index -> interleaved position
    std::vector<int> pos_eqn, pos_row;    // interleaved position -> (equation, row in equation)
    OpTables mass, stif;
    std::vector<NLTable> nl_terms;        // sorted by equation
    std::vector<int> nl_eqn_begin;        // [nvars+1]
    std::vector<double> nl_w;
    // Real form of the evolution operator: with variable v scaled by i^phi_v
    // and equation e by i^-psi_e (phi, psi in {0, 1}), every piece
    // (i kx)^a (i ky)^b becomes real when a + b + phi_v + psi_e is even for all
    // pieces.  rot[p] = psi of the equation at position p | phi of the variable
    // at position p << 1.
    bool real_form = false;
    std::vector<unsigned char> rot;
    // Structural blocks of the evolution operator M - c L: the bipartite graph
    // (row i, column j) of entries that are nonzero in some group of M or L
    // splits into nblk connected components of equal row and column counts, so
    // that after permutation the operator is block diagonal for every (kx, ky)
    // and c (for Coral's operators, even and odd Chebyshev degrees).  Block b
    // holds rows blk_row[b][0..] and columns blk_col[b][0..] in increasing
    // order, padded to nblk_n entries with -1; inside a block the rows and
    // columns form a band (blk_nl, blk_nu).  nblk = 1 when there is no split.
    int nblk = 1, nblk_n = 0, blk_nl = 0, blk_nu = 0;
    std::vector<int> blk_row, blk_col;   // [nblk][nblk_n]
    int band_width() const { return nl + nu + 1; }
};

// Extraction of a full-variable term: Chebyshev coefficient n of the term is
// sum_{k in [rowptr[n], rowptr[n+1])} val[k] * state[col[k]], times the
// horizontal monomial (i kx)^kx (i ky)^ky.  dsca and d/dz are folded in val.
struct ExtractTable {
    int full = 0, kind = 0, system = 0, kx = 0, ky = 0;
    std::vector<int> rowptr, col;
    std::vector<double> val;
};

struct Problem {
    int NX = 0, NY = 0, NZ = 0, NXAA = 0, NYAA = 0, NZAA = 0;
    double Lx = 0, Ly = 0, gap = 1, center = 0.5, kx_box = 0, ky_box = 0;
    int nstages = 0;
    std::vector<double> a_arr, a_hat;   // s x s and (s+1) x (s+1), row-major
    std::vector<SystemTables> kxky, zero;
    int nfull = 0, nquad = 0;
    std::vector<int> quad_a, quad_b, quad_filter;
    std::vector<ExtractTable> extract;  // kxky terms then mean terms
    // transforms
    std::vector<double> Tz;  // [NZAA][NZ]  T_n(x_j)
    std::vector<double> Cz;  // [NZ][NZAA]  forward DCT incl. 1/(NXAA NYAA) and the n=0 half
    std::vector<double> zweight;  // [NZAA] Gauss-Chebyshev midpoint weights
    // kept wavenumbers: s = ix * nky + jy, ky_int of jy, slot in the NYAA-long y line
    int nkx = 0, nky = 0, nsys = 0;
    std::vector<int> kx_int, ky_int, ky_slot;
};

// Build everything; throws std::runtime_error with a message on bad input.
Problem assemble(const coral_amd_problem &p);

}  // namespace coral
