// Host-side operator assembly (setup only).  Restates, in dense/banded form, the
// operator build of Coral's plane-layer solver:
//   Chebyshev matrices      chebyshev_elementary.f90:28-151
//   Galerkin stencils       chebyshev_galerkin_2.f90:7-440
//   I^a M^b building tools  PL_IMEX_timestepping.f90:726-791 (padding 15)
//   system shapes           PL_IMEX_timestepping.f90:870-936
//   degree shuffles         PL_IMEX_timestepping.f90:1344-1414
//   blocks + change of basis PL_IMEX_timestepping.f90:793-856, PL_algebra_z_3D.f90:40-53
//   extraction / insertion  PL_IMEX_timestepping.f90:1086-1218
//   NL insertion            PL_IMEX_timestepping_marching.f90:306-350
#include "assembly.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>

namespace coral {

Dense matmul(const Dense &A, const Dense &B) {
    if (A.cols != B.rows) throw std::runtime_error("matmul: shape mismatch");
    Dense C(A.rows, B.cols);
    for (int i = 0; i < A.rows; ++i)
        for (int k = 0; k < A.cols; ++k) {
            double a = A(i, k);
            if (a == 0.0) continue;
            for (int j = 0; j < B.cols; ++j) C(i, j) += a * B(k, j);
        }
    return C;
}

Dense identity(int n) {
    Dense I(n, n);
    for (int i = 0; i < n; ++i) I(i, i) = 1.0;
    return I;
}

void chebyshev_elementary(int N, double center, double gap, int padding, Dense &I, Dense &M) {
    const int Np = N + padding;
    M = Dense(Np, Np);
    M(1, 0) = gap / 2.0;                                   // x T_0 = T_1
    for (int i = 0; i < Np - 1; ++i) M(i, i + 1) = gap / 4.0;
    for (int i = 1; i < Np - 1; ++i) M(i + 1, i) = gap / 4.0;
    for (int i = 0; i < Np; ++i) M(i, i) = center;
    I = Dense(Np, Np);
    I(1, 0) = 2.0;                                         // int T_0 = T_1
    for (int r = 2; r < Np; ++r) I(r, r - 1) = 1.0 / double(r);       // T_{r-1} -> T_r / (2r)
    for (int r = 1; r < Np - 1; ++r) I(r, r + 1) = -1.0 / double(r);  // T_{r+1} -> -T_r / (2r)
    for (double &v : I.a) v *= gap / 4.0;
}

Dense cheb_IM(int NZ, double center, double gap, int iz, int mz) {
    Dense I, M;
    chebyshev_elementary(NZ, center, gap, 15, I, M);
    Dense A = identity(I.rows);
    for (int k = 0; k < iz; ++k) A = matmul(A, I);
    for (int k = 0; k < mz; ++k) A = matmul(A, M);
    Dense out(NZ, NZ);
    for (int i = 0; i < NZ; ++i)
        for (int j = 0; j < NZ; ++j) out(i, j) = A(i, j);
    return out;
}

Dense galerkin_stencil(int N, int bc) {
    if (bc == 0) return identity(N);
    const int ncol = N - bc / 10;
    if (ncol <= 0) throw std::runtime_error("NZ too small for boundary code " + std::to_string(bc));
    Dense S(N, ncol);
    for (int j0 = 0; j0 < ncol; ++j0) {
        const double j = j0 + 1;  // the reference loops are 1-based
        switch (bc) {
        case 10: S(j0, j0) = -1.0; S(j0 + 1, j0) = 1.0; break;
        case 11: S(j0, j0) = 1.0; S(j0 + 1, j0) = 1.0; break;
        case 12: S(j0, j0) = -(j + 1.0) * (j + 1.0); S(j0 + 1, j0) = j * j; break;
        case 13: S(j0, j0) = (j + 1.0) * (j + 1.0); S(j0 + 1, j0) = j * j; break;
        case 20: S(j0, j0) = -1.0; S(j0 + 2, j0) = 1.0; break;
        case 21: {
            double n = j + 1.0;
            S(j0, j0) = -1.0;
            S(j0 + 2, j0) = (n - 2.0) * (n - 2.0) / (n * n);
            break;
        }
        case 22:
        case 23: {
            double n = j - 1.0, s = (bc == 22) ? -1.0 : 1.0;
            S(j0, j0) = 1.0;
            S(j0 + 1, j0) = s * (4.0 * n + 4.0) / (2.0 * n * n + 6.0 * n + 5.0);
            S(j0 + 2, j0) = -(n * n + (n + 1.0) * (n + 1.0)) / ((n + 1.0) * (n + 1.0) + (n + 2.0) * (n + 2.0));
            break;
        }
        case 24: {
            if (j0 < 2) {
                S(j0, j0) = 1.0;
            } else {
                double n = j + 1.0;
                S(j0 + 2, j0) = 1.0;
                S(j0, j0) = -n * n * (n * n + 1.0) / ((n * n - 3.0 * n + 2.0) * (n * n - 5.0 * n + 6.0));
            }
            break;
        }
        case 40: {
            double n = j + 3.0;
            S(j0, j0) = -1.0;
            S(j0 + 2, j0) = (2.0 * n - 4.0) / (n - 1.0);
            S(j0 + 4, j0) = -(n - 3.0) / (n - 1.0);
            break;
        }
        case 41: {
            double n = j + 3.0;
            double m = 24.0 * n / (2.0 * n * n - 4.0 * n + 3.0) - 18.0 / (n - 1.0);
            S(j0, j0) = -1.0;
            S(j0 + 2, j0) = m + 2.0;
            S(j0 + 4, j0) = -m - 1.0;
            break;
        }
        case 42:
        case 43: {
            double n = j - 1.0, s = (bc == 42) ? -1.0 : 1.0;
            double n2 = n * n, n3 = n2 * n;
            S(j0, j0) = 1.0;
            S(j0 + 1, j0) = s * (2.0 * n + 2.0) / (n2 + 5.0 * n + 7.0);
            S(j0 + 2, j0) = -(2.0 * n3 + 12.0 * n2 + 28.0 * n + 24.0) / (n3 + 8.0 * n2 + 22.0 * n + 21.0);
            S(j0 + 3, j0) = -s * (2.0 * n + 2.0) / (n2 + 5.0 * n + 7.0);
            S(j0 + 4, j0) = (n3 + 4.0 * n2 + 6.0 * n + 3.0) / (n3 + 8.0 * n2 + 22.0 * n + 21.0);
            break;
        }
        default:
            throw std::runtime_error("unknown boundary-condition code " + std::to_string(bc));
        }
    }
    return S;
}

// Solve A X = B for dense square A (Gaussian elimination, partial pivoting).
static Dense dense_solve(Dense A, Dense B) {
    const int n = A.rows;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A(i, k)) > std::fabs(A(p, k))) p = i;
        if (A(p, k) == 0.0) throw std::runtime_error("singular Chebyshev integration matrix");
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
            for (int j = 0; j < B.cols; ++j) std::swap(B(k, j), B(p, j));
        }
        for (int i = k + 1; i < n; ++i) {
            double f = A(i, k) / A(k, k);
            if (f == 0.0) continue;
            for (int j = k; j < n; ++j) A(i, j) -= f * A(k, j);
            for (int j = 0; j < B.cols; ++j) B(i, j) -= f * B(k, j);
        }
    }
    for (int k = n - 1; k >= 0; --k)
        for (int j = 0; j < B.cols; ++j) {
            double s = B(k, j);
            for (int i = k + 1; i < n; ++i) s -= A(k, i) * B(i, j);
            B(k, j) = s / A(k, k);
        }
    return B;
}

Dense chebyshev_derivative(int NZ, double center, double gap) {
    Dense I, M;
    chebyshev_elementary(NZ, center, gap, 10, I, M);
    Dense It(NZ - 1, NZ - 1), shift(NZ - 1, NZ);
    for (int i = 0; i < NZ - 1; ++i)
        for (int j = 0; j < NZ - 1; ++j) It(i, j) = I(i + 1, j);
    for (int i = 0; i < NZ - 1; ++i) shift(i, i + 1) = 1.0;
    Dense X = dense_solve(It, shift);
    Dense D(NZ, NZ);
    for (int i = 0; i < NZ - 1; ++i)
        for (int j = 0; j < NZ; ++j) D(i, j) = X(i, j);
    return D;
}

static std::vector<int> degree_shuffle(const std::vector<int> &sizes) {
    const int nv = int(sizes.size());
    const int nmin = *std::min_element(sizes.begin(), sizes.end());
    std::vector<int> pos;
    int offset = 0;
    for (int v = 0; v < nv; ++v) {
        for (int c = 0; c < nmin; ++c) pos.push_back(v + c * nv);
        for (int r = 0; r < sizes[v] - nmin; ++r) pos.push_back(nmin * nv + offset++);
    }
    return pos;
}

// cos(pi * num / den) with exact integer argument reduction
static double cospi_frac(long long num, long long den) {
    long long m = num % (2 * den);
    if (m < 0) m += 2 * den;
    return std::cos(M_PI * double(m) / double(den));
}

// Parity assignment for the real form (see SystemTables::real_form): nodes
// 0..nvars-1 are equations, nvars..2nvars-1 variables; a piece of degree a+b
// joining equation e and variable v asks parity(e) xor parity(v) = (a+b) mod 2.
static void find_real_form(const coral_amd_system &s, SystemTables &t) {
    const int nv = s.nvars;
    std::vector<std::vector<std::pair<int, int>>> adj(2 * nv);
    for (const auto &grp : {std::make_pair(s.mass, s.n_mass), std::make_pair(s.stif, s.n_stif)})
        for (int k = 0; k < grp.second; ++k) {
            const coral_amd_piece &p = grp.first[k];
            const int par = (p.kx + p.ky) & 1;
            adj[p.eqn].push_back({nv + p.var, par});
            adj[nv + p.var].push_back({p.eqn, par});
        }
    std::vector<int> val(2 * nv, -1);
    bool ok = true;
    for (int root = 0; root < 2 * nv && ok; ++root) {
        if (val[root] >= 0) continue;
        val[root] = 0;
        std::vector<int> stack{root};
        while (!stack.empty() && ok) {
            const int u = stack.back();
            stack.pop_back();
            for (auto [w, par] : adj[u]) {
                const int want = val[u] ^ par;
                if (val[w] < 0) {
                    val[w] = want;
                    stack.push_back(w);
                } else if (val[w] != want) {
                    ok = false;
                    break;
                }
            }
        }
    }
    t.real_form = ok;
    t.rot.assign(t.total, 0);
    if (!ok) return;
    std::vector<int> var_at(t.total, -1);
    for (int v = 0; v < nv; ++v)
        for (int c = 0; c < t.var_size[v]; ++c) var_at[t.posP[t.var_first[v] + c]] = v;
    for (int p = 0; p < t.total; ++p)
        t.rot[p] = (unsigned char)(val[t.pos_eqn[p]] | (val[nv + var_at[p]] << 1));
}

// Structural blocks (see SystemTables::nblk): union-find over rows (0..n-1)
// and columns (n..2n-1) joined by every nonzero band entry of every group.
static void find_blocks(SystemTables &t) {
    const int n = t.total, w = t.band_width();
    std::vector<int> parent(2 * n);
    for (int i = 0; i < 2 * n; ++i) parent[i] = i;
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    std::vector<std::vector<int>> nzcols(n);   // structure, row -> columns
    for (const OpTables *op : {&t.mass, &t.stif})
        for (int g = 0; g < op->ngroups; ++g)
            for (int i = 0; i < n; ++i)
                for (int d = -t.nl; d <= t.nu; ++d) {
                    const int j = i + d;
                    if (j < 0 || j >= n || op->band[(size_t(g) * n + i) * w + d + t.nl] == 0.0) continue;
                    nzcols[i].push_back(j);
                    parent[find(i)] = find(n + j);
                }
    t.nblk = 1;
    t.nblk_n = n;
    t.blk_nl = t.nl;
    t.blk_nu = t.nu;
    t.blk_row.clear();
    t.blk_col.clear();
    std::map<int, int> id;   // root -> block
    std::vector<std::vector<int>> rows, cols;
    for (int x = 0; x < 2 * n; ++x) {
        const int r = find(x);
        auto it = id.find(r);
        if (it == id.end()) {
            it = id.emplace(r, int(rows.size())).first;
            rows.emplace_back();
            cols.emplace_back();
        }
        (x < n ? rows : cols)[it->second].push_back(x < n ? x : x - n);
    }
    if (rows.size() < 2) return;
    for (size_t b = 0; b < rows.size(); ++b)
        if (rows[b].size() != cols[b].size() || rows[b].empty()) return;   // structurally singular: no split
    const int nb = int(rows.size());
    int m = 0;
    for (const auto &r : rows) m = std::max(m, int(r.size()));
    std::vector<int> sub_of_col(n), blk_of_col(n);
    for (int b = 0; b < nb; ++b)
        for (size_t k = 0; k < cols[b].size(); ++k) {
            sub_of_col[cols[b][k]] = int(k);
            blk_of_col[cols[b][k]] = b;
        }
    int bl = 0, bu = 0;
    for (int b = 0; b < nb; ++b)
        for (size_t k = 0; k < rows[b].size(); ++k)
            for (int j : nzcols[rows[b][k]]) {
                if (blk_of_col[j] != b) throw std::runtime_error("inconsistent block structure");
                bl = std::max(bl, int(k) - sub_of_col[j]);
                bu = std::max(bu, sub_of_col[j] - int(k));
            }
    t.nblk = nb;
    t.nblk_n = m;
    t.blk_nl = bl;
    t.blk_nu = bu;
    t.blk_row.assign(size_t(nb) * m, -1);
    t.blk_col.assign(size_t(nb) * m, -1);
    for (int b = 0; b < nb; ++b)
        for (size_t k = 0; k < rows[b].size(); ++k) {
            t.blk_row[size_t(b) * m + k] = rows[b][k];
            t.blk_col[size_t(b) * m + k] = cols[b][k];
        }
}

static SystemTables build_system(const coral_amd_system &s, int NZ, double center, double gap, bool kxky,
                                 std::map<std::pair<int, int>, Dense> &im_cache, int nquad) {
    SystemTables t;
    t.kxky = kxky;
    t.nvars = s.nvars;
    if (s.nvars <= 0) throw std::runtime_error("coupled system without variables");
    for (int v = 0; v < s.nvars; ++v) {
        t.bc_codes.push_back(s.bc_codes[v]);
        t.eqn_orders.push_back(s.eqn_orders[v]);
        t.var_size.push_back(NZ - s.bc_codes[v] / 10);
        t.eqn_size.push_back(NZ - s.eqn_orders[v]);
    }
    int tv = 0, te = 0;
    for (int v = 0; v < s.nvars; ++v) {
        t.var_first.push_back(tv);
        t.eqn_first.push_back(te);
        tv += t.var_size[v];
        te += t.eqn_size[v];
    }
    if (tv != te)
        throw std::runtime_error("incompatible number of boundary conditions (" + std::to_string(tv) +
                                 ") and equation orders (" + std::to_string(te) + ")");
    t.total = tv;
    t.posP = degree_shuffle(t.var_size);
    t.posQ = degree_shuffle(t.eqn_size);
    t.pos_eqn.assign(t.total, -1);
    t.pos_row.assign(t.total, -1);
    for (int e = 0; e < s.nvars; ++e)
        for (int r = 0; r < t.eqn_size[e]; ++r) {
            int p = t.posQ[t.eqn_first[e] + r];
            t.pos_eqn[p] = e;
            t.pos_row[p] = r;
        }
    std::vector<Dense> stencils;
    for (int v = 0; v < s.nvars; ++v) stencils.push_back(galerkin_stencil(NZ, s.bc_codes[v]));
    auto IM = [&](int iz, int mz) -> const Dense & {
        auto key = std::make_pair(iz, mz);
        auto it = im_cache.find(key);
        if (it == im_cache.end()) it = im_cache.emplace(key, cheb_IM(NZ, center, gap, iz, mz)).first;
        return it->second;
    };
    // dense shuffled matrices per (kx, ky) exponent pair
    auto assemble_group = [&](const coral_amd_piece *pieces, int n, std::vector<std::pair<int, int>> &keys,
                              std::vector<Dense> &mats) {
        std::map<std::pair<int, int>, int> index;
        for (int k = 0; k < n; ++k) {
            const coral_amd_piece &p = pieces[k];
            if (p.var < 0 || p.var >= s.nvars || p.eqn < 0 || p.eqn >= s.nvars)
                throw std::runtime_error("piece refers to a variable or equation outside its system");
            if (p.iz > s.eqn_orders[p.eqn])
                throw std::runtime_error("Iz exponent exceeds the equation order");
            if (p.iz < 0 || p.mz < 0 || p.kx < 0 || p.ky < 0) throw std::runtime_error("negative operator exponent");
            if (!kxky && (p.kx || p.ky))
                throw std::runtime_error("horizontal derivative in a mean-field equation");
            auto key = std::make_pair(int(p.kx), int(p.ky));
            if (!index.count(key)) {
                index[key] = int(keys.size());
                keys.push_back(key);
                mats.emplace_back(t.total, t.total);
            }
            Dense &A = mats[index[key]];
            const Dense &im = IM(p.iz, p.mz);
            const Dense &S = stencils[p.var];
            const int order = s.eqn_orders[p.eqn];
            for (int r = 0; r < t.eqn_size[p.eqn]; ++r) {
                const int row = t.posQ[t.eqn_first[p.eqn] + r];
                for (int c = 0; c < t.var_size[p.var]; ++c) {
                    double acc = 0.0;
                    for (int m = 0; m < NZ; ++m) {
                        double sv = S(m, c);
                        if (sv != 0.0) acc += im(r + order, m) * sv;
                    }
                    if (acc != 0.0) A(row, t.posP[t.var_first[p.var] + c]) += p.dsca * acc;
                }
            }
        }
        std::vector<int> order(keys.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
        std::sort(order.begin(), order.end(), [&](int a, int b) { return keys[a] < keys[b]; });
        std::vector<std::pair<int, int>> k2;
        std::vector<Dense> m2;
        for (int i : order) {
            k2.push_back(keys[i]);
            m2.push_back(std::move(mats[i]));
        }
        keys.swap(k2);
        mats.swap(m2);
    };
    std::vector<std::pair<int, int>> mkeys, skeys;
    std::vector<Dense> mm, sm;
    assemble_group(s.mass, s.n_mass, mkeys, mm);
    assemble_group(s.stif, s.n_stif, skeys, sm);
    // band of the union of structures (z_csr2dia, sparse_formats_z.f90:812-829)
    int nl = 0, nu = 0;
    for (auto *v : {&mm, &sm})
        for (const Dense &A : *v)
            for (int i = 0; i < t.total; ++i)
                for (int j = 0; j < t.total; ++j)
                    if (A(i, j) != 0.0) {
                        nl = std::max(nl, i - j);
                        nu = std::max(nu, j - i);
                    }
    t.nl = nl;
    t.nu = nu;
    const int w = nl + nu + 1;
    auto build_op = [&](const std::vector<std::pair<int, int>> &keys, const std::vector<Dense> &mats, OpTables &op) {
        // merge exponent pairs whose matrices are identical
        std::vector<int> rep;                       // representative matrix of each group
        std::vector<std::vector<std::pair<int, int>>> monos;
        for (size_t k = 0; k < keys.size(); ++k) {
            int found = -1;
            for (size_t g = 0; g < rep.size(); ++g)
                if (mats[rep[g]].a == mats[k].a) {
                    found = int(g);
                    break;
                }
            if (found < 0) {
                rep.push_back(int(k));
                monos.push_back({keys[k]});
            } else {
                monos[found].push_back(keys[k]);
            }
        }
        op.ngroups = int(rep.size());
        op.mono_off.assign(1, 0);
        for (auto &ml : monos) {
            for (auto &m : ml) {
                op.mono_a.push_back(m.first);
                op.mono_b.push_back(m.second);
            }
            op.mono_off.push_back(int(op.mono_a.size()));
        }
        op.band.assign(rep.size() * size_t(t.total) * w, 0.0);
        for (size_t g = 0; g < rep.size(); ++g) {
            const Dense &A = mats[rep[g]];
            int ew = 1;
            for (int i = 0; i < t.total; ++i) {
                int cnt = 0;
                for (int d = -nl; d <= nu; ++d) {
                    int j = i + d;
                    if (j < 0 || j >= t.total) continue;
                    op.band[(g * t.total + i) * w + d + nl] = A(i, j);
                    cnt += A(i, j) != 0.0;
                }
                ew = std::max(ew, cnt);
            }
            op.ell_off.push_back(int(op.ell_col.size()));
            op.ell_w.push_back(ew);
            for (int i = 0; i < t.total; ++i) {
                int cnt = 0;
                for (int d = -nl; d <= nu; ++d) {
                    int j = i + d;
                    if (j < 0 || j >= t.total || A(i, j) == 0.0) continue;
                    op.ell_col.push_back(j);
                    op.ell_val.push_back(A(i, j));
                    ++cnt;
                }
                for (; cnt < ew; ++cnt) {
                    op.ell_col.push_back(i);
                    op.ell_val.push_back(0.0);
                }
            }
        }
    };
    build_op(mkeys, mm, t.mass);
    build_op(skeys, sm, t.stif);
    // NL terms, sorted by equation
    std::vector<coral_amd_nlterm> terms(s.nl, s.nl + s.n_nl);
    std::stable_sort(terms.begin(), terms.end(),
                     [](const coral_amd_nlterm &a, const coral_amd_nlterm &b) { return a.eqn < b.eqn; });
    t.nl_eqn_begin.assign(s.nvars + 1, 0);
    for (const auto &q : terms) {
        if (q.eqn < 0 || q.eqn >= s.nvars) throw std::runtime_error("NL term in an unknown equation");
        if (q.qvar < 0 || q.qvar >= nquad) throw std::runtime_error("NL term refers to an unknown quadratic variable");
        if (!kxky && (q.kx || q.ky)) throw std::runtime_error("horizontal derivative in a mean-field NL term");
        NLTable nt;
        nt.eqn = q.eqn;
        nt.qvar = q.qvar;
        nt.kx = q.kx;
        nt.ky = q.ky;
        nt.dsca = q.dsca;
        nt.order = s.eqn_orders[q.eqn];
        nt.w = q.iz + q.mz;
        nt.woff = int(t.nl_w.size());
        const Dense &im = IM(q.iz, q.mz);
        for (int r = 0; r < t.eqn_size[q.eqn]; ++r)
            for (int d = -nt.w; d <= nt.w; ++d) {
                int row = r + nt.order, col = row + d;
                t.nl_w.push_back((col >= 0 && col < NZ) ? im(row, col) : 0.0);
            }
        t.nl_terms.push_back(nt);
        t.nl_eqn_begin[q.eqn + 1]++;
    }
    for (int e = 0; e < s.nvars; ++e) t.nl_eqn_begin[e + 1] += t.nl_eqn_begin[e];
    find_real_form(s, t);
    find_blocks(t);
    return t;
}

Problem assemble(const coral_amd_problem &p) {
    Problem P;
    if (p.NX < 2 || p.NY < 2 || p.NZ < 6 || (p.NX % 2) || (p.NY % 2))
        throw std::runtime_error("resolution must have even NX, NY >= 2 and NZ >= 6");
    if (!(p.Lx > 0 && p.Ly > 0 && p.gap > 0)) throw std::runtime_error("box sizes must be positive");
    P.NX = p.NX;
    P.NY = p.NY;
    P.NZ = p.NZ;
    P.NXAA = 3 * p.NX / 2;
    P.NYAA = 3 * p.NY / 2;
    P.NZAA = 3 * p.NZ / 2;
    P.Lx = p.Lx;
    P.Ly = p.Ly;
    P.gap = p.gap;
    P.center = p.center;
    const double pi = 4.0 * std::atan(1.0);
    P.kx_box = 2.0 * pi / p.Lx;
    P.ky_box = 2.0 * pi / p.Ly;
    if (p.nstages < 1 || !p.a_arr || !p.a_hat) throw std::runtime_error("missing IMEX tableau");
    P.nstages = p.nstages;
    P.a_arr.assign(p.a_arr, p.a_arr + p.nstages * p.nstages);
    P.a_hat.assign(p.a_hat, p.a_hat + (p.nstages + 1) * (p.nstages + 1));
    P.nquad = p.n_quad;
    P.nfull = p.n_full;
    std::map<std::pair<int, int>, Dense> im_cache;
    if (p.n_kxky < 1) throw std::runtime_error("at least one kxky system is required");
    for (int i = 0; i < p.n_kxky; ++i)
        P.kxky.push_back(build_system(p.kxky[i], p.NZ, p.center, p.gap, true, im_cache, p.n_quad));
    for (int i = 0; i < p.n_zero; ++i)
        P.zero.push_back(build_system(p.zero[i], p.NZ, p.center, p.gap, false, im_cache, p.n_quad));
    for (int q = 0; q < p.n_quad; ++q) {
        const auto &qv = p.quad[q];
        if (qv.ivar1 < -1 || qv.ivar1 >= p.n_full || qv.ivar2 < -1 || qv.ivar2 >= p.n_full ||
            (qv.ivar1 == -1 && qv.ivar2 == -1))
            throw std::runtime_error("bad quadratic variable definition");
        P.quad_a.push_back(qv.ivar1);
        P.quad_b.push_back(qv.ivar2);
        P.quad_filter.push_back(qv.remove_z_integral ? 1 : 0);
    }
    // extraction tables (padStencilExtractShuffle, optional d/dz)
    Dense D;
    bool haveD = false;
    for (int f = 0; f < p.n_full; ++f) {
        for (int k = 0; k < p.full[f].nterms; ++k) {
            const auto &ft = p.full[f].terms[k];
            const std::vector<SystemTables> &pool = ft.kind == 0 ? P.kxky : P.zero;
            if (ft.kind != 0 && ft.kind != 1) throw std::runtime_error("full-variable term of unknown kind");
            if (ft.system < 0 || ft.system >= int(pool.size())) throw std::runtime_error("full-variable term: bad system");
            const SystemTables &st = pool[ft.system];
            if (ft.position < 0 || ft.position >= st.nvars) throw std::runtime_error("full-variable term: bad position");
            if (ft.kind == 1 && (ft.kx || ft.ky)) throw std::runtime_error("horizontal derivative of a mean variable");
            if (ft.mz) throw std::runtime_error("mz in a full-variable definition is not supported");
            Dense S = galerkin_stencil(p.NZ, st.bc_codes[ft.position]);
            if (ft.dz > 0) {
                if (!haveD) {
                    D = chebyshev_derivative(p.NZ, p.center, p.gap);
                    haveD = true;
                }
                for (int d = 0; d < ft.dz; ++d) S = matmul(D, S);
            }
            ExtractTable et;
            et.full = f;
            et.kind = ft.kind;
            et.system = ft.system;
            et.kx = ft.kx;
            et.ky = ft.ky;
            et.rowptr.push_back(0);
            for (int n = 0; n < p.NZ; ++n) {
                for (int m = 0; m < S.cols; ++m) {
                    double v = S(n, m);
                    if (v != 0.0) {
                        et.col.push_back(st.posP[st.var_first[ft.position] + m]);
                        et.val.push_back(ft.dsca * v);
                    }
                }
                et.rowptr.push_back(int(et.col.size()));
            }
            P.extract.push_back(std::move(et));
        }
    }
    std::stable_sort(P.extract.begin(), P.extract.end(),
                     [](const ExtractTable &a, const ExtractTable &b) { return a.kind < b.kind; });
    // transform matrices
    const int NZ = P.NZ, NZAA = P.NZAA;
    P.Tz.assign(size_t(NZAA) * NZ, 0.0);
    P.Cz.assign(size_t(NZ) * NZAA, 0.0);
    const double fnorm = 1.0 / (double(P.NXAA) * double(P.NYAA));
    for (int j = 0; j < NZAA; ++j)
        for (int n = 0; n < NZ; ++n) {
            // T_n(x_j), x_j = cos(pi (2j+1) / (2 NZAA))
            double c = cospi_frac((long long)n * (2 * j + 1), 2LL * NZAA);
            P.Tz[size_t(j) * NZ + n] = c;
            P.Cz[size_t(n) * NZAA + j] = (2.0 / NZAA) * c * (n == 0 ? 0.5 : 1.0) * fnorm;
        }
    P.zweight.resize(NZAA);
    {
        std::vector<double> g(NZAA), mid(NZAA + 1);
        for (int i = 0; i < NZAA; ++i) g[i] = std::cos(pi * double(2 * (i + 1) - 1) / double(2 * NZAA));
        mid[0] = 1.0;
        mid[NZAA] = -1.0;
        for (int i = 1; i < NZAA; ++i) mid[i] = 0.5 * (g[i] + g[i - 1]);
        for (int i = 0; i < NZAA; ++i) P.zweight[i] = (mid[i] - mid[i + 1]) / 2.0 * p.gap;
    }
    // kept wavenumbers
    P.nkx = P.NX / 2;
    P.nky = P.NY - 1;
    P.nsys = P.nkx * P.nky;
    for (int ix = 0; ix < P.nkx; ++ix) P.kx_int.push_back(ix);
    for (int jy = 0; jy < P.nky; ++jy) {
        if (jy < P.NY / 2) {
            P.ky_int.push_back(jy);
            P.ky_slot.push_back(jy);
        } else {
            int k = jy - (P.NY - 1);  // -(NY/2-1) .. -1
            P.ky_int.push_back(k);
            P.ky_slot.push_back(P.NYAA + k);
        }
    }
    return P;
}

}  // namespace coral
