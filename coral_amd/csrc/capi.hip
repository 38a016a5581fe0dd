// Solver context and C ABI (include/coral_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/coral_amd.h"
#include "assembly.hpp"
#include "kernels.hpp"
#include "nn.hpp"

using namespace coral;

namespace {

thread_local std::string g_err;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_OK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw HipError(std::string(#call) + ": " + hipGetErrorString(e_));                \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p) HIP_OK(hipFree(p));
        p = nullptr;
        n = count;
        if (count) HIP_OK(hipMalloc(&p, count * sizeof(T)));
    }
    void upload(const std::vector<T> &v) {
        alloc(v.size());
        if (!v.empty()) HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
};

struct SysDev {
    SystemTables *t = nullptr;
    bool kxky = true;
    int nsys = 0;
    DevBuf<double2> field;
    std::vector<std::unique_ptr<DevBuf<double2>>> Khat, X;   // per stage: K_hat, solution
    DevBuf<double2> R;   // right-hand side of the current stage (never read back by a later stage)
    DevBuf<double2> cmass, cstif;                              // [nsys][ngroups] group coefficients
    DevBuf<double2> Lf, Uf;      // complex factors
    bool real = false;           // factored in real form (Lr, Ur, rot)
    int ncol16 = 0;              // > 0: 16-lane solve layout (see solve16_ncol)
    // factored matrices per system: structural blocks of the 16-lane layout
    // (SystemTables::nblk; 1 = the whole system)
    int nblk = 1, blk_n = 0, blk_nl = 0, blk_nu = 0;
    DevBuf<int> blk_row, blk_col;
    DevBuf<unsigned> words;      // position words of the 16-lane solve (solve16_words)
    // trimmed U records of the 16-lane solve (solve16_ucol), rebuilt at every
    // factorization
    DevBuf<double> Urt;
    DevBuf<double2> Uft;
    DevBuf<int2> ucol;
    DevBuf<int> ulen;
    long long ustride = 0;
    double u_mean = 0.0;         // mean trimmed U record length over the block's columns
    DevBuf<double> Lr, Ur;
    DevBuf<unsigned char> rot;
    DevBuf<unsigned char> piv;
    DevBuf<int> info;
    // residual refinement: A x - rhs and the correction solved from it
    DevBuf<double2> Rres, Fcor;
    struct OpBufs {
        DevBuf<int> mono_off, mono_a, mono_b, ell_w, ell_off, ell_col;
        DevBuf<double> band, ell_val;
        int ngroups = 0;
        OpDev dev() const {
            return OpDev{ngroups, mono_off.p, mono_a.p, mono_b.p, band.p, ell_w.p, ell_off.p, ell_col.p, ell_val.p};
        }
    } mass, stif;
    DevBuf<int> pos_eqn, pos_row, eqn_begin, t_qvar, t_kx, t_ky, t_w, t_order, t_woff;
    DevBuf<double> t_dsca, wv;
    // mean systems: the sources' constant part of K_hat, one value per
    // position (zero without sources; coral_amd_set_source)
    DevBuf<double> src;
    // kxky NL program (NLProgArgs)
    DevBuf<int> nl_pos_ptr, nl_blk_term, nl_blk_ptr, nl_ent_qm;
    DevBuf<double> nl_ent_w;
    DevBuf<double2> nl_cmono;
    int nl_nterms = 0;
};

struct ExtractDev {
    int full, kind, system, kx, ky;
    DevBuf<int> rowptr, col;
    DevBuf<double> val;
};

struct TimerRec {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct coral_amd_ctx {
    Problem P;
    int device = 0;
    hipStream_t stream = nullptr;
    // the mean systems' NL terms, right-hand sides and solves run on a second
    // stream, overlapped with the kxky work: they wait for Qhat (ev_qhat) and
    // the kxky side waits for their solutions (ev_mean) before the next
    // stage's mean extraction
    hipStream_t side = nullptr;
    hipEvent_t ev_qhat = nullptr, ev_mean = nullptr;
    bool mean_pending = false;
    KeptModes km{};
    std::vector<std::unique_ptr<SysDev>> ks, zs;
    std::vector<std::unique_ptr<ExtractDev>> ext;
    std::vector<int> full_has_kxky;
    DevBuf<double2> spec, Fz, G, H, Hy, Qhat;
    DevBuf<double> physd;
    DevBuf<double> Tz, Cz, zweight;
    DevBuf<double> zfrag_b, zfrag_f;   // MFMA fragments of the folded z transforms
    ZFragsDev zdev{};
    DevBuf<double2> twx, twy;
    DevBuf<int> qa, qb;
    DevBuf<unsigned long long> cflmax;
    DevBuf<double2> sink;   // stores of padding positions in the 16-lane solve
    FFTPlan px, py;
    // Precision of the step's transform chain: 0 = FP64; 1 = FP32 transforms
    // and products (z products on FP32 MFMA with compensated accumulation),
    // NL sums, right-hand sides, factors and solves in FP64.
    int precision = 0;
    // FP64 residual-refinement sweeps after the solves: 0 none, 1 after the
    // last stage of each step, 2 after every stage
    int refine = 0;
    // replace the kx = 0 plane of every kxky state by its Hermitian part after
    // each step (coral_amd_set_kx0_projection)
    bool project_kx0 = false;
    // launch shapes (coral_amd_set_launch_shape; picked by the ensemble's tuner):
    // wavefronts per workgroup of the x pass (0 = the largest that fits), and
    // the z transforms as vector GEMMs instead of MFMA products
    int xpass_wpb = 0;
    bool z_plain = false;
    // per full variable: (boundary-value extraction, vertical decomposition),
    // coral_amd_set_full_variable_filter
    std::vector<std::pair<int, int>> full_filter;
    // the current step's stage coefficients (step_begin): M x^n, M x_j, L x_j, K_hat_j
    std::vector<std::vector<double>> stepRm, stepRx, stepRk;
    // mean-system sources: the terms (coral_amd_add_source_term) and the
    // sources' Chebyshev coefficients by index (coral_amd_set_source);
    // src_missing is the lowest index a term names with no definition (0: none)
    struct SourceTerm {
        int isys, eqn, iz, index;
        double dsca;
    };
    std::vector<SourceTerm> src_terms;
    std::map<int, std::vector<double>> src_cheb;
    int src_missing = 0;
    // full variables whose physical values do not come from the spectral
    // state: fixed fields (penalisation masks, coral_amd_set_fixed_physical;
    // fixed_slot[ifull] >= 0 is the field's plane block in fixed_all) and the
    // Smagorinsky viscosity (coral_amd_set_smagorinsky: the z-profile nu_prof
    // from the squares of the nu_grads variables).  The quadratic variables
    // with such a factor are formed in physical space (side_products).
    std::vector<int> fixed_slot;
    int nfixed = 0;
    DevBuf<double> fixed_all;
    int nu_full = -1;
    std::vector<int> nu_grads;
    double nu_prefactor = 1.0;
    DevBuf<double> nu_prof, nu_tmp;
    bool side_advance = true;   // a step keeps the new nuSmago profile; the products seam does not
    std::vector<int> side_q, side_reg;   // quadratic variables formed aside; full variables they transform
    DevBuf<int> side_desc;               // per side variable: kind and slot of both factors
    DevBuf<double> side_phys, side_tau;
    DevBuf<double2> side_Q;
    // learned subgrid closure (coral_amd_set_closure): the network, its
    // normalisation, inference workspace and the correction's buffers
    struct Closure {
        bool on = false;
        bool fused_io = false;   // lift and projection fused with the packing (k_closure_lift / _project)
        nn::NetParams net;
        nn::NetWork ws;
        int chunk = 0;   // planes per network launch
        DevBuf<float> w, mu_in, sd_in, mu_out, sd_out, Y, H, Cm, gwork, mwork;
        DevBuf<unsigned short> w16, wt16, basisT16, basis16p, X16, H16, Cs16, H16T;
        DevBuf<unsigned short> W0T;   // the lift's W0 transposed, W x kp (zero past cin), for k_closure_lift
        DevBuf<double> phys_full, phys_tau, zgrid;
        DevBuf<double2> Qtau;
    } closure;
    // incremented whenever the launches of a step change (dt, precision,
    // projection, launch shapes): graphs that hold this run's step (its own or
    // an ensemble's) are stale when it differs from their capture's
    unsigned long long gen = 0;
    // buffer swaps since creation (never reset): its parity is the state /
    // last-stage buffer assignment, which captured graphs bake in
    unsigned long long nswap = 0;
    DevBuf<float2> spec32, Fz32, G32, H32, Hy32, Qhat32, twx32, twy32;
    DevBuf<float> zfrag32_b, zfrag32_f;
    ZFrags32Dev z32{};
    double dt = 0.0;
    bool factored = false;
    bool timing = false;
    // HIP graphs of one step: one_step swaps each system's field and last
    // stage buffers, so two graphs (one per buffer assignment) alternate;
    // gpar is the graph whose assignment the host pointers hold now
    bool use_graphs = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int gpar = 0;
    void drop_graphs() {
        // an executable graph is destroyed only once its launches have ended
        if ((gexec[0] || gexec[1]) && stream) (void)hipStreamSynchronize(stream);
        for (auto &g : gexec)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        gpar = 0;
    }
    std::vector<TimerRec> recs;
    std::map<std::string, std::pair<double, int>> totals;

    ~coral_amd_ctx() {
        for (auto &r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        drop_graphs();
        if (ev_qhat) (void)hipEventDestroy(ev_qhat);
        if (ev_mean) (void)hipEventDestroy(ev_mean);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class F>
    void timed(const char *name, F &&f) {
        if (!timing) {
            f();
            return;
        }
        TimerRec r;
        r.name = name;
        HIP_OK(hipEventCreate(&r.a));
        HIP_OK(hipEventCreate(&r.b));
        HIP_OK(hipEventRecord(r.a, stream));
        f();
        HIP_OK(hipEventRecord(r.b, stream));
        recs.push_back(r);
    }

    void flush_timers() {
        if (recs.empty()) return;
        HIP_OK(hipStreamSynchronize(stream));
        for (auto &r : recs) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
            auto &t = totals[r.name];
            t.first += ms;
            t.second += 1;
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        recs.clear();
    }
};

namespace {

// Clears the CFL maxima with a kernel rather than hipMemsetAsync, so that a
// caller who captures the step into a HIP graph records no memset node.
__global__ void k_zero_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}
void zero_cflmax(hipStream_t st, unsigned long long *p) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, 3);
    check_launch("zero_cflmax");
}

FFTPlan make_plan(int N, DevBuf<double2> &tw) {
    FFTPlan p;
    p.N = N;
    int m = N;
    int nr = 0;
    while (m % 4 == 0) {
        p.radix[nr++] = 4;
        m /= 4;
    }
    while (m % 2 == 0) {
        p.radix[nr++] = 2;
        m /= 2;
    }
    while (m % 3 == 0) {
        p.radix[nr++] = 3;
        m /= 3;
    }
    if (m != 1 || nr > 16)
        throw std::runtime_error("FFT length " + std::to_string(N) + " is not of the form 2^a 3^b");
    p.nrad = nr;
    std::vector<double2> h(N);
    for (int k = 0; k < N; ++k) {
        // exp(-2 pi i k / N) with exact reduction of the angle
        long long num = 2LL * k;  // angle = -pi * num / N
        double a = M_PI * double(num % (2LL * N)) / double(N);
        h[k] = make_double2(std::cos(a), -std::sin(a));
    }
    tw.upload(h);
    p.tw = tw.p;
    return p;
}

void fill_system_device(SysDev &d, int nstages, const KeptModes &km, hipStream_t st, int NZ) {
    SystemTables &t = *d.t;
    const size_t n = size_t(d.nsys) * t.total;
    d.field.alloc(n);
    HIP_OK(hipMemset(d.field.p, 0, n * sizeof(double2)));
    d.R.alloc(n);
    for (int i = 0; i < nstages; ++i) {
        d.Khat.emplace_back(new DevBuf<double2>());
        d.X.emplace_back(new DevBuf<double2>());
        d.Khat.back()->alloc(n);
        d.X.back()->alloc(n);
    }
    auto up = [](SysDev::OpBufs &b, const OpTables &o) {
        b.ngroups = o.ngroups;
        b.mono_off.upload(o.mono_off);
        b.mono_a.upload(o.mono_a);
        b.mono_b.upload(o.mono_b);
        b.band.upload(o.band);
        b.ell_w.upload(o.ell_w);
        b.ell_off.upload(o.ell_off);
        b.ell_col.upload(o.ell_col);
        b.ell_val.upload(o.ell_val);
    };
    up(d.mass, t.mass);
    up(d.stif, t.stif);
    // group coefficients (i kx)^a (i ky)^b summed over each group, per system
    const KeptModes kmd = d.kxky ? km : KeptModes{1, 1, 2, 0.0, 0.0};
    d.cmass.alloc(size_t(d.nsys) * std::max(t.mass.ngroups, 1));
    d.cstif.alloc(size_t(d.nsys) * std::max(t.stif.ngroups, 1));
    launch_group_coefs(st, d.mass.dev(), kmd, d.nsys, d.cmass.p);
    launch_group_coefs(st, d.stif.dev(), kmd, d.nsys, d.cstif.p);
    d.pos_eqn.upload(t.pos_eqn);
    d.pos_row.upload(t.pos_row);
    d.eqn_begin.upload(t.nl_eqn_begin);
    std::vector<int> q, kx, ky, w, order, woff;
    std::vector<double> dsca;
    for (auto &nt : t.nl_terms) {
        q.push_back(nt.qvar);
        kx.push_back(nt.kx);
        ky.push_back(nt.ky);
        w.push_back(nt.w);
        order.push_back(nt.order);
        woff.push_back(nt.woff);
        dsca.push_back(nt.dsca);
    }
    // never upload empty arrays as null pointers that a kernel may index
    if (q.empty()) {
        q.push_back(0);
        kx.push_back(0);
        ky.push_back(0);
        w.push_back(0);
        order.push_back(0);
        woff.push_back(0);
        dsca.push_back(0.0);
    }
    d.t_qvar.upload(q);
    d.t_kx.upload(kx);
    d.t_ky.upload(ky);
    d.t_w.upload(w);
    d.t_order.upload(order);
    d.t_woff.upload(woff);
    d.t_dsca.upload(dsca);
    std::vector<double> wv = t.nl_w;
    if (wv.empty()) wv.push_back(0.0);
    d.wv.upload(wv);
    if (d.kxky) {
        // per-position NL program (see NLProgArgs)
        std::vector<int> pos_ptr(1, 0), blk_term, blk_ptr(1, 0), ent_qm;
        std::vector<double> ent_w;
        for (int p = 0; p < t.total; ++p) {
            const int e = t.pos_eqn[p], r = t.pos_row[p];
            for (int ti = t.nl_eqn_begin[e]; ti < t.nl_eqn_begin[e + 1]; ++ti) {
                const NLTable &nt = t.nl_terms[ti];
                const int w = nt.w, m0 = r + nt.order;
                for (int dd = -w; dd <= w; ++dd) {
                    const int m = m0 + dd;
                    if (m < 0 || m >= NZ) continue;
                    const double wt = t.nl_w[nt.woff + r * (2 * w + 1) + dd + w];
                    if (wt == 0.0) continue;
                    ent_qm.push_back(nt.qvar * NZ + m);
                    ent_w.push_back(nt.dsca * wt);
                }
                if (int(ent_qm.size()) > blk_ptr.back()) {
                    blk_term.push_back(ti);
                    blk_ptr.push_back(int(ent_qm.size()));
                }
            }
            pos_ptr.push_back(int(blk_term.size()));
        }
        if (blk_term.empty()) {
            blk_term.push_back(0);
            blk_ptr.push_back(0);
        }
        if (ent_qm.empty()) {
            ent_qm.push_back(0);
            ent_w.push_back(0.0);
        }
        d.nl_pos_ptr.upload(pos_ptr);
        d.nl_blk_term.upload(blk_term);
        d.nl_blk_ptr.upload(blk_ptr);
        d.nl_ent_qm.upload(ent_qm);
        d.nl_ent_w.upload(ent_w);
        d.nl_nterms = std::max(int(t.nl_terms.size()), 1);
        d.nl_cmono.alloc(size_t(d.nsys) * d.nl_nterms);
        launch_term_monomials(st, d.t_kx.p, d.t_ky.p, int(t.nl_terms.size()), km, d.nsys, d.nl_cmono.p);
    }
    const int kv = t.nl + t.nu;
    const char *force = std::getenv("CORAL_AMD_COMPLEX_FACTORS");
    d.real = t.real_form && !(force && force[0] == '1');
    // 16-lane solve layout when the band allows it; U arrays padded (zeroed) on
    // both sides for its out-of-record reads
    // (records of four matrices interleaved, in groups of four and blocks of 16
    // columns; the slots past the last matrix stay zero).  The 16-lane layout
    // factors the structural blocks of the operator separately when it splits
    // (CORAL_AMD_NO_BLOCK_SPLIT=1 keeps whole systems).
    const char *nosplit = std::getenv("CORAL_AMD_NO_BLOCK_SPLIT");
    const bool split = t.nblk > 1 && !(nosplit && nosplit[0] == '1') &&
                       solve16_ncol(t.nblk_n, t.blk_nl, t.blk_nl + t.blk_nu) > 0;
    d.ncol16 = split ? solve16_ncol(t.nblk_n, t.blk_nl, t.blk_nl + t.blk_nu) : solve16_ncol(t.total, t.nl, kv);
    if (split) {
        d.nblk = t.nblk;
        d.blk_n = t.nblk_n;
        d.blk_nl = t.blk_nl;
        d.blk_nu = t.blk_nu;
        d.blk_row.upload(t.blk_row);
        d.blk_col.upload(t.blk_col);
    } else {
        d.nblk = 1;
        d.blk_n = t.total;
        d.blk_nl = t.nl;
        d.blk_nu = t.nu;
    }
    if (d.ncol16 > 0)
        d.words.upload(solve16_words(t.total, d.nblk, d.blk_n, split ? t.blk_row.data() : nullptr,
                                     split ? t.blk_col.data() : nullptr, d.real ? t.rot.data() : nullptr));
    const int bkv = d.blk_nl + d.blk_nu;
    const size_t nmat = size_t(d.nsys) * d.nblk, nmat4 = (nmat + 3) / 4 * 4;
    const size_t nb = 16 * ((size_t(d.blk_n) + 15) / 16);
    const size_t lsize = d.ncol16 > 0 ? nmat4 * d.ncol16 * solve16_klp(d.blk_nl)
                                      : size_t(d.nsys) * t.total * std::max(t.nl, 1);
    const size_t usize =
        (d.ncol16 > 0 ? nmat4 * nb * (bkv + 1) : size_t(d.nsys) * t.total * (kv + 1)) + 2 * size_t(SOLVE16_PAD);
    if (d.real) {
        d.Lr.alloc(lsize);
        d.Ur.alloc(usize);
        HIP_OK(hipMemset(d.Lr.p, 0, lsize * sizeof(double)));
        HIP_OK(hipMemset(d.Ur.p, 0, usize * sizeof(double)));
        d.rot.upload(t.rot);
    } else {
        d.Lf.alloc(lsize);
        d.Uf.alloc(usize);
        HIP_OK(hipMemset(d.Lf.p, 0, lsize * sizeof(double2)));
        HIP_OK(hipMemset(d.Uf.p, 0, usize * sizeof(double2)));
    }
    // (16-lane layout: pivot records padded to whole blocks of 16, zero)
    const size_t pstride = d.ncol16 > 0 ? nb : size_t(d.blk_n);
    d.piv.alloc(nmat * pstride);
    HIP_OK(hipMemset(d.piv.p, 0, nmat * pstride));
    d.info.alloc(nmat);
}

void fill_zero(hipStream_t st, double2 *p, size_t n);

void factorize(coral_amd_ctx &c, SysDev &d, double dt) {
    SystemTables &t = *d.t;
    // matrices factored: nblk structural blocks of blk_n unknowns per system
    const int nmat = d.nsys * d.nblk, ldab = 2 * d.blk_nl + d.blk_nu + 1;
    const size_t per = size_t(d.blk_n) * ldab;
    // bound the working band to ~1 GiB per launch
    int chunk = int(std::max<size_t>(1, (size_t(1) << 30) / (per * sizeof(double2))));
    chunk = std::min(chunk, nmat);
    DevBuf<double2> work;
    work.alloc(size_t(chunk) * per);
    for (int s0 = 0; s0 < nmat; s0 += chunk) {
        FactorArgs a{};
        a.mass = d.mass.dev();
        a.stif = d.stif.dev();
        a.nl = t.nl;
        a.nu = t.nu;
        a.total = t.total;
        a.sdt = -dt * c.P.a_arr[0];
        a.s0 = s0;
        a.ns = std::min(chunk, nmat - s0);
        a.identity00 = d.kxky ? 1 : 0;
        a.km = d.kxky ? c.km : KeptModes{1, 1, 2, 0.0, 0.0};
        a.work = work.p;
        a.Lf = d.Lf.p;
        a.Uf = d.Uf.p ? d.Uf.p + SOLVE16_PAD : nullptr;
        a.rot = d.real ? d.rot.p : nullptr;
        a.Lr = d.Lr.p;
        a.Ur = d.Ur.p ? d.Ur.p + SOLVE16_PAD : nullptr;
        a.ncol16 = d.ncol16;
        a.nblk = d.nblk;
        a.blk_n = d.blk_n;
        a.blk_nl = d.blk_nl;
        a.blk_nu = d.blk_nu;
        a.blk_row = d.blk_row.p;
        a.blk_col = d.blk_col.p;
        a.piv = d.piv.p;
        a.info = d.info.p;
        c.timed("factor", [&] { launch_factor(c.stream, a); });
    }
    if (d.ncol16 > 0) {
        const int bkv = d.blk_nl + d.blk_nu, elem = d.real ? 8 : 16, nb = 16 * ((d.blk_n + 15) / 16);
        const void *Ufull = d.real ? (const void *)(d.Ur.p + SOLVE16_PAD) : (const void *)(d.Uf.p + SOLVE16_PAD);
        if (d.ulen.n < size_t(nb)) d.ulen.alloc(nb);
        std::vector<int2> ucol;
        solve16_ucol(c.stream, Ufull, elem, d.blk_n, bkv, nmat, d.ulen.p, ucol, &d.ustride);
        double lsum = 0.0;
        for (int j = 0; j < d.blk_n; ++j) lsum += ucol[j].y;
        d.u_mean = lsum / d.blk_n;
        const size_t need = size_t((nmat + 3) / 4) * size_t(d.ustride);
        void *Ut;
        if (d.real) {
            if (d.Urt.n < need) d.Urt.alloc(need);
            Ut = d.Urt.p;
        } else {
            if (d.Uft.n < need) d.Uft.alloc(need);
            Ut = d.Uft.p;
        }
        HIP_OK(hipMemsetAsync(Ut, 0, need * elem, c.stream));
        d.ucol.upload(ucol);
        c.timed("factor", [&] { solve16_trim_u_copy(c.stream, Ufull, elem, d.blk_n, bkv, nmat, d.ucol.p, d.ustride, Ut); });
    }
    HIP_OK(hipStreamSynchronize(c.stream));
    std::vector<int> info(nmat);
    HIP_OK(hipMemcpy(info.data(), d.info.p, nmat * sizeof(int), hipMemcpyDeviceToHost));
    for (int m = 0; m < nmat; ++m)
        if (info[m] != 0)
            throw std::runtime_error("singular evolution operator (zgbtrf info " + std::to_string(info[m]) +
                                     (d.nblk > 1 ? " in structural block " + std::to_string(m % d.nblk) : "") +
                                     ") at wavenumber index " + std::to_string(m / d.nblk));
}

// y = M (sum ma U) + L (sum xa X) + sum kb K
void rhs(coral_amd_ctx &c, SysDev &d, const std::vector<std::pair<const double2 *, double>> &ms,
         const std::vector<std::pair<const double2 *, double>> &xs,
         const std::vector<std::pair<const double2 *, double>> &ks, double2 *y) {
    SystemTables &t = *d.t;
    RhsArgs a{};
    a.mass = d.mass.dev();
    a.stif = d.stif.dev();
    a.cmass = d.cmass.p;
    a.cstif = d.cstif.p;
    a.total = t.total;
    a.nsys = d.nsys;
    a.bl = t.nl;
    a.bu = t.nu;
    if (ms.size() > 8 || xs.size() > 8 || ks.size() > 8) throw std::runtime_error("too many IMEX stages");
    a.nm = int(ms.size());
    for (size_t k = 0; k < ms.size(); ++k) {
        a.mv[k] = ms[k].first;
        a.ma[k] = ms[k].second;
    }
    a.nx = int(xs.size());
    for (size_t k = 0; k < xs.size(); ++k) {
        a.xv[k] = xs[k].first;
        a.xa[k] = xs[k].second;
    }
    a.nk = int(ks.size());
    for (size_t k = 0; k < ks.size(); ++k) {
        a.kv[k] = ks[k].first;
        a.kb[k] = ks[k].second;
    }
    a.zero00 = d.kxky ? 1 : 0;
    a.y = y;
    c.timed(d.kxky ? "rhs" : "rhs_mean", [&] { launch_rhs(c.stream, a); });
}

// z transforms: MFMA kernels where a variant covers the shape, plain GEMM otherwise
void zbwd(coral_amd_ctx &c, const double2 *spec, double2 *Fz, int nb) {
    const Problem &P = c.P;
    if (c.z_plain || !launch_zbwd_mfma(c.stream, c.zdev, spec, Fz, P.NZ, P.NZAA, P.nsys, nb))
        launch_zbwd(c.stream, c.Tz.p, spec, Fz, P.NZAA, P.NZ, P.nsys, nb);
}
void zfwd(coral_amd_ctx &c, const double2 *Hy, double2 *Qhat, int nb) {
    const Problem &P = c.P;
    if (c.z_plain || !launch_zfwd_mfma(c.stream, c.zdev, Hy, Qhat, P.NZ, P.NZAA, P.nsys, nb))
        launch_zfwd(c.stream, c.Cz.p, Hy, Qhat, P.NZ, P.NZAA, P.nsys, nb);
}

void nl_assemble(coral_amd_ctx &c, SysDev &d, double2 *Khat) {
    SystemTables &t = *d.t;
    if (d.kxky) {
        NLProgArgs a{};
        a.pos_ptr = d.nl_pos_ptr.p;
        a.blk_term = d.nl_blk_term.p;
        a.blk_ptr = d.nl_blk_ptr.p;
        a.ent_qm = d.nl_ent_qm.p;
        a.ent_w = d.nl_ent_w.p;
        a.cmono = d.nl_cmono.p;
        a.nterms = d.nl_nterms;
        a.Qhat = c.Qhat.p;
        a.Qhat32 = c.precision == 1 ? c.Qhat32.p : nullptr;
        a.nsys = c.P.nsys;
        a.total = t.total;
        a.Khat = Khat;
        c.timed("nl_assemble", [&] { launch_nl_prog(c.stream, a); });
        return;
    }
    NLArgs a{};
    a.Qhat = c.Qhat.p;
    a.Qhat32 = c.precision == 1 ? c.Qhat32.p : nullptr;
    a.nsys = c.P.nsys;
    a.NZ = c.P.NZ;
    a.total = t.total;
    a.mean = d.kxky ? 0 : 1;
    a.pos_eqn = d.pos_eqn.p;
    a.pos_row = d.pos_row.p;
    a.eqn_begin = d.eqn_begin.p;
    a.t_qvar = d.t_qvar.p;
    a.t_kx = d.t_kx.p;
    a.t_ky = d.t_ky.p;
    a.t_w = d.t_w.p;
    a.t_order = d.t_order.p;
    a.t_woff = d.t_woff.p;
    a.t_dsca = d.t_dsca.p;
    a.wv = d.wv.p;
    a.src = d.kxky ? nullptr : d.src.p;
    a.km = c.km;
    a.Khat = Khat;
    c.timed(d.kxky ? "nl_assemble" : "nl_mean", [&] { launch_nl(c.stream, a); });
}

// the main stream waits for the mean systems' pending work on the side stream
void join_mean(coral_amd_ctx &c) {
    if (!c.mean_pending) return;
    HIP_OK(hipStreamWaitEvent(c.stream, c.ev_mean, 0));
    c.mean_pending = false;
}

// x^{n+1} = x_s: the last stage's buffer becomes the state, the state's
// buffer the next step's last stage.  Every swap, eager or replayed, flips
// gpar, so the graph that matches the host pointers is known after any mix of
// eager steps (timers on) and replays.
void swap_step_buffers(coral_amd_ctx &c) {
    const int s = c.P.nstages;
    for (auto &d : c.ks) std::swap(d->field.p, d->X[s - 1]->p);
    for (auto &d : c.zs) std::swap(d->field.p, d->X[s - 1]->p);
    c.gpar ^= 1;
    ++c.nswap;
}

// f's launches (and timer events) go to the side stream
template <class F>
void on_side(coral_amd_ctx &c, F &&f) {
    std::swap(c.stream, c.side);
    try {
        f();
    } catch (...) {
        std::swap(c.stream, c.side);
        throw;
    }
    std::swap(c.stream, c.side);
}

// Chebyshev coefficients of all full variables from the given states, then the
// backward transforms up to G.
// fp32: the FP32 chain into spec32 / Fz32 / G32 (the step in precision mode 1)
void backward_to_G(coral_amd_ctx &c, const std::vector<const double2 *> &kst, const std::vector<const double2 *> &zst,
                   bool fp32 = false) {
    const Problem &P = c.P;
    const size_t fsz = size_t(P.nsys) * P.NZ;
    c.timed("extract", [&] {
        ExtractAllArgs a{};
        a.nfull = P.nfull;
        a.nsys = P.nsys;
        a.NZ = P.NZ;
        a.km = c.km;
        a.spec = c.spec.p;
        if (P.nfull > MAX_FULL) throw std::runtime_error("too many full variables");
        int nt = 0;
        for (int f = 0; f < P.nfull; ++f) {
            a.full_first[f] = nt;
            for (auto &e : c.ext) {
                if (e->kind != 0 || e->full != f) continue;
                if (nt >= MAX_EXTRACT_TERMS) throw std::runtime_error("too many full-variable terms");
                SysDev &d = *c.ks[e->system];
                int q = 0, qoff = 0;
                while (q < a.nstate && a.state[q] != kst[e->system]) qoff += a.stotal[q++];
                if (q == a.nstate && q < MAX_EXTRACT_STATES) {
                    a.state[q] = kst[e->system];
                    a.stotal[q] = d.t->total;
                    ++a.nstate;
                }
                a.term[nt++] = ExtractTerm{e->rowptr.p, e->col.p, e->val.p, kst[e->system], d.t->total, e->kx, e->ky,
                                           q < MAX_EXTRACT_STATES ? q : -1, qoff};
            }
        }
        a.full_first[P.nfull] = nt;
        a.nterms = nt;
        if (fp32) a.spec32 = c.spec32.p;
        if (P.nfull > 0) launch_extract_all(c.stream, a);
        join_mean(c);   // the mean states below come from the side stream
        for (auto &e : c.ext) {
            if (e->kind != 1) continue;
            if (fp32)
                launch_extract_mean32(c.stream, zst[e->system], P.NZ, P.nsys, e->rowptr.p, e->col.p, e->val.p,
                                      c.spec32.p + e->full * fsz);
            else
                launch_extract_mean(c.stream, zst[e->system], P.NZ, P.nsys, e->rowptr.p, e->col.p, e->val.p,
                                    c.spec.p + e->full * fsz);
        }
        // boundary values and vertical decompositions of the full variables
        // (marching.f90:544-577), on their complete coefficients
        for (int f = 0; f < P.nfull && f < (int)c.full_filter.size(); ++f) {
            const int ex = c.full_filter[f].first, zd = c.full_filter[f].second;
            if (!ex && !zd) continue;
            if (fp32)
                launch_full_zfilter32(c.stream, c.spec32.p + f * fsz, P.NZ, P.nsys, ex, zd);
            else
                launch_full_zfilter(c.stream, c.spec.p + f * fsz, P.NZ, P.nsys, ex, zd);
        }
    });
    if (fp32) {
        c.timed("zbwd", [&] { launch_zbwd32(c.stream, c.z32, c.spec32.p, c.Fz32.p, P.NZ, P.NZAA, P.nsys, P.nfull); });
        c.timed("ybwd", [&] {
            launch_ypass32(c.stream, false, c.py, c.Fz32.p, c.G32.p, P.nfull, P.NZAA, P.nkx, P.nky, P.NY);
        });
        return;
    }
    c.timed("zbwd", [&] { zbwd(c, c.spec.p, c.Fz.p, P.nfull); });
    c.timed("ybwd", [&] { launch_ybwd(c.stream, c.py, c.Fz.p, c.G.p, P.nfull, P.NZAA, P.nkx, P.nky, P.NY); });
}

// The quadratic variables of the given states in Qhat ([quad][NZ][s]): full
// variables to the 3/2 grid, products, forward transforms
// (marching.f90:279-300), z-integral removal (PL_IMEX_timestepping.f90:223-235);
// the CFL maxima of the full variables on the way.
// ---- full variables with physical values of their own -----------------------
// nuSmago's profile (compute_full_variables_in_physical_space,
// PL_IMEX_timestepping_marching.f90:514-541): out[z] = pf sqrt(prev[z] + the
// plane's mean of the summed squares of the gradient fields), prev its last
// values.  One workgroup per plane; a fixed-order tree keeps it deterministic.
__global__ __launch_bounds__(256) void k_smago_profile(const double *__restrict__ phys, int4 slots, int nslot,
                                                       long long plane, int N2, double pf,
                                                       const double *__restrict__ prev, double *__restrict__ out) {
    __shared__ double red[256];
    const int z = blockIdx.x;
    const int sl[4] = {slots.x, slots.y, slots.z, slots.w};
    double acc = 0.0;
    for (int i = threadIdx.x; i < N2; i += 256) {
        double t = 0.0;
        for (int k = 0; k < nslot; ++k) {
            const double v = phys[sl[k] * plane + (long long)z * N2 + i];
            t = fma(v, v, t);
        }
        acc += t;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[z] = pf * sqrt(prev[z] + red[0] / double(N2));
}

// tau[k] = a * b over a plane block for each side quadratic variable k; a
// factor is desc kind 0: a transformed full variable (its slot in reg), 1: a
// fixed field (its slot in fixed), 2: the nuSmago profile, 3: ONE
__device__ inline double side_factor(int kind, int slot, const double *reg, const double *fixed, const double *nu,
                                     long long plane, long long i, int N2) {
    if (kind == 0) return reg[slot * plane + i];
    if (kind == 1) return fixed[slot * plane + i];
    if (kind == 2) return nu[i / N2];
    return 1.0;
}
__global__ __launch_bounds__(256) void k_side_products(const double *__restrict__ reg, const double *__restrict__ fixed,
                                                       const double *__restrict__ nu, const int4 *__restrict__ desc,
                                                       long long plane, int N2, double *__restrict__ tau) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= plane) return;
    const int4 d = desc[blockIdx.y];
    const double a = side_factor(d.x, d.y, reg, fixed, nu, plane, i, N2);
    const double b = side_factor(d.z, d.w, reg, fixed, nu, plane, i, N2);
    tau[blockIdx.y * plane + i] = d.z == 3 ? a : (d.x == 3 ? b : a * b);
}

__global__ void k_copy_c(double2 *__restrict__ y, const double2 *__restrict__ x, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

__global__ void k_broadcast_profile(const double *__restrict__ prof, long long plane, int N2, double *__restrict__ out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < plane) out[i] = prof[i / N2];
}

int side_kind(const coral_amd_ctx &c, int f) {
    if (f < 0) return 3;
    if (c.fixed_slot.size() > size_t(f) && c.fixed_slot[f] >= 0) return 1;
    if (f == c.nu_full) return 2;
    return 0;
}

// The side path's tables and buffers for the current fixed fields and
// Smagorinsky setting (changes the launches of a step: graphs are dropped)
void rebuild_side(coral_amd_ctx &c) {
    const Problem &P = c.P;
    const long long plane = (long long)P.NZAA * P.NYAA * P.NXAA;
    c.side_q.clear();
    c.side_reg.clear();
    auto reg_slot = [&](int f) {
        for (size_t k = 0; k < c.side_reg.size(); ++k)
            if (c.side_reg[k] == f) return int(k);
        c.side_reg.push_back(f);
        return int(c.side_reg.size()) - 1;
    };
    std::vector<int4> desc;
    if (c.nu_full >= 0)
        for (int f : c.nu_grads) reg_slot(f);
    for (int q = 0; q < P.nquad; ++q) {
        const int ka = side_kind(c, P.quad_a[q]), kb = side_kind(c, P.quad_b[q]);
        if (ka != 1 && ka != 2 && kb != 1 && kb != 2) continue;
        auto slot = [&](int kind, int f) { return kind == 0 ? reg_slot(f) : kind == 1 ? c.fixed_slot[f] : 0; };
        const int sa = slot(ka, P.quad_a[q]), sb = slot(kb, P.quad_b[q]);
        c.side_q.push_back(q);
        desc.push_back(make_int4(ka, sa, kb, sb));
    }
    HIP_OK(hipDeviceSynchronize());
    c.drop_graphs();
    ++c.gen;
    const size_t ns = c.side_q.size();
    c.side_desc.upload(std::vector<int>(reinterpret_cast<const int *>(desc.data()),
                                        reinterpret_cast<const int *>(desc.data()) + 4 * ns));
    c.side_phys.alloc(c.side_reg.size() * plane);
    c.side_tau.alloc(ns * plane);
    c.side_Q.alloc(ns * size_t(P.nsys) * P.NZ);
}

// The quadratic variables with a fixed or Smagorinsky factor, formed in
// physical space as the reference forms every product
// (PL_IMEX_timestepping_marching.f90:279-300): x c2r of the full variables
// they use (from G, as the products left it), nuSmago's profile, the
// products, and the step's forward transforms; their spectra replace what
// the fused x pass produced for them.
void side_products(coral_amd_ctx &c) {
    if (c.side_q.empty() && c.nu_full < 0) return;
    const Problem &P = c.P;
    const long long nlines = (long long)P.NZAA * P.NYAA, plane = nlines * P.NXAA;
    const int N2 = P.NYAA * P.NXAA;
    c.timed("side_products", [&] {
        for (size_t k = 0; k < c.side_reg.size(); ++k)
            launch_xbwd(c.stream, c.px, c.G.p + (size_t)c.side_reg[k] * nlines * P.nkx, c.side_phys.p + k * plane, 1,
                        nlines, P.nkx);
        const double *nu = c.nu_prof.p;
        if (c.nu_full >= 0) {
            int sl[4] = {0, 0, 0, 0};
            for (size_t k = 0; k < c.nu_grads.size(); ++k) sl[k] = int(k);   // the gradients take the first slots
            double *out = c.side_advance ? c.nu_prof.p : c.nu_tmp.p;
            hipLaunchKernelGGL(k_smago_profile, dim3(unsigned(P.NZAA)), dim3(256), 0, c.stream, c.side_phys.p,
                               make_int4(sl[0], sl[1], sl[2], sl[3]), int(c.nu_grads.size()), plane, N2,
                               c.nu_prefactor, c.nu_prof.p, out);
            check_launch("smagorinsky profile");
            nu = out;
        }
        const int ns = int(c.side_q.size());
        if (!ns) return;
        hipLaunchKernelGGL(k_side_products, dim3(unsigned((plane + 255) / 256), unsigned(ns)), dim3(256), 0, c.stream,
                           c.side_phys.p, c.fixed_all.p, nu, reinterpret_cast<const int4 *>(c.side_desc.p), plane,
                           N2, c.side_tau.p);
        check_launch("side products");
        launch_xfwd(c.stream, c.px, c.side_tau.p, c.H.p, ns, nlines, P.nkx);
        launch_yfwd(c.stream, c.py, c.H.p, c.Hy.p, ns, P.NZAA, P.nkx, P.nky, P.NY);
        zfwd(c, c.Hy.p, c.side_Q.p, ns);
        const size_t per = size_t(P.nsys) * P.NZ;
        for (int k = 0; k < ns; ++k) {
            const int q = c.side_q[k];
            if (P.quad_filter[q]) launch_remove_zmean(c.stream, c.side_Q.p + k * per, P.NZ, P.nsys);
            hipLaunchKernelGGL(k_copy_c, dim3(unsigned((per + 255) / 256)), dim3(256), 0, c.stream,
                               c.Qhat.p + q * per, c.side_Q.p + k * per, per);
            check_launch("side copy");
        }
    });
}

void products_to_Qhat(coral_amd_ctx &c, const std::vector<const double2 *> &kst,
                      const std::vector<const double2 *> &zst, bool fp32 = false) {
    const Problem &P = c.P;
    backward_to_G(c, kst, zst, fp32);
    zero_cflmax(c.stream, c.cflmax.p);
    XPassArgs xa{};
    xa.G = fp32 ? reinterpret_cast<const double2 *>(c.G32.p) : c.G.p;
    xa.H = fp32 ? reinterpret_cast<double2 *>(c.H32.p) : c.H.p;
    xa.fp32 = fp32 ? 1 : 0;
    xa.nfull = P.nfull;
    xa.nquad = P.nquad;
    xa.NZAA = P.NZAA;
    xa.NYAA = P.NYAA;
    xa.nkx = P.nkx;
    xa.qa = c.qa.p;
    xa.qb = c.qb.p;
    xa.zweight = c.zweight.p;
    xa.cflmax = c.cflmax.p;
    xa.max_wpb = c.xpass_wpb;
    c.timed("xpass", [&] { launch_xpass(c.stream, c.px, xa); });
    if (fp32) {
        c.timed("yfwd", [&] {
            launch_ypass32(c.stream, true, c.py, c.H32.p, c.Hy32.p, P.nquad, P.NZAA, P.nkx, P.nky, P.NY);
        });
        c.timed("zfwd", [&] { launch_zfwd32(c.stream, c.z32, c.Hy32.p, c.Qhat32.p, P.NZ, P.NZAA, P.nsys, P.nquad); });
        for (int q = 0; q < P.nquad; ++q)
            if (P.quad_filter[q])
                launch_remove_zmean32(c.stream, c.Qhat32.p + size_t(q) * P.nsys * P.NZ, P.NZ, P.nsys);
        return;
    }
    c.timed("yfwd", [&] { launch_yfwd(c.stream, c.py, c.H.p, c.Hy.p, P.nquad, P.NZAA, P.nkx, P.nky, P.NY); });
    c.timed("zfwd", [&] { zfwd(c, c.Hy.p, c.Qhat.p, P.nquad); });
    for (int q = 0; q < P.nquad; ++q)
        if (P.quad_filter[q]) launch_remove_zmean(c.stream, c.Qhat.p + size_t(q) * P.nsys * P.NZ, P.NZ, P.nsys);
    side_products(c);
}

// ---- learned closure: the network's correction of the quadratic variables ----
__device__ inline unsigned short bf16_of(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// network input rows of planes z0 .. z0 + P - 1: the full variables on the
// point's 3 x 3 x 3 neighbourhood (periodic in x and y, clamped at the walls
// in z; offsets dz, dy, dx = -1, 0, 1 in that nesting, the variables
// innermost), the point's own values as they are and every neighbour as its
// difference from them (formed in FP64, so that the BF16 inputs keep the
// local differences' digits), normalised ((v - mean) * (1 / std) in FP64,
// rounded to FP32 then BF16), the height of the plane, normalised likewise, and zero
// columns up to cin (a multiple of 8, for 16-byte rows).  A workgroup fills
// PACK_ROWS consecutive rows (each neighbouring line read once into the
// caches): the rows' positions are decoded once into LDS, and each thread
// keeps one column (its neighbour offset and variable decoded once) over
// rows rs apart, so that the stores of a sweep are contiguous and no
// division remains per element.
constexpr int PACK_ROWS = 64;
__global__ __launch_bounds__(256) void k_closure_pack(const double *__restrict__ phys,
                                                      const double *__restrict__ zgrid, int z0, int P, int NXAA,
                                                      int NYAA, int NZAA, int nfull, int cin,
                                                      const float *__restrict__ mu, const float *__restrict__ sd,
                                                      unsigned short *__restrict__ X16) {
    __shared__ int rz[PACK_ROWS], ry[PACK_ROWS], rx[PACK_ROWS];
    const int N2 = NXAA * NYAA, ncore = 27 * nfull, nrow = P * N2;
    const int r0 = blockIdx.x * PACK_ROWS, nr = min(PACK_ROWS, nrow - r0);
    for (int rl = threadIdx.x; rl < nr; rl += blockDim.x) {
        const int r = r0 + rl, zl = r / N2, i = r - zl * N2, y = i / NXAA;
        rz[rl] = z0 + zl;
        ry[rl] = y;
        rx[rl] = i - y * NXAA;
    }
    __syncthreads();
    unsigned short *out = X16 + (size_t)r0 * cin;
    const int nt = blockDim.x, rs = cin <= nt ? nt / cin : 1;
    for (int cbase = 0; cbase < cin; cbase += nt) {
        int c, ro;
        if (cin <= nt) {
            c = threadIdx.x % cin;
            ro = threadIdx.x / cin;
            if (ro >= rs) break;
        } else {
            c = cbase + threadIdx.x;
            ro = 0;
            if (c >= cin) break;
        }
        const int o = c < ncore ? c / nfull : 0, f = c < ncore ? c - o * nfull : 0;
        const int dz = o / 9 - 1, dy = (o / 3) % 3 - 1, dx = o % 3 - 1;
        const double m = double(mu[c]), rinv = 1.0 / double(sd[c]);
        for (int rl = ro; rl < nr; rl += rs) {
            const int z = rz[rl], y = ry[rl], x = rx[rl];
            double d = 0.0;
            if (c < ncore) {
                const double v0 = phys[((size_t)f * NZAA + z) * N2 + (size_t)y * NXAA + x];
                if (o == 13) {
                    d = v0;
                } else {
                    const int zz = min(max(z + dz, 0), NZAA - 1);
                    const int yy = y + dy < 0 ? NYAA - 1 : (y + dy == NYAA ? 0 : y + dy);
                    const int xx = x + dx < 0 ? NXAA - 1 : (x + dx == NXAA ? 0 : x + dx);
                    d = phys[((size_t)f * NZAA + zz) * N2 + (size_t)yy * NXAA + xx] - v0;
                }
            } else if (c == ncore) {
                d = zgrid[z];
            }
            out[(size_t)rl * cin + c] = bf16_of(float((d - m) * rinv));
        }
        if (cin <= nt) break;
    }
}

// The same rows, one workgroup per x-line (z, y) of NXAA points: the 3 x 3
// neighbouring lines (dz, dy = -1, 0, 1, periodic in y, clamped in z) of
// every full variable staged in LDS with coalesced loads, then the line's
// NXAA rows written whole (each thread keeps one column over rows rs apart).
// Values and rounding are those of k_closure_pack.  LDS: 9 nfull NXAA doubles.
__global__ __launch_bounds__(256) void k_closure_pack_lines(const double *__restrict__ phys,
                                                            const double *__restrict__ zgrid, int z0, int NXAA,
                                                            int NYAA, int NZAA, int nfull, int cin,
                                                            const float *__restrict__ mu,
                                                            const float *__restrict__ sd,
                                                            unsigned short *__restrict__ X16) {
    extern __shared__ double tile[];   // [dz][dy][f][x]
    const int N2 = NXAA * NYAA, ncore = 27 * nfull;
    const int line = blockIdx.x, zl = line / NYAA, y = line - zl * NYAA, z = z0 + zl;
    for (int e = threadIdx.x; e < 9 * nfull * NXAA; e += blockDim.x) {
        const int x = e % NXAA, t = e / NXAA, f = t % nfull, o = t / nfull;   // o = 3 (dz+1) + (dy+1)
        const int dz = o / 3 - 1, dy = o % 3 - 1;
        const int zz = min(max(z + dz, 0), NZAA - 1);
        const int yy = y + dy < 0 ? NYAA - 1 : (y + dy == NYAA ? 0 : y + dy);
        tile[e] = phys[((size_t)f * NZAA + zz) * N2 + (size_t)yy * NXAA + x];
    }
    __syncthreads();
    unsigned short *out = X16 + (size_t)line * NXAA * cin;
    const int nt = blockDim.x, rs = cin <= nt ? nt / cin : 1;
    for (int cbase = 0; cbase < cin; cbase += nt) {
        int c, ro;
        if (cin <= nt) {
            c = threadIdx.x % cin;
            ro = threadIdx.x / cin;
            if (ro >= rs) break;
        } else {
            c = cbase + threadIdx.x;
            ro = 0;
            if (c >= cin) break;
        }
        const int o = c < ncore ? c / nfull : 0, f = c < ncore ? c - o * nfull : 0;
        const int dz = o / 9 - 1, dy = (o / 3) % 3 - 1, dx = o % 3 - 1;
        const double m = double(mu[c]), rinv = 1.0 / double(sd[c]);
        const double *lc = tile + (size_t)(4 * nfull + f) * NXAA;                       // dz = dy = 0
        const double *ln = tile + (size_t)((3 * (dz + 1) + (dy + 1)) * nfull + f) * NXAA;
        const double hz = zgrid[z];
        for (int x = ro; x < NXAA; x += rs) {
            double d = 0.0;
            if (c < ncore) {
                const double v0 = lc[x];
                if (o == 13) {
                    d = v0;
                } else {
                    const int xx = x + dx < 0 ? NXAA - 1 : (x + dx == NXAA ? 0 : x + dx);
                    d = ln[xx] - v0;
                }
            } else if (c == ncore) {
                d = hz;
            }
            out[(size_t)x * cin + c] = bf16_of(float((d - m) * rinv));
        }
        if (cin <= nt) break;
    }
}

// ---- the network's lift and projection fused with the packing and unpacking ----
typedef __bf16 cbf8 __attribute__((ext_vector_type(8)));
typedef float cf4 __attribute__((ext_vector_type(4)));
// FP32 -> BF16 with gfx950's conversion (as the network kernels' epilogues)
__device__ inline unsigned short f2bf_hw(float f) { return __builtin_bit_cast(unsigned short, static_cast<__bf16>(f)); }

// Rows per workgroup of the fused lift and projection: a tile lies inside one
// x-line, so NXAA % LIFT_ROWS == 0 is required.
constexpr int LIFT_ROWS = 64;

struct LiftShape {
    int kp = 0, ldk = 0;   // K rounded up to the MFMA depth; LDS row stride (bf16)
    size_t bytes = 0;      // dynamic LDS
};
LiftShape lift_shape(int nfull, int cin, int W) {
    LiftShape s;
    s.kp = (cin + 31) / 32 * 32;
    s.ldk = s.kp + 8;
    // the input tile (bf16 [LIFT_ROWS][ldk]), later reused by the epilogue's
    // FP32 tile [LIFT_ROWS][W + 4]
    const size_t ops = sizeof(unsigned short) * size_t(LIFT_ROWS) * s.ldk;
    const size_t epi = sizeof(float) * size_t(LIFT_ROWS) * (W + 4);
    s.bytes = sizeof(double) * (size_t(9) * nfull * (LIFT_ROWS + 2) + 2 * size_t(s.kp)) + std::max(ops, epi);
    return s;
}

// H = X W0 + b0 (FP32), H16 = bf16(H), H16T its tiled transposed copy (when
// given), for LIFT_ROWS rows of the chunk at a time, X formed on chip: the
// input rows of k_closure_pack_lines (same values and rounding) are built in
// LDS from the 3 x 3 neighbouring lines of every full variable and multiplied
// on MFMA with the K order, zero padding and epilogue of nn::gemm_bias, so H,
// H16 and H16T equal the pack-then-GEMM sequence's bit for bit without the
// input rows ever reaching HBM.  One wavefront per 16 rows, all W columns.
template <int W>
__global__ __launch_bounds__(256) void k_closure_lift(const double *__restrict__ phys,
                                                      const double *__restrict__ zgrid, int z0, int NXAA, int NYAA,
                                                      int NZAA, int nfull, int cin, int kp, int ldk,
                                                      const float *__restrict__ mu, const float *__restrict__ sd,
                                                      const unsigned short *__restrict__ W0T,
                                                      const float *__restrict__ b0, float *__restrict__ H,
                                                      unsigned short *__restrict__ H16,
                                                      unsigned short *__restrict__ H16T) {
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr int LR = LIFT_ROWS, NT = W / 16, span = LR + 2;
    const int N2 = NXAA * NYAA, ncore = 27 * nfull, ntile = 9 * nfull * span;
    double *tile = reinterpret_cast<double *>(lds);   // [dz][dy][f][x0-1 .. x0+LR]
    double *mv = tile + ntile, *rv = mv + kp;
    unsigned short *As = reinterpret_cast<unsigned short *>(rv + kp);   // [LR][ldk]
    const long long r0 = (long long)blockIdx.x * LR;
    const int zl = int(r0 / N2), rem = int(r0 - (long long)zl * N2), y = rem / NXAA, x0 = rem - y * NXAA;
    const int z = z0 + zl, tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6, fr = lane & 15, fk = 8 * (lane >> 4), rq = 4 * (lane >> 4);
    // the B fragments (W0^T rows, zero past cin) straight from memory, issued first
    constexpr int KMAX = 8;   // kp <= 256
    cbf8 bfr[KMAX][NT];
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
#pragma unroll
        for (int j = 0; j < NT; ++j)
            if (32 * k < kp) bfr[k][j] = *reinterpret_cast<const cbf8 *>(W0T + (size_t)(16 * j + fr) * kp + 32 * k + fk);
    // the neighbouring lines: wavefront wv takes lines wv, wv + 4, ... (line
    // L = o nfull + f, o = 3 (dz+1) + (dy+1)), a lane the points lane and 64 +
    // lane of the line's span; up to four lines' loads in flight before their
    // LDS stores
    {
        const int nlines = 9 * nfull;
        int xa = x0 - 1 + lane, xb = x0 + 63 + lane;
        xa = xa < 0 ? xa + NXAA : xa;
        xb = xb >= NXAA ? xb - NXAA : xb;
        const bool hasb = lane < span - 64;
        for (int L0 = wv; L0 < nlines; L0 += 16) {
            double va[4], vb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int L = L0 + 4 * i;
                if (L < nlines) {
                    const int o = L / nfull, f = L - o * nfull;
                    const int dz = o / 3 - 1, dy = o - 3 * (o / 3) - 1;
                    const int zz = min(max(z + dz, 0), NZAA - 1);
                    const int yy = y + dy < 0 ? NYAA - 1 : (y + dy == NYAA ? 0 : y + dy);
                    const double *row = phys + ((size_t)f * NZAA + zz) * N2 + (size_t)yy * NXAA;
                    va[i] = row[xa];
                    vb[i] = hasb ? row[xb] : 0.0;
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int L = L0 + 4 * i;
                if (L < nlines) {
                    tile[L * span + lane] = va[i];
                    if (hasb) tile[L * span + 64 + lane] = vb[i];
                }
            }
        }
    }
    for (int c = tid; c < kp; c += 256) {
        mv[c] = c < cin ? double(mu[c]) : 0.0;
        rv[c] = c < cin ? 1.0 / double(sd[c]) : 0.0;
    }
    __syncthreads();
    {
        // each thread keeps one column (its neighbour offset and variable
        // decoded once) over rows 256 / kp apart (kp <= 256)
        const int cpt = 256 / kp, c = tid % kp, rs = tid / kp;
        if (rs < cpt) {
            const bool core = c < ncore;
            const int o = core ? c / nfull : 13, f = core ? c - o * nfull : 0;
            const int dz = o / 9 - 1, dy = (o / 3) % 3 - 1, dx = o % 3 - 1;
            const double *lc = tile + (4 * nfull + f) * span + 1;
            const double *ln = tile + ((3 * (dz + 1) + (dy + 1)) * nfull + f) * span + 1 + dx;
            const double m = mv[c], ri = rv[c], hz = zgrid[z];
            for (int rl = rs; rl < LR; rl += cpt) {
                unsigned short h = 0;
                if (c < cin) {
                    double d = 0.0;
                    if (core) {
                        const double v0 = lc[rl];
                        d = o == 13 ? v0 : ln[rl] - v0;
                    } else if (c == ncore) {
                        d = hz;
                    }
                    h = bf16_of(float((d - m) * ri));
                }
                As[rl * ldk + c] = h;
            }
        }
    }
    __syncthreads();
    cf4 acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = cf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (32 * k < kp) {
            const cbf8 a = *reinterpret_cast<const cbf8 *>(&As[(16 * wv + fr) * ldk + 32 * k + fk]);
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bfr[k][j], acc[j], 0, 0, 0);
        }
    }
    // epilogue through LDS (the operand tiles are done with): the tile's H
    // rows, their BF16 copy and its two 32-row blocks of the tiled transposed
    // layout are each one contiguous run of memory, written with 16-byte stores
    __syncthreads();
    constexpr int LDH = W + 4;
    float *Hs = reinterpret_cast<float *>(As);   // [LR][LDH]
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int gn = 16 * j + fr;
        const float bias = b0[gn];
#pragma unroll
        for (int q = 0; q < 4; ++q) Hs[(16 * wv + rq + q) * LDH + gn] = acc[j][q] + bias;
    }
    __syncthreads();
    float4 *Ho = reinterpret_cast<float4 *>(H + r0 * W);
    for (int e = tid; e < LR * W / 4; e += 256) {
        const int rl = (4 * e) / W, n = 4 * e - rl * W;
        const float *v = Hs + rl * LDH + n;
        Ho[e] = make_float4(v[0], v[1], v[2], v[3]);
    }
    uint4 *H16o = reinterpret_cast<uint4 *>(H16 + r0 * W);
    for (int e = tid; e < LR * W / 8; e += 256) {
        const int rl = (8 * e) / W, n = 8 * e - rl * W;
        const float *v = Hs + rl * LDH + n;
        uint4 u;
        u.x = unsigned(f2bf_hw(v[0])) | (unsigned(f2bf_hw(v[1])) << 16);
        u.y = unsigned(f2bf_hw(v[2])) | (unsigned(f2bf_hw(v[3])) << 16);
        u.z = unsigned(f2bf_hw(v[4])) | (unsigned(f2bf_hw(v[5])) << 16);
        u.w = unsigned(f2bf_hw(v[6])) | (unsigned(f2bf_hw(v[7])) << 16);
        H16o[e] = u;
    }
    if (H16T) {
        // block t (32 rows) holds W columns of 32 consecutive rows
        uint4 *Ht = reinterpret_cast<uint4 *>(H16T + r0 * W);
        for (int e = tid; e < LR * W / 8; e += 256) {
            const int o = 8 * e, t = o / (32 * W), n = (o - t * 32 * W) / 32, rr = o & 31;
            const float *v = Hs + (32 * t + rr) * LDH + n;
            uint4 u;
            u.x = unsigned(f2bf_hw(v[0])) | (unsigned(f2bf_hw(v[LDH])) << 16);
            u.y = unsigned(f2bf_hw(v[2 * LDH])) | (unsigned(f2bf_hw(v[3 * LDH])) << 16);
            u.z = unsigned(f2bf_hw(v[4 * LDH])) | (unsigned(f2bf_hw(v[5 * LDH])) << 16);
            u.w = unsigned(f2bf_hw(v[6 * LDH])) | (unsigned(f2bf_hw(v[7 * LDH])) << 16);
            Ht[e] = u;
        }
    }
}

// tau[q][z][y][x] = (H16 Wp + bp) std + mean for LIFT_ROWS rows of the chunk:
// the network's projection on MFMA with the K order and epilogue of nn::gemm
// and nn::bias_act, the unpacking of k_closure_unpack, in one pass (the
// outputs of a tile are staged in LDS so that each variable's 64 values are
// one contiguous store).  cout <= 16.
template <int W>
__global__ __launch_bounds__(256) void k_closure_project(const unsigned short *__restrict__ H16,
                                                         const unsigned short *__restrict__ Wp,
                                                         const float *__restrict__ bp, int cout, int z0, int N2,
                                                         int NZAA, const float *__restrict__ mu,
                                                         const float *__restrict__ sd, double *__restrict__ tau) {
    constexpr int LR = LIFT_ROWS, LDB = W + 8, KS = W / 32;
    __shared__ __align__(16) unsigned short Bs[16 * LDB];
    __shared__ double outs[16 * LR];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, fr = lane & 15, fk = 8 * (lane >> 4);
    const int rq = 4 * (lane >> 4);
    const long long r0 = (long long)blockIdx.x * LR;
    cbf8 a[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k)
        a[k] = *reinterpret_cast<const cbf8 *>(H16 + (r0 + 16 * wv + fr) * W + 32 * k + fk);
    for (int e = tid; e < 16 * W; e += 256) {
        const int n = e / W, k = e - n * W;
        Bs[n * LDB + k] = n < cout ? Wp[k * cout + n] : (unsigned short)0;
    }
    __syncthreads();
    cf4 acc = cf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const cbf8 b = *reinterpret_cast<const cbf8 *>(&Bs[fr * LDB + 32 * k + fk]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[k], b, acc, 0, 0, 0);
    }
    if (fr < cout) {
        const float bias = bp[fr];
        const double s = double(sd[fr]), m = double(mu[fr]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float yv = acc[q] + bias;
            outs[fr * LR + 16 * wv + rq + q] = double(yv) * s + m;
        }
    }
    __syncthreads();
    const int zl = int(r0 / N2), i0 = int(r0 - (long long)zl * N2);
    for (int e = tid; e < cout * LR; e += 256) {
        const int q = e / LR, rr = e - q * LR;
        tau[((size_t)q * NZAA + z0 + zl) * N2 + i0 + rr] = outs[e];
    }
}

// the network's output rows of planes z0 .. back to physical corrections
// tau[q][z][y][x] = Y * std + mean
__global__ void k_closure_unpack(const float *__restrict__ Y, int z0, int P, int N2, int NZAA, int nquad,
                                 const float *__restrict__ mu, const float *__restrict__ sd, double *__restrict__ tau) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long long)P * N2) return;
    const int z = z0 + int(r / N2), i = int(r % N2);
    for (int q = 0; q < nquad; ++q)
        tau[((size_t)q * NZAA + z) * N2 + i] = double(Y[r * nquad + q]) * double(sd[q]) + double(mu[q]);
}

__global__ void k_add_c(double2 *__restrict__ y, const double2 *__restrict__ x, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = cadd(y[i], x[i]);
}

// Qhat += spectrum of the network's correction for the full variables in G
// (the stage's y-physical fields, as the products left them): x c2r of the
// full variables, the network plane by plane, the forward transforms of the
// corrections with the step's kernels (z-integral removal as for the
// variables themselves)
void closure_correction(coral_amd_ctx &c) {
    const Problem &P = c.P;
    auto &cl = c.closure;
    const long long nlines = (long long)P.NZAA * P.NYAA;
    const int N2 = P.NYAA * P.NXAA;
    c.timed("closure", [&] {
        launch_xbwd(c.stream, c.px, c.G.p, cl.phys_full.p, P.nfull, nlines, P.nkx);
        for (int z0 = 0; z0 < P.NZAA; z0 += cl.chunk) {
            const int np = std::min(cl.chunk, P.NZAA - z0);
            const unsigned g = unsigned(((long long)np * N2 + 255) / 256);
            if (cl.fused_io) {
                const LiftShape ls = lift_shape(P.nfull, cl.net.cin, cl.net.width);
                const unsigned nt = unsigned((long long)np * N2 / LIFT_ROWS);
                unsigned short *h16t = nn::net_tiled(cl.net, cl.ws) ? cl.ws.H16T : nullptr;
                const unsigned short *W0 = cl.W0T.p;
                const float *b0 = cl.net.w + cl.net.off_b0;
                if (cl.net.width == 32)
                    hipLaunchKernelGGL(k_closure_lift<32>, dim3(nt), dim3(256), ls.bytes, c.stream, cl.phys_full.p,
                                       cl.zgrid.p, z0, P.NXAA, P.NYAA, P.NZAA, P.nfull, cl.net.cin, ls.kp, ls.ldk,
                                       cl.mu_in.p, cl.sd_in.p, W0, b0, cl.ws.H, cl.ws.H16, h16t);
                else
                    hipLaunchKernelGGL(k_closure_lift<64>, dim3(nt), dim3(256), ls.bytes, c.stream, cl.phys_full.p,
                                       cl.zgrid.p, z0, P.NXAA, P.NYAA, P.NZAA, P.nfull, cl.net.cin, ls.kp, ls.ldk,
                                       cl.mu_in.p, cl.sd_in.p, W0, b0, cl.ws.H, cl.ws.H16, h16t);
                check_launch("closure lift");
                nn::net_forward(c.stream, cl.net, nullptr, np, nullptr, cl.ws);
                const unsigned short *Wp = cl.net.w16 + cl.net.off_Wp;
                const float *bp = cl.net.w + cl.net.off_bp;
                if (cl.net.width == 32)
                    hipLaunchKernelGGL(k_closure_project<32>, dim3(nt), dim3(256), 0, c.stream, cl.ws.H16, Wp, bp,
                                       cl.net.cout, z0, N2, P.NZAA, cl.mu_out.p, cl.sd_out.p, cl.phys_tau.p);
                else
                    hipLaunchKernelGGL(k_closure_project<64>, dim3(nt), dim3(256), 0, c.stream, cl.ws.H16, Wp, bp,
                                       cl.net.cout, z0, N2, P.NZAA, cl.mu_out.p, cl.sd_out.p, cl.phys_tau.p);
                check_launch("closure projection");
                continue;
            }
            const size_t tile = size_t(9) * P.nfull * P.NXAA * sizeof(double);
            if (tile <= 65536) {
                hipLaunchKernelGGL(k_closure_pack_lines, dim3(unsigned(np * P.NYAA)), dim3(256), tile, c.stream,
                                   cl.phys_full.p, cl.zgrid.p, z0, P.NXAA, P.NYAA, P.NZAA, P.nfull, cl.net.cin,
                                   cl.mu_in.p, cl.sd_in.p, cl.X16.p);
            } else {
                const unsigned gp = unsigned(((long long)np * N2 + PACK_ROWS - 1) / PACK_ROWS);
                hipLaunchKernelGGL(k_closure_pack, dim3(gp), dim3(256), 0, c.stream, cl.phys_full.p, cl.zgrid.p, z0,
                                   np, P.NXAA, P.NYAA, P.NZAA, P.nfull, cl.net.cin, cl.mu_in.p, cl.sd_in.p, cl.X16.p);
            }
            check_launch("closure pack");
            nn::net_forward(c.stream, cl.net, cl.X16.p, np, cl.Y.p, cl.ws);
            hipLaunchKernelGGL(k_closure_unpack, dim3(g), dim3(256), 0, c.stream, cl.Y.p, z0, np, N2, P.NZAA, P.nquad,
                               cl.mu_out.p, cl.sd_out.p, cl.phys_tau.p);
            check_launch("closure unpack");
        }
        launch_xfwd(c.stream, c.px, cl.phys_tau.p, c.H.p, P.nquad, nlines, P.nkx);
        launch_yfwd(c.stream, c.py, c.H.p, c.Hy.p, P.nquad, P.NZAA, P.nkx, P.nky, P.NY);
        zfwd(c, c.Hy.p, cl.Qtau.p, P.nquad);
        for (int q = 0; q < P.nquad; ++q)
            if (P.quad_filter[q]) launch_remove_zmean(c.stream, cl.Qtau.p + size_t(q) * P.nsys * P.NZ, P.NZ, P.nsys);
        const size_t n = size_t(P.nquad) * P.nsys * P.NZ;
        hipLaunchKernelGGL(k_add_c, dim3(unsigned((n + 255) / 256)), dim3(256), 0, c.stream, c.Qhat.p, cl.Qtau.p, n);
        check_launch("closure add");
    });
}

// the quadratic variables' spectra in Qhat (or Qhat32), with the closure's correction
void quad_spectra(coral_amd_ctx &c, const std::vector<const double2 *> &kst, const std::vector<const double2 *> &zst) {
    products_to_Qhat(c, kst, zst, c.precision == 1);
    if (c.closure.on) closure_correction(c);
}

void nl_terms(coral_amd_ctx &c, const std::vector<const double2 *> &kst, const std::vector<const double2 *> &zst,
              int stage) {
    quad_spectra(c, kst, zst);
    for (auto &d : c.ks) nl_assemble(c, *d, d->Khat[stage]->p);
    if (!c.zs.empty()) HIP_OK(hipEventRecord(c.ev_qhat, c.stream));
}

// NL terms of the mean systems from the Qhat of nl_terms (on the stream they
// are launched on, after ev_qhat)
void nl_terms_mean(coral_amd_ctx &c, int stage) {
    HIP_OK(hipStreamWaitEvent(c.stream, c.ev_qhat, 0));
    for (auto &d : c.zs) nl_assemble(c, *d, d->Khat[stage]->p);
}

// (M - dt a11 L) out = rhs for every wavenumber of a system, with the factors
// of the last set_dt (backsolve_to_aux, PL_IMEX_timestepping_marching.f90:57-58)
// the 16-lane solve's arguments for system d (false: d uses another kernel)
bool solve16_args_of(const coral_amd_ctx &c, const SysDev &d, const double2 *rhs, double2 *out, Solve16Args &sa) {
    if (d.ncol16 <= 0) return false;
    const SystemTables &t = *d.t;
    sa = Solve16Args{};
    sa.L16 = d.real ? (const void *)d.Lr.p : (const void *)d.Lf.p;
    sa.U = d.real ? (const void *)d.Urt.p : (const void *)d.Uft.p;
    sa.ucol = d.ucol.p;
    sa.ustride = d.ustride;
    sa.ncol16 = d.ncol16;
    sa.piv = d.piv.p;
    sa.words = d.words.p;
    sa.rotate = d.real ? 1 : 0;
    sa.ident0 = d.kxky ? 1 : 0;
    sa.total = t.total;
    sa.n = d.blk_n;
    sa.kl = d.blk_nl;
    sa.kv = d.blk_nl + d.blk_nu;
    sa.nsys = d.nsys;
    sa.nblk = d.nblk;
    sa.rhs = rhs;
    sa.x = out;
    sa.sink = c.sink.p;
    return true;
}

void solve_system(coral_amd_ctx &c, SysDev &d, const double2 *rhs, double2 *out, const char *name = nullptr) {
    const SystemTables &t = *d.t;
    c.timed(name ? name : (d.kxky ? "solve" : "solve_mean"), [&] {
        const int kv = t.nl + t.nu, id0 = d.kxky ? 1 : 0;
        bool done = true;
        Solve16Args sa;
        if (solve16_args_of(c, d, rhs, out, sa)) {
            done = d.real ? launch_solve16_real(c.stream, sa) : launch_solve16_complex(c.stream, sa);
        } else if (d.real)
            launch_solve_real(c.stream, d.Lr.p, d.Ur.p + SOLVE16_PAD, d.piv.p, d.rot.p, id0, t.total, t.nl, kv,
                              d.nsys, rhs, out);
        else
            launch_solve(c.stream, d.Lf.p, d.Uf.p + SOLVE16_PAD, d.piv.p, t.total, t.nl, kv, d.nsys, rhs, out);
        if (!done) throw std::runtime_error("16-lane solve rejected a band it was factored for");
    });
}

__global__ void k_sub_inplace(double2 *x, const double2 *f, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = csub(x[i], f[i]);
}

// One sweep of FP64 residual refinement of (M - dt a_ii L) x = rhs:
// e = (M - dt a_ii L) x - rhs (the right-hand-side kernel with x as both the
// mass and the stiffness operand), solve for f with the same factors, x -= f.
void refine_solution(coral_amd_ctx &c, SysDev &d, const double2 *rhs_i, double2 *x, double aii) {
    if (!d.Rres.p || !d.Fcor.p) throw std::runtime_error("refinement buffers missing");
    rhs(c, d, {{x, 1.0}}, {{x, -c.dt * aii}}, {{rhs_i, -1.0}}, d.Rres.p);
    solve_system(c, d, d.Rres.p, d.Fcor.p, d.kxky ? "refine_solve" : "refine_solve_mean");
    const size_t n = size_t(d.nsys) * d.t->total;
    c.timed(d.kxky ? "refine_update" : "refine_update_mean", [&] {
        hipLaunchKernelGGL(k_sub_inplace, dim3(unsigned((n + 255) / 256)), dim3(256), 0, c.stream, x, d.Fcor.p, n);
        check_launch("refine_update");
    });
}

// Hermitian part of the kx = 0 plane: for ky = m > 0 (kept slot jy = m) and its
// partner ky = -m (slot NY - 1 - m), c(m) <- (c(m) + conj c(-m)) / 2 and
// c(-m) <- conj c(m).  The physical fields see only this part (the x c2r
// drops the imaginary part of kx = 0 after the y transform), so it changes
// nothing they show; it removes the anti-Hermitian part, on which the
// Hermitian NL terms never act and which the linear operator alone evolves.
// One workgroup per ky pair, threads over the positions of the system.
__global__ void k_project_kx0(double2 *x, int total, int NY) {
    const int m = blockIdx.x + 1, jm = NY - 1 - m;
    double2 *a = x + (size_t)m * total, *b = x + (size_t)jm * total;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const double2 u = a[i], v = b[i];
        const double2 h = make_double2(0.5 * (u.x + v.x), 0.5 * (u.y - v.y));
        a[i] = h;
        b[i] = make_double2(h.x, -h.y);
    }
}

void project_kx0(coral_amd_ctx &c) {
    const Problem &P = c.P;
    if (P.NY / 2 - 1 < 1) return;
    for (auto &d : c.ks)
        c.timed("project_kx0", [&] {
            hipLaunchKernelGGL(k_project_kx0, dim3(unsigned(P.NY / 2 - 1)), dim3(256), 0, c.stream, d->field.p,
                               d->t->total, P.NY);
            check_launch("project_kx0");
        });
}

// one_step_beyond (PL_IMEX_timestepping_marching.f90:3-48): stage i takes the
// NL terms of the previous stage solution, rhs_i = M x^n + sum_{j<i} dt a_ij L x_j
// + sum_{j<=i} dt ah_{i+1,j} K_hat_j, and solves (M - dt a_ii L) x_i = rhs_i;
// x^{n+1} = x_s.
// A step in phases (one_step_beyond, PL_IMEX_timestepping_marching.f90:3-48):
// step_begin (the stage coefficients), per stage stage_nl (NL terms, the mean
// systems on the side stream) and per kxky system stage_rhs then
// stage_solve, and step_end.  one_step runs them for one run; the ensemble's
// lockstep step runs every run's stage_nl and stage_rhs, then their solves
// as one launch.
void step_begin(coral_amd_ctx &c) {
    if (c.src_missing)
        throw std::runtime_error("source" + std::to_string(c.src_missing / 10) + std::to_string(c.src_missing % 10) +
                                 " is used by a mean equation but has no definition (coral_amd_set_source)");
    const Problem &P = c.P;
    const int s = P.nstages;
    const double dt = c.dt;
    auto A = [&](int i, int j) { return P.a_arr[i * s + j]; };
    auto Ah = [&](int i, int j) { return P.a_hat[i * (s + 1) + j]; };
    // rhs_i = M x^n + sum_{j<i} dt a_ij L x_j + sum_{j<=i} dt ah_{i+1,j} K_hat_j.
    // Where a_jj != 0, dt L x_j = (M x_j - rhs_j) / a_jj, and rhs_j itself is
    // such a combination, so every rhs_i is a combination of M x^n, M x_j,
    // L x_j (only where a_jj = 0) and K_hat_j alone: no earlier rhs_j is read
    // back.  mc[0] multiplies M x^n, mc[1 + j] M x_j, xc[j] L x_j, kc[j] K_hat_j.
    auto &Rm = c.stepRm, &Rx = c.stepRx, &Rk = c.stepRk;
    Rm.assign(s, std::vector<double>(s + 1, 0.0));
    Rx.assign(s, std::vector<double>(s, 0.0));
    Rk.assign(s, std::vector<double>(s, 0.0));
    for (int i = 0; i < s; ++i) {
        std::vector<double> &mc = Rm[i], &xc = Rx[i], &kc = Rk[i];
        mc[0] = 1.0;
        for (int j = 0; j < i; ++j) {
            if (A(i, j) == 0.0) continue;
            if (A(j, j) != 0.0) {
                const double cij = A(i, j) / A(j, j);
                mc[1 + j] += cij;
                for (int q = 0; q <= s; ++q) mc[q] -= cij * Rm[j][q];
                for (int q = 0; q < s; ++q) xc[q] -= cij * Rx[j][q];
                for (int q = 0; q < s; ++q) kc[q] -= cij * Rk[j][q];
            } else {
                xc[j] += dt * A(i, j);
            }
        }
        for (int j = 0; j <= i; ++j) kc[j] += dt * Ah(i + 1, j);
    }
}

// stage i's right-hand side of system d into d.R
void stage_rhs(coral_amd_ctx &c, int i, SysDev &d) {
    std::vector<std::pair<const double2 *, double>> ms, xs, ks;
    if (c.stepRm[i][0] != 0.0) ms.push_back({d.field.p, c.stepRm[i][0]});
    for (int j = 0; j < i; ++j) {
        if (c.stepRm[i][1 + j] != 0.0) ms.push_back({d.X[j]->p, c.stepRm[i][1 + j]});
        if (c.stepRx[i][j] != 0.0) xs.push_back({d.X[j]->p, c.stepRx[i][j]});
    }
    for (int j = 0; j <= i; ++j)
        if (c.stepRk[i][j] != 0.0) ks.push_back({d.Khat[j]->p, c.stepRk[i][j]});
    rhs(c, d, ms, xs, ks, d.R.p);
}

// stage i's solution of system d into d.X[i] (with the refinement sweep when set)
void stage_solve(coral_amd_ctx &c, int i, SysDev &d) {
    solve_system(c, d, d.R.p, d.X[i]->p);
    const int s = c.P.nstages;
    if (c.refine == 2 || (c.refine == 1 && i == s - 1)) refine_solution(c, d, d.R.p, d.X[i]->p, c.P.a_arr[i * s + i]);
}

// stage i's NL terms, and the mean systems' whole stage on the side stream
void stage_nl(coral_amd_ctx &c, int i) {
    std::vector<const double2 *> kst, zst;
    for (auto &d : c.ks) kst.push_back(i == 0 ? d->field.p : d->X[i - 1]->p);
    for (auto &d : c.zs) zst.push_back(i == 0 ? d->field.p : d->X[i - 1]->p);
    nl_terms(c, kst, zst, i);
    if (!c.zs.empty()) {
        on_side(c, [&] {
            nl_terms_mean(c, i);
            for (auto &d : c.zs) {
                stage_rhs(c, i, *d);
                stage_solve(c, i, *d);
            }
            HIP_OK(hipEventRecord(c.ev_mean, c.stream));
        });
        c.mean_pending = true;
    }
}

void step_end(coral_amd_ctx &c) {
    join_mean(c);
    swap_step_buffers(c);
    if (c.project_kx0) project_kx0(c);
}

void one_step(coral_amd_ctx &c) {
    step_begin(c);
    for (int i = 0; i < c.P.nstages; ++i) {
        stage_nl(c, i);
        for (auto &d : c.ks) {
            stage_rhs(c, i, *d);
            stage_solve(c, i, *d);
        }
    }
    step_end(c);
}

// Two graphs of one step each, captured from the stream (the side stream
// joins through its events).  Capturing runs one_step's host logic, which
// swaps the step buffers; after both captures they are back where they were.
// The step holds no memset and no allocation, so the graphs replay it
// exactly; they are dropped when dt changes.
void capture_step_graphs(coral_amd_ctx &c) {
    c.drop_graphs();
    for (int p = 0; p < 2; ++p) {
        HIP_OK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
        hipGraph_t g = nullptr;
        try {
            one_step(c);
        } catch (...) {
            (void)hipStreamEndCapture(c.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            c.drop_graphs();
            throw;
        }
        HIP_OK(hipStreamEndCapture(c.stream, &g));
        const hipError_t e = hipGraphInstantiate(&c.gexec[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c.drop_graphs();
            throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    c.gpar = 0;
}

void ensure_phys(coral_amd_ctx &c) {
    const size_t n = size_t(c.P.NZAA) * c.P.NYAA * c.P.NXAA;
    if (c.physd.n != n) c.physd.alloc(n);
}

// device [z][y][x] -> host [x][y][z]
void download_phys(coral_amd_ctx &c, double *out) {
    const Problem &P = c.P;
    HIP_OK(hipStreamSynchronize(c.stream));
    std::vector<double> h(c.physd.n);
    HIP_OK(hipMemcpy(h.data(), c.physd.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int z = 0; z < P.NZAA; ++z)
        for (int y = 0; y < P.NYAA; ++y)
            for (int x = 0; x < P.NXAA; ++x)
                out[(size_t(x) * P.NYAA + y) * P.NZAA + z] = h[(size_t(z) * P.NYAA + y) * P.NXAA + x];
}

int fail(const std::exception &e) {
    g_err = e.what();
    return -1;
}

}  // namespace

// ---- small helper kernels used from the host side -------------------------
namespace {
__global__ void k_gather_strided(const double2 *src, size_t stride, int count, double2 *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[(size_t)i * stride];
}

__global__ void k_fill_zero(double2 *p, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}
void fill_zero(hipStream_t st, double2 *p, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_zero, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, p, n);
    check_launch("fill_zero");
}

}  // namespace

namespace coral {
// error message of the calling thread, for entry points in other sources
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace coral

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int coral_amd_last_error(char *buf, int32_t len) {
    if (!buf || len <= 0) return -1;
    std::strncpy(buf, g_err.c_str(), size_t(len) - 1);
    buf[len - 1] = 0;
    return 0;
}

int coral_amd_create(const coral_amd_problem *prob, int device, coral_amd_ctx **out) {
    if (!prob || !out) {
        g_err = "null argument";
        return -1;
    }
    try {
        std::unique_ptr<coral_amd_ctx> c(new coral_amd_ctx());
        c->P = assemble(*prob);
        Problem &P = c->P;
        c->device = device;
        HIP_OK(hipSetDevice(device));
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&c->ev_qhat, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&c->ev_mean, hipEventDisableTiming));
        c->km = KeptModes{P.nkx, P.nky, P.NY, P.kx_box, P.ky_box};
        for (auto &t : P.kxky) {
            c->ks.emplace_back(new SysDev());
            SysDev &d = *c->ks.back();
            d.t = &t;
            d.kxky = true;
            d.nsys = P.nsys;
            fill_system_device(d, P.nstages, c->km, c->stream, P.NZ);
        }
        for (auto &t : P.zero) {
            c->zs.emplace_back(new SysDev());
            SysDev &d = *c->zs.back();
            d.t = &t;
            d.kxky = false;
            d.nsys = 1;
            fill_system_device(d, P.nstages, c->km, c->stream, P.NZ);
            d.src.upload(std::vector<double>(std::max(t.total, 1), 0.0));
        }
        c->full_has_kxky.assign(P.nfull, 0);
        for (auto &e : P.extract) {
            c->ext.emplace_back(new ExtractDev());
            ExtractDev &x = *c->ext.back();
            x.full = e.full;
            x.kind = e.kind;
            x.system = e.system;
            x.kx = e.kx;
            x.ky = e.ky;
            x.rowptr.upload(e.rowptr);
            std::vector<int> col = e.col;
            std::vector<double> val = e.val;
            if (col.empty()) {
                col.push_back(0);
                val.push_back(0.0);
            }
            x.col.upload(col);
            x.val.upload(val);
            if (e.kind == 0) c->full_has_kxky[e.full] = 1;
        }
        const size_t nfull = std::max(P.nfull, 1), nquad = std::max(P.nquad, 1);
        c->spec.alloc(nfull * P.nsys * P.NZ);
        c->Fz.alloc(nfull * P.NZAA * P.nsys);
        c->G.alloc(nfull * size_t(P.NZAA) * P.NYAA * P.nkx);
        c->H.alloc(nquad * size_t(P.NZAA) * P.NYAA * P.nkx);
        c->Hy.alloc(nquad * size_t(P.NZAA) * P.nsys);
        c->Qhat.alloc(nquad * size_t(P.nsys) * P.NZ);
        HIP_OK(hipMemset(c->Qhat.p, 0, c->Qhat.n * sizeof(double2)));
        c->Tz.upload(P.Tz);
        c->Cz.upload(P.Cz);
        {
            ZFrags zf;
            make_zfrags(P.NZ, P.NZAA, P.Tz, P.Cz, device, zf);
            if (zf.bwd.empty()) zf.bwd.push_back(0.0);
            if (zf.fwd.empty()) zf.fwd.push_back(0.0);
            c->zfrag_b.upload(zf.bwd);
            c->zfrag_f.upload(zf.fwd);
            int ncu = 0;
            HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
            c->zdev = ZFragsDev{zf.TE, zf.TO, zf.NEK, zf.NOK, zf.RE, zf.RO, zf.KEK, zf.KOK,
                                zf.bwd_ok && !std::getenv("CORAL_AMD_NO_MFMA"), zf.fwd_ok && !std::getenv("CORAL_AMD_NO_MFMA"),
                                std::max(ncu, 1), c->zfrag_b.p, c->zfrag_f.p, zf.nbp, zf.nfp, {}, {}};
            for (int p = 0; p < ZMAXPARTS; ++p) {
                c->zdev.bp[p] = zf.bp[p];
                c->zdev.fp[p] = zf.fp[p];
            }
        }
        c->zweight.upload(P.zweight);
        std::vector<int> qa = P.quad_a, qb = P.quad_b;
        if (qa.empty()) {
            qa.push_back(-1);
            qb.push_back(-1);
        }
        c->qa.upload(qa);
        c->qb.upload(qb);
        c->cflmax.alloc(3);
        c->sink.alloc(64);
        HIP_OK(hipMemset(c->cflmax.p, 0, 3 * sizeof(unsigned long long)));
        c->px = make_plan(P.NXAA, c->twx);
        c->py = make_plan(P.NYAA, c->twy);
        HIP_OK(hipDeviceSynchronize());
        *out = c.release();
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_destroy(coral_amd_ctx *ctx) {
    try {
        if (ctx) {
            HIP_OK(hipSetDevice(ctx->device));
            if (ctx->stream) HIP_OK(hipStreamSynchronize(ctx->stream));
            delete ctx;
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_dt(coral_amd_ctx *c, double dt) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw std::runtime_error("dt must be positive and finite");
        HIP_OK(hipSetDevice(c->device));
        c->drop_graphs();
        ++c->gen;   // the factors may be reallocated and dt enters the launches
        for (auto &d : c->ks) factorize(*c, *d, dt);
        for (auto &d : c->zs) factorize(*c, *d, dt);
        c->dt = dt;
        c->factored = true;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_upload_state(coral_amd_ctx *c, int32_t kind, int32_t isys, const double *data) {
    try {
        if (!c || !data) throw std::runtime_error("null argument");
        HIP_OK(hipSetDevice(c->device));
        // the copies below go through the null stream, which does not wait
        // for the context's non-blocking streams: finish queued steps first
        HIP_OK(hipStreamSynchronize(c->stream));
        HIP_OK(hipStreamSynchronize(c->side));
        const Problem &P = c->P;
        if (kind == 0) {
            if (isys < 0 || isys >= int(c->ks.size())) throw std::runtime_error("bad kxky system index");
            SysDev &d = *c->ks[isys];
            const int total = d.t->total;
            std::vector<double2> h(size_t(P.nsys) * total);
            for (int ix = 0; ix < P.nkx; ++ix)
                for (int jy = 0; jy < P.nky; ++jy) {
                    const int s = ix * P.nky + jy;
                    const size_t src = (size_t(ix) * P.NYAA + P.ky_slot[jy]) * total;
                    for (int i = 0; i < total; ++i)
                        h[size_t(s) * total + i] =
                            (s == 0) ? make_double2(0.0, 0.0) : make_double2(data[2 * (src + i)], data[2 * (src + i) + 1]);
                }
            HIP_OK(hipMemcpy(d.field.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
        } else if (kind == 1) {
            if (isys < 0 || isys >= int(c->zs.size())) throw std::runtime_error("bad mean system index");
            SysDev &d = *c->zs[isys];
            std::vector<double2> h(d.t->total);
            for (int i = 0; i < d.t->total; ++i) h[i] = make_double2(data[i], 0.0);
            HIP_OK(hipMemcpy(d.field.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
        } else {
            throw std::runtime_error("kind must be 0 (kxky) or 1 (mean)");
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_download_state(coral_amd_ctx *c, int32_t kind, int32_t isys, double *data) {
    try {
        if (!c || !data) throw std::runtime_error("null argument");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        const Problem &P = c->P;
        if (kind == 0) {
            if (isys < 0 || isys >= int(c->ks.size())) throw std::runtime_error("bad kxky system index");
            SysDev &d = *c->ks[isys];
            const int total = d.t->total;
            std::vector<double2> h(size_t(P.nsys) * total);
            HIP_OK(hipMemcpy(h.data(), d.field.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
            std::memset(data, 0, sizeof(double) * 2 * size_t(P.nkx) * P.NYAA * total);
            for (int ix = 0; ix < P.nkx; ++ix)
                for (int jy = 0; jy < P.nky; ++jy) {
                    const int s = ix * P.nky + jy;
                    const size_t dst = (size_t(ix) * P.NYAA + P.ky_slot[jy]) * total;
                    for (int i = 0; i < total; ++i) {
                        data[2 * (dst + i)] = h[size_t(s) * total + i].x;
                        data[2 * (dst + i) + 1] = h[size_t(s) * total + i].y;
                    }
                }
        } else if (kind == 1) {
            if (isys < 0 || isys >= int(c->zs.size())) throw std::runtime_error("bad mean system index");
            SysDev &d = *c->zs[isys];
            std::vector<double2> h(d.t->total);
            HIP_OK(hipMemcpy(h.data(), d.field.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
            for (int i = 0; i < d.t->total; ++i) data[i] = h[i].x;
        } else {
            throw std::runtime_error("kind must be 0 (kxky) or 1 (mean)");
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_step(coral_amd_ctx *c, int32_t nsteps) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (!c->factored) throw std::runtime_error("coral_amd_set_dt must be called before stepping");
        if (nsteps < 0) throw std::runtime_error("nsteps must be >= 0");
        HIP_OK(hipSetDevice(c->device));
        if (c->use_graphs && !c->timing && nsteps > 0) {
            if (!c->gexec[0]) capture_step_graphs(*c);
            for (int n = 0; n < nsteps; ++n) {
                HIP_OK(hipGraphLaunch(c->gexec[c->gpar], c->stream));
                swap_step_buffers(*c);
            }
        } else {
            for (int n = 0; n < nsteps; ++n) one_step(*c);
        }
        HIP_OK(hipGetLastError());
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_synchronize(coral_amd_ctx *c) {
    try {
        if (!c) throw std::runtime_error("null context");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_get_cfl_dt(coral_amd_ctx *c, double *out) {
    try {
        if (!c || !out) throw std::runtime_error("null argument");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        unsigned long long h[3];
        HIP_OK(hipMemcpy(h, c->cflmax.p, sizeof(h), hipMemcpyDeviceToHost));
        double m[3];
        std::memcpy(m, h, sizeof(m));
        const Problem &P = c->P;
        double u = m[0] * P.NXAA / P.Lx, v = m[1] * P.NYAA / P.Ly;
        double w = std::max(v, m[2]);
        double mx = std::max(u, std::max(v, w));
        *out = mx > 0 ? 1.0 / mx : INFINITY;
        if (c->nu_full >= 0) {
            // the Smagorinsky bound (PL_IMEX_timestepping_marching.f90:241-253)
            std::vector<double> prof(P.NZAA);
            HIP_OK(hipMemcpy(prof.data(), c->nu_prof.p, prof.size() * sizeof(double), hipMemcpyDeviceToHost));
            const double numax = *std::max_element(prof.begin(), prof.end());
            const double kx = P.NXAA / P.Lx, ky = P.NYAA / P.Ly;
            *out = std::min(*out, 0.95 / ((2 * 3.141592) * (2 * 3.141592) * numax * (kx * kx + ky * ky)));
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_wavenumbers(coral_amd_ctx *c, int32_t *kx_int, int32_t *ky_int, uint8_t *dealias_x,
                          uint8_t *dealias_y) {
    try {
        if (!c) throw std::runtime_error("null context");
        const Problem &P = c->P;
        for (int ix = 0; ix < P.NX / 2; ++ix) {
            int k = ix + 1;
            int ki = (k >= P.NXAA / 2 + 1) ? -(P.NXAA + 1 - k) : k - 1;
            if (kx_int) kx_int[ix] = ki;
            if (dealias_x) dealias_x[ix] = (std::fabs(ki * P.kx_box) >= P.kx_box * P.NX / 2.0) ? 1 : 0;
        }
        for (int iy = 0; iy < P.NYAA; ++iy) {
            int k = iy + 1;
            int ki = (k >= P.NYAA / 2 + 1) ? -(P.NYAA + 1 - k) : k - 1;
            if (ky_int) ky_int[iy] = ki;
            if (dealias_y) dealias_y[iy] = (std::fabs(ki * P.ky_box) >= P.ky_box * P.NY / 2.0) ? 1 : 0;
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_system_info(coral_amd_ctx *c, int32_t isys, int32_t *total, int32_t *nl, int32_t *nu,
                          int32_t *nkept) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (isys < 0 || isys >= int(c->P.kxky.size())) throw std::runtime_error("bad kxky system index");
        const SystemTables &t = c->P.kxky[isys];
        if (total) *total = t.total;
        if (nl) *nl = t.nl;
        if (nu) *nu = t.nu;
        if (nkept) *nkept = c->P.nsys;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_factor_kind(coral_amd_ctx *c, int32_t isys, int32_t *is_real) {
    try {
        if (!c || !is_real) throw std::runtime_error("null argument");
        if (isys < 0 || isys >= int(c->ks.size())) throw std::runtime_error("bad kxky system index");
        *is_real = c->ks[isys]->real ? 1 : 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_factor_blocks(coral_amd_ctx *c, int32_t kind, int32_t isys, int32_t *nblk, int32_t *blk_n,
                            int32_t *blk_nl, int32_t *blk_nu, double *u_entries) {
    try {
        if (!c || !nblk || !blk_n || !blk_nl || !blk_nu) throw std::runtime_error("null argument");
        const auto &v = kind == 0 ? c->ks : c->zs;
        if (kind < 0 || kind > 1 || isys < 0 || isys >= int(v.size())) throw std::runtime_error("bad system index");
        const SysDev &d = *v[isys];
        *nblk = d.nblk;
        *blk_n = d.blk_n;
        *blk_nl = d.blk_nl;
        *blk_nu = d.blk_nu;
        if (u_entries) *u_entries = d.ncol16 > 0 ? d.u_mean : double(d.blk_nl + d.blk_nu + 1);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_get_physical(coral_amd_ctx *c, int32_t ifull, double *out) {
    try {
        if (!c || !out) throw std::runtime_error("null argument");
        const Problem &P = c->P;
        if (ifull < 0 || ifull >= P.nfull) throw std::runtime_error("bad full-variable index");
        HIP_OK(hipSetDevice(c->device));
        std::vector<const double2 *> kst, zst;
        for (auto &d : c->ks) kst.push_back(d->field.p);
        for (auto &d : c->zs) zst.push_back(d->field.p);
        const long long nlines = (long long)P.NZAA * P.NYAA, plane = nlines * P.NXAA;
        ensure_phys(*c);
        const int kind = side_kind(*c, ifull);
        if (kind == 1) {
            HIP_OK(hipMemcpyAsync(c->physd.p, c->fixed_all.p + (size_t)c->fixed_slot[ifull] * plane,
                                  plane * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        } else if (kind == 2) {
            hipLaunchKernelGGL(k_broadcast_profile, dim3(unsigned((plane + 255) / 256)), dim3(256), 0, c->stream,
                               c->nu_prof.p, plane, P.NYAA * P.NXAA, c->physd.p);
            check_launch("broadcast profile");
        } else {
            backward_to_G(*c, kst, zst);
            launch_xbwd(c->stream, c->px, c->G.p + (size_t)ifull * nlines * P.nkx, c->physd.p, 1, nlines, P.nkx);
        }
        download_phys(*c, out);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

namespace {
// Maps of one coupled variable: Chebyshev coefficients c = S g of its Galerkin
// coefficients g (stencil S, NZ x ncol, chebyshev_galerkin_2.f90) at the
// interleaved positions pos, and the inverse of the square truncation
// S[0:ncol, 0:ncol] used to go back (prepare_stencils, PL_IMEX_timestepping.f90:663-723).
struct VarMaps {
    int ncol = 0;
    std::vector<int> pos;
    Dense S;
    std::vector<double> Sinv;  // ncol x ncol, row-major
};

VarMaps var_maps(const SystemTables &t, int NZ, int v) {
    if (v < 0 || v >= t.nvars) throw std::runtime_error("bad variable index");
    VarMaps m;
    m.ncol = t.var_size[v];
    for (int j = 0; j < m.ncol; ++j) m.pos.push_back(t.posP[t.var_first[v] + j]);
    m.S = galerkin_stencil(NZ, t.bc_codes[v]);
    // inverse of the square truncation by Gauss-Jordan elimination with partial pivoting
    const int n = m.ncol;
    std::vector<double> A(size_t(n) * n), B(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A[size_t(i) * n + j] = m.S(i, j);
    for (int i = 0; i < n; ++i) B[size_t(i) * n + i] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A[size_t(r) * n + col]) > std::fabs(A[size_t(piv) * n + col])) piv = r;
        if (A[size_t(piv) * n + col] == 0.0) throw std::runtime_error("singular square stencil");
        if (piv != col)
            for (int j = 0; j < n; ++j) {
                std::swap(A[size_t(piv) * n + j], A[size_t(col) * n + j]);
                std::swap(B[size_t(piv) * n + j], B[size_t(col) * n + j]);
            }
        const double d = A[size_t(col) * n + col];
        for (int j = 0; j < n; ++j) {
            A[size_t(col) * n + j] /= d;
            B[size_t(col) * n + j] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = A[size_t(r) * n + col];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                A[size_t(r) * n + j] -= f * A[size_t(col) * n + j];
                B[size_t(r) * n + j] -= f * B[size_t(col) * n + j];
            }
        }
    }
    m.Sinv = B;
    return m;
}
}  // namespace

int coral_amd_get_variable_physical(coral_amd_ctx *c, int32_t kind, int32_t isys, int32_t ivar, double *out) {
    try {
        if (!c || !out) throw std::runtime_error("null argument");
        const Problem &P = c->P;
        HIP_OK(hipSetDevice(c->device));
        if (kind == 0) {
            if (isys < 0 || isys >= int(c->ks.size())) throw std::runtime_error("bad kxky system index");
            SysDev &d = *c->ks[isys];
            const VarMaps m = var_maps(*d.t, P.NZ, ivar);
            std::vector<int> rowptr(1, 0), col;
            std::vector<double> val;
            for (int n = 0; n < P.NZ; ++n) {
                for (int j = 0; j < m.ncol; ++j)
                    if (m.S(n, j) != 0.0) {
                        col.push_back(m.pos[j]);
                        val.push_back(m.S(n, j));
                    }
                rowptr.push_back(int(col.size()));
            }
            DevBuf<int> drow, dcol;
            DevBuf<double> dval;
            drow.upload(rowptr);
            dcol.upload(col);
            dval.upload(val);
            launch_extract(c->stream, d.field.p, d.t->total, P.nsys, P.NZ, drow.p, dcol.p, dval.p, 0, 0, c->km,
                           c->spec.p, 0);
            zbwd(*c, c->spec.p, c->Fz.p, 1);
            launch_ybwd(c->stream, c->py, c->Fz.p, c->G.p, 1, P.NZAA, P.nkx, P.nky, P.NY);
            ensure_phys(*c);
            launch_xbwd(c->stream, c->px, c->G.p, c->physd.p, 1, (long long)P.NZAA * P.NYAA, P.nkx);
            download_phys(*c, out);
        } else if (kind == 1) {
            if (isys < 0 || isys >= int(c->zs.size())) throw std::runtime_error("bad mean system index");
            SysDev &d = *c->zs[isys];
            const VarMaps m = var_maps(*d.t, P.NZ, ivar);
            HIP_OK(hipStreamSynchronize(c->stream));
            std::vector<double2> f(d.t->total);
            HIP_OK(hipMemcpy(f.data(), d.field.p, f.size() * sizeof(double2), hipMemcpyDeviceToHost));
            std::vector<double> cz(P.NZ, 0.0);
            for (int n = 0; n < P.NZ; ++n)
                for (int j = 0; j < m.ncol; ++j) cz[n] += m.S(n, j) * f[m.pos[j]].x;
            for (int z = 0; z < P.NZAA; ++z) {
                double v = 0.0;
                for (int n = 0; n < P.NZ; ++n) v += P.Tz[size_t(z) * P.NZ + n] * cz[n];
                out[z] = v;
            }
        } else {
            throw std::runtime_error("kind must be 0 (kxky) or 1 (mean)");
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_variable_physical(coral_amd_ctx *c, int32_t kind, int32_t isys, int32_t ivar, const double *in) {
    try {
        if (!c || !in) throw std::runtime_error("null argument");
        const Problem &P = c->P;
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        if (kind == 0) {
            if (isys < 0 || isys >= int(c->ks.size())) throw std::runtime_error("bad kxky system index");
            SysDev &d = *c->ks[isys];
            const VarMaps m = var_maps(*d.t, P.NZ, ivar);
            ensure_phys(*c);
            std::vector<double> h(size_t(P.NXAA) * P.NYAA * P.NZAA);
            for (int x = 0; x < P.NXAA; ++x)
                for (int y = 0; y < P.NYAA; ++y)
                    for (int z = 0; z < P.NZAA; ++z)
                        h[(size_t(z) * P.NYAA + y) * P.NXAA + x] = in[(size_t(x) * P.NYAA + y) * P.NZAA + z];
            HIP_OK(hipMemcpy(c->physd.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
            launch_xfwd(c->stream, c->px, c->physd.p, c->H.p, 1, (long long)P.NZAA * P.NYAA, P.nkx);
            launch_yfwd(c->stream, c->py, c->H.p, c->Hy.p, 1, P.NZAA, P.nkx, P.nky, P.NY);
            zfwd(*c, c->Hy.p, c->Qhat.p, 1);
            DevBuf<double> dsinv;
            DevBuf<int> dpos;
            dsinv.upload(m.Sinv);
            dpos.upload(m.pos);
            launch_insert_var(c->stream, c->Qhat.p, P.nsys, m.ncol, dsinv.p, dpos.p, d.field.p, d.t->total);
            HIP_OK(hipStreamSynchronize(c->stream));
        } else if (kind == 1) {
            if (isys < 0 || isys >= int(c->zs.size())) throw std::runtime_error("bad mean system index");
            SysDev &d = *c->zs[isys];
            const VarMaps m = var_maps(*d.t, P.NZ, ivar);
            // Chebyshev coefficients of the profile: the z part of the forward transform
            const double scale = double(P.NXAA) * double(P.NYAA);
            std::vector<double> cz(P.NZ, 0.0);
            for (int n = 0; n < P.NZ; ++n)
                for (int z = 0; z < P.NZAA; ++z) cz[n] += scale * P.Cz[size_t(n) * P.NZAA + z] * in[z];
            std::vector<double2> f(d.t->total);
            HIP_OK(hipMemcpy(f.data(), d.field.p, f.size() * sizeof(double2), hipMemcpyDeviceToHost));
            for (int j = 0; j < m.ncol; ++j) {
                double g = 0.0;
                for (int n = 0; n < m.ncol; ++n) g += m.Sinv[size_t(j) * m.ncol + n] * cz[n];
                f[m.pos[j]] = make_double2(g, 0.0);
            }
            HIP_OK(hipMemcpy(d.field.p, f.data(), f.size() * sizeof(double2), hipMemcpyHostToDevice));
        } else {
            throw std::runtime_error("kind must be 0 (kxky) or 1 (mean)");
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_horizontal_averages(coral_amd_ctx *c, double *full_avg, double *quad_avg) {
    try {
        if (!c) throw std::runtime_error("null context");
        const Problem &P = c->P;
        HIP_OK(hipSetDevice(c->device));
        std::vector<const double2 *> kst, zst;
        for (auto &d : c->ks) kst.push_back(d->field.p);
        for (auto &d : c->zs) zst.push_back(d->field.p);
        DevBuf<double2> col;
        const int nf = std::max(P.nfull, 1), nq = std::max(P.nquad, 1);
        col.alloc(size_t(nf) * P.NZ + size_t(nq) * P.NZAA);
        if (quad_avg) {
            // products of the current state up to the y transform; column s = 0 is the (0,0) mode
            backward_to_G(*c, kst, zst);
            zero_cflmax(c->stream, c->cflmax.p);
            XPassArgs xa{};
            xa.G = c->G.p;
            xa.H = c->H.p;
            xa.nfull = P.nfull;
            xa.nquad = P.nquad;
            xa.NZAA = P.NZAA;
            xa.NYAA = P.NYAA;
            xa.nkx = P.nkx;
            xa.qa = c->qa.p;
            xa.qb = c->qb.p;
            xa.zweight = c->zweight.p;
            xa.cflmax = c->cflmax.p;
            launch_xpass(c->stream, c->px, xa);
            launch_yfwd(c->stream, c->py, c->H.p, c->Hy.p, P.nquad, P.NZAA, P.nkx, P.nky, P.NY);
            const int n = P.nquad * P.NZAA;
            hipLaunchKernelGGL(k_gather_strided, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->Hy.p,
                               size_t(P.nsys), n, col.p + size_t(nf) * P.NZ);
            check_launch("horizontal_averages");
        } else {
            backward_to_G(*c, kst, zst);
        }
        // full variables: (0,0) Chebyshev coefficients spec[f][n][0]
        const int m = P.nfull * P.NZ;
        if (m > 0)
            hipLaunchKernelGGL(k_gather_strided, dim3((m + 255) / 256), dim3(256), 0, c->stream, c->spec.p,
                               size_t(P.nsys), m, col.p);
        if (m > 0) check_launch("horizontal_averages");
        HIP_OK(hipStreamSynchronize(c->stream));
        std::vector<double2> h(col.n);
        HIP_OK(hipMemcpy(h.data(), col.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
        if (full_avg)
            for (int f = 0; f < P.nfull; ++f)
                for (int z = 0; z < P.NZAA; ++z) {
                    double v = 0.0;
                    for (int n2 = 0; n2 < P.NZ; ++n2) v += P.Tz[size_t(z) * P.NZ + n2] * h[size_t(f) * P.NZ + n2].x;
                    full_avg[size_t(f) * P.NZAA + z] = v;
                }
        if (quad_avg) {
            const double inv = 1.0 / (double(P.NXAA) * double(P.NYAA));
            for (int q = 0; q < P.nquad; ++q)
                for (int z = 0; z < P.NZAA; ++z)
                    quad_avg[size_t(q) * P.NZAA + z] = h[size_t(nf) * P.NZ + size_t(q) * P.NZAA + z].x * inv;
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_products(coral_amd_ctx *c, double *out) {
    try {
        if (!c || !out) throw std::runtime_error("null argument");
        const Problem &P = c->P;
        HIP_OK(hipSetDevice(c->device));
        if (P.nquad == 0) return 0;
        std::vector<const double2 *> kst, zst;
        for (auto &d : c->ks) kst.push_back(d->field.p);
        for (auto &d : c->zs) zst.push_back(d->field.p);
        const bool fp32 = c->precision == 1;
        c->side_advance = false;
        try {
            quad_spectra(*c, kst, zst);
        } catch (...) {
            c->side_advance = true;
            throw;
        }
        c->side_advance = true;
        join_mean(*c);
        HIP_OK(hipStreamSynchronize(c->stream));
        std::vector<double2> h(size_t(P.nquad) * P.NZ * P.nsys);
        if (fp32) {
            std::vector<float2> hf(h.size());
            HIP_OK(hipMemcpy(hf.data(), c->Qhat32.p, hf.size() * sizeof(float2), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < h.size(); ++i) h[i] = make_double2(hf[i].x, hf[i].y);
        } else {
            HIP_OK(hipMemcpy(h.data(), c->Qhat.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
        }
        // [q][ix][y slot][n], dealiased ky slots zero
        const size_t per_q = size_t(P.nkx) * P.NYAA * P.NZ;
        std::fill(out, out + 2 * per_q * P.nquad, 0.0);
        for (int q = 0; q < P.nquad; ++q)
            for (int ix = 0; ix < P.nkx; ++ix)
                for (int jy = 0; jy < P.nky; ++jy) {
                    const size_t sy = size_t(ix) * P.nky + jy;
                    const size_t dst = size_t(q) * per_q + (size_t(ix) * P.NYAA + P.ky_slot[jy]) * P.NZ;
                    for (int n = 0; n < P.NZ; ++n) {
                        const double2 v = h[(size_t(q) * P.NZ + n) * P.nsys + sy];
                        out[2 * (dst + n)] = v.x;
                        out[2 * (dst + n) + 1] = v.y;
                    }
                }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_bwd_transform(coral_amd_ctx *c, const double *spec, double *phys) {
    try {
        if (!c || !spec || !phys) throw std::runtime_error("null argument");
        const Problem &P = c->P;
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        std::vector<double2> h(size_t(P.nsys) * P.NZ);
        for (int ix = 0; ix < P.nkx; ++ix)
            for (int jy = 0; jy < P.nky; ++jy) {
                const size_t src = (size_t(ix) * P.NYAA + P.ky_slot[jy]) * P.NZ;
                const size_t sy = size_t(ix) * P.nky + jy;
                for (int n = 0; n < P.NZ; ++n)
                    h[size_t(n) * P.nsys + sy] = make_double2(spec[2 * (src + n)], spec[2 * (src + n) + 1]);
            }
        HIP_OK(hipMemcpy(c->spec.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
        zbwd(*c, c->spec.p, c->Fz.p, 1);
        launch_ybwd(c->stream, c->py, c->Fz.p, c->G.p, 1, P.NZAA, P.nkx, P.nky, P.NY);
        const long long nlines = (long long)P.NZAA * P.NYAA;
        ensure_phys(*c);
        launch_xbwd(c->stream, c->px, c->G.p, c->physd.p, 1, nlines, P.nkx);
        download_phys(*c, phys);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_fwd_transform(coral_amd_ctx *c, const double *phys, double *spec) {
    try {
        if (!c || !spec || !phys) throw std::runtime_error("null argument");
        const Problem &P = c->P;
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        ensure_phys(*c);
        // host [x][y][z] -> device [z][y][x]
        std::vector<double> h(size_t(P.NXAA) * P.NYAA * P.NZAA);
        for (int x = 0; x < P.NXAA; ++x)
            for (int y = 0; y < P.NYAA; ++y)
                for (int z = 0; z < P.NZAA; ++z)
                    h[(size_t(z) * P.NYAA + y) * P.NXAA + x] = phys[(size_t(x) * P.NYAA + y) * P.NZAA + z];
        HIP_OK(hipMemcpy(c->physd.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        const long long nlines = (long long)P.NZAA * P.NYAA;
        launch_xfwd(c->stream, c->px, c->physd.p, c->H.p, 1, nlines, P.nkx);
        launch_yfwd(c->stream, c->py, c->H.p, c->Hy.p, 1, P.NZAA, P.nkx, P.nky, P.NY);
        zfwd(*c, c->Hy.p, c->Qhat.p, 1);
        HIP_OK(hipStreamSynchronize(c->stream));
        std::vector<double2> q(size_t(P.nsys) * P.NZ);
        HIP_OK(hipMemcpy(q.data(), c->Qhat.p, q.size() * sizeof(double2), hipMemcpyDeviceToHost));
        std::memset(spec, 0, sizeof(double) * 2 * size_t(P.nkx) * P.NYAA * P.NZ);
        for (int ix = 0; ix < P.nkx; ++ix)
            for (int jy = 0; jy < P.nky; ++jy) {
                const size_t dst = (size_t(ix) * P.NYAA + P.ky_slot[jy]) * P.NZ;
                const size_t sy = size_t(ix) * P.nky + jy;
                for (int n = 0; n < P.NZ; ++n) {
                    spec[2 * (dst + n)] = q[size_t(n) * P.nsys + sy].x;
                    spec[2 * (dst + n) + 1] = q[size_t(n) * P.nsys + sy].y;
                }
            }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

namespace {
// The stream of the banded-solve seam, destroyed on every path out.
struct OwnedStream {
    hipStream_t s = nullptr;
    ~OwnedStream() {
        if (s) (void)hipStreamDestroy(s);
    }
};
}  // namespace

int coral_amd_banded_solve(int32_t device, int32_t n, int32_t kl, int32_t ku, int32_t nsys, const double *ab,
                           double *x, int32_t *info) {
    try {
        if (!ab || !x || !info) throw std::runtime_error("null argument");
        if (n < 1 || kl < 0 || ku < 0 || nsys < 1 || kl > 255) throw std::runtime_error("bad band dimensions");
        HIP_OK(hipSetDevice(device));
        OwnedStream os;
        HIP_OK(hipStreamCreateWithFlags(&os.s, hipStreamNonBlocking));
        hipStream_t st = os.s;
        const int ldab = 2 * kl + ku + 1, kv = kl + ku;
        DevBuf<double2> work, xb, rb;
        DevBuf<int> inf;
        work.alloc(size_t(nsys) * n * ldab);
        xb.alloc(size_t(nsys) * n);
        inf.alloc(nsys);
        HIP_OK(hipMemcpy(work.p, ab, work.n * sizeof(double2), hipMemcpyHostToDevice));
        const int ncol16 = solve16_ncol(n, kl, kv);
        if (ncol16 > 0) {
            // the records, U trimming and 16-lane solve kernel of the time step
            // (one block per system, complex, no rotation)
            const size_t nmat4 = (size_t(nsys) + 3) / 4 * 4, nb = 16 * ((size_t(n) + 15) / 16);
            DevBuf<double2> L16, Uf, Ut, sink;
            DevBuf<unsigned char> piv;
            DevBuf<int> ulen;
            DevBuf<int2> ucol;
            DevBuf<unsigned> words;
            L16.alloc(nmat4 * ncol16 * solve16_klp(kl));
            Uf.alloc(nmat4 * nb * (kv + 1) + 2 * size_t(SOLVE16_PAD));
            piv.alloc(size_t(nsys) * nb);
            HIP_OK(hipMemsetAsync(L16.p, 0, L16.n * sizeof(double2), st));
            HIP_OK(hipMemsetAsync(Uf.p, 0, Uf.n * sizeof(double2), st));
            HIP_OK(hipMemsetAsync(piv.p, 0, piv.n, st));
            launch_factor_lapack16(st, work.p, n, kl, ku, nsys, ncol16, L16.p, Uf.p + SOLVE16_PAD, piv.p, inf.p);
            ulen.alloc(nb);
            std::vector<int2> uc;
            long long ustride = 0;
            solve16_ucol(st, Uf.p + SOLVE16_PAD, 16, n, kv, nsys, ulen.p, uc, &ustride);
            Ut.alloc(size_t((nsys + 3) / 4) * size_t(ustride));
            HIP_OK(hipMemsetAsync(Ut.p, 0, Ut.n * sizeof(double2), st));
            ucol.upload(uc);
            solve16_trim_u_copy(st, Uf.p + SOLVE16_PAD, 16, n, kv, nsys, ucol.p, ustride, Ut.p);
            words.upload(solve16_words(n, 1, n, nullptr, nullptr, nullptr));
            rb.alloc(size_t(nsys) * n);
            sink.alloc(64);
            HIP_OK(hipMemcpy(rb.p, x, rb.n * sizeof(double2), hipMemcpyHostToDevice));
            Solve16Args sa{};
            sa.L16 = L16.p;
            sa.U = Ut.p;
            sa.ucol = ucol.p;
            sa.ustride = ustride;
            sa.ncol16 = ncol16;
            sa.piv = piv.p;
            sa.words = words.p;
            sa.rotate = 0;
            sa.ident0 = 0;
            sa.total = n;
            sa.n = n;
            sa.kl = kl;
            sa.kv = kv;
            sa.nsys = nsys;
            sa.nblk = 1;
            sa.rhs = rb.p;
            sa.x = xb.p;
            sa.sink = sink.p;
            if (!launch_solve16_complex(st, sa)) throw std::runtime_error("16-lane solve rejected the band");
            HIP_OK(hipStreamSynchronize(st));
        } else {
            DevBuf<double2> Lf, Uf;
            DevBuf<unsigned char> piv;
            Lf.alloc(size_t(nsys) * n * std::max(kl, 1));
            Uf.alloc(size_t(nsys) * n * (kv + 1));
            piv.alloc(size_t(nsys) * n);
            HIP_OK(hipMemcpy(xb.p, x, xb.n * sizeof(double2), hipMemcpyHostToDevice));
            launch_factor_lapack(st, work.p, n, kl, ku, nsys, Lf.p, Uf.p, piv.p, inf.p);
            launch_solve(st, Lf.p, Uf.p, piv.p, n, kl, kv, nsys, xb.p, xb.p);
            HIP_OK(hipStreamSynchronize(st));
        }
        HIP_OK(hipMemcpy(x, xb.p, xb.n * sizeof(double2), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(info, inf.p, nsys * sizeof(int), hipMemcpyDeviceToHost));
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_system_solve(coral_amd_ctx *c, int32_t isys, const double *rhs, double *x) {
    try {
        if (!c || !rhs || !x) throw std::runtime_error("null argument");
        if (!c->factored) throw std::runtime_error("coral_amd_set_dt must be called before solving");
        if (isys < 0 || isys >= int(c->ks.size())) throw std::runtime_error("bad kxky system index");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        HIP_OK(hipStreamSynchronize(c->side));
        const Problem &P = c->P;
        SysDev &d = *c->ks[isys];
        const int total = d.t->total;
        std::vector<double2> h(size_t(P.nsys) * total);
        for (int ix = 0; ix < P.nkx; ++ix)
            for (int jy = 0; jy < P.nky; ++jy) {
                const size_t s = size_t(ix) * P.nky + jy, src = (size_t(ix) * P.NYAA + P.ky_slot[jy]) * total;
                for (int i = 0; i < total; ++i)
                    h[s * total + i] = make_double2(rhs[2 * (src + i)], rhs[2 * (src + i) + 1]);
            }
        HIP_OK(hipMemcpy(d.R.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
        solve_system(*c, d, d.R.p, d.X[0]->p);
        HIP_OK(hipStreamSynchronize(c->stream));
        HIP_OK(hipMemcpy(h.data(), d.X[0]->p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
        std::memset(x, 0, sizeof(double) * 2 * size_t(P.nkx) * P.NYAA * total);
        for (int ix = 0; ix < P.nkx; ++ix)
            for (int jy = 0; jy < P.nky; ++jy) {
                const size_t s = size_t(ix) * P.nky + jy, dst = (size_t(ix) * P.NYAA + P.ky_slot[jy]) * total;
                for (int i = 0; i < total; ++i) {
                    x[2 * (dst + i)] = h[s * total + i].x;
                    x[2 * (dst + i) + 1] = h[s * total + i].y;
                }
            }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_assembled_operator(const coral_amd_problem *prob, int32_t isys, int32_t which, int32_t kx_int,
                                 int32_t ky_int, double *out, int32_t *total, int32_t *nl, int32_t *nu) {
    try {
        if (!prob) throw std::runtime_error("null problem");
        Problem P = assemble(*prob);
        if (isys < 0 || isys >= int(P.kxky.size())) throw std::runtime_error("bad kxky system index");
        const SystemTables &t = P.kxky[isys];
        if (total) *total = t.total;
        if (nl) *nl = t.nl;
        if (nu) *nu = t.nu;
        if (!out) return 0;
        const OpTables &op = which == 0 ? t.mass : t.stif;
        const int n = t.total, w = t.nl + t.nu + 1;
        std::memset(out, 0, sizeof(double) * 2 * size_t(n) * n);
        const double kx = kx_int * P.kx_box, ky = ky_int * P.ky_box;
        for (int g = 0; g < op.ngroups; ++g) {
            double2 cf = make_double2(0.0, 0.0);
            for (int m = op.mono_off[g]; m < op.mono_off[g + 1]; ++m)
                cf = cadd(cf, monomial(op.mono_a[m], op.mono_b[m], kx, ky));
            for (int i = 0; i < n; ++i)
                for (int d = -t.nl; d <= t.nu; ++d) {
                    int j = i + d;
                    if (j < 0 || j >= n) continue;
                    double b = op.band[(size_t(g) * n + i) * w + d + t.nl];
                    out[2 * (size_t(i) * n + j)] += b * cf.x;
                    out[2 * (size_t(i) * n + j) + 1] += b * cf.y;
                }
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_precision(coral_amd_ctx *c, int32_t mode, int32_t refine) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (mode < 0 || mode > 1) throw std::runtime_error("precision mode must be 0 (FP64) or 1 (FP32 transforms)");
        if (refine < 0 || refine > 2) throw std::runtime_error("refine must be 0, 1 (last stage) or 2 (every stage)");
        if (mode == 1 && c->closure.on) throw std::runtime_error("the closure runs with the FP64 transform chain only");
        if (mode == 1 && (c->nfixed > 0 || c->nu_full >= 0))
            throw std::runtime_error("fixed physical fields and the Smagorinsky viscosity run with the FP64 transform chain only");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        HIP_OK(hipStreamSynchronize(c->side));
        Problem &P = c->P;
        if (mode == 1 && !c->spec32.p) {
            ZFrags32 zf;
            if (!make_zfrags32(P.NZ, P.NZAA, P.Tz, P.Cz, lds_limit_current(), zf))
                throw std::runtime_error("no FP32 MFMA plan for the z transforms at NZ = " + std::to_string(P.NZ));
            auto f32 = [](const FFTPlan &pl, DevBuf<float2> &buf) {
                std::vector<double2> h(pl.N);
                HIP_OK(hipMemcpy(h.data(), pl.tw, pl.N * sizeof(double2), hipMemcpyDeviceToHost));
                std::vector<float2> f(pl.N);
                for (int k = 0; k < pl.N; ++k) f[k] = make_float2(float(h[k].x), float(h[k].y));
                buf.upload(f);
            };
            f32(c->px, c->twx32);
            f32(c->py, c->twy32);
            c->px.twf = c->twx32.p;
            c->py.twf = c->twy32.p;
            const size_t nfull = std::max(P.nfull, 1), nquad = std::max(P.nquad, 1);
            c->spec32.alloc(nfull * P.nsys * P.NZ);
            c->Fz32.alloc(nfull * P.NZAA * P.nsys);
            c->G32.alloc(nfull * size_t(P.NZAA) * P.NYAA * P.nkx);
            c->H32.alloc(nquad * size_t(P.NZAA) * P.NYAA * P.nkx);
            c->Hy32.alloc(nquad * size_t(P.NZAA) * P.nsys);
            c->Qhat32.alloc(nquad * size_t(P.nsys) * P.NZ);
            HIP_OK(hipMemset(c->Qhat32.p, 0, c->Qhat32.n * sizeof(float2)));
            c->zfrag32_b.upload(zf.bwd);
            c->zfrag32_f.upload(zf.fwd);
            int ncu = 0;
            HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
            ZFrags32Dev z{};
            z.TE = zf.TE, z.TO = zf.TO, z.NEK = zf.NEK, z.NOK = zf.NOK;
            z.RE = zf.RE, z.RO = zf.RO, z.KEK = zf.KEK, z.KOK = zf.KOK;
            z.bwd_ok = zf.bwd_ok, z.fwd_ok = zf.fwd_ok;
            z.ncu = std::max(ncu, 1);
            z.bwd = c->zfrag32_b.p, z.fwd = c->zfrag32_f.p;
            z.nbp = zf.nbp, z.nfp = zf.nfp;
            for (int q = 0; q < ZMAXPARTS; ++q) {
                z.bp[q] = zf.bp[q];
                z.fp[q] = zf.fp[q];
            }
            c->z32 = z;
        }
        if (refine > 0)
            for (auto *v : {&c->ks, &c->zs})
                for (auto &d : *v)
                    if (!d->Rres.p) {
                        d->Rres.alloc(size_t(d->nsys) * d->t->total);
                        d->Fcor.alloc(size_t(d->nsys) * d->t->total);
                    }
        c->drop_graphs();
        ++c->gen;
        c->precision = mode;
        c->refine = refine;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_kx0_projection(coral_amd_ctx *c, int32_t enable) {
    try {
        if (!c) throw std::runtime_error("null context");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->drop_graphs();
        ++c->gen;
        c->project_kx0 = enable != 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_use_graphs(coral_amd_ctx *c, int32_t enable) {
    try {
        if (!c) throw std::runtime_error("null context");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->use_graphs = enable != 0;
        c->drop_graphs();
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_timers_enable(coral_amd_ctx *c, int32_t enable) {
    try {
        if (!c) throw std::runtime_error("null context");
        c->flush_timers();
        c->totals.clear();
        c->timing = enable != 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_timer_read(coral_amd_ctx *c, int32_t i, char *name, int32_t name_len, double *total_ms,
                         int32_t *count) {
    try {
        if (!c) throw std::runtime_error("null context");
        c->flush_timers();
        int k = 0;
        for (auto &kv : c->totals) {
            if (k == i) {
                if (name && name_len > 0) {
                    std::strncpy(name, kv.first.c_str(), size_t(name_len) - 1);
                    name[name_len - 1] = 0;
                }
                if (total_ms) *total_ms = kv.second.first;
                if (count) *count = kv.second.second;
                break;
            }
            ++k;
        }
        return int(c->totals.size());
    } catch (const std::exception &e) {
        return fail(e);
    }
}


int coral_amd_set_closure(coral_amd_ctx *c, const coral_amd_closure *cl) {
    try {
        if (!c) throw std::runtime_error("null context");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->drop_graphs();
        ++c->gen;
        auto &C = c->closure;
        C.on = false;
        if (!cl) return 0;
        const Problem &P = c->P;
        if (c->precision == 1) throw std::runtime_error("the closure runs with the FP64 transform chain only");
        if (P.nfull < 1 || P.nquad < 1) throw std::runtime_error("the closure needs full and quadratic variables");
        if (!cl->weights || !cl->in_mean || !cl->in_std || !cl->out_mean || !cl->out_std)
            throw std::runtime_error("closure: null array");
        if (cl->layers < 0 || cl->modes < 0 || !nn::mlp_supported(cl->width, cl->hidden))
            throw std::runtime_error("closure: bad network shape (width 32 or 64, hidden 128 or 256)");
        const int NXAA = P.NXAA, NYAA = P.NYAA, N2 = NXAA * NYAA, M = cl->modes;
        if (2 * M >= NXAA || 2 * M >= NYAA) throw std::runtime_error("closure: modes must stay below the grid's Nyquist");
        // the plane's Fourier columns (coral_amd.closure.fourier_basis): cos and
        // sin of kx x + ky y over the half plane |kx|, |ky| <= M, each of unit
        // norm, rounded to FP32 and then BF16 like the trainer's copy
        std::vector<std::vector<double>> cols;
        for (int kx = 0; kx <= M; ++kx)
            for (int ky = -M; ky <= M; ++ky) {
                if (kx == 0 && ky < 0) continue;
                for (int sn = 0; sn < 2; ++sn) {
                    if (sn == 1 && kx == 0 && ky == 0) continue;
                    std::vector<double> v(N2);
                    double nrm = 0.0;
                    for (int y = 0; y < NYAA; ++y)
                        for (int x = 0; x < NXAA; ++x) {
                            const double ph = kx * (2.0 * M_PI * x / NXAA) + ky * (2.0 * M_PI * y / NYAA);
                            const double e = sn == 0 ? std::cos(ph) : std::sin(ph);
                            v[size_t(y) * NXAA + x] = e;
                            nrm += e * e;
                        }
                    nrm = std::sqrt(nrm);
                    for (auto &e : v) e /= nrm;
                    cols.push_back(std::move(v));
                }
            }
        const int nb = int(cols.size()), nbp = (nb + 31) / 32 * 32;
        std::vector<unsigned short> b16(size_t(N2) * nbp, 0);   // transposed: nbp x N2, zero rows past nb
        for (int k = 0; k < nb; ++k)
            for (int i = 0; i < N2; ++i) b16[size_t(k) * N2 + i] = nn::host_bf16(float(cols[k][i]));
        std::vector<unsigned short> b16p(size_t(N2) * nbp, 0);   // N2 x nbp, the fused synthesis' operand
        for (int i = 0; i < N2; ++i)
            for (int k = 0; k < nb; ++k) b16p[size_t(i) * nbp + k] = b16[size_t(k) * N2 + i];
        nn::NetParams &n = C.net;
        n = nn::NetParams{};
        n.cin = (27 * P.nfull + 1 + 7) / 8 * 8;
        n.cout = P.nquad;
        n.width = cl->width;
        n.hidden = cl->hidden;
        n.layers = cl->layers;
        n.nb = nb;
        n.N2 = N2;
        nn::net_layout(n);
        if (cl->nweights != n.size)
            throw std::runtime_error("closure: " + std::to_string(cl->nweights) + " weights given, the network has " +
                                     std::to_string(n.size));
        std::vector<float> wh(cl->weights, cl->weights + n.size);
        std::vector<unsigned short> w16(n.size);
        for (long long i = 0; i < n.size; ++i) w16[i] = nn::host_bf16(wh[i]);
        // per layer W1^T (Hd x W) and W2^T (W x Hd), the fused MLP's operands
        const int W = n.width, Hd = n.hidden;
        std::vector<unsigned short> wt(size_t(n.layers) * 2 * W * Hd);
        for (int l = 0; l < n.layers; ++l) {
            unsigned short *t1 = wt.data() + size_t(l) * 2 * W * Hd, *t2 = t1 + size_t(W) * Hd;
            for (int k = 0; k < W; ++k)
                for (int j = 0; j < Hd; ++j) {
                    t1[size_t(j) * W + k] = w16[n.off_W1[l] + size_t(k) * Hd + j];
                    t2[size_t(k) * Hd + j] = w16[n.off_W2[l] + size_t(j) * W + k];
                }
        }
        C.w.upload(wh);
        C.w16.upload(w16);
        C.wt16.upload(wt);
        C.basisT16.upload(b16);
        C.basis16p.upload(b16p);
        C.mu_in.upload(std::vector<float>(cl->in_mean, cl->in_mean + n.cin));
        C.sd_in.upload(std::vector<float>(cl->in_std, cl->in_std + n.cin));
        C.mu_out.upload(std::vector<float>(cl->out_mean, cl->out_mean + n.cout));
        C.sd_out.upload(std::vector<float>(cl->out_std, cl->out_std + n.cout));
        n.w = C.w.p;
        n.w16 = C.w16.p;
        n.wt16 = C.wt16.p;
        n.basisT16 = C.basisT16.p;
        n.basis16p = C.basis16p.p;
        n.nbp = nbp;
        std::vector<double> zg(P.NZAA);
        for (int k = 0; k < P.NZAA; ++k) zg[k] = P.center + 0.5 * P.gap * std::cos(M_PI * (2 * k + 1) / (2.0 * P.NZAA));
        C.zgrid.upload(zg);
        // planes per network launch: about 4M rows at a time (the fused MLP
        // keeps the hidden activations on chip; a row costs 2 cin + 4 cout +
        // 6 width bytes)
        C.chunk = std::max(1, std::min(P.NZAA, int(4000000 / N2)));
        const size_t R = size_t(C.chunk) * N2;
        C.X16.alloc(R * n.cin);
        C.Y.alloc(R * n.cout);
        C.H.alloc(R * n.width);
        C.H16.alloc(R * n.width);
        C.Cm.alloc(size_t(C.chunk) * nb * n.width);
        C.Cs16.alloc(size_t(C.chunk) * nbp * n.width);
        long long ngw = 0;
        for (int p = 1; p <= C.chunk; ++p) ngw = std::max(ngw, nn::gemm_work(nb, n.width, N2, p));
        C.gwork.alloc(size_t(std::max(ngw, 1LL)));
        C.ws = nn::NetWork{};
        C.ws.H = C.H.p;
        C.ws.Cm = C.Cm.p;
        C.ws.gwork = C.gwork.p;
        C.ws.ngwork = ngw;
        C.ws.H16 = C.H16.p;
        C.ws.Cs16 = C.Cs16.p;
        if (N2 % nn::MLP_ROWS == 0 && nn::modes_supported(nbp, N2, n.width)) {
            C.H16T.alloc(R * n.width);
            C.mwork.alloc(size_t(nn::modes_work(nbp, N2, n.width, C.chunk)));
            C.ws.H16T = C.H16T.p;
            C.ws.mwork = C.mwork.p;
            C.ws.nmwork = nn::modes_work(nbp, N2, n.width, C.chunk);
        }
        // the fused lift / projection: tiles inside x-lines, at most 16 outputs,
        // their dynamic LDS within the device limit (CORAL_AMD_CLOSURE_UNFUSED=1
        // keeps the pack, GEMM and unpack kernels)
        {
            const char *ev = std::getenv("CORAL_AMD_CLOSURE_UNFUSED");
            const LiftShape ls = lift_shape(P.nfull, n.cin, n.width);
            C.fused_io = !(ev && ev[0] == '1') && P.NXAA % LIFT_ROWS == 0 && n.cout <= 16 &&
                         (n.width == 32 || n.width == 64) && ls.kp <= 256 && ls.bytes <= 65536;
            if (C.fused_io && ls.bytes > 49152) {
                if (n.width == 32)
                    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void *>(k_closure_lift<32>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, int(ls.bytes)));
                else
                    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void *>(k_closure_lift<64>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, int(ls.bytes)));
            }
            if (C.fused_io) {   // no packed input rows or FP32 outputs to keep
                C.X16.alloc(0);
                C.Y.alloc(0);
                std::vector<unsigned short> w0t(size_t(n.width) * ls.kp, 0);
                for (int k = 0; k < n.cin; ++k)
                    for (int j = 0; j < n.width; ++j) w0t[size_t(j) * ls.kp + k] = w16[n.off_W0 + size_t(k) * n.width + j];
                C.W0T.upload(w0t);
            }
        }
        C.phys_full.alloc(size_t(P.nfull) * P.NZAA * N2);
        C.phys_tau.alloc(size_t(P.nquad) * P.NZAA * N2);
        C.Qtau.alloc(size_t(P.nquad) * P.nsys * P.NZ);
        C.on = true;
        return 0;
    } catch (const std::exception &e) {
        if (c) c->closure.on = false;
        return fail(e);
    }
}

namespace {
// The sources' part of each mean system's K_hat, the same at every stage:
// for each term, its source's first NZ Chebyshev coefficients times dsca,
// through trunc(I^iz)[order:, :] into the rows of its equation
// (PL_IMEX_timestepping_marching.f90:399-416).  Terms whose source has no
// definition are left out and the lowest such index kept in src_missing, on
// which stepping stops.
void rebuild_sources(coral_amd_ctx &c) {
    const Problem &P = c.P;
    const int NZ = P.NZ;
    c.src_missing = 0;
    std::vector<std::vector<double>> v(c.zs.size());
    for (size_t i = 0; i < c.zs.size(); ++i) v[i].assign(c.zs[i]->src.n, 0.0);
    std::map<int, Dense> im;
    for (const auto &st : c.src_terms) {
        const auto it = c.src_cheb.find(st.index);
        if (it == c.src_cheb.end()) {
            if (!c.src_missing || st.index < c.src_missing) c.src_missing = st.index;
            continue;
        }
        const SystemTables &t = *c.zs[st.isys]->t;
        auto m = im.find(st.iz);
        if (m == im.end()) m = im.emplace(st.iz, cheb_IM(NZ, P.center, P.gap, st.iz, 0)).first;
        const Dense &I = m->second;
        const int order = t.eqn_orders[st.eqn];
        std::vector<double> b(NZ);
        for (int k = 0; k < NZ; ++k) b[k] = it->second[k] * st.dsca;
        for (int p = 0; p < t.total; ++p) {
            if (t.pos_eqn[p] != st.eqn) continue;
            const int r = t.pos_row[p];
            double acc = 0.0;
            for (int k = 0; k < NZ; ++k) acc += I(order + r, k) * b[k];
            v[st.isys][p] += acc;
        }
    }
    // the vectors are read by every step in flight, including an ensemble's
    HIP_OK(hipDeviceSynchronize());
    for (size_t i = 0; i < c.zs.size(); ++i)
        HIP_OK(hipMemcpy(c.zs[i]->src.p, v[i].data(), v[i].size() * sizeof(double), hipMemcpyHostToDevice));
}
}  // namespace

int coral_amd_add_source_term(coral_amd_ctx *c, int32_t isys, int32_t eqn, int32_t iz, double dsca,
                              int32_t source_index) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (isys < 0 || isys >= (int)c->zs.size()) throw std::runtime_error("mean system index out of range");
        if (eqn < 0 || eqn >= c->zs[isys]->t->nvars) throw std::runtime_error("equation index out of range");
        if (iz < 0) throw std::runtime_error("negative Iz exponent");
        if (source_index < 1 || source_index > 99) throw std::runtime_error("source index must be in 1..99");
        HIP_OK(hipSetDevice(c->device));
        c->src_terms.push_back({isys, eqn, iz, source_index, dsca});
        rebuild_sources(*c);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_clear_source_terms(coral_amd_ctx *c) {
    try {
        if (!c) throw std::runtime_error("null context");
        HIP_OK(hipSetDevice(c->device));
        c->src_terms.clear();
        rebuild_sources(*c);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_source(coral_amd_ctx *c, int32_t source_index, const double *physical) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (source_index < 1 || source_index > 99) throw std::runtime_error("source index must be in 1..99");
        HIP_OK(hipSetDevice(c->device));
        if (!physical) {
            c->src_cheb.erase(source_index);
        } else {
            // Compute_1dCheby_transform (PL_IMEX_timestepping_marching.f90:845-866):
            // REDFT10 of the NZAA Gauss-point values over NZAA, the first halved;
            // the first NZ coefficients are used.  In the cosine's argument
            // pi k (2j+1) / (2 NZAA), k (2j+1) is reduced modulo 4 NZAA (its period).
            const int N = c->P.NZAA, NZ = c->P.NZ;
            std::vector<double> coef(NZ, 0.0);
            const double h = 3.14159265358979323846 / (2.0 * N);
            for (int k = 0; k < NZ; ++k) {
                double acc = 0.0;
                for (int j = 0; j < N; ++j) {
                    const long long m = (long long)k * (2 * j + 1) % (4LL * N);
                    acc += physical[j] * std::cos(h * double(m));
                }
                coef[k] = 2.0 * acc / N;
            }
            coef[0] *= 0.5;
            c->src_cheb[source_index] = coef;
        }
        rebuild_sources(*c);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_fixed_physical(coral_amd_ctx *c, int32_t ifull, const double *phys) {
    try {
        if (!c) throw std::runtime_error("null context");
        const Problem &P = c->P;
        if (ifull < 0 || ifull >= P.nfull) throw std::runtime_error("full variable index out of range");
        if (ifull < 3 && phys)
            throw std::runtime_error("the first three full variables are the velocities of the CFL bound and cannot be fixed");
        if (ifull == c->nu_full) throw std::runtime_error("the Smagorinsky variable cannot be fixed");
        if (phys && c->precision == 1)
            throw std::runtime_error("fixed physical fields run with the FP64 transform chain only");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipDeviceSynchronize());
        const size_t plane = size_t(P.NZAA) * P.NYAA * P.NXAA;
        if ((int)c->fixed_slot.size() < P.nfull) c->fixed_slot.resize(P.nfull, -1);
        if (!phys) {
            c->fixed_slot[ifull] = -1;   // its block stays allocated, unused
        } else {
            if (c->fixed_slot[ifull] < 0) {
                // grow the block array by one plane block, keeping the others
                DevBuf<double> grown;
                grown.alloc((c->nfixed + 1) * plane);
                if (c->nfixed)
                    HIP_OK(hipMemcpy(grown.p, c->fixed_all.p, c->nfixed * plane * sizeof(double), hipMemcpyDeviceToDevice));
                std::swap(grown.p, c->fixed_all.p);
                std::swap(grown.n, c->fixed_all.n);
                c->fixed_slot[ifull] = c->nfixed++;
            }
            // host [x][y][z] -> device [z][y][x]
            std::vector<double> h(plane);
            for (int z = 0; z < P.NZAA; ++z)
                for (int y = 0; y < P.NYAA; ++y)
                    for (int x = 0; x < P.NXAA; ++x)
                        h[(size_t(z) * P.NYAA + y) * P.NXAA + x] = phys[(size_t(x) * P.NYAA + y) * P.NZAA + z];
            HIP_OK(hipMemcpy(c->fixed_all.p + size_t(c->fixed_slot[ifull]) * plane, h.data(), plane * sizeof(double),
                             hipMemcpyHostToDevice));
        }
        rebuild_side(*c);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_smagorinsky(coral_amd_ctx *c, int32_t ifull, const int32_t *grads, int32_t ngrads,
                              double prefactor) {
    try {
        if (!c) throw std::runtime_error("null context");
        const Problem &P = c->P;
        if (ifull >= P.nfull) throw std::runtime_error("full variable index out of range");
        if (ifull >= 0 && ifull != P.nfull - 1)
            throw std::runtime_error("the Smagorinsky variable is the last full variable");
        if (ngrads < 0 || ngrads > 4 || (ngrads && !grads)) throw std::runtime_error("0 to 4 gradient variables");
        if (ifull >= 0 && c->precision == 1)
            throw std::runtime_error("the Smagorinsky viscosity runs with the FP64 transform chain only");
        std::vector<int> g;
        for (int k = 0; k < ngrads; ++k) {
            if (grads[k] < 0 || grads[k] >= P.nfull || grads[k] == ifull)
                throw std::runtime_error("bad gradient variable index");
            if (c->fixed_slot.size() > size_t(grads[k]) && c->fixed_slot[grads[k]] >= 0)
                throw std::runtime_error("a gradient variable has fixed physical values");
            g.push_back(grads[k]);
        }
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipDeviceSynchronize());
        c->nu_full = ifull;
        c->nu_grads = g;
        c->nu_prefactor = prefactor;
        // the profile starts at zero, and restarts at zero when set again
        c->nu_prof.upload(std::vector<double>(P.NZAA, 0.0));
        c->nu_tmp.upload(std::vector<double>(P.NZAA, 0.0));
        rebuild_side(*c);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_full_variable_filter(coral_amd_ctx *c, int32_t ifull, int32_t extract, int32_t zdecomp) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (ifull < 0 || ifull >= c->P.nfull) throw std::runtime_error("full variable index out of range");
        if (extract < 0 || extract > 2) throw std::runtime_error("extract must be 0 (none), 1 (bottom) or 2 (top)");
        if (zdecomp < 0 || zdecomp > 2)
            throw std::runtime_error("vertical decomposition must be 0 (none), 1 (barotropic) or 2 (baroclinic)");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->drop_graphs();
        ++c->gen;
        if ((int)c->full_filter.size() < c->P.nfull) c->full_filter.resize(c->P.nfull, {0, 0});
        c->full_filter[ifull] = {extract, zdecomp};
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_set_launch_shape(coral_amd_ctx *c, int32_t xpass_wpb, int32_t z_plain) {
    try {
        if (!c) throw std::runtime_error("null context");
        if (xpass_wpb < 0 || xpass_wpb > 4) throw std::runtime_error("x pass wavefronts per workgroup must be 0..4");
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->drop_graphs();
        ++c->gen;
        c->xpass_wpb = xpass_wpb;
        c->z_plain = z_plain != 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_get_launch_shape(coral_amd_ctx *c, int32_t *xpass_wpb, int32_t *z_plain) {
    if (!c) {
        g_err = "null context";
        return -1;
    }
    if (xpass_wpb) *xpass_wpb = c->xpass_wpb;
    if (z_plain) *z_plain = c->z_plain ? 1 : 0;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ensembles: independent runs (each its own context, operators, dt and state)
// stepped together.  One step of every run is one HIP graph: the runs are
// spread over nbranch parallel branches (a fork from the ensemble's stream,
// each branch a stream of its own, joined at the end), so the small kernels
// of different runs overlap on the device; inside a branch its runs step one
// after the other.  Each run's step swaps its state and last-stage buffers,
// so two graphs alternate, as for a single run.
// ---------------------------------------------------------------------------
struct coral_amd_ensemble {
    std::vector<coral_amd_ctx *> runs;
    int device = 0;
    hipStream_t master = nullptr;
    std::vector<hipStream_t> branch;
    std::vector<int> assign;          // run -> branch
    int nbranch = 1;
    hipEvent_t ev_fork = nullptr, ev_done = nullptr;
    std::vector<hipEvent_t> ev_join, ev_run;
    // per branch, two graphs (one per buffer assignment) of one step of the
    // branch's runs; each branch replays on its own stream
    std::vector<hipGraphExec_t> gexec[2];
    int epar = 0;                     // graph to replay next
    std::vector<unsigned long long> cap_nswap;   // each run's swap count after the capture
    std::vector<unsigned long long> cap_gen;
    std::vector<double> run_ms;       // per-run step time measured by the tuner (0 = not measured)
    double tuned_ms = 0.0;            // ensemble step time of the tuned schedule
    // lockstep schedule (set_schedule with 0 branches): one graph per step
    // holding every run's NL terms and right-hand sides of a stage, then the
    // stage's kxky solves of all runs as one launch per kernel variant; the
    // launches' argument tables are filled during the capture and uploaded
    // after it
    bool lockstep = false;
    std::vector<Solve16Args> lk_args_h;
    std::vector<int> lk_start_h;
    DevBuf<Solve16Args> lk_args;
    DevBuf<int> lk_start;

    void drop() {
        // an executable graph is destroyed only once its launches have ended
        if (!gexec[0].empty()) {
            for (auto &b : branch) (void)hipStreamSynchronize(b);
            if (master) (void)hipStreamSynchronize(master);
        }
        for (auto &v : gexec) {
            for (auto &g : v)
                if (g) (void)hipGraphExecDestroy(g);
            v.clear();
        }
        epar = 0;
    }
    ~coral_amd_ensemble() {
        drop();
        for (auto &b : branch) (void)hipStreamDestroy(b);
        for (auto &ev : ev_join) (void)hipEventDestroy(ev);
        for (auto &ev : ev_run) (void)hipEventDestroy(ev);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_done) (void)hipEventDestroy(ev_done);
        if (master) (void)hipStreamDestroy(master);
    }
};

namespace {

void ens_set_branches(coral_amd_ensemble &e, int nb, const std::vector<int> &assign) {
    e.drop();
    e.lockstep = false;
    while (int(e.branch.size()) < nb) {
        hipStream_t st = nullptr;
        HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        e.branch.push_back(st);
        hipEvent_t ev = nullptr;
        HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        e.ev_join.push_back(ev);
    }
    e.nbranch = nb;
    e.assign = assign;
}

// Longest-processing-time schedule: runs in decreasing step time, each to the
// branch with the least time so far.
std::vector<int> lpt_assign(const std::vector<double> &ms, int nb) {
    std::vector<int> order(ms.size()), assign(ms.size(), 0);
    for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ms[a] > ms[b]; });
    std::vector<double> load(nb, 0.0);
    for (int r : order) {
        const int b = int(std::min_element(load.begin(), load.end()) - load.begin());
        assign[r] = b;
        load[b] += ms[r] > 0.0 ? ms[r] : 1.0;
    }
    return assign;
}

void ens_capture(coral_amd_ensemble &e) {
    e.drop();
    for (auto *c : e.runs) {
        if (c->timing) throw std::runtime_error("ensemble steps need timers off on every run");
        if (!c->factored) throw std::runtime_error("coral_amd_set_dt must be called on every run before stepping");
    }
    for (int p = 0; p < 2; ++p) {
        e.gexec[p].assign(e.nbranch, nullptr);
        for (int b = 0; b < e.nbranch; ++b) {
            hipStream_t st = e.branch[b];
            HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            hipGraph_t g = nullptr;
            try {
                for (size_t r = 0; r < e.runs.size(); ++r) {
                    if (e.assign[r] != b) continue;
                    coral_amd_ctx &c = *e.runs[r];
                    hipStream_t own = c.stream;
                    c.stream = st;
                    try {
                        one_step(c);
                    } catch (...) {
                        c.stream = own;
                        throw;
                    }
                    c.stream = own;
                }
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                e.drop();
                throw;
            }
            HIP_OK(hipStreamEndCapture(st, &g));
            const hipError_t err = hipGraphInstantiate(&e.gexec[p][b], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (err != hipSuccess) {
                e.drop();
                throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
            }
        }
    }
    // each capture swapped every run's buffers once: they are back, parity too
    e.cap_nswap.resize(e.runs.size());
    e.cap_gen.resize(e.runs.size());
    for (size_t r = 0; r < e.runs.size(); ++r) {
        e.cap_nswap[r] = e.runs[r]->nswap;
        e.cap_gen[r] = e.runs[r]->gen;
    }
    e.epar = 0;
}

// whether the runs can step in lockstep: the same number of stages and no
// residual refinement (the batched solves leave it out)
bool ens_lockstep_ok(const coral_amd_ensemble &e, std::string *why = nullptr) {
    for (auto *c : e.runs) {
        if (c->P.nstages != e.runs[0]->P.nstages) {
            if (why) *why = "lockstep steps need the same IMEX scheme on every run";
            return false;
        }
        if (c->refine) {
            if (why) *why = "lockstep steps do not take residual refinement";
            return false;
        }
    }
    return true;
}

// The lockstep graphs (one per buffer assignment): per stage every run's NL
// terms (its mean systems forked onto its side stream) and kxky right-hand
// sides, then the stage's kxky solves grouped by kernel variant, each group
// one k_solve16_batch launch; systems the 16-lane solve does not take are
// solved on their own.  Every run's arithmetic is its own step's.
void ens_capture_lockstep(coral_amd_ensemble &e) {
    e.drop();
    std::string why;
    if (!ens_lockstep_ok(e, &why)) throw std::runtime_error(why);
    for (auto *c : e.runs) {
        if (c->timing) throw std::runtime_error("ensemble steps need timers off on every run");
        if (!c->factored) throw std::runtime_error("coral_amd_set_dt must be called on every run before stepping");
    }
    const int s = e.runs[0]->P.nstages;
    size_t nks = 0;
    for (auto *c : e.runs) nks += c->ks.size();
    const size_t cap_args = 2 * size_t(s) * std::max<size_t>(nks, 1), cap_start = 2 * size_t(s) * (nks + 8);
    e.lk_args_h.assign(cap_args, Solve16Args{});
    e.lk_start_h.assign(cap_start, 0);
    e.lk_args.alloc(cap_args);
    e.lk_start.alloc(cap_start);
    size_t ao = 0, so = 0;
    hipStream_t st = e.branch[0];
    for (int p = 0; p < 2; ++p) {
        e.gexec[p].assign(1, nullptr);
        std::vector<hipStream_t> own(e.runs.size());
        for (size_t r = 0; r < e.runs.size(); ++r) own[r] = e.runs[r]->stream;
        auto restore = [&] {
            for (size_t r = 0; r < e.runs.size(); ++r) e.runs[r]->stream = own[r];
        };
        HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        hipGraph_t g = nullptr;
        try {
            for (auto *c : e.runs) {
                c->stream = st;
                step_begin(*c);
            }
            for (int i = 0; i < s; ++i) {
                // the runs' NL terms and right-hand sides, branch b's runs on
                // its stream (branch 0 on the capture's own), joined before the solves
                HIP_OK(hipEventRecord(e.ev_fork, st));
                for (int b = 0; b < e.nbranch; ++b) {
                    hipStream_t bs = b == 0 ? st : e.branch[b];
                    if (b > 0) HIP_OK(hipStreamWaitEvent(bs, e.ev_fork, 0));
                    for (size_t r = 0; r < e.runs.size(); ++r) {
                        if (e.assign[r] != b) continue;
                        coral_amd_ctx &c = *e.runs[r];
                        c.stream = bs;
                        stage_nl(c, i);
                        for (auto &d : c.ks) stage_rhs(c, i, *d);
                    }
                    if (b > 0) {
                        HIP_OK(hipEventRecord(e.ev_join[b], bs));
                        HIP_OK(hipStreamWaitEvent(st, e.ev_join[b], 0));
                    }
                }
                for (auto *c : e.runs) c->stream = st;
                // (real form, NQ) -> the group's arguments, workgroups and LDS
                std::map<std::pair<int, int>, std::vector<std::pair<Solve16Args, int>>> groups;
                std::map<std::pair<int, int>, size_t> glds;
                for (auto *c : e.runs)
                    for (auto &d : c->ks) {
                        Solve16Args sa;
                        int nwg = 0, nq = 0;
                        size_t lds = 0;
                        if (solve16_args_of(*c, *d, d->R.p, d->X[i]->p, sa) && solve16_shape(sa, nwg, lds, nq)) {
                            if (!d->real) sa.rotate = 0;
                            const auto key = std::make_pair(d->real ? 1 : 0, nq);
                            groups[key].push_back({sa, nwg});
                            glds[key] = std::max(glds[key], lds);
                        } else {
                            stage_solve(*c, i, *d);
                        }
                    }
                for (auto &gp : groups) {
                    const auto &v = gp.second;
                    if (ao + v.size() > cap_args || so + v.size() + 1 > cap_start)
                        throw std::runtime_error("lockstep argument tables too small");
                    int tot = 0;
                    for (size_t k = 0; k < v.size(); ++k) {
                        e.lk_args_h[ao + k] = v[k].first;
                        e.lk_start_h[so + k] = tot;
                        tot += v[k].second;
                    }
                    e.lk_start_h[so + v.size()] = tot;
                    if (!launch_solve16_batch(st, gp.first.first != 0, gp.first.second, tot, glds[gp.first],
                                              e.lk_args.p + ao, e.lk_start.p + so, int(v.size())))
                        throw std::runtime_error("batched 16-lane solve rejected its launches");
                    ao += v.size();
                    so += v.size() + 1;
                }
            }
            for (auto *c : e.runs) step_end(*c);
        } catch (...) {
            restore();
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            e.drop();
            throw;
        }
        restore();
        HIP_OK(hipStreamEndCapture(st, &g));
        const hipError_t err = hipGraphInstantiate(&e.gexec[p][0], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) {
            e.drop();
            throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
        }
    }
    // the graphs read their argument tables from now on
    HIP_OK(hipMemcpy(e.lk_args.p, e.lk_args_h.data(), ao * sizeof(Solve16Args), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(e.lk_start.p, e.lk_start_h.data(), so * sizeof(int), hipMemcpyHostToDevice));
    e.cap_nswap.resize(e.runs.size());
    e.cap_gen.resize(e.runs.size());
    for (size_t r = 0; r < e.runs.size(); ++r) {
        e.cap_nswap[r] = e.runs[r]->nswap;
        e.cap_gen[r] = e.runs[r]->gen;
    }
    e.epar = 0;
}

// one ensemble step: every branch's graph on its own stream, forked from and
// joined into the ensemble's stream (lockstep: its one graph on that stream)
void ens_launch(coral_amd_ensemble &e) {
    if (e.lockstep) {
        HIP_OK(hipGraphLaunch(e.gexec[e.epar][0], e.master));
        for (auto *c : e.runs) swap_step_buffers(*c);
        e.epar ^= 1;
        return;
    }
    HIP_OK(hipEventRecord(e.ev_fork, e.master));
    for (int b = 0; b < e.nbranch; ++b) {
        HIP_OK(hipStreamWaitEvent(e.branch[b], e.ev_fork, 0));
        HIP_OK(hipGraphLaunch(e.gexec[e.epar][b], e.branch[b]));
        HIP_OK(hipEventRecord(e.ev_join[b], e.branch[b]));
        HIP_OK(hipStreamWaitEvent(e.master, e.ev_join[b], 0));
    }
    for (auto *c : e.runs) swap_step_buffers(*c);
    e.epar ^= 1;
}

// the captured graphs still hold every run's current launches and buffer assignment
bool ens_valid(const coral_amd_ensemble &e) {
    if (e.gexec[0].empty()) return false;
    for (size_t r = 0; r < e.runs.size(); ++r)
        if (e.runs[r]->gen != e.cap_gen[r] || int((e.runs[r]->nswap - e.cap_nswap[r]) & 1) != e.epar) return false;
    return true;
}

// nsteps ensemble steps on the ensemble's stream, ordered after each run's
// queued work and before anything queued on the runs afterwards
void ens_steps(coral_amd_ensemble &e, int nsteps) {
    for (size_t r = 0; r < e.runs.size(); ++r) {
        HIP_OK(hipEventRecord(e.ev_run[r], e.runs[r]->stream));
        HIP_OK(hipStreamWaitEvent(e.master, e.ev_run[r], 0));
    }
    for (int n = 0; n < nsteps; ++n) {
        if (!ens_valid(e)) {
            if (e.lockstep)
                ens_capture_lockstep(e);
            else
                ens_capture(e);
        }
        ens_launch(e);
    }
    HIP_OK(hipEventRecord(e.ev_done, e.master));
    for (auto *c : e.runs) HIP_OK(hipStreamWaitEvent(c->stream, e.ev_done, 0));
}

// device copies of every system's state, to run tuning steps and put the state back
struct StateBackup {
    std::vector<std::unique_ptr<DevBuf<double2>>> bufs;
    void save(coral_amd_ctx &c) {
        bufs.clear();
        for (auto *v : {&c.ks, &c.zs})
            for (auto &d : *v) {
                bufs.emplace_back(new DevBuf<double2>());
                const size_t n = size_t(d->nsys) * d->t->total;
                bufs.back()->alloc(n);
                HIP_OK(hipMemcpyAsync(bufs.back()->p, d->field.p, n * sizeof(double2), hipMemcpyDeviceToDevice,
                                      c.stream));
            }
        HIP_OK(hipStreamSynchronize(c.stream));
    }
    void restore(coral_amd_ctx &c) {
        HIP_OK(hipStreamSynchronize(c.stream));
        size_t k = 0;
        for (auto *v : {&c.ks, &c.zs})
            for (auto &d : *v) {
                const size_t n = size_t(d->nsys) * d->t->total;
                HIP_OK(hipMemcpyAsync(d->field.p, bufs[k++]->p, n * sizeof(double2), hipMemcpyDeviceToDevice,
                                      c.stream));
            }
        HIP_OK(hipStreamSynchronize(c.stream));
    }
};

// milliseconds per step of `steps` graph replays of one run (an even count
// leaves its buffer assignment where it was)
double time_run_steps(coral_amd_ctx &c, int steps) {
    const bool g0 = c.use_graphs;
    c.use_graphs = true;
    c.drop_graphs();
    capture_step_graphs(c);
    for (int n = 0; n < 2; ++n) {   // warm-up
        HIP_OK(hipGraphLaunch(c.gexec[c.gpar], c.stream));
        swap_step_buffers(c);
    }
    hipEvent_t a = nullptr, b = nullptr;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    HIP_OK(hipEventRecord(a, c.stream));
    for (int n = 0; n < steps; ++n) {
        HIP_OK(hipGraphLaunch(c.gexec[c.gpar], c.stream));
        swap_step_buffers(c);
    }
    HIP_OK(hipEventRecord(b, c.stream));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c.drop_graphs();
    c.use_graphs = g0;
    return double(ms) / steps;
}

double time_ens_steps(coral_amd_ensemble &e, int steps) {
    ens_steps(e, 2);   // capture and warm-up
    hipEvent_t a = nullptr, b = nullptr;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    HIP_OK(hipEventRecord(a, e.master));
    for (int n = 0; n < steps; ++n) ens_launch(e);
    HIP_OK(hipEventRecord(b, e.master));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms) / steps;
}

}  // namespace

extern "C" {

int coral_amd_ensemble_create(coral_amd_ctx *const *runs, int32_t nrun, coral_amd_ensemble **out) {
    try {
        if (!runs || !out || nrun < 1) throw std::runtime_error("null argument or empty ensemble");
        std::unique_ptr<coral_amd_ensemble> e(new coral_amd_ensemble());
        for (int r = 0; r < nrun; ++r) {
            if (!runs[r]) throw std::runtime_error("null run context");
            for (int q = 0; q < r; ++q)
                if (runs[q] == runs[r]) throw std::runtime_error("a context appears twice in the ensemble");
            if (runs[r]->device != runs[0]->device) throw std::runtime_error("every run must be on one device");
            e->runs.push_back(runs[r]);
        }
        e->device = runs[0]->device;
        HIP_OK(hipSetDevice(e->device));
        HIP_OK(hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&e->ev_done, hipEventDisableTiming));
        e->ev_run.resize(nrun, nullptr);
        for (auto &ev : e->ev_run) HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        e->run_ms.assign(nrun, 0.0);
        ens_set_branches(*e, 1, std::vector<int>(nrun, 0));
        *out = e.release();
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int coral_amd_ensemble_destroy(coral_amd_ensemble *e) {
    try {
        if (!e) return 0;
        HIP_OK(hipSetDevice(e->device));
        HIP_OK(hipStreamSynchronize(e->master));
        delete e;
        return 0;
    } catch (const std::exception &ex) {
        return fail(ex);
    }
}

int coral_amd_ensemble_set_schedule(coral_amd_ensemble *e, int32_t nbranch, const int32_t *assign) {
    try {
        if (!e) throw std::runtime_error("null ensemble");
        if (nbranch < -int(e->runs.size()) || nbranch > int(e->runs.size()))
            throw std::runtime_error("nbranch must be -nrun..nrun");
        HIP_OK(hipSetDevice(e->device));
        HIP_OK(hipStreamSynchronize(e->master));
        if (nbranch <= 0) {
            // lockstep, the stage fronts over max(1, -nbranch) branches
            std::string why;
            if (!ens_lockstep_ok(*e, &why)) throw std::runtime_error(why);
            const int nb = std::max(1, -nbranch);
            std::vector<int> a(e->runs.size());
            if (assign) {
                for (size_t r = 0; r < a.size(); ++r) {
                    if (assign[r] < 0 || assign[r] >= nb) throw std::runtime_error("branch index out of range");
                    a[r] = assign[r];
                }
            } else {
                a = lpt_assign(e->run_ms, nb);
            }
            ens_set_branches(*e, nb, a);
            e->lockstep = true;
            return 0;
        }
        std::vector<int> a(e->runs.size());
        if (assign) {
            for (size_t r = 0; r < a.size(); ++r) {
                if (assign[r] < 0 || assign[r] >= nbranch) throw std::runtime_error("branch index out of range");
                a[r] = assign[r];
            }
        } else {
            a = lpt_assign(e->run_ms, nbranch);
        }
        ens_set_branches(*e, nbranch, a);
        return 0;
    } catch (const std::exception &ex) {
        return fail(ex);
    }
}

int coral_amd_ensemble_step(coral_amd_ensemble *e, int32_t nsteps) {
    try {
        if (!e) throw std::runtime_error("null ensemble");
        if (nsteps < 0) throw std::runtime_error("nsteps must be >= 0");
        HIP_OK(hipSetDevice(e->device));
        ens_steps(*e, nsteps);
        HIP_OK(hipGetLastError());
        return 0;
    } catch (const std::exception &ex) {
        return fail(ex);
    }
}

int coral_amd_ensemble_synchronize(coral_amd_ensemble *e) {
    try {
        if (!e) throw std::runtime_error("null ensemble");
        HIP_OK(hipSetDevice(e->device));
        HIP_OK(hipStreamSynchronize(e->master));
        return 0;
    } catch (const std::exception &ex) {
        return fail(ex);
    }
}

int coral_amd_ensemble_tune(coral_amd_ensemble *e, int32_t trial_steps, int32_t *nbranch_out, double *ms_out) {
    try {
        if (!e) throw std::runtime_error("null ensemble");
        const int steps = std::max(2, (trial_steps + 1) / 2 * 2);
        HIP_OK(hipSetDevice(e->device));
        HIP_OK(hipStreamSynchronize(e->master));
        const size_t nr = e->runs.size();
        std::vector<StateBackup> saved(nr);
        for (size_t r = 0; r < nr; ++r) {
            coral_amd_ctx &c = *e->runs[r];
            if (!c.factored) throw std::runtime_error("coral_amd_set_dt must be called on every run before tuning");
            if (c.timing) throw std::runtime_error("tuning needs timers off on every run");
            HIP_OK(hipStreamSynchronize(c.stream));
            saved[r].save(c);
        }
        // launch shapes of each run: x pass wavefronts per workgroup and the
        // z transforms' arithmetic (MFMA products or vector GEMMs)
        for (size_t r = 0; r < nr; ++r) {
            coral_amd_ctx &c = *e->runs[r];
            double best = 1e300;
            int bw = 0, bz = 0;
            for (int z = 0; z < 2; ++z)
                for (int w : {4, 2, 1}) {
                    c.xpass_wpb = w;
                    c.z_plain = z != 0;
                    ++c.gen;
                    const double ms = time_run_steps(c, steps);
                    if (ms < best) {
                        best = ms;
                        bw = w;
                        bz = z;
                    }
                }
            c.xpass_wpb = bw;
            c.z_plain = bz != 0;
            ++c.gen;
            e->run_ms[r] = best;
            saved[r].restore(c);
        }
        // branches: LPT schedules over 1, 2, 4, ... branches
        double best = 1e300;
        int bnb = 1;
        for (int nb = 1; nb <= int(nr) && nb <= 16; nb *= 2) {
            ens_set_branches(*e, nb, lpt_assign(e->run_ms, nb));
            const double ms = time_ens_steps(*e, steps);
            if (ms < best) {
                best = ms;
                bnb = nb;
            }
        }
        // and lockstep schedules with their stage fronts over 1, 2, 4, ...
        // branches (reported as minus that count)
        if (ens_lockstep_ok(*e)) {
            for (int nb = 1; nb <= int(nr) && nb <= 16; nb *= 2) {
                ens_set_branches(*e, nb, lpt_assign(e->run_ms, nb));
                e->lockstep = true;
                const double ms = time_ens_steps(*e, steps);
                if (ms < best) {
                    best = ms;
                    bnb = -nb;
                }
            }
        }
        HIP_OK(hipStreamSynchronize(e->master));
        ens_set_branches(*e, std::abs(bnb), lpt_assign(e->run_ms, std::abs(bnb)));
        e->lockstep = bnb < 0;
        e->tuned_ms = best;
        for (size_t r = 0; r < nr; ++r) saved[r].restore(*e->runs[r]);
        if (nbranch_out) *nbranch_out = bnb;
        if (ms_out) *ms_out = best;
        return 0;
    } catch (const std::exception &ex) {
        return fail(ex);
    }
}

}  // extern "C"
