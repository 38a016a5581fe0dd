"""The closure's HIP kernels (coral_amd_nn_*, nn.hip) through the C ABI: the
BF16-MFMA GEMM against numpy on the same BF16 operands, the network's
forward and backward against the CPU reference (oracle/closure_ref.py), and a
short training run.

Tolerances: the GEMM's products of BF16 operands are exact in FP32 and only
the FP32 summation differs from numpy's float64 (relative 1e-5).  The network
rounds every GEMM operand to BF16 on both sides, but the values rounded differ
in their last FP32 bits, so a few roundings land one BF16 ulp (2^-8) apart:
outputs and gradients agree to a few 1e-3 in relative L2, far below the
size of any gradient error (a wrong term or sign is an O(1) difference)."""
import numpy as np
import pytest

from oracle import closure_ref as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nnmod(native_lib):
    from coral_amd import closure, devmem
    return closure, devmem


def to_dev_bf16(closure, devmem, a):
    x = devmem.DevArray.from_numpy(a)
    y = devmem.DevArray(x.n, "bf16")
    closure.check(closure.lib.coral_amd_nn_cast_bf16(x.ptr, y.ptr, x.n, None), "cast")
    return y


@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_bf16_gemm_matches_numpy(nnmod, tA, tB):
    closure, devmem = nnmod
    rng = np.random.default_rng(tA * 2 + tB)
    M, N, K, batch = 70, 33, 45, 3
    A = rng.normal(size=(batch, K, M) if tA else (batch, M, K))
    B = rng.normal(size=(batch, N, K) if tB else (batch, K, N))
    C0 = rng.normal(size=(batch, M, N))
    Ad, Bd = to_dev_bf16(closure, devmem, A), to_dev_bf16(closure, devmem, B)
    Cd = devmem.DevArray.from_numpy(C0)
    lda, ldb = A.shape[2], B.shape[2]
    closure.check(closure.lib.coral_amd_nn_gemm(tA, tB, M, N, K, 0.5, Ad.ptr, lda, A[0].size, Bd.ptr,
                                                ldb, B[0].size, 1.0, Cd.ptr, N, M * N, batch, None, 0, None), "gemm")
    got = Cd.numpy().astype(np.float64).reshape(batch, M, N)
    Aq, Bq = C.bf16(A), C.bf16(B)
    for b in range(batch):
        opA = Aq[b].T if tA else Aq[b]
        opB = Bq[b].T if tB else Bq[b]
        ref = 0.5 * opA @ opB + C0[b].astype(np.float32)
        assert np.linalg.norm(got[b] - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("tA", [0, 1])
def test_split_k_gemm_matches_numpy(nnmod, tA):
    """A long K with few output tiles (the per-plane projection onto the
    Fourier modes): split along K into partials, summed in a fixed order;
    equal to numpy to FP32 summation, the same bits on a second call, and
    refused without its work space."""
    closure, devmem = nnmod
    rng = np.random.default_rng(10 + tA)
    M, N, K, batch = 81, 64, 9216, 3
    A = rng.normal(size=(batch, K, M) if tA else (batch, M, K))
    B = rng.normal(size=(batch, K, N))
    C0 = rng.normal(size=(batch, M, N))
    need = int(closure.lib.coral_amd_nn_gemm_work(M, N, K, batch))
    assert need > 0
    Ad, Bd = to_dev_bf16(closure, devmem, A), to_dev_bf16(closure, devmem, B)
    work = devmem.DevArray(need)
    outs = []
    for _ in range(2):
        Cd = devmem.DevArray.from_numpy(C0)
        closure.check(closure.lib.coral_amd_nn_gemm(tA, 0, M, N, K, 0.5, Ad.ptr, A.shape[2], A[0].size, Bd.ptr, N,
                                                    B[0].size, 1.0, Cd.ptr, N, M * N, batch, work.ptr, need, None),
                      "gemm")
        outs.append(Cd.numpy())
    assert np.array_equal(outs[0], outs[1])
    got = outs[0].astype(np.float64).reshape(batch, M, N)
    Aq, Bq = C.bf16(A), C.bf16(B)
    for b in range(batch):
        ref = 0.5 * (Aq[b].T if tA else Aq[b]) @ Bq[b] + C0[b].astype(np.float32)
        assert np.linalg.norm(got[b] - ref) / np.linalg.norm(ref) < 1e-5
    Cd = devmem.DevArray.from_numpy(C0)
    rc = closure.lib.coral_amd_nn_gemm(tA, 0, M, N, K, 0.5, Ad.ptr, A.shape[2], A[0].size, Bd.ptr, N, B[0].size, 1.0,
                                       Cd.ptr, N, M * N, batch, None, 0, None)
    assert rc < 0


@pytest.mark.parametrize("W,Hd,R,spectral", [(64, 256, 1000, False), (32, 128, 64 * 7, False),
                                              (64, 256, 3 * 288, True), (32, 128, 2 * 288, True)])
def test_fused_mlp_matches_unfused_sequence(nnmod, W, Hd, R, spectral):
    """coral_amd_nn_mlp gives the bits of gemm, bias_act (GELU), gemm,
    bias_act, axpy on the same operands, its caches those of the unfused
    pre-activation and activation, also in place (Hout16 = Hin16) and with a
    ragged last tile of rows; with the spectral synthesis (planes of N2 = 288
    points, 17 modes padded to 32) the bits of the unfused H += B Cs GEMM
    (beta = 1) before the MLP."""
    closure, devmem = nnmod
    lib, check = closure.lib, closure.check
    rng = np.random.default_rng(W + Hd + R)
    H0 = rng.normal(size=(R, W)).astype(np.float32)
    W1, W2 = rng.normal(size=(W, Hd)) / 8, rng.normal(size=(Hd, W)) / 16
    b1, b2 = rng.normal(size=Hd).astype(np.float32), rng.normal(size=W).astype(np.float32)
    W1d, W2d = to_dev_bf16(closure, devmem, W1), to_dev_bf16(closure, devmem, W2)
    W1t, W2t = devmem.DevArray(W * Hd, "bf16"), devmem.DevArray(W * Hd, "bf16")
    check(lib.coral_amd_nn_transpose_bf16(W1d.ptr, W, Hd, W1t.ptr, None), "t")
    check(lib.coral_amd_nn_transpose_bf16(W2d.ptr, Hd, W, W2t.ptr, None), "t")
    b1d, b2d = devmem.DevArray.from_numpy(b1), devmem.DevArray.from_numpy(b2)
    Hin16 = to_dev_bf16(closure, devmem, H0)
    N2, nb, nbp = 288, 17, 32
    P = R // N2
    Bp16 = CsT16 = None
    if spectral:
        B = rng.normal(size=(N2, nb)) / 10
        Cs = rng.normal(size=(P, nb, W))
        BT16 = to_dev_bf16(closure, devmem, B.T.copy())
        Cs16 = to_dev_bf16(closure, devmem, Cs)
        Bpad = np.zeros((N2, nbp))
        Bpad[:, :nb] = B
        CsT = np.zeros((P, W, nbp))
        CsT[:, :, :nb] = np.transpose(Cs, (0, 2, 1))
        Bp16, CsT16 = to_dev_bf16(closure, devmem, Bpad), to_dev_bf16(closure, devmem, CsT)
    # unfused
    Hu = devmem.DevArray.from_numpy(H0)
    Z, A16, U = devmem.DevArray(R * Hd), devmem.DevArray(R * Hd, "bf16"), devmem.DevArray(R * W)
    Hu16 = devmem.DevArray(R * W, "bf16")
    if spectral:
        check(lib.coral_amd_nn_gemm(1, 0, N2, W, nb, 1.0, BT16.ptr, N2, 0, Cs16.ptr, W, nb * W, 1.0, Hu.ptr, W, N2 * W,
                                    P, None, 0, None), "g")
    check(lib.coral_amd_nn_gemm(0, 0, R, Hd, W, 1.0, Hin16.ptr, W, 0, W1d.ptr, Hd, 0, 0.0, Z.ptr, Hd, 0, 1, None, 0,
                                None), "g")
    check(lib.coral_amd_nn_bias_act(Z.ptr, b1d.ptr, Hd, R * Hd, 1, None, A16.ptr, None), "b")
    check(lib.coral_amd_nn_gemm(0, 0, R, W, Hd, 1.0, A16.ptr, Hd, 0, W2d.ptr, W, 0, 0.0, U.ptr, W, 0, 1, None, 0,
                                None), "g")
    check(lib.coral_amd_nn_bias_act(U.ptr, b2d.ptr, W, R * W, 0, None, None, None), "b")
    check(lib.coral_amd_nn_axpy(Hu.ptr, U.ptr, 1.0, R * W, Hu16.ptr, None), "a")
    sp = (Bp16.ptr, CsT16.ptr, N2, nbp) if spectral else (None, None, 0, 0)
    # fused, caches on, then in place without caches
    Hf, Hf16 = devmem.DevArray.from_numpy(H0), devmem.DevArray(R * W, "bf16")
    Zf, Af16 = devmem.DevArray(R * Hd), devmem.DevArray(R * Hd, "bf16")
    HT = devmem.DevArray(R * W, "bf16") if spectral else None
    check(lib.coral_amd_nn_mlp(R, W, Hd, Hin16.ptr, W1t.ptr, b1d.ptr, W2t.ptr, b2d.ptr, Hf.ptr, Hf16.ptr, Zf.ptr,
                               Af16.ptr, *sp, HT.ptr if spectral else None, None), "mlp")
    Hi, Hi16 = devmem.DevArray.from_numpy(H0), to_dev_bf16(closure, devmem, H0)
    check(lib.coral_amd_nn_mlp(R, W, Hd, Hi16.ptr, W1t.ptr, b1d.ptr, W2t.ptr, b2d.ptr, Hi.ptr, Hi16.ptr, None, None,
                               *sp, None, None), "mlp")
    if spectral:   # the new BF16 values in the tiled transposed layout
        tiled = Hu16.numpy().reshape(R // 32, 32, W).transpose(0, 2, 1).ravel()
        assert np.array_equal(HT.numpy(), tiled)
    assert np.array_equal(Hf.numpy(), Hu.numpy())
    assert np.array_equal(Hf16.numpy(), Hu16.numpy())
    assert np.array_equal(Zf.numpy(), Z.numpy())
    assert np.array_equal(Af16.numpy(), A16.numpy())
    assert np.array_equal(Hi.numpy(), Hu.numpy())
    assert np.array_equal(Hi16.numpy(), Hu16.numpy())
    # and the math: float64 on the BF16 operands
    ref = H0 + (C.bf16(C.gelu(C.bf16(H0) @ C.bf16(W1) + b1)) @ C.bf16(W2) + b2)
    if spectral:
        ref = ref + np.einsum("in,pnc->pic", C.bf16(B), C.bf16(Cs)).reshape(R, W)
    assert np.linalg.norm(Hu.numpy().reshape(R, W) - ref) / np.linalg.norm(ref) < 1e-3


def test_gemm_bias_matches_unfused_sequence(nnmod):
    """coral_amd_nn_gemm_bias (the lift: product, bias, BF16 copy, tiled
    transposed copy in one kernel) gives the bits of coral_amd_nn_gemm,
    _bias_act and _tile_transpose."""
    closure, devmem = nnmod
    lib, check = closure.lib, closure.check
    rng = np.random.default_rng(21)
    M, N, K = 32 * 45, 32, 112
    A = to_dev_bf16(closure, devmem, rng.normal(size=(M, K)))
    B = to_dev_bf16(closure, devmem, rng.normal(size=(K, N)) / 10)
    bias = devmem.DevArray.from_numpy(rng.normal(size=N))
    C1, C2 = devmem.DevArray(M * N), devmem.DevArray(M * N)
    h1, h2 = devmem.DevArray(M * N, "bf16"), devmem.DevArray(M * N, "bf16")
    t1, t2 = devmem.DevArray(M * N, "bf16"), devmem.DevArray(M * N, "bf16")
    check(lib.coral_amd_nn_gemm(0, 0, M, N, K, 1.0, A.ptr, K, 0, B.ptr, N, 0, 0.0, C1.ptr, N, 0, 1, None, 0, None), "g")
    check(lib.coral_amd_nn_bias_act(C1.ptr, bias.ptr, N, M * N, 0, None, h1.ptr, None), "b")
    check(lib.coral_amd_nn_tile_transpose(h1.ptr, M, N, t1.ptr, None), "t")
    check(lib.coral_amd_nn_gemm_bias(M, N, K, A.ptr, K, B.ptr, N, bias.ptr, C2.ptr, h2.ptr, t2.ptr, None), "gb")
    for a, b in ((C1, C2), (h1, h2), (t1, t2)):
        assert np.array_equal(a.numpy(), b.numpy())


@pytest.mark.parametrize("W,nb,N2,P", [(32, 81, 36864 // 16, 3), (64, 17, 9216, 2)])
def test_mode_projection_matches_numpy(nnmod, W, nb, N2, P):
    """coral_amd_nn_modes: Cm[p] = B^T H[p] per plane from the transposed
    operands (chunk partials added in a fixed order), against numpy on the same
    BF16 values to FP32 summation, and the same bits on a second call."""
    closure, devmem = nnmod
    lib, check = closure.lib, closure.check
    nbp = (nb + 31) // 32 * 32
    assert lib.coral_amd_nn_modes_supported(nbp, N2, W) == 1
    rng = np.random.default_rng(nb + W)
    B = rng.normal(size=(N2, nb)) / 30
    H = rng.normal(size=(P, N2, W))
    BT = np.zeros((nbp, N2))
    BT[:nb] = B.T
    BTd = to_dev_bf16(closure, devmem, BT)
    Hd16 = to_dev_bf16(closure, devmem, H.reshape(P * N2, W))
    HTd = devmem.DevArray(P * N2 * W, "bf16")
    check(lib.coral_amd_nn_tile_transpose(Hd16.ptr, P * N2, W, HTd.ptr, None), "tile transpose")
    assert np.array_equal(HTd.numpy(), Hd16.numpy().reshape(P * N2 // 32, 32, W).transpose(0, 2, 1).ravel())
    nwork = int(lib.coral_amd_nn_modes_work(nbp, N2, W, P))
    work = devmem.DevArray(nwork)
    outs = []
    for _ in range(2):
        Cm = devmem.DevArray(P * nb * W)
        check(lib.coral_amd_nn_modes(BTd.ptr, nbp, nb, N2, HTd.ptr, W, P, Cm.ptr, work.ptr, nwork, None), "modes")
        outs.append(Cm.numpy())
    assert np.array_equal(outs[0], outs[1])
    ref = np.einsum("in,pic->pnc", C.bf16(B), C.bf16(H))
    got = outs[0].astype(np.float64).reshape(P, nb, W)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


def test_network_forward_backward_match_reference(nnmod):
    closure, devmem = nnmod
    NX, NY, M, P = 16, 16, 3, 3
    basis = closure.fourier_basis(NX, NY, M)
    lay = closure.Layout(cin=5, cout=9, width=64, hidden=256, layers=2, nb=basis.shape[1])
    net = closure.ClosureNet(lay, basis, seed=7)
    flat = net.params()
    rng = np.random.default_rng(1)
    X = rng.normal(size=(P * NX * NY, 5))
    T = rng.normal(size=(P * NX * NY, 9)) * 0.3
    X16 = to_dev_bf16(closure, devmem, X)
    Td = devmem.DevArray.from_numpy(T)
    Y = net.forward(X16, P)
    loss = net.loss_grad(Y, Td)
    net.backward()
    g = net.g.numpy().astype(np.float64)
    pd = net.param_dict(flat)
    ref = {"W0": pd["W0"], "b0": pd["b0"], "Wp": pd["Wp"], "bp": pd["bp"],
           "W1": [pd[f"W1.{l}"] for l in range(2)], "b1": [pd[f"b1.{l}"] for l in range(2)],
           "W2": [pd[f"W2.{l}"] for l in range(2)], "b2": [pd[f"b2.{l}"] for l in range(2)],
           "D": [pd[f"D.{l}"] for l in range(2)]}
    lref, gref, Yref = C.loss_and_grads(ref, X, T, basis, P)
    Yg = Y.numpy().reshape(Yref.shape)
    print("loss", loss, lref, "Y rel", np.linalg.norm(Yg - Yref) / np.linalg.norm(Yref), flush=True)
    assert loss == pytest.approx(lref, rel=2e-3)
    assert np.linalg.norm(Yg - Yref) / np.linalg.norm(Yref) < 5e-3
    gd = net.param_dict(g)
    worst = {}
    for name, (off, shp) in net.off.items():
        base, _, l = name.partition(".")
        r = gref[base][int(l)] if l else gref[base]
        e = np.linalg.norm(gd[name] - r) / max(np.linalg.norm(r), 1e-30)
        worst[name] = e
        assert e < 2e-2, (name, e)
    print("gradient rel L2 per block:", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)


def test_training_lowers_the_loss(nnmod):
    """A synthetic regression whose target couples each point to the plane's
    low modes (so both branches are needed): 300 Adam steps lower the MSE by
    more than 10x."""
    closure, devmem = nnmod
    NX, NY, M, P = 24, 24, 3, 8
    basis = closure.fourier_basis(NX, NY, M)
    lay = closure.Layout(cin=5, cout=9, width=64, hidden=256, layers=4, nb=basis.shape[1])
    net = closure.ClosureNet(lay, basis, seed=3)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(P, NX * NY, 5))
    Wt = rng.normal(size=(5, 9)) * 0.5
    low = np.einsum("in,pnc->pic", basis, np.einsum("in,pic->pnc", basis, X))   # low-mode part per plane
    T = np.tanh(X @ Wt) + 0.5 * (low @ Wt)
    X16 = to_dev_bf16(closure, devmem, X.reshape(-1, 5))
    Td = devmem.DevArray.from_numpy(T.reshape(-1, 9))
    losses = [net.train_step(X16, Td, P, 2e-3) for _ in range(300)]
    print("loss first/last:", losses[0], losses[-1], flush=True)
    assert np.isfinite(losses[-1]) and losses[-1] < losses[0] / 10


@pytest.mark.parametrize("NX", [32, 64], ids=["x48", "x96"])
def test_step_closure_matches_network(nnmod, NX):
    """coral_amd_set_closure: the correction the step adds to the quadratic
    variables (coral_amd_products with the closure minus without) equals the
    forward transform of the trainer's network applied to the same physical
    full variables (ClosureNet.forward, then coral_amd_fwd_transform), within
    FP64 rounding of the transforms (both paths run the same BF16 network
    kernels on the same BF16 inputs).  At 96-point x lines the step's own x
    transforms take two variables per complex sequence (four inputs, nine
    corrections) while coral_amd_physical and coral_amd_fwd_transform take
    one, so the paired kernels are checked against the single ones (which
    test_transforms_match_oracle checks against the oracle)."""
    closure, devmem = nnmod
    from conftest import product_recipe, seeded_state
    from coral_amd.diagnostics import gauss_chebyshev_grid
    from coral_amd.solver import Solver
    r = product_recipe("coral.equations.rbc_ra1e5")
    NZ, M = 17, 3
    NXAA, NZAA = 3 * NX // 2, 3 * NZ // 2
    basis = closure.fourier_basis(NXAA, NXAA, M)
    cin = closure.input_width(4)
    lay = closure.Layout(cin=cin, cout=9, width=64, hidden=256, layers=4, nb=basis.shape[1])
    rng = np.random.default_rng(4)
    model = closure.Model(lay, M, closure.init_params(lay, 11) * 3.0, rng.normal(size=cin) * 0.1,
                          1.0 + rng.random(cin), rng.normal(size=9) * 0.1, 0.5 + rng.random(9))
    k0, z0 = seeded_state(r, NX, NX, NZ, 6, 6, 9, amps=(("u", 1.0), ("v", 1.0), ("w", 1.0), ("t", 0.1)))
    with Solver(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_state(k0, z0)
        q0 = g.products()
        fields = np.stack([g.physical(i) for i in range(4)])
        g.set_closure(model)
        q1 = g.products()
        # the network on the host side's copy of the same fields
        rows = closure.plane_rows(fields, gauss_chebyshev_grid(NZAA, 0.5, 1.0))
        net = closure.ClosureNet(lay, basis, model.params)
        X16 = to_dev_bf16(closure, devmem, closure.normalise(rows, model.in_mean, model.in_std))
        Y = net.forward(X16, NZAA).numpy().reshape(NZAA, NXAA, NXAA, 9).astype(np.float64)
        tau = Y * model.out_std.astype(np.float32) + model.out_mean.astype(np.float32)
        worst = 0.0
        for q in range(9):
            tq = np.transpose(tau[..., q], (2, 1, 0))            # [x][y][z]
            spec = g.fwd_transform(tq)
            d = q1[q] - q0[q]
            worst = max(worst, float(np.linalg.norm(d - spec) / np.linalg.norm(spec)))
        g.set_closure(None)
        q2 = g.products()
    print("closure correction vs network + transform, worst rel L2:", worst, flush=True)
    assert worst <= 1e-10
    assert np.array_equal(q2, q0)


def test_closure_learns_filtered_subgrid_terms(nnmod):
    """End to end at small size: Rayleigh-Benard at Ra = 1e6, Pr = 1 on
    64^2 x 33 spun up from noise (adaptive dt), 48 snapshots top-hat filtered
    onto 32^2 x 17; the network trained on the first 40 predicts the subgrid
    part of the quadratic variables of the last 8 with a relative L2 error
    well below 1, the error of no closure (a model that learned nothing, or
    inputs and targets paired wrongly, stays near 1)."""
    closure, devmem = nnmod
    import math
    from coral_amd import ic
    from coral_amd.diagnostics import gauss_chebyshev_grid
    from coral_amd.ensemble import rbc_recipe
    from coral_amd.solver import Solver
    from coral_amd.timekeeping import Timings
    r = rbc_recipe(1e6, 1.0)
    L = 2 * math.pi
    (NXf, NZf), (NXc, NZc) = (64, 33), (32, 17)
    tk = Timings(1e-3, 0.3, 0.2)

    def advance(g, n):
        for _ in range(n):
            tk.check_for_update()
            if tk.refactorization_is_due:
                g.factorize_operators(tk.dt_current)
                tk.refactorization_is_due = False
            g.march_forward(tk.dt_current, 1)
            tk.apply_cfl(g.cfl_dt())
            tk.increment()
    with Solver(r, NXf, NXf, NZf, L, L) as gf, Solver(r, NXc, NXc, NZc, L, L) as gc:
        gf.use_graphs(True)
        gf.set_kx0_projection(True)
        k0, z0 = ic.noisy_initial_state(r, NXf, NXf, NZf, L, L, 1e-2, 1e-2, 5)
        gf.set_state(k0, z0)
        advance(gf, 6000)
        print(f"spun up to t = {tk.absolute_time:.4f}", flush=True)
        box = closure.BoxFilter(r, (NXf, NXf, NZf), (NXc, NXc, NZc))
        zg = gauss_chebyshev_grid(3 * NZc // 2, 0.5, 1.0)
        X, T = [], []
        for _ in range(48):
            advance(gf, 20)
            xi, ti = closure.subgrid_sample(gf, gc, box, zg)
            X.append(xi)
            T.append(ti)
    assert all(np.isfinite(x).all() for x in X + T)
    N2 = (3 * NXc // 2) ** 2
    basis = closure.fourier_basis(3 * NXc // 2, 3 * NXc // 2, 4)
    model, hist = closure.train(np.concatenate(X[:40]), np.concatenate(T[:40]), N2, basis, 4, width=32, hidden=128,
                                epochs=10, planes_per_batch=16, lr=2e-3)
    Tva = np.concatenate(T[40:])
    pred = closure.predict(model, np.concatenate(X[40:]), N2, basis)
    err = np.linalg.norm(pred - Tva, axis=0) / np.linalg.norm(Tva, axis=0)
    overall = float(np.linalg.norm(pred - Tva) / np.linalg.norm(Tva))
    print("training loss by epoch:", np.round(hist, 4), "held-out error:", overall, np.round(err, 3), flush=True)
    assert hist[-1] < hist[0]
    assert overall < 0.85


@pytest.mark.parametrize("width,hidden", [(32, 128), (64, 256)])
def test_fused_lift_and_projection_match_the_pack_gemm_sequence(nnmod, monkeypatch, width, hidden):
    """On grids whose x-lines split into 64-point tiles (here 128^2 x 17, a
    192-point padded line) the closure's lift runs fused with the packing of
    its inputs (k_closure_lift) and its projection with the unpacking
    (k_closure_project), and the MLP layers on the persistent pipelined kernel
    (k_mlp_pipe).  The corrected quadratic spectra equal those of the pack,
    GEMM, bias and unpack kernels and the one-tile MLP kernel
    (CORAL_AMD_CLOSURE_UNFUSED=1, CORAL_AMD_MLP_UNPIPELINED=1) bit for bit,
    and the network's own forward pass on the host's rows (within FP64
    rounding of the transforms)."""
    closure, devmem = nnmod
    from conftest import product_recipe, seeded_state
    from coral_amd.diagnostics import gauss_chebyshev_grid
    from coral_amd.solver import Solver
    r = product_recipe("coral.equations.rbc_ra1e5")
    NX, NZ, M = 128, 17, 4
    NXAA, NZAA = 3 * NX // 2, 3 * NZ // 2
    basis = closure.fourier_basis(NXAA, NXAA, M)
    cin = closure.input_width(4)
    lay = closure.Layout(cin=cin, cout=9, width=width, hidden=hidden, layers=4, nb=basis.shape[1])
    rng = np.random.default_rng(5)
    model = closure.Model(lay, M, closure.init_params(lay, 12) * 3.0, rng.normal(size=cin) * 0.1,
                          1.0 + rng.random(cin), rng.normal(size=9) * 0.1, 0.5 + rng.random(9))
    k0, z0 = seeded_state(r, NX, NX, NZ, 6, 6, 9, amps=(("u", 1.0), ("v", 1.0), ("w", 1.0), ("t", 0.1)))
    out = {}
    for mode in ("fused", "unfused"):
        if mode == "unfused":
            monkeypatch.setenv("CORAL_AMD_CLOSURE_UNFUSED", "1")
            monkeypatch.setenv("CORAL_AMD_MLP_UNPIPELINED", "1")
        with Solver(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi) as g:
            g.set_state(k0, z0)
            if mode == "fused":
                q0 = g.products()
                fields = np.stack([g.physical(i) for i in range(4)])
            g.set_closure(model)
            out[mode] = g.products()
            if mode == "fused":
                rows = closure.plane_rows(fields, gauss_chebyshev_grid(NZAA, 0.5, 1.0))
                net = closure.ClosureNet(lay, basis, model.params)
                X16 = to_dev_bf16(closure, devmem, closure.normalise(rows, model.in_mean, model.in_std))
                Y = net.forward(X16, NZAA).numpy().reshape(NZAA, NXAA, NXAA, 9).astype(np.float64)
                tau = Y * model.out_std.astype(np.float32) + model.out_mean.astype(np.float32)
                worst = 0.0
                for q in range(9):
                    spec = g.fwd_transform(np.transpose(tau[..., q], (2, 1, 0)))
                    d = out[mode][q] - q0[q]
                    worst = max(worst, float(np.linalg.norm(d - spec) / np.linalg.norm(spec)))
    monkeypatch.delenv("CORAL_AMD_CLOSURE_UNFUSED")
    monkeypatch.delenv("CORAL_AMD_MLP_UNPIPELINED")
    print("fused closure vs network + transform, worst rel L2:", worst, flush=True)
    assert worst <= 1e-10
    assert np.array_equal(out["fused"], out["unfused"])


def test_closure_steps_as_graph_replays_match_eager_steps(nnmod):
    """Steps with the closure (fused lift and projection, pipelined MLP,
    LDS-staged mode projections at 128^2 x 17) replayed as HIP graphs equal
    the same steps launched one by one, bit for bit, replay after replay."""
    closure, devmem = nnmod
    from conftest import product_recipe, seeded_state
    from coral_amd.solver import Solver
    r = product_recipe("coral.equations.rbc_ra1e5")
    NX, NZ, M = 128, 17, 4
    NXAA = 3 * NX // 2
    basis = closure.fourier_basis(NXAA, NXAA, M)
    cin = closure.input_width(4)
    lay = closure.Layout(cin=cin, cout=9, width=32, hidden=128, layers=4, nb=basis.shape[1])
    rng = np.random.default_rng(6)
    model = closure.Model(lay, M, closure.init_params(lay, 13) * 0.3, rng.normal(size=cin) * 0.1,
                          1.0 + rng.random(cin), rng.normal(size=9) * 0.01, 0.05 + 0.1 * rng.random(9))
    k0, z0 = seeded_state(r, NX, NX, NZ, 6, 6, 9, amps=(("u", 1e-2), ("v", 1e-2), ("w", 1e-2), ("t", 1e-3)))
    with Solver(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_closure(model)
        g.set_state(k0, z0)
        eager = []
        for _ in range(4):
            g.march_forward(1e-5)
            eager.append(g.get_state())
        g.set_state(k0, z0)
        g.use_graphs(True)
        for n in range(4):
            g.march_forward(1e-5)
            k, z = g.get_state()
            for a, b in zip(k + z, eager[n][0] + eager[n][1]):
                assert np.array_equal(a, b), f"replay {n}"
        assert all(np.isfinite(a).all() for a in eager[-1][0])
