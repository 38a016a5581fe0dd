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
                                                ldb, B[0].size, 1.0, Cd.ptr, N, M * N, batch, None), "gemm")
    got = Cd.numpy().astype(np.float64).reshape(batch, M, N)
    Aq, Bq = C.bf16(A), C.bf16(B)
    for b in range(batch):
        opA = Aq[b].T if tA else Aq[b]
        opB = Bq[b].T if tB else Bq[b]
        ref = 0.5 * opA @ opB + C0[b].astype(np.float32)
        assert np.linalg.norm(got[b] - ref) / np.linalg.norm(ref) < 1e-5


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
