"""ORACLE -- test infrastructure only.

CPU restatement (numpy, float64 arithmetic) of the subgrid closure network of
``coral_amd/closure.py``, used by the tests to check the HIP forward and
backward kernels.  The reference Coral has no learned closure, so this network
has no reference counterpart; it is pinned by finite differences of its own
loss (tests/test_closure.py) and the GPU is checked against it.

BF16 rounding is applied where the GPU rounds: every GEMM operand (weights,
activations, gradients) is rounded to BF16 (round to nearest even) before the
product; products, sums, biases and the optimizer run in higher precision
(float64 here, FP32 on the GPU).

Network (per horizontal plane of P planes, N2 points each, rows r = p N2 + i):
    H0 = X W0 + b0
    layer l = 1..L:  A = gelu(H W1_l + b1_l);  U = A W2_l + b2_l
                     S = B ((B^T H) * D_l)      (per plane; B: N2 x nb basis)
                     H <- H + U + S
    Y = H Wp + bp;   loss = sum (Y - T)^2 / n
"""
from __future__ import annotations

import numpy as np

GELU_C = 0.7978845608028654


def bf16(x):
    """Round to BF16 (nearest even) and return as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where((u & 0x7F800000) == 0x7F800000, u & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * z * z)


def forward(params, X, basis, P, round_bf16=True):
    """params: dict W0, b0, W1[l], b1[l], W2[l], b2[l], D[l], Wp, bp (float64);
    X: (P*N2, Cin); basis: (N2, nb).  Returns (Y, cache)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    N2, nb = basis.shape
    Bq = r(basis)
    cache = {"X": r(X), "layers": []}
    H = cache["X"] @ r(params["W0"]) + params["b0"]
    for l in range(len(params["W1"])):
        Hq = r(H)
        Z = Hq @ r(params["W1"][l]) + params["b1"][l]
        A = gelu(Z)
        Aq = r(A)
        U = Aq @ r(params["W2"][l]) + params["b2"][l]
        Cm = np.einsum("in,pic->pnc", Bq, Hq.reshape(P, N2, -1))
        Cs = Cm * params["D"][l][None]
        S = np.einsum("in,pnc->pic", Bq, r(Cs)).reshape(P * N2, -1)
        cache["layers"].append({"Hq": Hq, "Z": Z, "Aq": Aq, "Cm": Cm})
        H = H + U + S
    Hq = r(H)
    cache["Hq_out"] = Hq
    Y = Hq @ r(params["Wp"]) + params["bp"]
    return Y, cache


def loss_and_grads(params, X, T, basis, P, round_bf16=True):
    """MSE loss sum (Y - T)^2 / n and its gradients (same keys as params)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    N2, nb = basis.shape
    Bq = r(basis)
    Y, cache = forward(params, X, basis, P, round_bf16)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    dY = r(2.0 * (Y - T) / n)
    g = {"W1": [None] * len(params["W1"]), "b1": [None] * len(params["W1"]), "W2": [None] * len(params["W1"]),
         "b2": [None] * len(params["W1"]), "D": [None] * len(params["W1"])}
    g["Wp"] = cache["Hq_out"].T @ dY
    g["bp"] = dY.sum(axis=0)
    dH = dY @ r(params["Wp"]).T
    for l in reversed(range(len(params["W1"]))):
        c = cache["layers"][l]
        dHq = r(dH)
        # U branch
        g["W2"][l] = c["Aq"].T @ dHq
        g["b2"][l] = dH.sum(axis=0)
        dA = dHq @ r(params["W2"][l]).T
        dZ = dA * gelu_grad(c["Z"])
        dZq = r(dZ)
        g["W1"][l] = c["Hq"].T @ dZq
        g["b1"][l] = dZ.sum(axis=0)
        # spectral branch
        dCs = np.einsum("in,pic->pnc", Bq, dHq.reshape(P, N2, -1))
        g["D"][l] = np.sum(c["Cm"] * dCs, axis=0)
        dCm = r(dCs * params["D"][l][None])
        dH = dH + dZq @ r(params["W1"][l]).T + np.einsum("in,pnc->pic", Bq, dCm).reshape(P * N2, -1)
    dHq = r(dH)
    g["W0"] = cache["X"].T @ dHq
    g["b0"] = dH.sum(axis=0)
    return loss, g, Y
