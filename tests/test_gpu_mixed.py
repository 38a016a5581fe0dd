"""Mixed precision (coral_amd_set_precision): FP32 transforms and products,
with the z products on FP32 MFMA with compensated accumulation; FP64 NL sums,
right-hand sides, banded factors and solves, and FP64 residual refinement.

Tolerances: the FP32 chain rounds the Chebyshev coefficients to FP32 (relative
6e-8) and runs three transforms each way in FP32 (lengths <= 768; error grows
like log N per FFT stage); the quadratic spectra are therefore compared with
the FP64 oracle at relative L2 <= 2e-6.  The state after a few steps moves by
dt times the NL terms, so its error against FP64 is smaller still (<= 1e-7
relative here).  FP64 with refinement stays within 1e-12 of FP64 without."""
import numpy as np
import pytest

from conftest import product_recipe
from coral_amd import ic
from oracle import coral_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver_mod(native_lib):
    from coral_amd import solver
    return solver


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


PRODUCT_CASES = [
    ("coral.equations.rbc_ra1e5", 64, 64, 17, 2 * np.pi, 2 * np.pi),
    ("coral.equations.rrbc_ra5e7_ek1e-4", 128, 96, 33, 2 * np.pi, 3.0),
    ("coral.equations.rrbc_ra5e7_ek1e-4", 64, 64, 129, 2 * np.pi, 2 * np.pi),
]


@pytest.mark.parametrize("case", PRODUCT_CASES, ids=[f"{c[1]}x{c[2]}x{c[3]}" for c in PRODUCT_CASES])
def test_mixed_products_match_oracle(solver_mod, case):
    """The quadratic variables' spectra of the FP32 chain (broadband state,
    every kept mode seeded) against the FP64 oracle."""
    name, NX, NY, NZ, Lx, Ly = case
    print("case", case, flush=True)
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 10 * max(NX, NY), 8, 11)
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    qref = O.OracleSolver(r, geo, "ARS_222").quad_spectra([a.copy() for a in k0], [a.copy() for a in z0])
    with solver_mod.Solver(r, NX, NY, NZ, Lx, Ly) as g:
        g.set_state(k0, z0)
        g.set_precision("mixed")
        q = g.products()
        g.set_precision("fp64")
        q64 = g.products()
    wn = np.zeros(q.shape[2], bool)
    wn[: NY // 2] = True
    wn[q.shape[2] - (NY // 2 - 1):] = True
    worst = max(rel(q[i][:, wn], qref[i][:, wn]) for i in range(q.shape[0]))
    worst64 = max(rel(q64[i][:, wn], qref[i][:, wn]) for i in range(q.shape[0]))
    print("FP32 chain worst rel L2", worst, "FP64", worst64, flush=True)
    assert worst <= 2e-6
    assert worst > 10 * worst64          # the FP32 chain really ran
    assert worst64 <= 1e-10


def test_mixed_steps_close_to_oracle(solver_mod):
    """Five ARS_222 steps of the 64^2 x 33 case in mixed precision: the state
    stays within 1e-7 (relative L2) of the FP64 oracle; with one refinement
    sweep per step the same."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NZ, dt, L = 64, 33, 2e-5, 2 * np.pi
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-2, 10 * NX, 8, 11)
    ora = O.OracleSolver(r, O.Geometry(NX, NX, NZ, L, L), "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    ora.run(dt, 5)
    errs = []
    for refine in (0, 1):
        with solver_mod.Solver(r, NX, NX, NZ, L, L) as g:
            g.set_state(k0, z0)
            g.set_precision("mixed", refine)
            g.march_forward(dt, 5)
            k, z = g.get_state()
        errs.append((rel(k[0], ora.field[0]), max(rel(a, b) for a, b in zip(z, ora.zfield))))
    print("mixed 5 steps rel L2 (kxky, mean) without / with refinement:", errs, flush=True)
    for ek, ez in errs:
        assert ek <= 1e-7 and ez <= 1e-7
    assert errs[0][0] > 1e-13


def test_refinement_keeps_fp64_result(solver_mod):
    """FP64 steps with a refinement sweep after every stage: the correction is
    at rounding level, the state agrees with plain FP64 steps to 1e-12 and
    with the oracle to 1e-10."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NZ, dt, L = 32, 25, 2e-5, 2 * np.pi
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-2, 6, 6, 3)
    out = []
    for refine in (0, 2):
        with solver_mod.Solver(r, NX, NX, NZ, L, L) as g:
            g.set_state(k0, z0)
            g.set_precision("fp64", refine)
            g.march_forward(dt, 4)
            out.append(g.get_state())
    ora = O.OracleSolver(r, O.Geometry(NX, NX, NZ, L, L), "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    ora.run(dt, 4)
    (ka, za), (kb, zb) = out
    assert rel(kb[0], ka[0]) <= 1e-12
    assert rel(kb[0], ora.field[0]) <= 1e-10
    for a, b in zip(zb, za):
        assert rel(a, b) <= 1e-12


def test_mixed_graph_replay_matches_eager(solver_mod):
    """Mixed-precision steps (with refinement) replayed as HIP graphs equal the
    eager mixed steps bit for bit."""
    r = product_recipe("coral.equations.rbc_ra1e5")
    NX, NZ, dt, L = 64, 33, 1e-4, 2 * np.pi
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-2, 10 * NX, 8, 5)
    res = []
    for graphs in (False, True):
        with solver_mod.Solver(r, NX, NX, NZ, L, L) as g:
            g.set_precision("mixed", 1)
            if graphs:
                g.use_graphs(True)
            g.set_state(k0, z0)
            g.march_forward(dt, 3)
            res.append(g.get_state())
    for a, b in zip(res[0][0] + res[0][1], res[1][0] + res[1][1]):
        assert np.array_equal(a, b)


def test_mixed_at_production_size_tracks_fp64(solver_mod):
    """256^2 x 129 rotating workload, three steps from the stated amplitude
    1e-2: the mixed-precision state against the FP64 GPU state (itself within
    1e-10 of the oracle, test_gpu_parity) stays within 1e-7."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NZ, dt, L = 256, 129, 2e-5, 2 * np.pi
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-2, 32, 8, 2)
    out = []
    for mode in ("fp64", "mixed"):
        with solver_mod.Solver(r, NX, NX, NZ, L, L) as g:
            g.set_precision(mode, 1 if mode == "mixed" else 0)
            g.set_state(k0, z0)
            g.march_forward(dt, 3)
            out.append(g.get_state())
    e = rel(out[1][0][0], out[0][0][0])
    print("256^2 x 129 mixed vs FP64, 3 steps:", e, flush=True)
    assert 0 < e <= 1e-7
