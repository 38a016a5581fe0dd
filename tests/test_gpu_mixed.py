"""Mixed precision (coral_amd_set_precision): FP32 transforms and products,
with the z products on FP32 MFMA with compensated accumulation; FP64 NL sums,
right-hand sides, banded factors and solves, and FP64 residual refinement.

The yardstick of the FP32 chain is the oracle's own transform chain run in
single precision (scipy's FP32 FFTs and DCTs, products rounded to FP32,
``OracleSolver.transform_dtype = np.float32``): against the FP64 oracle the
GPU's FP32 chain must be no worse than twice that error.  The states seed the
velocities at unit amplitude and the temperature a thousand times smaller, as
in convection (velocities in thermal-diffusion units are large), so the
packed pairs of the x pass (w beside t, velocity products beside
temperature products) are far from balanced.  FP64 with refinement stays
within 1e-12 of FP64 without."""
import numpy as np
import pytest

from conftest import product_recipe, seeded_state
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
    # 768-point x lines (two wavefronts per line in the x pass)
    ("coral.equations.rrbc_ra5e7_ek1e-4", 512, 512, 17, 2 * np.pi, 2 * np.pi),
]


@pytest.mark.parametrize("case", PRODUCT_CASES, ids=[f"{c[1]}x{c[2]}x{c[3]}" for c in PRODUCT_CASES])
def test_mixed_products_match_oracle(solver_mod, case):
    """The quadratic variables' spectra of the FP32 chain (every kept mode of
    u, v, w, t seeded) against the FP64 oracle, per variable, within twice the
    error of the oracle's FP32 chain."""
    name, NX, NY, NZ, Lx, Ly = case
    print("case", case, flush=True)
    r = product_recipe(name)
    k0, z0 = seeded_state(r, NX, NY, NZ, 10 * max(NX, NY), 8, 11)
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    ora = O.OracleSolver(r, geo, "ARS_222")
    qref = ora.quad_spectra([a.copy() for a in k0], [a.copy() for a in z0])
    ora.transform_dtype = np.float32
    q32 = ora.quad_spectra([a.copy() for a in k0], [a.copy() for a in z0])
    with solver_mod.Solver(r, NX, NY, NZ, Lx, Ly) as g:
        g.set_state(k0, z0)
        g.set_precision("mixed")
        q = g.products()
        g.set_precision("fp64")
        q64 = g.products()
    wn = np.zeros(q.shape[2], bool)
    wn[: NY // 2] = True
    wn[q.shape[2] - (NY // 2 - 1):] = True
    for i in range(q.shape[0]):
        assert np.linalg.norm(qref[i][:, wn]) > 0
        e, e32, e64 = (rel(x[i][:, wn], qref[i][:, wn]) for x in (q, q32, q64))
        print(f"quad {i}: FP32 chain {e:.3e}, oracle FP32 {e32:.3e}, FP64 {e64:.3e}", flush=True)
        assert e <= 2.0 * e32
        assert e > 10 * e64          # the FP32 chain really ran
        assert e64 <= 1e-10


STEP_CASES = [
    # recipe, NX, NZ, dt, amplitude of u, v, w and t, steps
    ("coral.equations.rbc_ra1e5", 64, 33, 1e-4, 1e-2, 8),
    ("coral.equations.rrbc_ra5e7_ek1e-4", 64, 33, 2e-5, 1e-4, 6),
]


@pytest.mark.parametrize("case", STEP_CASES, ids=[c[0].split(".")[-1] for c in STEP_CASES])
def test_mixed_steps_close_to_oracle(solver_mod, case):
    """Mixed-precision steps against the FP64 oracle: the state's error is at
    most twice that of the oracle stepped with its FP32 transform chain, with
    and without a refinement sweep per step.  The cases stay well inside their
    CFL limit (the oracle's CFL time step is over 100 times dt), where the
    FP32 error of the NL terms accumulates linearly; past that limit any
    rounding difference is amplified by the integration itself."""
    name, NX, NZ, dt, amp, nsteps = case
    r = product_recipe(name)
    L = 2 * np.pi
    k0, z0 = seeded_state(r, NX, NX, NZ, 10 * NX, 8, 11, amps=(("u", amp), ("v", amp), ("w", amp), ("t", amp)))
    ref = {}
    for dtype in (np.float64, np.float32):
        ora = O.OracleSolver(r, O.Geometry(NX, NX, NZ, L, L), "ARS_222")
        ora.transform_dtype = dtype
        ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
        ora.run(dt, nsteps)
        ref[dtype] = (ora.field, ora.zfield, ora.cfl_dt)
    kref, zref, cfl = ref[np.float64]
    assert cfl > 100 * dt
    e32 = (rel(ref[np.float32][0][0], kref[0]), max(rel(a, b) for a, b in zip(ref[np.float32][1], zref)))
    errs = []
    for refine in (0, 1):
        with solver_mod.Solver(r, NX, NX, NZ, L, L) as g:
            g.set_state(k0, z0)
            g.set_precision("mixed", refine)
            g.march_forward(dt, nsteps)
            k, z = g.get_state()
        errs.append((rel(k[0], kref[0]), max(rel(a, b) for a, b in zip(z, zref))))
    print("oracle FP32 chain (kxky, mean):", e32, flush=True)
    print(f"mixed {nsteps} steps rel L2 (kxky, mean) without / with refinement:", errs, flush=True)
    for ek, ez in errs:
        assert ek <= 2.0 * e32[0] and ez <= 2.0 * e32[1]
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
