"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs.  Tolerance of the spectral state after N steps: relative L2
<= 1e-10 (FP64 path).  Integer indexing and masks are compared bit for bit."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg.lapack as lap

from conftest import GOLDEN, product_recipe, recipe, seeded_state
from coral_amd import ic
from oracle import coral_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def solver_mod(native_lib):
    from coral_amd import solver
    return solver


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


CASES = [
    # name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps
    ("coral.equations.rbc_ra1e5", 16, 16, 17, 2 * np.pi, 2 * np.pi, "ARS_222", 1e-4, 6),
    ("coral.equations.rbc_ra1e5", 32, 48, 17, 2.0, 3.0, "ARS_222", 1e-4, 5),
    ("coral.equations.rrbc_ra5e7_ek1e-4", 32, 32, 25, 2 * np.pi, 2 * np.pi, "ARS_222", 2e-5, 5),
    ("coral.equations.rbc_ra1e5", 16, 24, 17, 2 * np.pi, 2 * np.pi, "ARS_111", 1e-4, 4),
    ("coral.equations.rbc_ra1e5", 16, 16, 17, 2 * np.pi, 2 * np.pi, "ARS_443", 1e-4, 3),
    ("coral.equations.rrbc_ra5e7_ek1e-4", 16, 16, 21, 1.5, 1.5, "ARS_343", 1e-5, 3),
    ("coral.equations.rbc_ra1e5", 16, 16, 17, 2 * np.pi, 2 * np.pi, "ARS_232", 1e-4, 3),
    ("coral.equations.rbc_ra1e5", 16, 16, 17, 2 * np.pi, 2 * np.pi, "ARS_233", 1e-4, 3),
    ("coral.equations.rbc_ra1e5", 16, 16, 17, 2 * np.pi, 2 * np.pi, "ARS_122", 1e-4, 3),
    # vertical resolutions of the production grids (z transform tile shapes)
    ("coral.equations.rrbc_ra5e7_ek1e-4", 16, 16, 129, 2 * np.pi, 2 * np.pi, "ARS_222", 2e-5, 2),
    ("coral.equations.rbc_ra1e5", 32, 32, 65, 2 * np.pi, 2 * np.pi, "ARS_222", 1e-4, 2),
    ("coral.equations.rbc_ra1e5", 16, 32, 49, 2.0, 2 * np.pi, "ARS_343", 1e-4, 2),
    ("coral.equations.rrbc_ra5e7_ek1e-4", 16, 16, 257, 2 * np.pi, 2 * np.pi, "ARS_222", 5e-6, 2),
    # the case of Coral's published timing (etc/benchmarks/02_RRBC_julien_JFM96:
    # 96x96x48 modes, 144-point horizontal lines, box 2x2x1, Ra=2.81e5, 1/Ek=707.1)
    ("coral.equations.rrbc_ra2.81e5_iek707", 96, 96, 48, 2.0, 2.0, "ARS_222", 1e-3, 2),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0].split('.')[-1]}-{c[1]}x{c[2]}x{c[3]}-{c[6]}" for c in CASES])
def test_step_parity(solver_mod, case):
    name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps = case
    print("case", case, flush=True)
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 6, 6, 3)
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    ora = O.OracleSolver(r, geo, scheme)
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    g = solver_mod.Solver(r, NX, NY, NZ, Lx, Ly, scheme=scheme)
    try:
        g.set_state(k0, z0)
        assert g.factor_is_real(0)
        for it in range(nsteps):
            ora.step(dt)
            g.march_forward(dt)
        k, z = g.get_state()
        assert rel(k[0], ora.field[0]) <= TOL
        for a, b in zip(z, ora.zfield):
            assert rel(a, b) <= TOL
        assert abs(g.cfl_dt() / ora.cfl_dt - 1.0) <= 1e-10
        # physical full variables
        phys = ora.physical_full_variables()
        for f in range(len(phys)):
            assert rel(g.physical(f), phys[f]) <= TOL
    finally:
        g.close()


def test_complex_factor_path_parity(solver_mod, monkeypatch):
    """The complex-arithmetic factor/solve (used for recipes without a real
    form) on a recipe that has one, forced through CORAL_AMD_COMPLEX_FACTORS."""
    name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps = CASES[2]
    monkeypatch.setenv("CORAL_AMD_COMPLEX_FACTORS", "1")
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 6, 6, 3)
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    ora = O.OracleSolver(r, geo, scheme)
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    g = solver_mod.Solver(r, NX, NY, NZ, Lx, Ly, scheme=scheme)
    try:
        assert not g.factor_is_real(0)
        g.set_state(k0, z0)
        for it in range(nsteps):
            ora.step(dt)
            g.march_forward(dt)
        k, z = g.get_state()
        assert rel(k[0], ora.field[0]) <= TOL
        for a, b in zip(z, ora.zfield):
            assert rel(a, b) <= TOL
    finally:
        g.close()


def test_plain_gemm_z_transform_parity(solver_mod, monkeypatch):
    """The vector-FMA z transforms (used when no MFMA variant covers NZ),
    forced through CORAL_AMD_NO_MFMA."""
    name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps = CASES[1]
    monkeypatch.setenv("CORAL_AMD_NO_MFMA", "1")
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 6, 6, 3)
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    ora = O.OracleSolver(r, geo, scheme)
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    g = solver_mod.Solver(r, NX, NY, NZ, Lx, Ly, scheme=scheme)
    try:
        g.set_state(k0, z0)
        for it in range(nsteps):
            ora.step(dt)
            g.march_forward(dt)
        k, z = g.get_state()
        assert rel(k[0], ora.field[0]) <= TOL
        for a, b in zip(z, ora.zfield):
            assert rel(a, b) <= TOL
    finally:
        g.close()


BROADBAND = [
    # name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps: every kept mode of t seeded
    ("coral.equations.rrbc_ra5e7_ek1e-4", 256, 256, 17, 2 * np.pi, 2 * np.pi, "ARS_222", 2e-5, 2),
    ("coral.equations.rbc_ra1e5", 128, 96, 17, 2 * np.pi, 3.0, "ARS_222", 1e-4, 2),
    ("coral.equations.rbc_ra1e5", 64, 256, 17, 2.0, 2 * np.pi, "ARS_443", 1e-4, 1),
    # horizontal transform lengths of the 512x512 grid (768-point lines)
    ("coral.equations.rrbc_ra5e7_ek1e-4", 512, 512, 17, 2 * np.pi, 2 * np.pi, "ARS_222", 5e-6, 1),
]


@pytest.mark.parametrize("case", BROADBAND, ids=[f"{c[0].split('.')[-1]}-{c[1]}x{c[2]}x{c[3]}" for c in BROADBAND])
def test_broadband_step_parity(solver_mod, case):
    """All kept horizontal modes populated, at the transform lengths of the
    production grids (NXAA, NYAA in 96 .. 384)."""
    name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps = case
    print("case", case, flush=True)
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 10 * max(NX, NY), 8, 11)
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    ora = O.OracleSolver(r, geo, scheme)
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    g = solver_mod.Solver(r, NX, NY, NZ, Lx, Ly, scheme=scheme)
    try:
        g.set_state(k0, z0)
        for it in range(nsteps):
            ora.step(dt)
            g.march_forward(dt)
        k, z = g.get_state()
        assert rel(k[0], ora.field[0]) <= TOL
        for a, b in zip(z, ora.zfield):
            assert rel(a, b) <= TOL
        assert abs(g.cfl_dt() / ora.cfl_dt - 1.0) <= 1e-10
    finally:
        g.close()


def test_config1_geometry_parity(solver_mod):
    """64^2 x 33 Fourier-Chebyshev, Ra = 1e5, dt = 1e-4 (the small CPU-runnable case), 10 steps."""
    r = product_recipe("coral.equations.rbc_ra1e5")
    NX, NZ, dt = 64, 33, 1e-4
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-3, 8, 8, 1)
    geo = O.Geometry(NX, NX, NZ, 2 * np.pi, 2 * np.pi)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    with solver_mod.Solver(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_state(k0, z0)
        for it in range(10):
            ora.step(dt)
            g.march_forward(dt)
            if it in (0, 9):
                k, z = g.get_state()
                print("step", it, rel(k[0], ora.field[0]), flush=True)
                assert rel(k[0], ora.field[0]) <= TOL
                assert all(rel(a, b) <= TOL for a, b in zip(z, ora.zfield))


def test_heat_equation_table_on_gpu(solver_mod, reference_numbers):
    """The reference's heat-equation error table (error_report.py:4-15), through the GPU mean systems."""
    r = recipe("heat_equation.coral.equations")
    setup = reference_numbers["heat_equation_setup"]
    table = reference_numbers["heat_equation_error_after_10_steps"]
    NZ = 128
    geo = O.Geometry(4, 4, NZ, 1.0, 1.0)
    zg = geo.gauss_chebyshev_grid()
    n = setup["modes"]["u"]
    u0 = setup["amplitudes"]["u"] * np.sin((2 * n + 1) * (zg - 1) * np.pi / 2)
    rate = setup["Prandtl"] * ((2 * n + 1) * np.pi / 2) ** 2
    ops = O.SystemOps(r.zero_systems[0], NZ, 0.5, 1.0)
    import scipy.fft as sfft
    c = sfft.dct(u0, type=2)[:NZ] / geo.NZAA
    c[0] *= 0.5
    nc = ops.var_size[0]
    zf = np.zeros(ops.total)
    zf[ops.posP[:nc]] = np.linalg.solve(ops.stencils[0][:nc, :], c[:nc])
    for scheme in ("ARS_111", "ARS_222", "ARS_443"):
        for idx in (3, 5, 6):
            dt, ref = table["dt"][idx], table[scheme][idx]
            with solver_mod.Solver(r, 4, 4, NZ, 1.0, 1.0, scheme=scheme) as g:
                kz = [np.zeros((2, 6, t), np.complex128) for t in g.totals]
                g.set_state(kz, [zf, np.zeros(g.zero_totals[1]), np.zeros(g.zero_totals[2])])
                p0 = g.physical(0)[0, 0, :].copy()
                g.march_forward(dt, 10)
                p1 = g.physical(0)[0, 0, :]
            err = math.sqrt(np.sum((p1 - p0 * np.exp(-rate * 10 * dt)) ** 2) / np.sum(p0 ** 2))
            print(scheme, dt, err, ref, flush=True)
            assert abs(err / ref - 1.0) < 2e-3


def test_banded_solve_matches_lapack(solver_mod):
    rng = np.random.default_rng(9)
    n, kl, ku, ns = 190, 16, 27, 6
    mats, abs_ = [], []
    for s in range(ns):
        A = np.zeros((n, n), np.complex128)
        for i in range(n):
            for j in range(max(0, i - kl), min(n, i + ku + 1)):
                A[i, j] = rng.normal() + 1j * rng.normal()
        mats.append(A)
        abs_.append(O.dense_to_lapack_band(A, kl, ku))
    ab = np.stack(abs_)
    b = rng.normal(size=(ns, n)) + 1j * rng.normal(size=(ns, n))
    x, info = solver_mod.banded_solve(ab, kl, ku, b)
    assert not info.any()
    for s in range(ns):
        lu, piv, _ = lap.zgbtrf(abs_[s].T.copy(), kl, ku)
        xr, _ = lap.zgbtrs(lu, kl, ku, b[s], piv)
        assert rel(x[s], xr) <= 1e-12


def test_banded_solve_known_answer_and_singular(solver_mod, reference_numbers):
    kat = reference_numbers["driver_algebra_5x5"]
    A = np.zeros((5, 5), np.complex128)
    for i, j, v in kat["entries_1based"]:
        A[i - 1, j - 1] = v
    ab = O.dense_to_lapack_band(A, 1, 2)[None]
    x, info = solver_mod.banded_solve(ab, 1, 2, np.array([kat["rhs"]], np.complex128))
    assert info[0] == 0
    np.testing.assert_allclose(x[0].real, [-1, 2, -3, 4, -5], atol=1e-13)
    S = A.copy()
    S[:, 2] = 0.0   # singular: zgbtrf reports info = 3
    _, info = solver_mod.banded_solve(O.dense_to_lapack_band(S, 1, 2)[None], 1, 2, np.ones((1, 5), np.complex128))
    assert info[0] == 3


def test_factorization_at_production_size(solver_mod):
    """Full-size (256^2 x 129, rotating Ra = 5e7) operators: GPU LU solve vs LAPACK on sampled wavenumbers."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NZ, dt = 256, 129, 2e-5
    A = 1.0 - 1.0 / math.sqrt(2.0)
    rng = np.random.default_rng(2)
    abs_, bs, refs = [], [], []
    for kxi, kyi in [(0, 1), (1, 0), (5, -7), (40, 33), (127, -127), (64, 100)]:
        M, nl, nu = solver_mod.assembled_operator(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi, 0, "mass", kxi, kyi)
        L, _, _ = solver_mod.assembled_operator(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi, 0, "stif", kxi, kyi)
        Aop = M - dt * A * L
        ab = O.dense_to_lapack_band(Aop, nl, nu)
        b = rng.normal(size=Aop.shape[0]) + 1j * rng.normal(size=Aop.shape[0])
        lu, piv, _ = lap.zgbtrf(ab.T.copy(), nl, nu)
        xr, _ = lap.zgbtrs(lu, nl, nu, b, piv)
        abs_.append(ab), bs.append(b), refs.append(xr)
    x, info = solver_mod.banded_solve(np.stack(abs_), nl, nu, np.stack(bs))
    assert not info.any()
    for s in range(len(refs)):
        assert rel(x[s], refs[s]) <= 1e-11


def test_system_solve_matches_lapack_at_production_size(solver_mod):
    """coral_amd_system_solve: the factors the time step uses (real form,
    even/odd structural blocks, U records trimmed to the longest prefix over
    all 32,640 wavenumbers) and its 16-lane solve kernel, at 256^2 x 129
    (rotating Ra = 5e7, dt = 2e-5), against LAPACK zgbtrf/zgbtrs of the
    assembled complex operator M - dt a11 L (PL_algebra_z_3D.f90:189-276) at
    sampled wavenumbers: the corners of the kept set, the axes and random ones."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NZ, dt, L = 256, 129, 2e-5, 2 * np.pi
    a11 = 1.0 - 1.0 / math.sqrt(2.0)
    kx, ky, kept = ic.wavenumber_indices(NX, NX)
    with solver_mod.Solver(r, NX, NX, NZ, L, L) as g:
        g.factorize_operators(dt)
        rng = np.random.default_rng(5)
        shape = (NX // 2, 3 * NX // 2, g.totals[0])
        b = rng.normal(size=shape) + 1j * rng.normal(size=shape)
        x = g.solve_system(0, b)
    np.testing.assert_array_equal(x[0, 0], b[0, 0])     # the (0,0) system is the identity
    assert not x[~kept].any()
    slots = [(0, 1), (1, 0), (0, 127), (0, 384 - 127), (127, 1), (127, 127), (127, 384 - 127), (64, 0), (1, 1)]
    kept_idx = np.argwhere(kept)
    slots += [tuple(kept_idx[i]) for i in rng.choice(len(kept_idx), 15, replace=False)]
    worst = 0.0
    for ix, iy in slots:
        if (ix, iy) == (0, 0):
            continue
        M, nl, nu = solver_mod.assembled_operator(r, NX, NX, NZ, L, L, 0, "mass", int(kx[ix]), int(ky[iy]))
        S, _, _ = solver_mod.assembled_operator(r, NX, NX, NZ, L, L, 0, "stif", int(kx[ix]), int(ky[iy]))
        ab = O.dense_to_lapack_band(M - dt * a11 * S, nl, nu)
        lu, piv, info = lap.zgbtrf(ab.T.copy(), nl, nu)
        assert info == 0
        xr, _ = lap.zgbtrs(lu, nl, nu, b[ix, iy], piv)
        worst = max(worst, rel(x[ix, iy], xr))
    print("worst rel L2 over", len(slots), "wavenumbers:", worst, flush=True)
    assert worst <= 1e-11


def test_production_workload_two_steps_match_oracle(solver_mod):
    """The 256^2 x 129 rotating workload (Ek = 1e-4, Ra = 5e7, dt = 2e-5) at
    its stated amplitude 1e-2 on |k| <= 32: two ARS_222 steps of the HIP path
    against the oracle, every kept mode, relative L2 <= 1e-10 (the full-size
    solve, trimmed U records, transforms and products together)."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NZ, dt, L = 256, 129, 2e-5, 2 * np.pi
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-2, 32, 8, 2)
    with solver_mod.Solver(r, NX, NX, NZ, L, L) as g:
        g.set_state(k0, z0)
        g.march_forward(dt, 2)
        k, z = g.get_state()
        cfl = g.cfl_dt()
    geo = O.Geometry(NX, NX, NZ, L, L)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    ora.run(dt, 2)
    e = rel(k[0], ora.field[0])
    print("256^2 x 129, 2 steps: rel L2", e, "cfl", cfl, ora.cfl_dt, flush=True)
    assert np.isfinite(k[0]).all() and np.linalg.norm(k[0] - k0[0]) > 0
    assert e <= TOL
    for a, b in zip(z, ora.zfield):
        assert rel(a, b) <= TOL
    assert abs(cfl / ora.cfl_dt - 1.0) <= 1e-10


def _zfiltered_recipe(filtered=("uw", "tw", "vv")):
    """The RBC recipe with some quadratic variables declared as
    zFiltered_quadratic_variable (PL_equations.f90:1884-1905), whose z-integral
    the step removes (PL_IMEX_timestepping.f90:223-235)."""
    from coral_amd.equations import parse_equations
    from conftest import ROOT
    text = open(os.path.join(ROOT, "coral_amd", "inputs", "coral.equations.rbc_ra1e5")).read()
    for q in filtered:
        text = text.replace(f">>quadratic_variable :: {q} :=", f">>zFiltered_quadratic_variable :: {q} :=")
    r = parse_equations(text)
    assert sum(q.remove_z_integral for q in r.quad_vars) == len(filtered)
    return r


@pytest.mark.parametrize("NX,NY,NZ", [(16, 16, 17), (32, 24, 32)])
def test_zfiltered_quadratics_match_oracle(solver_mod, NX, NY, NZ):
    """z-integral removal of zFiltered quadratic variables (k_remove_zmean):
    the products seam and three steps against the oracle, at odd and even NZ;
    the filtered spectra differ from the unfiltered ones in row 0 only."""
    r = _zfiltered_recipe()
    r0 = product_recipe("coral.equations.rbc_ra1e5")
    # velocities and temperature seeded: every product has a nonzero row 0
    k0, z0 = seeded_state(r, NX, NY, NZ, 6, 6, 7, amps=(("u", 1e-2), ("v", 1e-2), ("w", 1e-2), ("t", 1e-2)))
    rng = np.random.default_rng(NZ)
    z0 = [a + 1e-2 * rng.normal(size=a.shape) for a in z0]
    geo = O.Geometry(NX, NY, NZ, 2 * np.pi, 2 * np.pi)
    ora = O.OracleSolver(r, geo, "ARS_222")
    qref = ora.quad_spectra([a.copy() for a in k0], [a.copy() for a in z0])
    with solver_mod.Solver(r0, NX, NY, NZ, 2 * np.pi, 2 * np.pi) as g0:
        g0.set_state(k0, z0)
        q0 = g0.products()
    with solver_mod.Solver(r, NX, NY, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_state(k0, z0)
        q = g.products()
        wn = np.zeros(q.shape[2], bool)
        wn[: NY // 2] = True
        wn[q.shape[2] - (NY // 2 - 1):] = True
        for i, qv in enumerate(r.quad_vars):
            assert rel(q[i][:, wn], qref[i][:, wn]) <= TOL
            if qv.remove_z_integral:
                assert np.array_equal(q[i][..., 1:], q0[i][..., 1:])
                assert not np.allclose(q[i][..., 0], q0[i][..., 0])
            else:
                assert np.array_equal(q[i], q0[i])
        ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
        for it in range(3):
            ora.step(1e-4)
            g.march_forward(1e-4)
        k, z = g.get_state()
    assert rel(k[0], ora.field[0]) <= TOL
    for a, b in zip(z, ora.zfield):
        assert rel(a, b) <= TOL


def test_transforms_match_oracle(solver_mod):
    r = product_recipe("coral.equations.rbc_ra1e5")
    # x lines of 24 (generic kernels), 96, 144, 192, 384 and 768 points
    # (wavefront kernels)
    for NX, NY, NZ, Lx, Ly in [(16, 16, 17, 2 * np.pi, 2 * np.pi), (64, 48, 33, 2.0, 3.0), (96, 16, 9, 2.0, 1.0),
                               (128, 16, 9, 2.0, 1.0), (256, 16, 9, 2.0, 1.0), (512, 16, 9, 2.0, 1.0)]:
        geo = O.Geometry(NX, NY, NZ, Lx, Ly)
        rng = np.random.default_rng(NX)
        S = rng.normal(size=(geo.nkx, geo.nky, NZ)) + 1j * rng.normal(size=(geo.nkx, geo.nky, NZ))
        S[geo.dealias] = 0.0
        S[0, 0] = S[0, 0].real
        for iy in range(geo.nky):
            k = geo.ky_int[iy]
            if k < 0 and not geo.dealias_y[iy]:
                S[0, iy] = np.conj(S[0, list(geo.ky_int).index(-k)])
        with solver_mod.Solver(r, NX, NY, NZ, Lx, Ly) as g:
            P = g.bwd_transform(S)
            Pr = O.spec_to_phys(S, geo)
            assert rel(P, Pr) <= 1e-13
            Q = P * P + 0.3 * P
            Sg = g.fwd_transform(Q)
            So = O.phys_to_spec(Q, geo)
            keep = ~geo.dealias
            assert rel(Sg[keep], So[keep]) <= 1e-13
            assert not Sg[geo.dealias].any()


def test_wavenumbers_bit_exact(solver_mod):
    gold = json.load(open(os.path.join(GOLDEN, "wavenumbers.json")))
    r = product_recipe("coral.equations.rbc_ra1e5")
    for key in ("64x64", "96x64", "16x32"):
        NX, NY = map(int, key.split("x"))
        with solver_mod.Solver(r, NX, NY, 9, 2 * np.pi, 3.0) as g:
            kx, ky, dx, dy = g.wavenumbers()
        g0 = gold[key]
        assert kx.tolist() == g0["kx_int"] and ky.tolist() == g0["ky_int"]
        assert dx.astype(int).tolist() == g0["dealias_x"] and dy.astype(int).tolist() == g0["dealias_y"]


def test_state_round_trip_and_masks(solver_mod):
    r = product_recipe("coral.equations.rbc_ra1e5")
    NX, NY, NZ = 32, 32, 17
    rng = np.random.default_rng(1)
    with solver_mod.Solver(r, NX, NY, NZ, 2 * np.pi, 2 * np.pi) as g:
        t = g.totals[0]
        a = rng.normal(size=(NX // 2, 3 * NY // 2, t)) + 1j * rng.normal(size=(NX // 2, 3 * NY // 2, t))
        g.set_field("kxky", 0, a)
        b = g.get_field("kxky", 0)
        kx, ky, kept = ic.wavenumber_indices(NX, NY)
        kept00 = kept.copy()
        kept00[0, 0] = False
        np.testing.assert_array_equal(b[kept00], a[kept00])
        assert not b[~kept00].any()
        zr = rng.normal(size=g.zero_totals[0])
        g.set_field("zero", 0, zr)
        np.testing.assert_array_equal(g.get_field("zero", 0), zr)


def test_error_paths(solver_mod):
    from coral_amd._lib import CoralError, check, lib
    r = product_recipe("coral.equations.rbc_ra1e5")
    with solver_mod.Solver(r, 16, 16, 17, 2 * np.pi, 2 * np.pi) as g:
        with pytest.raises(CoralError, match="set_dt"):
            check(lib.coral_amd_step(g.handle, 1), "coral_amd_step")
        with pytest.raises(CoralError, match="positive"):
            g.factorize_operators(-1.0)
        with pytest.raises(ValueError):
            g.set_field("kxky", 0, np.zeros((3, 3, 3)))
        with pytest.raises(CoralError, match="full-variable index"):
            g.physical(99)


def test_production_size_properties(solver_mod):
    """256^2 x 129 rotating Ra = 5e7: size-independent properties of the step --
    determinism, Hermitian kx = 0 plane, finiteness and linearity in the small-amplitude limit."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NZ, dt = 256, 129, 2e-5
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-2, 32, 8, 2)
    runs = []
    for rep in range(2):
        with solver_mod.Solver(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi) as g:
            g.set_state(k0, z0)
            g.march_forward(dt, 3)
            runs.append(g.get_state())
            if rep == 0:
                # linearity: the quadratic terms scale with the square of the amplitude
                outs = []
                for eps in (1e-9, 2e-9):
                    g.set_state([a * (eps / 1e-2) for a in k0], z0)
                    g.march_forward(dt, 1)
                    outs.append(g.get_state()[0][0])
                assert rel(outs[1], 2.0 * outs[0]) <= 1e-7
    (ka, za), (kb, zb) = runs
    np.testing.assert_array_equal(ka[0], kb[0])
    assert np.isfinite(ka[0]).all()
    kx, ky, kept = ic.wavenumber_indices(NX, NX)
    plane = ka[0][0]
    for iy in range(ky.size):
        if ky[iy] < 0 and kept[0, iy]:
            jy = list(ky).index(-ky[iy])
            assert np.max(np.abs(plane[iy] - np.conj(plane[jy]))) <= 1e-12 * np.max(np.abs(plane))


@pytest.mark.parametrize("NX,NY,NZ,nsteps", [(16, 16, 17, 3), (32, 24, 33, 2)])
def test_toroidal_poloidal_recipe_parity(solver_mod, NX, NY, NZ, nsteps):
    """The reference's toroidal-poloidal RRBC recipe
    (etc/benchmarks/02_RRBC_julien_JFM96/coral.equations.toroidalPoloidal): BC 40,
    fourth-order equations, full variables built with dz and dx dy factors."""
    r = recipe("rrbc_julien.coral.equations.toroidalPoloidal")
    k0 = None
    for seed, var in enumerate(("psi", "phi", "theta")):
        kv, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-3, 6, 6, 20 + seed, var_name=var)
        k0 = kv if k0 is None else [a + b for a, b in zip(k0, kv)]
    rng = np.random.default_rng(NZ)
    z0 = [rng.normal(size=a.shape) * 1e-3 for a in z0]
    geo = O.Geometry(NX, NY, NZ, 2.0, 2.0)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    dt = 1e-4
    g = solver_mod.Solver(r, NX, NY, NZ, 2.0, 2.0, scheme="ARS_222")
    try:
        g.set_state(k0, z0)
        for it in range(nsteps):
            ora.step(dt)
            g.march_forward(dt)
        k, z = g.get_state()
        assert np.linalg.norm(ora.field[0]) > 0
        assert rel(k[0], ora.field[0]) <= TOL
        for a, b in zip(z, ora.zfield):
            assert rel(a, b) <= TOL
        phys = ora.physical_full_variables()
        for f in range(len(phys)):
            assert rel(g.physical(f), phys[f]) <= TOL
    finally:
        g.close()


def test_structural_blocks_match_whole_system_factors(solver_mod, monkeypatch):
    """Factoring the even- and odd-degree blocks of the evolution operator on
    their own against factoring whole systems (CORAL_AMD_NO_BLOCK_SPLIT): the
    same eliminations reach every row, so the states agree to rounding; the
    whole-system path also matches the oracle."""
    name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps = CASES[2]
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 6, 6, 3)
    out = []
    for nosplit in ("0", "1"):
        monkeypatch.setenv("CORAL_AMD_NO_BLOCK_SPLIT", nosplit)
        g = solver_mod.Solver(r, NX, NY, NZ, Lx, Ly, scheme=scheme)
        try:
            nblk, blk_n, blk_nl, blk_nu = g.factor_blocks(0)
            total, nl, nu, _ = g.system_info(0)
            print("blocks", nosplit, nblk, blk_n, blk_nl, blk_nu, "of", total, nl, nu, flush=True)
            if nosplit == "1":
                assert (nblk, blk_n, blk_nl, blk_nu) == (1, total, nl, nu)
            else:
                assert nblk == 2 and total <= 2 * blk_n < total + 8 and blk_nl < nl and blk_nu < nu
            g.set_state(k0, z0)
            g.march_forward(dt, nsteps)
            # U records are cut to their longest nonzero prefix over all wavenumbers
            ue = g.u_entries_per_column(0)
            assert 1.0 <= ue <= blk_nl + blk_nu + 1
            out.append(g.get_state())
        finally:
            g.close()
    (ka, za), (kb, zb) = out
    assert rel(ka[0], kb[0]) <= 1e-12
    for a, b in zip(za, zb):
        assert rel(a, b) <= 1e-12
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    ora = O.OracleSolver(r, geo, scheme)
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    for it in range(nsteps):
        ora.step(dt)
    assert rel(kb[0], ora.field[0]) <= TOL


def test_graph_replay_matches_eager(solver_mod):
    """Steps replayed as HIP graphs (Solver.use_graphs) against eager steps, bit
    for bit: 2 + 3 replays (both buffer assignments), an eager step with the
    timers on between replays, state downloads between them, a re-upload, then
    a change of dt (which drops and recaptures the graphs); the mean systems'
    side stream is part of every graph."""
    name, NX, NY, NZ, L, dt = "coral.equations.rrbc_ra5e7_ek1e-4", 64, 64, 33, 2 * np.pi, 2e-5
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 10 * NX, 8, 11)
    # timers on make steps run eagerly between replays (one step: the buffer
    # assignment flips), which the replays that follow must see
    plan = [("run", dt, 2), ("run", dt, 3), ("timers", True, None), ("run", dt, 1), ("timers", False, None),
            ("run", dt, 3), ("reset", None, None), ("run", dt, 1), ("run", dt / 2, 2)]
    results = []
    for graphs in (False, True):
        out = []
        with solver_mod.Solver(r, NX, NY, NZ, L, L, scheme="ARS_222") as g:
            if graphs:
                g.use_graphs(True)
            g.set_state(k0, z0)
            for what, h, n in plan:
                if what == "reset":
                    g.set_state(k0, z0)
                    continue
                if what == "timers":
                    g.timers_enable(h)
                    continue
                print("graphs", graphs, "dt", h, "steps", n, flush=True)
                g.march_forward(h, n)
                k, z = g.get_state()
                out.append((k, z, g.cfl_dt()))
        results.append(out)
    for (ke, ze, ce), (kg, zg, cg) in zip(*results):
        for a, b in zip(ke, kg):
            assert np.array_equal(a, b)
        for a, b in zip(ze, zg):
            assert np.array_equal(a, b)
        assert ce == cg
    assert np.all(np.isfinite(results[1][-1][0][0]))


@pytest.mark.parametrize("case", [CASES[2], BROADBAND[1]], ids=["rrbc-32x32x25", "rbc-128x96x17"])
def test_products_seam_matches_oracle(solver_mod, case):
    """coral_amd_products: the quadratic variables' spectra of a state (full
    variables to the 3/2 grid, products, forward transforms, z-integral
    removal) against the oracle's, at every kept mode; dealiased ky slots are
    zero."""
    name, NX, NY, NZ, Lx, Ly, scheme, dt, nsteps = case
    r = product_recipe(name)
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 10 * max(NX, NY), 8, 11)
    geo = O.Geometry(NX, NY, NZ, Lx, Ly)
    ora = O.OracleSolver(r, geo, scheme)
    qref = ora.quad_spectra([a.copy() for a in k0], [a.copy() for a in z0])
    with solver_mod.Solver(r, NX, NY, NZ, Lx, Ly, scheme=scheme) as g:
        g.set_state(k0, z0)
        q = g.products()
    wn = np.zeros(q.shape[2], bool)
    wn[: NY // 2] = True
    wn[q.shape[2] - (NY // 2 - 1):] = True
    assert len(qref) == q.shape[0] == len(r.quad_vars)
    for i in range(q.shape[0]):
        assert rel(q[i][:, wn], qref[i][:, wn]) <= TOL
        assert not np.any(q[i][:, ~wn])


def test_config1_200_steps_graph_replays(solver_mod):
    """The 64^2 x 33 Ra = 1e5 case over 200 steps of dt = 1e-4 (a fifth of the
    1000-step run), the GPU side as HIP graph replays in blocks of 50: the
    spectral state stays within 1e-10 (relative L2) of the oracle's."""
    r = product_recipe("coral.equations.rbc_ra1e5")
    NX, NZ, dt = 64, 33, 1e-4
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-3, 8, 8, 1)
    geo = O.Geometry(NX, NX, NZ, 2 * np.pi, 2 * np.pi)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    with solver_mod.Solver(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.use_graphs(True)
        g.set_state(k0, z0)
        for block in range(4):
            g.march_forward(dt, 50)
            ora.run(dt, 50)
            k, z = g.get_state()
            e = rel(k[0], ora.field[0])
            print("steps", 50 * (block + 1), "rel L2", e, flush=True)
            assert e <= TOL
            assert all(rel(a, b) <= TOL for a, b in zip(z, ora.zfield))


def test_kx0_projection_matches_oracle(solver_mod):
    """coral_amd_set_kx0_projection: from a state whose kx = 0 plane carries an
    anti-Hermitian part, five steps with the projection match the oracle's
    projected steps (<= 1e-10), leave the plane exactly Hermitian, and give the
    same physical fields as steps without it."""
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NY, NZ, dt = 32, 24, 17, 2e-5
    geo = O.Geometry(NX, NY, NZ, 2 * np.pi, 3.0)
    k0, z0 = seeded_state(r, NX, NY, NZ, 8, 6, 4, amps=(("u", 1e-2), ("v", 1e-2), ("w", 1e-2), ("t", 1e-2)))
    rng = np.random.default_rng(12)
    noise = 1e-2 * (rng.normal(size=k0[0][0].shape) + 1j * rng.normal(size=k0[0][0].shape))
    noise[geo.dealias[0]] = 0.0     # kept modes only, (0, 0) excluded, as an upload keeps them
    noise[0] = 0.0
    k0[0][0] += noise
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.project_kx0 = True
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    ora.run(dt, 5)
    phys = []
    for proj in (True, False):
        with solver_mod.Solver(r, NX, NY, NZ, 2 * np.pi, 3.0) as g:
            g.set_state(k0, z0)
            g.set_kx0_projection(proj)
            g.march_forward(dt, 5)
            k, z = g.get_state()
            phys.append([g.physical(i) for i in range(4)])
            if proj:
                print("projected 5 steps vs oracle:", rel(k[0], ora.field[0]), flush=True)
                assert rel(k[0], ora.field[0]) <= 1e-10
                for a, b in zip(z, ora.zfield):
                    assert rel(a, b) <= 1e-10
                ky = geo.ky_int
                for iy in range(geo.nky):
                    if ky[iy] > 0 and not geo.dealias[0, iy]:
                        jy = int(np.nonzero(ky == -ky[iy])[0][0])
                        assert np.array_equal(k[0][0, jy], np.conj(k[0][0, iy]))
    for a, b in zip(phys[0], phys[1]):
        assert rel(a, b) <= 1e-12


@pytest.mark.parametrize("extract,zdecomp,NZ", [("TopSurf", "baroclinic", 17), ("BotSurf", "barotropic", 24)])
def test_full_variable_filters_match_oracle(solver_mod, extract, zdecomp, NZ):
    """Boundary-value extraction (linear_variable_{top,bottom}_boundary_value)
    and vertical decompositions (vertical_decomposition :: baroclinic /
    barotropic) of full variables (k_full_zfilter, coral_amd_set_full_variable_filter;
    PL_IMEX_timestepping_marching.f90:544-577), in a recipe whose temperature
    equation takes products of the filtered variables: the physical full
    variables, the quadratic spectra and three steps against the oracle; the
    extracted variable is constant in z and the decomposed one has the
    z-structure the filter leaves (zero vertical mean, or none but the mean)."""
    from conftest import full_filter_recipe
    r = full_filter_recipe(extract, zdecomp)
    NX, NY = 16, 16
    k0, z0 = seeded_state(r, NX, NY, NZ, 6, 6, 5, amps=(("u", 1e-2), ("v", 1e-2), ("w", 1e-2), ("t", 1e-2)))
    rng = np.random.default_rng(NZ)
    z0 = [a + 1e-2 * rng.normal(size=a.shape) for a in z0]
    geo = O.Geometry(NX, NY, NZ, 2 * np.pi, 2 * np.pi)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    phys_ref = ora.physical_full_variables()
    qref = ora.quad_spectra([a.copy() for a in k0], [a.copy() for a in z0])
    with solver_mod.Solver(r, NX, NY, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_state(k0, z0)
        phys = [g.physical(f) for f in range(len(r.full_vars))]
        q = g.products()
        for it in range(3):
            ora.step(1e-4)
            g.march_forward(1e-4)
        k, z = g.get_state()
    for f in range(len(phys)):
        assert rel(phys[f], phys_ref[f]) <= TOL, r.full_vars[f].name
    tx, uz = phys[-2], phys[-1]
    assert np.abs(tx - tx[:, :, :1]).max() <= 1e-12 * np.abs(tx).max()     # constant in z
    assert np.abs(tx).max() > 0 and np.abs(uz).max() > 0
    wn = np.zeros(q.shape[2], bool)
    wn[: NY // 2] = True
    wn[q.shape[2] - (NY // 2 - 1):] = True
    for i in range(len(r.quad_vars)):
        assert rel(q[i][:, wn], qref[i][:, wn]) <= TOL, r.quad_vars[i].name
    assert rel(k[0], ora.field[0]) <= TOL
    for a, b in zip(z, ora.zfield):
        assert rel(a, b) <= TOL


# --- mean-system sources ------------------------------------------------------
SOURCE_CASES = [
    # reference equation file, definitions, kxky perturbations
    ("stratified_channel.coral.equations.velocityPressure", "SHIPPED",
     (("u", 1e-2), ("v", 1e-2), ("w", 1e-2), ("theta", 1e-2))),
    ("radiative_heating.coral.equations", "SHIPPED", (("psi", 1e-2), ("phi", 1e-2), ("theta", 1e-2))),
    ("baroclinic_turbulence.coral.equations", "BAROCLINIC_TURBULENCE", (("psi", 1e-2), ("phi", 1e-2), ("b", 1e-2))),
]


@pytest.mark.parametrize("case", SOURCE_CASES, ids=[c[0].split(".")[0] for c in SOURCE_CASES])
def test_mean_sources_match_oracle(solver_mod, case):
    """Recipes of the reference with add_rhs_source terms (the stratified
    channel of etc/benchmarks, the radiative-heating example, the baroclinic
    turbulence case with its own definitions file), their sources from the
    definitions files restated in coral_amd.sources: three steps from a
    seeded state against the oracle, every system."""
    from coral_amd import sources as S
    name, defs, amps = case
    print("case", name, flush=True)
    r = recipe(name)
    defs = getattr(S, defs)
    NX, NY, NZ = 16, 16, 17
    k0, z0 = seeded_state(r, NX, NY, NZ, 6, 6, 5, amps=amps)
    rng = np.random.default_rng(7)
    z0 = [a + 1e-3 * rng.normal(size=a.shape) for a in z0]
    geo = O.Geometry(NX, NY, NZ, 2 * np.pi, 2 * np.pi)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    zg = geo.gauss_chebyshev_grid()
    for t in r.sources:
        ora.set_source(t.index, S.evaluate(defs[t.index], zg, S.SourceParams(r.source_params)))
    with solver_mod.Solver(r, NX, NY, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_state(k0, z0)
        g.set_sources(defs)
        for it in range(3):
            ora.step(1e-4)
            g.march_forward(1e-4)
        k, z = g.get_state()
    for a, b in zip(k, ora.field):
        assert rel(a, b) <= TOL
    for i, (a, b) in enumerate(zip(z, ora.zfield)):
        assert rel(a, b) <= TOL, r.zero_systems[i].var_names


def test_mean_sources_closed_forms_on_gpu(solver_mod):
    """Without fluctuations the stratified channel's mean flow steps to
    u = (z - z^2/2) / nu (bottom no-slip, top stress-free, unit forcing from
    source02 = -1); the baroclinic case's algebraic mean equations
    0 = -UShear + sourceNN hold the sources' Chebyshev coefficients,
    exp(-(1 - z)/L) = e^{-1/(2L)} (I_0(a) + 2 sum I_k(a) T_k(2z - 1)), a = 1/(2L),
    times 1/L and 1/L^2 for sources 4 and 5.  Stepping with a source that has
    no definition fails."""
    import scipy.special as ss
    from coral_amd import sources as S
    r = recipe("stratified_channel.coral.equations.velocityPressure")
    NX, NY, NZ = 8, 8, 17
    with solver_mod.Solver(r, NX, NY, NZ, 1.0, 1.0) as g:
        with pytest.raises(Exception, match="source02"):
            g.march_forward(2e-3)
        g.set_sources(S.SHIPPED)
        g.march_forward(2e-3, 80)
        u = g.physical(0)
    zg = O.Geometry(NX, NY, NZ, 1.0, 1.0).gauss_chebyshev_grid()
    want = (zg - zg ** 2 / 2) / r.params["invReTau"]
    assert np.abs(u - want[None, None, :]).max() <= 1e-10 * np.abs(want).max()
    r = recipe("baroclinic_turbulence.coral.equations")
    L = r.source_params["shearL"]
    a = 1.0 / (2 * L)
    k = np.arange(NZ)
    series = np.exp(-a) * 2 * ss.iv(k, a)
    series[0] /= 2
    with solver_mod.Solver(r, NX, NY, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_sources(S.BAROCLINIC_TURBULENCE)
        g.march_forward(1e-4)
        _, z = g.get_state()
    for isys, scale in ((1, 1.0), (2, 1.0 / L), (3, 1.0 / L ** 2)):
        np.testing.assert_allclose(z[isys], scale * series, rtol=1e-12, atol=1e-15)


# --- penalisation and the horizontal Smagorinsky viscosity ----------------------
def test_penalisation_matches_oracle(solver_mod):
    """The reference's penalised RBC recipe (coral.equations.RBC_penalised)
    with the shipped masks (PL_user_penalisation_mask.f90): GAMMA's physical
    values are the mask field as set, the products Gu, Gv, Gw (formed in
    physical space with it) and three steps match the oracle."""
    from coral_amd import penalisation as PN
    r = recipe("rbc_penalised.coral.equations")
    NX, NY, NZ = 16, 16, 17
    L = 2 * np.pi
    k0, z0 = seeded_state(r, NX, NY, NZ, 6, 6, 5, amps=(("u", 1e-2), ("v", 1e-2), ("w", 1e-2), ("t", 1e-2)))
    geo = O.Geometry(NX, NY, NZ, L, L)
    gvar = r.full_vars[4]
    field = PN.mask_field(PN.SHIPPED, gvar.penalisation_strength, gvar.penalisation_width, geo.NXAA, geo.NYAA,
                          geo.NZAA, L, L, geo.center, geo.gap)
    assert 0 <= field.min() < field.max() <= 1e3 and 0 < np.median(field) < 1e3
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    ora.set_fixed_physical(4, field)
    qref = ora.quad_spectra([a.copy() for a in k0], [a.copy() for a in z0])
    with solver_mod.Solver(r, NX, NY, NZ, L, L) as g:
        g.set_state(k0, z0)
        g.set_penalisation_mask()
        assert np.array_equal(g.physical(4), field)
        q = g.products()
        for it in range(3):
            ora.step(1e-4)
            g.march_forward(1e-4)
        k, z = g.get_state()
    wn = np.zeros(q.shape[2], bool)
    wn[: NY // 2] = True
    wn[q.shape[2] - (NY // 2 - 1):] = True
    for i in range(len(r.quad_vars)):
        assert rel(q[i][:, wn], qref[i][:, wn]) <= TOL, r.quad_vars[i].name
    for a, b in zip(k, ora.field):
        assert rel(a, b) <= TOL
    for a, b in zip(z, ora.zfield):
        assert rel(a, b) <= TOL


def test_smagorinsky_matches_oracle(solver_mod):
    """An RBC recipe with the horizontal velocity gradients as full variables
    and eddy-viscosity terms built on nuSmago (read with the
    --withHorizontalSmagorinsky flag, prefactor 0.7): the products seam leaves
    the profile as it was; two eager steps and two steps replayed as HIP
    graphs match the oracle, and so do nuSmago's physical values (the profile
    of the last stage) and the CFL step with the Smagorinsky bound."""
    from conftest import smagorinsky_recipe
    r = smagorinsky_recipe()
    NX, NY, NZ = 16, 16, 17
    L = 2 * np.pi
    nu = len(r.full_vars) - 1
    k0, z0 = seeded_state(r, NX, NY, NZ, 6, 6, 5, amps=(("u", 1e-1), ("v", 1e-1), ("w", 1e-1), ("t", 1e-2)))
    geo = O.Geometry(NX, NY, NZ, L, L)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.smagorinsky_prefactor = 0.7
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    qref = ora.quad_spectra([a.copy() for a in k0], [a.copy() for a in z0], advance=False)
    with solver_mod.Solver(r, NX, NY, NZ, L, L, smagorinsky_prefactor=0.7) as g:
        g.set_state(k0, z0)
        q = g.products()
        assert np.array_equal(g.products(), q)
        assert not g.physical(nu).any()          # the profile starts at zero
        for it in range(2):
            ora.step(1e-4)
            g.march_forward(1e-4)
        g.use_graphs(True)
        for it in range(2):
            ora.step(1e-4)
            g.march_forward(1e-4)
        k, z = g.get_state()
        pnu = g.physical(nu)
        cdt = g.cfl_dt()
    wn = np.zeros(q.shape[2], bool)
    wn[: NY // 2] = True
    wn[q.shape[2] - (NY // 2 - 1):] = True
    for i in range(len(r.quad_vars)):
        assert rel(q[i][:, wn], qref[i][:, wn]) <= TOL, r.quad_vars[i].name
    for a, b in zip(k, ora.field):
        assert rel(a, b) <= TOL
    for a, b in zip(z, ora.zfield):
        assert rel(a, b) <= TOL
    assert ora.nu_prof.max() > 0
    assert rel(pnu, np.broadcast_to(ora.nu_prof, pnu.shape)) <= TOL
    assert cdt == pytest.approx(ora.cfl_dt, rel=1e-10)
