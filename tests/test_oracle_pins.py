"""Pins of the oracle (and of the product's CPU-side assembly) against what the
reference's own files hold and what the mathematics fixes.  No GPU needed."""
import json
import math
import os

import numpy as np
import numpy.polynomial.chebyshev as C
import pytest
import scipy.linalg as sl
import scipy.linalg.lapack as lap

from conftest import GOLDEN, recipe
from coral_amd import ic
from oracle import coral_oracle as O


# --- integer wavenumbers and dealiasing masks (bit exact) ---------------------

def test_wavenumber_golden_vectors_are_fft_frequencies():
    gold = json.load(open(os.path.join(GOLDEN, "wavenumbers.json")))
    for key, g in gold.items():
        NX, NY = map(int, key.split("x"))
        fy = np.rint(np.fft.fftfreq(3 * NY // 2) * (3 * NY // 2)).astype(int)
        fx = np.rint(np.fft.fftfreq(3 * NX // 2) * (3 * NX // 2)).astype(int)[: NX // 2]
        assert g["ky_int"] == fy.tolist()
        assert g["kx_int"] == fx.tolist()
        assert g["dealias_y"] == [int(2 * abs(k) >= NY) for k in fy]


def test_oracle_and_host_indexing_match_golden():
    gold = json.load(open(os.path.join(GOLDEN, "wavenumbers.json")))
    for key, g in gold.items():
        NX, NY = map(int, key.split("x"))
        geo = O.Geometry(NX, NY, 9, 2 * np.pi, 3.0)
        assert geo.kx_int.tolist() == g["kx_int"] and geo.ky_int.tolist() == g["ky_int"]
        assert geo.dealias_x.astype(int).tolist() == g["dealias_x"]
        assert geo.dealias_y.astype(int).tolist() == g["dealias_y"]
        kx, ky, kept = ic.wavenumber_indices(NX, NY)
        assert kx.tolist() == g["kx_int"] and ky.tolist() == g["ky_int"]
        assert (~kept).astype(int).tolist() == (np.array(g["dealias_x"])[:, None] | np.array(g["dealias_y"])[None, :]).tolist()


# --- Chebyshev integration / multiplication matrices ---------------------------

@pytest.mark.parametrize("center,gap", [(0.5, 1.0), (0.0, 2.0), (1.3, 0.7)])
def test_integration_and_multiplication_match_numpy_chebyshev(center, gap):
    N, pad = 24, 6
    I, M = O.chebyshev_elementary_matrices(N, center, gap, pad)
    rng = np.random.default_rng(1)
    c = np.zeros(N + pad)
    c[: N - 2] = rng.normal(size=N - 2)
    # d/dz = (2/gap) d/dx, so int dz = (gap/2) int dx; T_0 coefficient is free
    ref = (gap / 2.0) * C.chebint(c[: N - 2])
    got = I @ c
    np.testing.assert_allclose(got[1 : N - 1], ref[1 : N - 1], rtol=0, atol=1e-13)
    # z p(z) with z = center + gap/2 x
    ref = center * c[: N - 1] + (gap / 2.0) * C.chebmulx(c[: N - 2])
    np.testing.assert_allclose((M @ c)[: N - 1], ref, atol=1e-13)


def test_derivative_operator_differentiates_polynomials():
    NZ, center, gap = 20, 0.5, 1.0
    D = O.chebyshev_derivative_matrix(NZ, center, gap)
    c = np.random.default_rng(2).normal(size=NZ)
    np.testing.assert_allclose(D @ c, np.concatenate([(2.0 / gap) * C.chebder(c), [0.0]]), atol=1e-9)


# --- Galerkin stencils: columns satisfy their boundary conditions -------------

def _functionals(N):
    n = np.arange(N, dtype=float)
    e = lambda x, d: np.array([C.chebval(x, C.chebder(np.eye(N)[k], d)) if d else C.chebval(x, np.eye(N)[k])
                               for k in range(N)])
    return {"f+": e(1.0, 0), "f-": e(-1.0, 0), "d+": e(1.0, 1), "d-": e(-1.0, 1), "dd+": e(1.0, 2), "dd-": e(-1.0, 2)}


@pytest.mark.parametrize("bc,conds", [(10, ["f+"]), (11, ["f-"]), (20, ["f+", "f-"]), (21, ["d+", "d-"]),
                                      (22, ["f+", "d-"]), (23, ["d+", "f-"]), (40, ["f+", "f-", "d+", "d-"]),
                                      (41, ["f+", "f-", "dd+", "dd-"]), (42, ["f+", "d+", "f-", "dd-"]),
                                      (43, ["f+", "dd+", "f-", "d-"])])
def test_stencil_columns_satisfy_boundary_conditions(bc, conds):
    N = 18
    S = O.galerkin_stencil(N, bc)
    assert S.shape == (N, N - bc // 10)
    F = _functionals(N)
    for c in conds:
        scale = max(1.0, float(np.max(np.abs(F[c]))))
        assert np.max(np.abs(F[c] @ S)) < 1e-11 * scale, (bc, c)


# --- structure of the RBC primitive-variable system ---------------------------

@pytest.mark.parametrize("NZ", [17, 33, 49, 65, 129])
def test_rbc_system_structure(NZ):
    r = recipe("rbc_vanilla.coral.equations")
    ops = O.SystemOps(r.kxky_systems[0], NZ, 0.5, 1.0)
    mats = [ops.piece_matrix(p) for p in ops.sys.mass + ops.sys.stif]
    assert ops.total == 6 * NZ - 8
    assert O.bandwidth(mats) == (16, 27)
    assert sorted(ops.posP.tolist()) == list(range(ops.total))
    assert sorted(ops.posQ.tolist()) == list(range(ops.total))


@pytest.mark.parametrize("name,NZ", [("rbc_vanilla.coral.equations", 17),
                                     ("rrbc_julien.coral.equations.velocityPressure", 25),
                                     ("stressfree_rb.coral.equations", 24),
                                     ("heat_equation.coral.equations", 16)])
def test_product_assembly_matches_oracle(native_lib, name, NZ):
    from coral_amd.solver import assembled_operator
    r = recipe(name)
    Lx, Ly = 2.0, 3.0
    for isys, sysr in enumerate(r.kxky_systems):
        ops = O.SystemOps(sysr, NZ, 0.5, 1.0)
        for which, pieces in (("mass", sysr.mass), ("stif", sysr.stif)):
            groups = ops.grouped_pieces(pieces)
            for kxi, kyi in [(0, 0), (1, 0), (3, -2), (7, 5)]:
                kx, ky = kxi * 2 * np.pi / Lx, kyi * 2 * np.pi / Ly
                ref = sum((((1j * kx) ** a) * ((1j * ky) ** b)) * m for (a, b), m in groups)
                got, nl, nu = assembled_operator(r, 16, 16, NZ, Lx, Ly, isys, which, kxi, kyi)
                assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
                mats = [ops.piece_matrix(p) for p in sysr.mass + sysr.stif]
                assert (nl, nu) == O.bandwidth(mats)


@pytest.mark.parametrize("name,NZ", [("rbc_vanilla.coral.equations", 17),
                                     ("rrbc_julien.coral.equations.velocityPressure", 25),
                                     ("stressfree_rb.coral.equations", 24)])
def test_evolution_operator_has_real_form(native_lib, name, NZ):
    """The GPU factors M - dt a L in real arithmetic after scaling unknowns and
    equations by powers of i.  Found here from the matrix entries alone: every
    nonzero of the assembled complex operator is real or imaginary, and a
    parity labelling of rows and columns (breadth-first over the nonzeros)
    makes every entry of D_row^-1 A D_col real."""
    from coral_amd.solver import assembled_operator
    r = recipe(name)
    for kxi, kyi in [(1, 0), (3, -2), (7, 5)]:
        M, nl, nu = assembled_operator(r, 16, 16, NZ, 2.0, 3.0, 0, "mass", kxi, kyi)
        L, _, _ = assembled_operator(r, 16, 16, NZ, 2.0, 3.0, 0, "stif", kxi, kyi)
        A = M - 0.37 * L
        n = A.shape[0]
        nz = np.argwhere(np.abs(A) > 0)
        imag = np.abs(A.imag) > np.abs(A.real)
        # each entry is (numerically) purely real or purely imaginary
        small = np.where(imag, np.abs(A.real), np.abs(A.imag))
        assert np.max(small) <= 1e-14 * np.max(np.abs(A))
        row, col = -np.ones(n, int), -np.ones(n, int)
        adj_r = [[] for _ in range(n)]
        adj_c = [[] for _ in range(n)]
        for i, j in nz:
            adj_r[i].append(j)
            adj_c[j].append(i)
        for start in range(n):
            if row[start] >= 0 or not adj_r[start]:
                continue
            row[start] = 0
            stack = [("r", start)]
            while stack:
                kind, u = stack.pop()
                if kind == "r":
                    for j in adj_r[u]:
                        want = row[u] ^ int(imag[u, j])
                        if col[j] < 0:
                            col[j] = want
                            stack.append(("c", j))
                        assert col[j] == want
                else:
                    for i in adj_c[u]:
                        want = col[u] ^ int(imag[i, u])
                        if row[i] < 0:
                            row[i] = want
                            stack.append(("r", i))
                        assert row[i] == want
        row[row < 0] = 0
        col[col < 0] = 0
        Ar = (1j ** -row)[:, None] * A * (1j ** col)[None, :]
        assert np.max(np.abs(Ar.imag)) <= 1e-14 * np.max(np.abs(A))


# --- stress-free Rayleigh-Benard growth rate ----------------------------------

def test_stressfree_growth_rate(reference_numbers):
    ref = reference_numbers["stressfree_rb_max_growth_rate"]
    Ra, Pr, k = ref["Rayleigh"], ref["Prandtl"], ref["k_opt"]
    # closed form of linearGrowthrate_numerical_vs_analytical.py:52-55 at the notebook's k_opt
    D2 = k * k + math.pi ** 2
    s = 0.5 * (-D2 * (Pr + 1) + math.sqrt((Pr + 1) ** 2 * D2 ** 2 - 4 * (Pr * D2 ** 2 - Ra * Pr * k * k / D2)))
    assert abs(s - ref["s_max"]) < 1e-12
    # discrete operator of the benchmark's equations at kx = k_opt
    r = recipe("stressfree_rb.coral.equations")
    ops = O.SystemOps(r.kxky_systems[0], 24, 0.5, 1.0)
    mono = lambda a, b: ((1j * k) ** a) * (0j ** b)   # kx = k_opt, ky = 0
    M = sum(mono(a, b) * m for (a, b), m in ops.grouped_pieces(ops.sys.mass))
    L = sum(mono(a, b) * m for (a, b), m in ops.grouped_pieces(ops.sys.stif))
    w = sl.eig(L, M, right=False)
    w = w[np.isfinite(w)]
    assert abs(np.max(w.real) - ref["s_max"]) < 1e-9


def stability_rate(scheme, dt, lam):
    """Growth rate per unit time of the ARS implicit stages for dx/dt = lam x:
    x_i = x^n + dt sum_{j<=i} a_ij lam x_j, x^{n+1} = x_s, i.e.
    log(e_s^T (I - dt lam A)^-1 1) / dt (IMEX_schemes.f90 A_arr)."""
    from coral_amd.schemes import tableau
    A, _ = tableau(scheme)
    R = np.linalg.solve(np.eye(A.shape[0]) - dt * lam * A, np.ones(A.shape[0]))[-1]
    return math.log(R) / dt


def stepped_growth_rate(step, state_norm, dt, n0, n1):
    """Growth rate fitted from the norm of one wavenumber's state between
    steps n0 and n1 of a run (step() advances one step)."""
    norms = {}
    for n in range(1, n1 + 1):
        step()
        if n in (n0, n1):
            norms[n] = state_norm()
    return math.log(norms[n1] / norms[n0]) / ((n1 - n0) * dt)


@pytest.mark.parametrize("scheme,dt,n0,n1", [("ARS_222", 1e-3, 400, 800), ("ARS_443", 2e-3, 200, 400)])
def test_stressfree_growth_rate_by_time_stepping(reference_numbers, scheme, dt, n0, n1):
    """The oracle's kxky time stepping (factorisation, right-hand sides, solves
    and the nonlinear path at negligible amplitude) on the stress-free
    benchmark (Ra = 2000, Pr = 7) in a box whose first x wavenumber is the
    notebook's k_opt: the temperature perturbation of that wavenumber grows
    at exactly the rate the scheme's stability function gives for the
    published growth rate 24.137487 (benchmark_01...ipynb:68).  Other
    eigenvalues of the pencil (the next is -39.4) have decayed by the fit
    window.  A wrong tableau entry, sign or operator term moves the rate."""
    ref = reference_numbers["stressfree_rb_max_growth_rate"]
    r = recipe("stressfree_rb.coral.equations")
    L = 2 * math.pi / ref["k_opt"]
    NX, NZ = 8, 24
    geo = O.Geometry(NX, NX, NZ, L, L)
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-30, 1, 8, 5, var_name="t")
    ora = O.OracleSolver(r, geo, scheme)
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    rate = stepped_growth_rate(lambda: ora.step(dt), lambda: np.linalg.norm(ora.field[0][1, 0]), dt, n0, n1)
    expect = stability_rate(scheme, dt, ref["s_max"])
    print(scheme, "stepped", rate, "stability function", expect, "continuous", ref["s_max"])
    assert abs(rate / expect - 1.0) < 1e-10
    # and the scheme's time error is what separates it from the continuous rate
    assert abs(rate - ref["s_max"]) < 1e-3 * ref["s_max"]


# --- z-integral removal of zFiltered quadratic variables ----------------------

@pytest.mark.parametrize("NZ", [16, 17, 32, 33])
def test_z_integral_removal_zeroes_the_integral(NZ):
    """remove_z_integral (PL_IMEX_timestepping.f90:223-235): the Chebyshev
    series integrates to zero over [-1, 1] afterwards (numpy's Chebyshev
    integral), except that for odd NZ the reference's loop stops at row
    NZ - 3, so row NZ - 1 keeps its share of the integral: 2 c / (1 - (NZ-1)^2)."""
    rng = np.random.default_rng(NZ)
    s = rng.normal(size=(3, NZ)) + 1j * rng.normal(size=(3, NZ))
    f = O.remove_z_integral(s.copy(), NZ)
    assert np.array_equal(f[:, 1:], s[:, 1:])
    for row in f:
        integ = C.chebint(row)
        total = C.chebval(1.0, integ) - C.chebval(-1.0, integ)
        left = 0.0 if NZ % 2 == 0 else 2.0 * row[NZ - 1] / (1.0 - (NZ - 1) ** 2)
        assert abs(total - left) <= 1e-13 * np.abs(row).sum()


# --- heat-equation error table (IMEX tableaux + mean operators + solve) --------

def heat_error(scheme, dt, nsteps=10):
    r = recipe("heat_equation.coral.equations")
    setup = json.load(open(os.path.join(GOLDEN, "reference_numbers.json")))["heat_equation_setup"]
    geo = O.Geometry(4, 4, 128, 1.0, 1.0)
    z = geo.gauss_chebyshev_grid()
    n = setup["modes"]["u"]
    u0 = setup["amplitudes"]["u"] * np.sin((2 * n + 1) * (z - 1) * np.pi / 2)
    rate = setup["Prandtl"] * ((2 * n + 1) * np.pi / 2) ** 2
    s = O.OracleSolver(r, geo, scheme)
    ops = s.zops[0]
    import scipy.fft as sfft
    c = sfft.dct(u0, type=2)[: geo.NZ] / geo.NZAA
    c[0] *= 0.5
    nc = ops.var_size[0]
    s.zfield[0][ops.posP[:nc]] = np.linalg.solve(ops.stencils[0][:nc, :], c[:nc])
    p0 = s.physical_full_variables()[0][0, 0, :].copy()
    s.run(dt, nsteps)
    p1 = s.physical_full_variables()[0][0, 0, :]
    return float(np.sqrt(np.sum((p1 - p0 * np.exp(-rate * nsteps * dt)) ** 2) / np.sum(p0 ** 2)))


@pytest.mark.parametrize("scheme", ["ARS_111", "ARS_222", "ARS_443"])
def test_heat_equation_error_table(scheme, reference_numbers):
    table = reference_numbers["heat_equation_error_after_10_steps"]
    for idx in (3, 6):  # dt = 1e-4 and 1e-3
        dt, ref = table["dt"][idx], table[scheme][idx]
        err = heat_error(scheme, dt)
        assert abs(err / ref - 1.0) < 2e-3, (scheme, dt, err, ref)


# --- banded LU against LAPACK and the 5x5 known answer ------------------------

def _random_band(rng, n, kl, ku, pivoting):
    A = np.zeros((n, n), np.complex128)
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            A[i, j] = rng.normal() + 1j * rng.normal()
        if not pivoting:
            A[i, i] += 50.0
    return A


@pytest.mark.parametrize("pivoting", [True, False])
def test_banded_lu_matches_lapack(pivoting):
    rng = np.random.default_rng(3)
    n, kl, ku, ns = 120, 16, 27, 6
    mats = [_random_band(rng, n, kl, ku, pivoting) for _ in range(ns)]
    ab = np.stack([O.dense_to_lapack_band(A, kl, ku) for A in mats])
    b = rng.normal(size=(ns, n)) + 1j * rng.normal(size=(ns, n))
    ipiv, info = O.zgbtrf_batch(ab, kl, ku)
    assert not info.any()
    x = b.copy()
    O.zgbtrs_batch(ab, kl, ku, ipiv, x)
    for s in range(ns):
        lu, piv, inf = lap.zgbtrf(O.dense_to_lapack_band(mats[s], kl, ku).T.copy(), kl, ku)
        np.testing.assert_array_equal(ipiv[s] - 1, piv)
        xr, _ = lap.zgbtrs(lu, kl, ku, b[s], piv)
        assert np.linalg.norm(x[s] - xr) <= 1e-12 * np.linalg.norm(xr)
        assert np.linalg.norm(mats[s] @ x[s] - b[s]) <= 1e-10 * np.linalg.norm(b[s])


def test_banded_lu_5x5_known_answer(reference_numbers):
    kat = reference_numbers["driver_algebra_5x5"]
    A = np.zeros((5, 5), np.complex128)
    for i, j, v in kat["entries_1based"]:
        A[i - 1, j - 1] = v
    ab = O.dense_to_lapack_band(A, 1, 2)[None].copy()
    ipiv, info = O.zgbtrf_batch(ab, 1, 2)
    x = np.array(kat["rhs"], np.complex128)[None].copy()
    O.zgbtrs_batch(ab, 1, 2, ipiv, x)
    np.testing.assert_allclose(x[0].real, [-1, 2, -3, 4, -5], atol=1e-13)
    assert ipiv[0, 0] == 2  # A(1,1) = 0 forces the first interchange


# --- transforms ------------------------------------------------------------------

def test_transform_normalisation_single_mode():
    geo = O.Geometry(8, 8, 9, 2.0, 3.0)
    x = np.arange(geo.NXAA) * geo.Lx / geo.NXAA
    y = np.arange(geo.NYAA) * geo.Ly / geo.NYAA
    z = geo.gauss_chebyshev_grid()
    t = (z - geo.center) * 2.0 / geo.gap
    for n in (0, 3):
        for kxi, kyi in ((0, 0), (2, 0), (1, -3)):
            kx, ky = kxi * geo.kx_box, kyi * geo.ky_box
            P = (np.cos(kx * x)[:, None, None] * np.cos(ky * y)[None, :, None] if kyi else
                 np.cos(kx * x)[:, None, None] * np.ones_like(y)[None, :, None]) * C.chebval(t, np.eye(n + 1)[n])[None, None, :]
            S = O.phys_to_spec(P, geo)
            expected = np.zeros_like(S)
            iy = kyi % geo.NYAA
            if kxi == 0 and kyi == 0:
                expected[0, 0, n] = 1.0
            elif kyi == 0:
                expected[kxi, 0, n] = 0.5
            else:
                # cos(a)cos(b) = [cos(a+b) + cos(a-b)]/2 -> coefficients 1/4 at (kx, ky), (kx, -ky)
                expected[kxi, iy, n] = 0.25
                expected[kxi, (-kyi) % geo.NYAA, n] = 0.25
            np.testing.assert_allclose(S, expected, atol=1e-14)


def _random_spectral(geo, rng, hermitian=True):
    S = rng.normal(size=(geo.nkx, geo.nky, geo.NZ)) + 1j * rng.normal(size=(geo.nkx, geo.nky, geo.NZ))
    S[geo.dealias] = 0.0
    if hermitian:
        S[0, 0, :] = S[0, 0, :].real
        for iy in range(geo.nky):
            k = geo.ky_int[iy]
            if k < 0 and not geo.dealias_y[iy]:
                S[0, iy, :] = np.conj(S[0, list(geo.ky_int).index(-k), :])
    return S


def test_transform_round_trip():
    geo = O.Geometry(12, 8, 7, 2.0, 1.0)
    S = _random_spectral(geo, np.random.default_rng(4))
    np.testing.assert_allclose(O.phys_to_spec(O.spec_to_phys(S, geo), geo), S, atol=1e-13)


def test_fp32_transform_chain_is_single_precision():
    """The oracle's FP32 transform chain (the yardstick of the mixed-precision
    step): the brute-force dealiased product of the test below, computed with
    complex64 spectra, stays within a few FP32 ulps of the exact product, well
    above FP64 rounding (so it really ran in single precision), and its round
    trip returns the spectra to FP32 accuracy."""
    geo = O.Geometry(4, 4, 5, 2 * np.pi, 2 * np.pi)
    rng = np.random.default_rng(5)
    A, B = _random_spectral(geo, rng), _random_spectral(geo, rng)
    exact = O.phys_to_spec(O.spec_to_phys(A, geo) * O.spec_to_phys(B, geo), geo)
    pa, pb = O.spec_to_phys(A.astype(np.complex64), geo), O.spec_to_phys(B.astype(np.complex64), geo)
    assert pa.dtype == np.float32
    got = O.phys_to_spec(pa * pb, geo)
    assert got.dtype == np.complex64
    keep = ~geo.dealias
    err = np.linalg.norm(got[keep] - exact[keep]) / np.linalg.norm(exact[keep])
    assert 1e-9 < err < 1e-6
    S = _random_spectral(geo, np.random.default_rng(4)).astype(np.complex64)
    rt = O.phys_to_spec(O.spec_to_phys(S, geo), geo)
    assert np.linalg.norm(rt[keep] - S[keep]) / np.linalg.norm(S[keep]) < 1e-6


def test_dealiased_product_equals_truncated_convolution():
    """Brute force on a tiny grid: the 3/2-padded pseudo-spectral product of two
    band-limited fields equals their exact product truncated to the kept modes."""
    geo = O.Geometry(4, 4, 5, 2 * np.pi, 2 * np.pi)
    rng = np.random.default_rng(5)
    A, B = _random_spectral(geo, rng), _random_spectral(geo, rng)
    got = O.phys_to_spec(O.spec_to_phys(A, geo) * O.spec_to_phys(B, geo), geo)

    def full(S):  # dict over all (kx, ky) with Hermitian extension
        out = {}
        for ix in range(geo.nkx):
            for iy in range(geo.nky):
                if geo.dealias[ix, iy]:
                    continue
                kx, ky = int(geo.kx_int[ix]), int(geo.ky_int[iy])
                out[(kx, ky)] = S[ix, iy]
                if kx > 0:
                    out[(-kx, -ky)] = np.conj(S[ix, iy])
        return out

    fa, fb = full(A), full(B)
    exp = np.zeros_like(got)
    for (ax, ay), ca in fa.items():
        for (bx, by), cb in fb.items():
            kx, ky = ax + bx, ay + by
            if kx < 0 or kx >= geo.nkx or 2 * abs(ky) >= geo.NY:
                continue
            iy = ky % geo.NYAA
            # T_m T_n = (T_{m+n} + T_{|m-n|}) / 2
            prod = np.zeros(2 * geo.NZ, np.complex128)
            for m in range(geo.NZ):
                for n in range(geo.NZ):
                    v = ca[m] * cb[n] / 2.0
                    prod[m + n] += v
                    prod[abs(m - n)] += v
            exp[kx, iy] += prod[: geo.NZ]
    keep = ~geo.dealias   # the reference zeroes the 2/3-rule modes of every NL term
    np.testing.assert_allclose(got[keep], exp[keep], atol=1e-12)


# --- seeded initial condition ---------------------------------------------------

def test_initial_condition_properties():
    r = recipe("rbc_vanilla.coral.equations")
    k1, z1 = ic.rbc_initial_state(r, 32, 32, 17, 1e-3, 8, 8, 11)
    k2, _ = ic.rbc_initial_state(r, 32, 32, 17, 1e-3, 8, 8, 11)
    np.testing.assert_array_equal(k1[0], k2[0])
    s = k1[0]
    kx, ky, kept = ic.wavenumber_indices(32, 32)
    assert not s[~kept].any() and not s[0, 0].any()
    for iy in range(ky.size):
        if ky[iy] < 0 and kept[0, iy]:
            np.testing.assert_array_equal(s[0, iy], np.conj(s[0, list(ky).index(-ky[iy])]))
    norm = math.sqrt(np.sum(np.abs(s[0]) ** 2) + 2 * np.sum(np.abs(s[1:]) ** 2))
    assert abs(norm - 1e-3) < 1e-15
    assert all(not z.any() for z in z1)


def test_kx0_projection_keeps_physical_fields():
    """The Hermitian projection of the kx = 0 plane (OracleSolver.project_kx0)
    leaves the physical fields unchanged and the plane exactly Hermitian: the
    anti-Hermitian part it removes is invisible to the c2r transform."""
    r = recipe("rbc_vanilla.coral.equations")
    NX, NY, NZ = 16, 12, 13
    geo = O.Geometry(NX, NY, NZ, 2 * np.pi, 3.0)
    ora = O.OracleSolver(r, geo, "ARS_222")
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 6, 6, 3)
    rng = np.random.default_rng(8)
    k0[0][0] += 1e-2 * (rng.normal(size=k0[0][0].shape) + 1j * rng.normal(size=k0[0][0].shape))
    k0[0][geo.dealias] = 0.0
    k0[0][0, 0] = 0.0
    ora.field, ora.zfield = [k0[0].copy()], [a.copy() for a in z0]
    before = ora.physical_full_variables()
    f = O.project_kx0_plane(k0[0].copy(), geo)
    ora.field = [f]
    after = ora.physical_full_variables()
    for a, b in zip(after, before):
        np.testing.assert_allclose(a, b, atol=1e-15 * max(1.0, np.abs(b).max()))
    ky = geo.ky_int
    for iy in range(geo.nky):
        if ky[iy] > 0 and not geo.dealias[0, iy]:
            jy = int(np.nonzero(ky == -ky[iy])[0][0])
            assert np.array_equal(f[0, jy], np.conj(f[0, iy]))
    assert not np.allclose(f[0], k0[0][0])


# --- boundary values and vertical decompositions of full variables -----------
@pytest.mark.parametrize("NZ", [17, 16])
def test_full_variable_boundary_values_and_decompositions(NZ):
    """The oracle's boundary_value / keep_z_integral_only / remove_z_integral
    against the mathematics of a Chebyshev series: the extracted row 0 is the
    series' value at z = -1 (BotSurf) or z = +1 (TopSurf) (numpy chebval);
    keeping only the z-integral leaves the series' mean over [-1, 1] (numpy
    chebint), removing it leaves a series of mean zero, for series whose
    degree stays below 2 (NZ/2 - 1), where the reference's loops end."""
    from numpy.polynomial import chebyshev as C
    rng = np.random.default_rng(NZ)
    top = 2 * (NZ // 2 - 1)
    c = np.zeros((3, NZ), complex)
    c[:, :top + 1] = rng.normal(size=(3, top + 1)) + 1j * rng.normal(size=(3, top + 1))

    def mean(cf):
        i = C.chebint(cf)
        return (C.chebval(1.0, i) - C.chebval(-1.0, i)) / 2.0
    for where, z in (("BotSurf", -1.0), ("TopSurf", 1.0)):
        full = rng.normal(size=(3, NZ)) + 1j * rng.normal(size=(3, NZ))   # every row, the loops run to NZ - 1
        s = O.boundary_value(full.copy(), NZ, where)
        for k in range(3):
            assert s[k, 0] == pytest.approx(C.chebval(z, full[k]), rel=1e-13)
        assert not s[:, 1:].any()
    s = O.keep_z_integral_only(c.copy(), NZ)
    for k in range(3):
        assert s[k, 0] == pytest.approx(mean(c[k]), rel=1e-13)
    assert not s[:, 1:top + 1].any()
    s = O.remove_z_integral(c.copy(), NZ)
    for k in range(3):
        assert abs(mean(s[k])) < 1e-13 * np.abs(c[k]).max()
        np.testing.assert_array_equal(s[k, 1:], c[k, 1:])


def test_parser_reads_boundary_values_and_decompositions():
    from conftest import full_filter_recipe
    from coral_amd.equations import EquationsError, parse_equations
    r = full_filter_recipe("BotSurf", "barotropic")
    assert r.full_vars[-2].extract == "BotSurf" and r.full_vars[-1].zdecomp == "barotropic"
    assert [len(f.terms) for f in r.full_vars[-2:]] == [2, 2]
    with pytest.raises(EquationsError):
        parse_equations(">>linear_variable_full :: a <<\n>>vertical_decomposition :: baroclinicX <<\n")
