"""GPU tests of the state files (QuickSave volumes), through the C ABI."""
import math
import os
import shutil

import numpy as np
import pytest

from conftest import product_recipe, recipe
from coral_amd import ic, state_io
from oracle import coral_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver_mod(native_lib):
    from coral_amd import solver
    return solver


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def test_quicksave_import_heat_equation(solver_mod, reference_numbers, tmp_path):
    """Restart files laid out as etc/benchmarks/00_heat_equation/generate_initial_condition.py
    writes them (zero kxky volumes, sine profiles of the mean variables), imported and
    checked against the oracle's projection and against the reference's error table."""
    r = recipe("heat_equation.coral.equations")
    setup = reference_numbers["heat_equation_setup"]
    table = reference_numbers["heat_equation_error_after_10_steps"]
    NX, NZ = 4, 128
    geo = O.Geometry(NX, NX, NZ, 1.0, 1.0)
    zg = geo.gauss_chebyshev_grid()
    restart = tmp_path / "Restart"
    restart.mkdir()
    zero = np.zeros((geo.NXAA, geo.NYAA, geo.NZAA))
    for isys, s in enumerate(r.kxky_systems):
        for ivar in range(s.nvars):
            zero.tofile(restart / state_io.quicksave_name("kxky", isys, ivar, 1))
    prof = {"u": setup["amplitudes"]["u"] * np.sin((2 * setup["modes"]["u"] + 1) * (zg - 1) * np.pi / 2),
            "v": setup["amplitudes"]["v"] * np.sin((2 * setup["modes"]["v"] + 1) * zg * np.pi / 2),
            "w": setup["amplitudes"]["w"] * np.sin(setup["modes"]["w"] * zg * np.pi)}
    names = ["u", "v", "w"]
    for isys, s in enumerate(r.zero_systems):
        prof[names[isys]].tofile(restart / state_io.quicksave_name("zero", isys, 0, 1))
    # the oracle's projection of the first profile (DCT-II of the Gauss values, stencil solve)
    ops = O.SystemOps(r.zero_systems[0], NZ, 0.5, 1.0)
    import scipy.fft as sfft
    c = sfft.dct(prof["u"], type=2)[:NZ] / geo.NZAA
    c[0] *= 0.5
    nc = ops.var_size[0]
    zref = np.zeros(ops.total)
    zref[ops.posP[:nc]] = np.linalg.solve(ops.stencils[0][:nc, :], c[:nc])
    n = setup["modes"]["u"]
    rate = setup["Prandtl"] * ((2 * n + 1) * np.pi / 2) ** 2
    idx = 3
    dt, ref = table["dt"][idx], table["ARS_222"][idx]
    with solver_mod.Solver(r, NX, NX, NZ, 1.0, 1.0, scheme="ARS_222") as g:
        assert state_io.import_quicksave(g, str(tmp_path)) == 1
        k, z = g.get_state()
        assert rel(z[0], zref) <= 1e-12
        for a in k:
            assert not np.any(a)
        p0 = g.physical(0)[0, 0, :].copy()
        assert rel(p0, prof["u"]) <= 1e-12
        g.march_forward(dt, 10)
        p1 = g.physical(0)[0, 0, :]
    err = math.sqrt(np.sum((p1 - p0 * np.exp(-rate * 10 * dt)) ** 2) / np.sum(p0 ** 2))
    assert abs(err / ref - 1.0) < 2e-3


def test_checkpoint_round_trip_and_consistency(solver_mod, tmp_path):
    name = "coral.equations.rrbc_ra5e7_ek1e-4"
    r = product_recipe(name)
    NX, NY, NZ = 32, 24, 17
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 40, 8, 4)
    rng = np.random.default_rng(4)
    # populate the mean variables too
    z0 = [rng.normal(size=a.shape) * 1e-3 for a in z0]
    with solver_mod.Solver(r, NX, NY, NZ, 2.0, 3.0) as g:
        g.set_state(k0, z0)
        g.factorize_operators(2e-5)
        g.march_forward(2e-5, 2)
        k1, z1 = g.get_state()
        paths = state_io.export_checkpoints(g, str(tmp_path), rolling=2)
        assert len(paths) == sum(s.nvars for s in r.kxky_systems + r.zero_systems)
        # a full variable is the sum of its kxky and mean parts
        for ifull, fv in enumerate(r.full_vars):
            total = np.zeros((g.NXAA, g.NYAA, g.NZAA))
            for t in fv.terms:
                if (t.kx, t.ky, t.dz, t.mz) != (0, 0, 0, 0):
                    break
                part = g.variable_physical(t.kind, t.system, t.position)
                total += t.dsca * (part if t.kind == "kxky" else part[None, None, :])
            else:
                assert rel(total, g.physical(ifull)) <= 1e-12
        # state unchanged by the export
        k2, z2 = g.get_state()
        assert all(np.array_equal(a, b) for a, b in zip(k1, k2))
    shutil.move(str(tmp_path / "CheckPoints"), str(tmp_path / "Restart"))
    with solver_mod.Solver(r, NX, NY, NZ, 2.0, 3.0) as h:
        assert state_io.import_quicksave(h, str(tmp_path)) == 2
        k3, z3 = h.get_state()
    for a, b in zip(k3, k1):
        assert rel(a, b) <= 1e-12
    for a, b in zip(z3, z1):
        assert rel(a, b) <= 1e-12


def test_horizontal_averages_match_oracle(solver_mod):
    """Horizontal averages of full and quadratic variables (the timeseries
    quantities) against the oracle's physical fields averaged over x and y."""
    r = product_recipe("coral.equations.rbc_ra1e5")
    NX, NY, NZ = 32, 24, 17
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 40, 8, 6)
    rng = np.random.default_rng(6)
    z0 = [rng.normal(size=a.shape) * 1e-2 for a in z0]
    geo = O.Geometry(NX, NY, NZ, 2.0, 3.0)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    phys = ora.physical_full_variables()
    full_ref = [p.mean(axis=(0, 1)) for p in phys]
    quad_ref = []
    for q in r.quad_vars:
        a = phys[q.ivar1] if q.ivar1 >= 0 else 1.0
        b = phys[q.ivar2] if q.ivar2 >= 0 else 1.0
        quad_ref.append(np.broadcast_to(a * b, phys[0].shape).mean(axis=(0, 1)))
    with solver_mod.Solver(r, NX, NY, NZ, 2.0, 3.0) as g:
        g.set_state(k0, z0)
        full, quad = g.horizontal_averages()
    for got, ref in zip(full, full_ref):
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    for got, ref in zip(quad, quad_ref):
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


class _OracleAsSolver:
    """The oracle behind the solver's time-stepping interface."""

    def __init__(self, ora):
        self.ora = ora

    def factorize_operators(self, dt):
        self.ora.factorize(dt)

    def march_forward(self, dt, n):
        for _ in range(n):
            self.ora.step(dt)

    def cfl_dt(self):
        return self.ora.cfl_dt


def test_adaptive_time_step_matches_oracle(solver_mod):
    """CFL-driven dt with tolerance-triggered refactorization (timeKeeping.f90),
    the same rules driving the GPU solver and the oracle."""
    from coral_amd.timekeeping import Timings, march_adaptive
    r = product_recipe("coral.equations.rrbc_ra5e7_ek1e-4")
    NX, NY, NZ = 32, 32, 17
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 10, 8, 9)
    geo = O.Geometry(NX, NY, NZ, 2 * np.pi, 2 * np.pi)
    ora = O.OracleSolver(r, geo, "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    t_ora = Timings(dt_max=4e-5, tolerance_factor=0.3, CFL_C=0.1)
    used_ora = march_adaptive(_OracleAsSolver(ora), t_ora, 24)
    with solver_mod.Solver(r, NX, NY, NZ, 2 * np.pi, 2 * np.pi) as g:
        g.set_state(k0, z0)
        t_gpu = Timings(dt_max=4e-5, tolerance_factor=0.3, CFL_C=0.1)
        used_gpu = march_adaptive(g, t_gpu, 24)
        k, z = g.get_state()
    # the same steps change dt (CFL maxima agree to round-off, far from the 30 % threshold)
    change = lambda u: [i for i in range(1, len(u)) if u[i] != u[i - 1]]
    assert change(used_gpu) == change(used_ora)
    assert np.allclose(used_gpu, used_ora, rtol=1e-10, atol=0.0)
    assert len(change(used_gpu)) >= 2, "the run should refactorize at least twice"
    assert rel(k[0], ora.field[0]) <= 1e-10
    assert abs(t_gpu.absolute_time / t_ora.absolute_time - 1.0) <= 1e-10


def test_user_outputs_from_the_device_state(solver_mod, tmp_path):
    """coral.usrOutput objects written from a GPU run's state: a volume of a
    full variable holds the oracle's physical field of the same state (in the
    reference's z-fastest order), a quadratic volume the product of its full
    variables, a zFiltered quadratic the back transform of its filtered
    spectrum, a profile the x-y mean, a z slice the level; the
    Geometry records carry the run's sizes."""
    from coral_amd import output as OUT
    from conftest import seeded_state
    r = product_recipe("coral.equations.rbc_ra1e5")
    r.quad_vars[8].remove_z_integral = True      # tw, filtered
    NX, NZ = 16, 17
    k0, z0 = seeded_state(r, NX, NX, NZ, 4, 4, 7, amps=(("u", 1.0), ("v", 1.0), ("w", 1.0), ("t", 0.1)))
    text = (">>Volume :: tFull ;; Period= 1 <<\n>>Volume :: uw ;; Period= 1 <<\n>>Volume :: tw ;; Period= 1 <<\n"
            ">>Profile:: wFull ;; Period= 1 <<\n>>ZSlice :: tFull ;; Period= 1 ;; Position= 3 <<\n>>EOF<<\n")
    with solver_mod.Solver(r, NX, NX, NZ, 2 * math.pi, 2 * math.pi) as g:
        g.set_state(k0, z0)
        objs = OUT.parse_usr_output(text, r, g.NXAA, g.NYAA, g.NZAA)
        out = OUT.UserOutput(g, objs, str(tmp_path))
        files = out.before_step()
        OUT.write_geometry(g, str(tmp_path))
        full = [g.physical(i) for i in range(4)]
        tw_ref = g.bwd_transform(g.products()[8])
    assert len(files) == 5
    ora = O.OracleSolver(r, O.Geometry(NX, NX, NZ, 2 * math.pi, 2 * math.pi), "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    ref = ora.physical_full_variables()
    NXAA, NZAA = 3 * NX // 2, 3 * NZ // 2

    def vol(name):
        return np.fromfile(os.path.join(tmp_path, "Volumes", name), "<f8").reshape(NXAA, NXAA, NZAA)   # [x][y][z]
    assert rel(vol("linear_var04_time00000001_full.dat"), ref[3]) < 1e-12
    assert np.array_equal(vol("quadra_var03_time00000001_full.dat"), full[0] * full[2])
    tw = vol("quadra_var09_time00000001_full.dat")
    assert np.array_equal(tw, tw_ref) and np.abs(tw).max() > 0
    assert not np.allclose(tw, full[3] * full[2])        # the filtered product, not the raw one
    prof = np.fromfile(os.path.join(tmp_path, "Profiles", "linear_var03_time00000001_full.dat"), "<f8")
    assert rel(prof, full[2].mean(axis=(0, 1))) < 1e-14
    zs = np.fromfile(os.path.join(tmp_path, "Slices", "linear_var04_z00003_time00000001_full.dat"), "<f8")
    assert np.array_equal(zs, full[3][:, :, 2].ravel())
    ints = np.fromfile(os.path.join(tmp_path, "Geometry", "domDecmp.core0000"), "<i4")
    assert list(ints[:6]) == [NX, NX, NZ, NXAA, NXAA, NZAA]
