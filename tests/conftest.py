import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
INPUTS = os.path.join(GOLDEN, "inputs")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: takes more than ~20 s")


def _ensure_oracle_built():
    so = os.path.join(ROOT, "oracle", "_build", "liboracle_banded.so")
    src = os.path.join(ROOT, "oracle", "banded.c")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def _ensure_library_built():
    so = os.path.join(ROOT, "coral_amd", "lib", "libcoral_amd.so")
    if not os.path.exists(so):
        from coral_amd import build
        build.build()


_ensure_oracle_built()


@pytest.fixture(scope="session")
def reference_numbers():
    with open(os.path.join(GOLDEN, "reference_numbers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def native_lib():
    _ensure_library_built()
    from coral_amd import _lib
    return _lib


def input_path(name):
    return os.path.join(INPUTS, name)


def recipe(name):
    from coral_amd.equations import read_equations
    return read_equations(input_path(name))


def product_recipe(name):
    from coral_amd.equations import read_equations
    return read_equations(os.path.join(ROOT, "coral_amd", "inputs", name))


def seeded_state(r, NX, NY, NZ, kmax, nmodes, seed, amps=(("u", 1.0), ("v", 1.0), ("w", 1.0), ("t", 1e-3))):
    """Seeded perturbations of several variables of the first kxky system at
    the given amplitudes (by default the velocities at unit amplitude and the
    temperature a thousand times smaller, as in convection, where velocities
    in thermal-diffusion units are large); mean fields zero."""
    from coral_amd import ic
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 0.0, kmax, nmodes, seed)
    s = r.kxky_systems[0]
    for i, (name, amp) in enumerate(amps):
        k0[0] = k0[0] + ic.seeded_perturbation(NX, NY, NZ, s.bc_codes, s.var_names.index(name), amp,
                                               kmax, nmodes, seed + 17 * i)
    return k0, z0


def full_filter_recipe(extract="TopSurf", zdecomp="baroclinic"):
    """The RBC recipe (coral_amd/inputs/coral.equations.rbc_ra1e5) with two
    more full variables: tX, the boundary value (extract: TopSurf or BotSurf)
    of tMean + t, and uZ, uMean + u under the given vertical decomposition
    (barotropic or baroclinic); two quadratic variables built on them
    (wFull.tX, uZ.wFull) enter the temperature equation with a factor 0.1."""
    from coral_amd.equations import parse_equations
    text = open(os.path.join(ROOT, "coral_amd", "inputs", "coral.equations.rbc_ra1e5")).read()
    key = "top" if extract == "TopSurf" else "bottom"
    text = text.replace(">>Param :: RayleighPr := 1.d5 <<",
                        ">>Param :: RayleighPr := 1.d5 <<\n>>Param :: small := 1.d-1 <<")
    text = text.replace(">>quadratic_variable :: uu := uFull.uFull <<",
                        f">>linear_variable_{key}_boundary_value :: tX <<\n"
                        ">>linear_variable_full_build :: + tMean <<\n>>linear_variable_full_build :: + t <<\n"
                        ">>====================================<<\n"
                        ">>linear_variable_full       :: uZ <<\n"
                        f">>vertical_decomposition :: {zdecomp} <<\n"
                        ">>linear_variable_full_build :: + uMean <<\n>>linear_variable_full_build :: + u <<\n"
                        ">>====================================<<\n"
                        ">>quadratic_variable :: uu := uFull.uFull <<")
    text = text.replace(">>quadratic_variable :: tw := tFull.wFull <<",
                        ">>quadratic_variable :: tw := tFull.wFull <<\n>>quadratic_variable :: wtX := wFull.tX <<\n"
                        ">>quadratic_variable :: uZw := uZ.wFull <<")
    text = text.replace(">>add_rhs_NL     :: - Iz.Iz.dx.tu - Iz.Iz.dy.tv - Iz.tw <<",
                        ">>add_rhs_NL     :: - Iz.Iz.dx.tu - Iz.Iz.dy.tv - Iz.tw <<\n"
                        ">>add_rhs_NL     :: - small.Iz.wtX - small.Iz.Iz.dx.uZw <<")
    r = parse_equations(text)
    assert [f.name for f in r.full_vars][-2:] == ["tX", "uZ"]
    assert r.full_vars[-2].extract == extract and r.full_vars[-1].zdecomp == zdecomp
    return r


def smagorinsky_recipe(prefactor_param="0.3d0"):
    """The RBC recipe (coral_amd/inputs/coral.equations.rbc_ra1e5) with the
    horizontal velocity gradients as full variables (dxUfull, dyUfull,
    dxVfull, dyVfull, the names the reference's Smagorinsky code reads), read
    with smagorinsky=True (nuSmago added), and an eddy-viscosity term
    nu dx(nuSmago dxUfull) + nu dy(nuSmago dyUfull) in the u equation and the
    like in the v equation, nu = 0.3."""
    from coral_amd.equations import parse_equations
    text = open(os.path.join(ROOT, "coral_amd", "inputs", "coral.equations.rbc_ra1e5")).read()
    text = text.replace(">>Param :: RayleighPr := 1.d5 <<",
                        f">>Param :: RayleighPr := 1.d5 <<\n>>Param :: nuT := {prefactor_param} <<")
    grads = "".join(f">>linear_variable_full :: d{a}{b.upper()}full <<\n>>linear_variable_full_build :: + d{a}.{b} <<\n"
                    for b in ("u", "v") for a in ("x", "y"))
    text = text.replace(">>quadratic_variable :: uu := uFull.uFull <<",
                        grads + ">>quadratic_variable :: uu := uFull.uFull <<")
    text = text.replace(">>quadratic_variable :: tw := tFull.wFull <<",
                        ">>quadratic_variable :: tw := tFull.wFull <<\n"
                        ">>quadratic_variable :: nxu := nuSmago.dxUfull <<\n"
                        ">>quadratic_variable :: nyu := nuSmago.dyUfull <<\n"
                        ">>quadratic_variable :: nxv := dxVfull.nuSmago <<\n"
                        ">>quadratic_variable :: nyv := nuSmago.dyVfull <<\n"
                        ">>quadratic_variable :: nn := nuSmago.ONE <<")
    text = text.replace(">>add_rhs_NL     :: - Iz.Iz.dx.uu - Iz.Iz.dy.uv - Iz.uw <<",
                        ">>add_rhs_NL     :: - Iz.Iz.dx.uu - Iz.Iz.dy.uv - Iz.uw <<\n"
                        ">>add_rhs_NL     :: + nuT.Iz.Iz.dx.nxu + nuT.Iz.Iz.dy.nyu <<", 1)
    text = text.replace(">>add_rhs_NL     :: - Iz.Iz.dx.uv - Iz.Iz.dy.vv - Iz.vw <<",
                        ">>add_rhs_NL     :: - Iz.Iz.dx.uv - Iz.Iz.dy.vv - Iz.vw <<\n"
                        ">>add_rhs_NL     :: + nuT.Iz.Iz.dx.nxv + nuT.Iz.Iz.dy.nyv - nuT.Iz.Iz.nn <<", 1)
    r = parse_equations(text, smagorinsky=True)
    assert [f.name for f in r.full_vars][-5:] == ["dxUfull", "dyUfull", "dxVfull", "dyVfull", "nuSmago"]
    assert sum(len(s.nl) for s in r.kxky_systems) > 0
    return r
