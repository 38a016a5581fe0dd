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
