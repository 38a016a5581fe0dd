"""Pins of the HIP time stepper against numbers the reference publishes,
reached by time stepping on the GPU (through the C ABI): the stress-free
Rayleigh-Benard growth rate (etc/benchmarks/01_Rayleigh_Benard_stressFree)."""
import math

import numpy as np
import pytest

from conftest import recipe
from coral_amd import ic
from test_oracle_pins import stability_rate, stepped_growth_rate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver_mod(native_lib):
    from coral_amd import solver
    return solver


@pytest.mark.parametrize("scheme,dt,n0,n1", [("ARS_222", 1e-3, 400, 800), ("ARS_443", 2e-3, 200, 400)])
def test_stressfree_growth_rate_on_gpu(solver_mod, reference_numbers, scheme, dt, n0, n1):
    """The stress-free benchmark (Ra = 2000, Pr = 7) stepped by the HIP path in
    a box whose first x wavenumber is the notebook's k_opt: the temperature
    perturbation of that wavenumber grows at the rate the scheme's stability
    function gives for the published 24.137487 (benchmark_01...ipynb:68)."""
    ref = reference_numbers["stressfree_rb_max_growth_rate"]
    r = recipe("stressfree_rb.coral.equations")
    L = 2 * math.pi / ref["k_opt"]
    NX, NZ = 16, 25
    k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-30, 1, 8, 5, var_name="t")
    with solver_mod.Solver(r, NX, NX, NZ, L, L, scheme=scheme) as g:
        g.set_state(k0, z0)
        g.factorize_operators(dt)
        rate = stepped_growth_rate(lambda: g.march_forward(dt, 1),
                                   lambda: np.linalg.norm(g.get_field("kxky", 0)[1, 0]), dt, n0, n1)
    expect = stability_rate(scheme, dt, ref["s_max"])
    print(scheme, "stepped", rate, "stability function", expect, flush=True)
    assert abs(rate / expect - 1.0) < 1e-10
