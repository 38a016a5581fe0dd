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


@pytest.mark.parametrize("form", ["velocityPressure", "toroidalPoloidal"])
def test_published_rrbc_nusselt(solver_mod, reference_numbers, form):
    """Coral's published rotating convection case (Julien et al. 1996, third
    no-slip case; etc/benchmarks/02_RRBC_julien_JFM96: 96x96x48 modes, box
    2x2x1, Pr = 1, Ra = 2.81e5, 1/Ek = 707.1, ARS_222, dt_max 1e-3, tolerance
    0.3, CFL 0.2) stepped on the GPU with the reference's adaptive time step
    from its noisy initial condition (seeded phases, amplitude 1e-3), 40,000
    steps (0.55 diffusive times).  The horizontal mean temperature at the
    first and last Gauss points is sampled before every step, as the
    reference's timeseries are, and compute_nusselt.py's quarter averages
    (coral_amd.diagnostics.nusselt_quarters) give the Nusselt number, which
    must lie within 1 % of the published value (4.76905 velocity-pressure,
    4.77566 toroidal-poloidal; benchmark_02_RRBC_julien96.ipynb:97,114).
    The kx = 0 plane is projected on its Hermitian part after every step: its
    anti-Hermitian part, invisible to the fields, otherwise grows at the
    linear growth rate of the ky = +-3 modes and ends the run near step
    15,000."""
    from conftest import input_path
    from coral_amd import diagnostics
    from coral_amd.equations import read_equations
    from coral_amd.params import read_parameters
    from coral_amd.timekeeping import Timings
    r = read_equations(input_path("rrbc_julien.coral.equations." + form))
    p = read_parameters(input_path("rrbc_julien.coral.parameters.in"))
    k0, z0 = ic.noisy_initial_state(r, p.NX, p.NY, p.NZ, p.Lx, p.Ly, 1e-3, 1e-3, 1)
    tsys = [i for i, s in enumerate(r.zero_systems) if "tMean" in s.var_names][0]
    tvar = r.zero_systems[tsys].var_names.index("tMean")
    nsteps = 40000
    times, ttop, tbot = [], [], []
    with solver_mod.Solver(r, p.NX, p.NY, p.NZ, p.Lx, p.Ly, p.gap, p.center, p.scheme) as g:
        g.use_graphs(True)
        g.set_kx0_projection(True)
        g.set_state(k0, z0)
        tk = Timings(p.dt_max, p.tolerance, p.cfl_c)
        for i in range(nsteps):
            prof = g.variable_physical("zero", tsys, tvar)
            assert np.all(np.isfinite(prof)), f"non-finite mean temperature at step {i}"
            times.append(tk.absolute_time)
            ttop.append(prof[0])
            tbot.append(prof[-1])
            tk.check_for_update()
            if tk.refactorization_is_due:
                g.factorize_operators(tk.dt_current)
                tk.refactorization_is_due = False
            g.march_forward(tk.dt_current, 1)
            tk.apply_cfl(g.cfl_dt())
            tk.increment()
            if (i + 1) % 10000 == 0:
                print(f"{form}: step {i + 1}, t = {tk.absolute_time:.4f}, dt = {tk.dt_current:.3e}", flush=True)
    quarters, nu, spread = diagnostics.nusselt_quarters(np.array(times), np.array(ttop), np.array(tbot),
                                                        3 * p.NZ // 2, p.gap)
    pub, pub_spread = reference_numbers["rrbc_nusselt"]["velocity_pressure" if form == "velocityPressure"
                                                        else "toroidal_poloidal"]
    print(f"{form}: quarters {quarters}", flush=True)
    print(f"{form}: Nu = {nu:.5f} +- {spread:.5f} (published {pub} +- {pub_spread}), "
          f"{100 * abs(nu / pub - 1):.2f} % off", flush=True)
    assert math.isfinite(nu)
    assert abs(nu / pub - 1.0) <= 0.01
