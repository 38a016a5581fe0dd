"""Penalisation variables and the horizontal Smagorinsky viscosity on the
CPU: the parser on the reference's penalised recipe and with the
--withHorizontalSmagorinsky flag, the mask field against its definition
point by point, and the oracle's nuSmago profile and penalised products
against what the mathematics fixes (Parseval's identity for the horizontal
mean of squares, linearity for a constant mask).  No GPU needed."""
import math

import numpy as np
import pytest

from conftest import recipe, seeded_state, smagorinsky_recipe
from coral_amd import penalisation as PN
from coral_amd.equations import EquationsError, parse_equations
from oracle import coral_oracle as O


def test_parser_reads_the_penalised_recipe():
    r = recipe("rbc_penalised.coral.equations")
    g = r.full_vars[4]
    assert (g.name, g.penalisation, g.penalisation_strength, g.penalisation_width, g.terms) == \
        ("GAMMA", True, 1e3, 1e2, [])
    assert [(q.name, q.ivar1, q.ivar2) for q in r.quad_vars[-3:]] == [("Gu", 4, 0), ("Gv", 4, 1), ("Gw", 4, 2)]
    assert not r.smagorinsky and not any(f.smagorinsky for f in r.full_vars)
    with pytest.raises(EquationsError):
        parse_equations(">>linear_variable_full :: f <<\n>>penalisation_width :: 2. <<\n>>EOF<<\n")


def test_smagorinsky_flag_adds_nusmago_before_the_quadratic_variables():
    """PL_equations.f90:519-525: nuSmago joins the full variables at the first
    quadratic_variable line, so quadratic variables can name it."""
    r = smagorinsky_recipe()
    assert r.smagorinsky and r.full_vars[-1].name == "nuSmago" and r.full_vars[-1].smagorinsky
    assert r.full_vars[-1].terms == []
    with pytest.raises(EquationsError):
        # without the flag nuSmago is no variable
        text = open(__import__("conftest").input_path("rbc_penalised.coral.equations")).read()
        parse_equations(text.replace(">>quadratic_variable :: Gu := GAMMA.uFull <<",
                                     ">>quadratic_variable :: Gu := nuSmago.uFull <<"))


def test_mask_field_follows_the_reference_loop():
    """init_penalisation (marching.f90:778-822) restated with explicit loops
    over 1-based indices: x = (ix + 1) Lx / NXAA, y = (iy + 1) Ly / NYAA (the
    reference's xGrid / yGrid), z its Gauss grid, pm = max of the shipped
    masks, strength (1 + tanh(width pm)) / 2."""
    NXAA, NYAA, NZAA, Lx, Ly, c, gap = 6, 9, 5, 2.0, 3.0, 0.5, 1.0
    s, w = 7.0, 3.0
    f = PN.mask_field(PN.SHIPPED, s, w, NXAA, NYAA, NZAA, Lx, Ly, c, gap)
    pi = 3.14159265358979311599796346854
    for ix in range(1, NXAA + 1):
        for iy in range(1, NYAA + 1):
            for iz in range(1, NZAA + 1):
                x, y = (1 + ix) * Lx / NXAA, (1 + iy) * Ly / NYAA
                z = c + math.cos(pi * (2 * iz - 1) / (2 * NZAA)) * 0.5 * gap
                m1 = z - (0.9 + 0.1 * math.cos(5 * (x + y) * 2 * pi / Lx))
                m2 = 0.4 * math.exp(-((6.0 + math.cos(2 * pi * x / Lx)) ** 2)) / math.exp(-25.0) - z
                want = s * (1.0 + math.tanh(w * max(m1, m2))) / 2.0
                assert f[ix - 1, iy - 1, iz - 1] == pytest.approx(want, rel=1e-14, abs=1e-14)


def _horizontal_mean_of_square(spec, g):
    """Parseval from the spectrum alone: at each Gauss point the Chebyshev
    series (numpy chebval) gives the horizontal coefficients F(kx, ky), and the
    x-y mean of f^2 is sum |H0(ky)|^2 + 2 sum_{kx>0} |F(kx, ky)|^2, H0 the
    Hermitian part of the kx = 0 plane (the part a real field keeps)."""
    from numpy.polynomial import chebyshev as C
    x = np.cos(np.pi * (2 * np.arange(g.NZAA) + 1) / (2 * g.NZAA))
    F = C.chebval(x, np.moveaxis(spec, 2, 0))          # (nkx, NYAA, NZAA)
    neg = (-np.arange(g.NYAA)) % g.NYAA                # slot of -ky (FFT order)
    H0 = (F[0] + np.conj(F[0][neg])) / 2
    return np.sum(np.abs(H0) ** 2, axis=0) + 2 * np.sum(np.abs(F[1:]) ** 2, axis=(0, 1))


def test_oracle_smagorinsky_profile_is_the_mean_of_squares():
    """One stage: nuSmago = pf sqrt(0 + <dxU^2 + dyU^2 + dxV^2 + dyV^2>_xy),
    the mean taken by Parseval from the planes' 2-D spectra; a second stage
    on the same state: pf sqrt(previous + the same mean) (the reference's
    nuSmago array is not reset between calls)."""
    r = smagorinsky_recipe()
    NX, NY, NZ = 8, 8, 9
    k0, z0 = seeded_state(r, NX, NY, NZ, 3, 3, 5, amps=(("u", 1e-1), ("v", 1e-1), ("w", 1e-1), ("t", 1e-2)))
    g = O.Geometry(NX, NY, NZ, 2 * np.pi, 3.0)
    s = O.OracleSolver(r, g)
    s.smagorinsky_prefactor = 0.7
    specs = s.full_variables_spec(k0, z0)
    names = [f.name for f in r.full_vars]
    S = sum(_horizontal_mean_of_square(specs[names.index(n)], g) for n in ("dxUfull", "dyUfull", "dxVfull", "dyVfull"))
    assert S.max() > 0
    s.quad_spectra(k0, z0)
    np.testing.assert_allclose(s.nu_prof, 0.7 * np.sqrt(S), rtol=1e-12)
    p1 = s.nu_prof.copy()
    s.quad_spectra(k0, z0, advance=False)
    np.testing.assert_array_equal(s.nu_prof, p1)
    s.quad_spectra(k0, z0)
    np.testing.assert_allclose(s.nu_prof, 0.7 * np.sqrt(p1 + S), rtol=1e-12)
    # the CFL bound of marching.f90:241-253
    bound = 0.95 / ((2 * 3.141592) ** 2 * s.nu_prof.max() * ((g.NXAA / g.Lx) ** 2 + (g.NYAA / g.Ly) ** 2))
    assert s.cfl_dt <= bound


def test_oracle_constant_mask_scales_the_products():
    """A mask field that is the constant c makes Gu = c uFull: its spectrum is
    c times uFull's (products formed with the fixed field, not with the
    variable's own empty spectrum)."""
    r = recipe("rbc_penalised.coral.equations")
    NX, NY, NZ = 8, 8, 9
    k0, z0 = seeded_state(r, NX, NY, NZ, 3, 3, 4, amps=(("u", 1e-1), ("v", 1e-1), ("w", 1e-1), ("t", 1e-2)))
    g = O.Geometry(NX, NY, NZ, 2 * np.pi, 2 * np.pi)
    s = O.OracleSolver(r, g)
    s.set_fixed_physical(4, np.full((g.NXAA, g.NYAA, g.NZAA), 3.5))
    q = s.quad_spectra(k0, z0)
    specs = s.full_variables_spec(k0, z0)
    u = O.phys_to_spec(O.spec_to_phys(specs[0], g), g)
    names = [x.name for x in r.quad_vars]
    np.testing.assert_allclose(q[names.index("Gu")], 3.5 * u, rtol=1e-12, atol=1e-15 * np.abs(u).max())
    assert np.array_equal(s.physical_full_variables()[4], np.full((g.NXAA, g.NYAA, g.NZAA), 3.5))
