"""The closure network's CPU reference (oracle/closure_ref.py) against finite
differences of its own loss, and the Fourier basis (no GPU)."""
import numpy as np
import pytest

from oracle import closure_ref as C


def small_problem(seed=0, P=2, N2=16, nb=5, cin=3, cout=2, width=4, hidden=6, layers=2):
    rng = np.random.default_rng(seed)
    p = {"W0": rng.normal(size=(cin, width)) * 0.5, "b0": rng.normal(size=width) * 0.1,
         "W1": [rng.normal(size=(width, hidden)) * 0.5 for _ in range(layers)],
         "b1": [rng.normal(size=hidden) * 0.1 for _ in range(layers)],
         "W2": [rng.normal(size=(hidden, width)) * 0.5 for _ in range(layers)],
         "b2": [rng.normal(size=width) * 0.1 for _ in range(layers)],
         "D": [rng.normal(size=(nb, width)) * 0.3 for _ in range(layers)],
         "Wp": rng.normal(size=(width, cout)) * 0.5, "bp": rng.normal(size=cout) * 0.1}
    basis = np.linalg.qr(rng.normal(size=(N2, nb)))[0]
    X = rng.normal(size=(P * N2, cin))
    T = rng.normal(size=(P * N2, cout))
    return p, basis, X, T


def test_reference_gradients_match_finite_differences():
    """Central differences of the float64 loss (no BF16 rounding) against
    every parameter block's analytic gradient."""
    p, basis, X, T = small_problem()
    P = 2
    _, g, _ = C.loss_and_grads(p, X, T, basis, P, round_bf16=False)
    rng = np.random.default_rng(3)
    h = 1e-6
    for key in ("W0", "b0", "Wp", "bp", "W1", "b1", "W2", "b2", "D"):
        arrs = p[key] if isinstance(p[key], list) else [p[key]]
        grads = g[key] if isinstance(g[key], list) else [g[key]]
        for a, ga in zip(arrs, grads):
            for _ in range(3):
                idx = tuple(rng.integers(0, s) for s in a.shape)
                old = a[idx]
                a[idx] = old + h
                lp = C.loss_and_grads(p, X, T, basis, P, round_bf16=False)[0]
                a[idx] = old - h
                lm = C.loss_and_grads(p, X, T, basis, P, round_bf16=False)[0]
                a[idx] = old
                fd = (lp - lm) / (2 * h)
                assert fd == pytest.approx(ga[idx], rel=1e-6, abs=1e-9), (key, idx)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    r = C.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even (mantissa bit 0)
    assert r[2] == np.float32(1.0 + 2 ** -6)       # tie to even (rounds up)
    assert r[3] == -2.5
    assert np.isfinite(r[4])


def test_fourier_basis_is_orthonormal(native_lib):
    from coral_amd.closure import fourier_basis
    B = fourier_basis(24, 18, 4)
    assert B.shape == (24 * 18, 2 * (4 * 9 + 5) - 1)
    np.testing.assert_allclose(B.T @ B, np.eye(B.shape[1]), atol=1e-12)


def test_box_filter_is_exact_on_coarse_states(native_lib):
    """A coarse state embedded in the fine grid (same integer wavenumbers,
    same leading Galerkin coefficients of every variable) filters back to
    itself; the filter of a fine state zeroes nothing the coarse grid keeps
    that the fine state has, and quadratic spectra keep exactly the coarse
    wavenumbers and degrees."""
    from coral_amd import ic
    from coral_amd.closure import BoxFilter
    from coral_amd.equations import read_equations
    from conftest import ROOT
    import os
    r = read_equations(os.path.join(ROOT, "coral_amd", "inputs", "coral.equations.rbc_ra1e5"))
    fine, coarse = (32, 24, 17), (16, 12, 9)
    box = BoxFilter(r, fine, coarse, kind="sharp")
    kc, zc = ic.rbc_initial_state(r, *coarse, 1.0, 50, 50, 3)
    rng = np.random.default_rng(2)
    zc = [rng.normal(size=a.shape) for a in zc]
    # embed: inverse of the filter's index maps
    kf = [np.zeros((fine[0] // 2, 3 * fine[1] // 2, sum(fine[2] - bc // 10 for bc in s.bc_codes)), np.complex128)
          for s in r.kxky_systems]
    zf = [np.zeros(sum(fine[2] - bc // 10 for bc in s.bc_codes)) for s in r.zero_systems]
    for (pf, pc, _), a, b in zip(box.kpos, kf, kc):
        for i, ix in enumerate(box.ix):
            for j, iy in enumerate(box.iy):
                a[ix, iy, pf] = b[i, j, pc]
    for (pf, pc, _), a, b in zip(box.zpos, zf, zc):
        a[pf] = b[pc]
    k2, z2 = box.state(kf, zf)
    for a, b in zip(k2 + z2, kc + zc):
        assert np.array_equal(a, b)
    # embed is that embedding (dealiased coarse slots stay zero)
    ke, ze = box.embed(kc, zc)
    for a, b in zip(ke + ze, kf + zf):
        assert np.array_equal(a, b)
    q = rng.normal(size=(9, fine[0] // 2, 3 * fine[1] // 2, fine[2])) + 0j
    qc = box.spectra(q)
    assert qc.shape == (9, coarse[0] // 2, 3 * coarse[1] // 2, coarse[2])
    assert np.array_equal(qc[:, box.keep], q[:, box.ix][:, :, box.iy][..., :coarse[2]][:, box.keep])
    assert not qc[:, ~box.keep].any()


def test_tophat_filter_is_the_cell_average(native_lib):
    """The top-hat kind weights mode (kx, ky) by the average of exp(i (kx x +
    ky y)) over one coarse cell centred on the point, computed here by
    quadrature of the cell (sinc closed form pinned independently), and is
    otherwise the sharp filter: the mean passes unchanged, every kept mode is
    the sharp filter's times its weight."""
    from coral_amd import ic
    from coral_amd.closure import BoxFilter
    from coral_amd.equations import read_equations
    from conftest import ROOT
    import os
    r = read_equations(os.path.join(ROOT, "coral_amd", "inputs", "coral.equations.rbc_ra1e5"))
    fine, coarse = (32, 24, 17), (16, 12, 9)
    Lx, Ly = 2.0, 3.0
    top, sharp = BoxFilter(r, fine, coarse), BoxFilter(r, fine, coarse, kind="sharp")
    kxc, kyc, _ = ic.wavenumber_indices(coarse[0], coarse[1])
    dx, dy = Lx / coarse[0], Ly / coarse[1]
    s = (np.arange(4000) + 0.5) / 4000 - 0.5           # midpoint rule over the cell
    for i, kx in enumerate(kxc):
        for j, ky in enumerate(kyc):
            ax = np.mean(np.exp(1j * 2 * np.pi * kx / Lx * s * dx))
            ay = np.mean(np.exp(1j * 2 * np.pi * ky / Ly * s * dy))
            assert abs(top.weight[i, j] - (ax * ay).real) < 1e-6 and abs((ax * ay).imag) < 1e-12
    assert top.weight[0, list(kyc).index(0)] == 1.0
    rng = np.random.default_rng(5)
    kf = [rng.normal(size=(fine[0] // 2, 3 * fine[1] // 2, sum(fine[2] - bc // 10 for bc in sy.bc_codes)))
          + 1j * rng.normal(size=(fine[0] // 2, 3 * fine[1] // 2, sum(fine[2] - bc // 10 for bc in sy.bc_codes)))
          for sy in r.kxky_systems]
    zf = [rng.normal(size=sum(fine[2] - bc // 10 for bc in sy.bc_codes)) for sy in r.zero_systems]
    (kt, zt), (ks, zs) = top.state(kf, zf), sharp.state(kf, zf)
    for a, b in zip(kt, ks):
        np.testing.assert_allclose(a, b * top.weight[:, :, None], rtol=0, atol=1e-15)
    for a, b in zip(zt, zs):
        assert np.array_equal(a, b)
    q = rng.normal(size=(9, fine[0] // 2, 3 * fine[1] // 2, fine[2])) + 0j
    np.testing.assert_allclose(top.spectra(q), sharp.spectra(q) * top.weight[None, :, :, None], rtol=0, atol=1e-15)
