"""Diagnostics on the host: timeseries objects, weights and wall fluxes (no GPU)."""
import math
import os

import numpy as np
import pytest

from conftest import INPUTS
from coral_amd import diagnostics as D


def test_parse_reference_timeseries_file():
    text = open(os.path.join(INPUTS, "rrbc_julien.coral.timeseries")).read()
    objs = D.parse_timeseries(text)
    assert [(o.kind, o.name, o.position) for o in objs] == [
        ("volume", "uu", 0), ("volume", "vv", 0), ("volume", "ww", 0),
        ("zSlice", "tFull", 1), ("zSlice", "tFull", 24), ("zSlice", "tFull", 48), ("zSlice", "tFull", 72)]
    # the file names compute_nusselt.py reads (etc/benchmarks/02_RRBC_julien_JFM96/compute_nusselt.py:8-9)
    assert objs[3].filename == "tFull_XYavg_z00001.dat"
    assert objs[6].filename == "tFull_XYavg_z00072.dat"
    assert objs[0].filename == "uu_volAvg.dat"
    with pytest.raises(ValueError):
        D.parse_timeseries(">>Volume_sum :: uu <<")


@pytest.mark.parametrize("N", [25, 72, 193])
def test_weights_integrate_low_polynomials(N):
    # midpoint weights sum to the gap and integrate linear functions exactly
    gap, center = 1.3, 0.2
    w = D.gauss_chebyshev_weights(N, gap)
    z = D.gauss_chebyshev_grid(N, center, gap)
    assert abs(w.sum() - gap) < 1e-14
    assert abs(np.dot(w, z) - gap * center) < 1e-13


def test_wall_gradients_of_polynomials():
    N = 96
    z = D.gauss_chebyshev_grid(N, 0.5, 1.0)
    bot, top = D.wall_gradients(z ** 3 - 2 * z, 1.0)
    assert abs(bot + 2.0) < 1e-9 and abs(top - 1.0) < 1e-9
    bot, top = D.wall_gradients(np.sin(np.pi * z), 1.0)
    assert abs(bot - np.pi) < 1e-8 and abs(top + np.pi) < 1e-8


def test_wall_differences_match_a_conduction_profile():
    # theta = a (1 - z) near the top wall and a z near the bottom wall: the
    # one-sided difference over the first Gauss point is exact for linear profiles
    N, a = 72, 0.7
    z = D.gauss_chebyshev_grid(N, 0.5, 1.0)
    top, _ = D.nusselt_wall_differences(a * (1 - z), 1.0)
    _, bot = D.nusselt_wall_differences(-a * z, 1.0)
    assert abs(top - a) < 1e-12 and abs(bot - a) < 1e-12
    # dz of the reference script for NZAA = 72
    assert abs(0.5 * (1 - math.cos(math.pi / 144)) - (1. - 0.5 - 0.5 * np.cos(np.pi / 2 / 72.))) < 1e-16


def test_nusselt_quarters_follows_compute_nusselt():
    """compute_nusselt.py:16-57 on series whose answer is known: a constant
    profile gives T/dz in every quarter; a temperature that jumps between
    quarters gives each quarter's own value, time-weighted, and the mean and
    population spread of quarters 2-4 only."""
    NZAA, gap = 72, 1.0
    dz = 1.0 - 0.5 - 0.5 * np.cos(np.pi / 2 / 72.0)   # the script's own expression
    n = 401
    time = np.cumsum(np.linspace(1e-5, 3e-5, n))
    quarters, mean, std = D.nusselt_quarters(time, np.full(n, 0.5 * dz), np.full(n, -0.25 * dz), NZAA, gap)
    for top, bot in quarters:
        assert top == pytest.approx(0.5, rel=1e-12) and bot == pytest.approx(0.25, rel=1e-12)
    q = (n - 1) // 4
    level = np.append(np.repeat([9.0, 4.0, 5.0, 6.0], q), 0.0) * dz
    quarters, mean, std = D.nusselt_quarters(time, level, -level, NZAA, gap)
    assert [round(t, 9) for t, _ in quarters] == [9.0, 4.0, 5.0, 6.0]
    assert mean == pytest.approx(5.0, rel=1e-12)
    assert std == pytest.approx(np.sqrt(2.0 / 3.0), rel=1e-12)
    # the time weights: one large step in a quarter dominates its mean
    t2 = time.copy()
    t2[q + 1:] += 1.0                      # dt_q (the sample at index q) is huge
    ramp = np.arange(n, dtype=float) * dz
    quarters, _, _ = D.nusselt_quarters(t2, ramp, -ramp, NZAA, gap)
    assert quarters[1][0] == pytest.approx(float(q), rel=1e-3)


def test_horizontal_spectrum_shells():
    """Single modes land in the shell of their |k| with weight 2 for kx > 0
    and 1 for kx = 0; dealiased slots are excluded; other positions do not
    count; a box with Ly = 2 Lx halves the y wavenumbers."""
    NX, NY, tot = 16, 16, 5
    kx, ky, kept = __import__("coral_amd.ic", fromlist=["x"]).wavenumber_indices(NX, NY)
    iy = {int(k): i for i, k in enumerate(ky)}
    s = np.zeros((NX // 2, 3 * NY // 2, tot), complex)
    s[3, iy[-4], 1] = 2.0 + 1.0j          # |k| = 5, kx > 0: 2 * 5 = 10
    s[0, iy[2], 3] = 3.0                  # |k| = 2, kx = 0: 9
    s[1, iy[1], 0] = 100.0                # position not counted
    s[0, iy[-8], 1] = 50.0                # |ky| = 8 = NY / 2: dealiased
    sp = D.horizontal_spectrum(s, [1, 3], NX, NY, 2 * np.pi, 2 * np.pi)
    assert sp[5] == pytest.approx(10.0) and sp[2] == pytest.approx(9.0)
    assert sp.sum() == pytest.approx(19.0)
    sp2 = D.horizontal_spectrum(s, [1, 3], NX, NY, 2 * np.pi, 4 * np.pi)
    assert sp2[1] == pytest.approx(9.0)   # ky = 2 on a box twice as long in y: |k| = 1
    assert sp2[round(np.hypot(3, 2))] == pytest.approx(10.0)
