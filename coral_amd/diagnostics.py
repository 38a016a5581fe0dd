"""Diagnostics of a running solver: Coral's timeseries objects, kinetic energy
and wall heat flux, computed from horizontal averages taken on the GPU.

Timeseries objects follow ``output_global_quantities``
(``src/plane_layer/PL_IMEX_timestepping_output.f90:5-231``):

* ``Volume_average :: name``  ->  sum_z w_z <f>_xy(z) with the Gauss-Chebyshev
  weights of ``PL_cheby_misc.f90:28-55`` (midpoint spacing times gap/2),
  written to ``Timeseries/<name>_volAvg.dat``;
* ``Horizontal_avg :: name ;; Position= k``  ->  <f>_xy at Gauss point k
  (1-based, k = 1 next to the top wall), written to
  ``Timeseries/<name>_XYavg_z%05d.dat``;
* ``KE_display`` = the sum of the volume averages of ``uu``, ``vv``, ``ww``.

Files are raw little-endian FP64 appended one value per sample; ``time.dat``
holds the sample times.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np


@dataclass
class TimeseriesObject:
    kind: str          # "volume" or "zSlice"
    name: str          # full or quadratic variable
    position: int = 0  # 1-based Gauss point of a zSlice

    @property
    def filename(self) -> str:
        if self.kind == "volume":
            return f"{self.name}_volAvg.dat"
        return f"{self.name}_XYavg_z{self.position:05d}.dat"


def parse_timeseries(text: str) -> List[TimeseriesObject]:
    """Parse a ``coral.timeseries`` file (``>>Volume_average :: uu <<``,
    ``>>Horizontal_avg :: tFull ;; Position= 24 <<``, ``>>EOF<<``)."""
    objs = []
    for line in text.splitlines():
        m = re.search(r">>(.*)<<", line)
        if not m:
            continue
        body = m.group(1).strip()
        if body == "EOF":
            break
        key, _, rest = body.partition("::")
        key = key.strip()
        if key == "Volume_average":
            objs.append(TimeseriesObject("volume", rest.strip()))
        elif key == "Horizontal_avg":
            name, _, pos = rest.partition(";;")
            pm = re.search(r"Position\s*=\s*(\d+)", pos)
            if not pm:
                raise ValueError(f"missing Position in: {line.strip()}")
            objs.append(TimeseriesObject("zSlice", name.strip(), int(pm.group(1))))
        else:
            raise ValueError(f"unknown timeseries object: {line.strip()}")
    return objs


def gauss_chebyshev_weights(NZAA: int, gap: float) -> np.ndarray:
    """Midpoint spacing of the Gauss-Chebyshev grid on [-1, 1], times gap/2."""
    x = np.cos(np.pi * (2 * np.arange(NZAA) + 1) / (2 * NZAA))
    mid = np.concatenate(([1.0], 0.5 * (x[1:] + x[:-1]), [-1.0]))
    return (mid[:-1] - mid[1:]) / 2.0 * gap


def gauss_chebyshev_grid(NZAA: int, center: float, gap: float) -> np.ndarray:
    return center + 0.5 * gap * np.cos(np.pi * (2 * np.arange(NZAA) + 1) / (2 * NZAA))


class Diagnostics:
    """Samples the timeseries objects of a solver."""

    def __init__(self, solver, objects: Optional[List[TimeseriesObject]] = None):
        self.solver = solver
        r = solver.recipe
        self.full_names = [f.name for f in r.full_vars]
        self.quad_names = [q.name for q in r.quad_vars]
        self.objects = objects or []
        for o in self.objects:
            if o.name not in self.full_names and o.name not in self.quad_names:
                raise ValueError(f"timeseries object on unknown variable {o.name}")
            if o.kind == "zSlice" and not 1 <= o.position <= solver.NZAA:
                raise ValueError(f"Position {o.position} outside 1..{solver.NZAA}")
        self.weights = gauss_chebyshev_weights(solver.NZAA, solver.gap)

    def profiles(self) -> Dict[str, np.ndarray]:
        full, quad = self.solver.horizontal_averages()
        out = {n: full[i] for i, n in enumerate(self.full_names)}
        out.update({n: quad[i] for i, n in enumerate(self.quad_names)})
        return out

    def sample(self) -> Tuple[Dict[str, float], float]:
        """Values of every object, and KE_display."""
        prof = self.profiles()
        vals = {}
        for o in self.objects:
            p = prof[o.name]
            vals[o.filename] = float(np.dot(self.weights, p)) if o.kind == "volume" else float(p[o.position - 1])
        ke = sum(float(np.dot(self.weights, prof[n])) for n in ("uu", "vv", "ww") if n in prof)
        return vals, ke

    def append(self, directory: str, time: float) -> Dict[str, float]:
        """Append one sample to ``<directory>/Timeseries``."""
        d = os.path.join(directory, "Timeseries")
        os.makedirs(d, exist_ok=True)
        vals, _ = self.sample()
        with open(os.path.join(d, "time.dat"), "ab") as f:
            f.write(np.float64(time).tobytes())
        for fname, v in vals.items():
            with open(os.path.join(d, fname), "ab") as f:
                f.write(np.float64(v).tobytes())
        return vals


def chebyshev_coefficients(profile: np.ndarray) -> np.ndarray:
    """Chebyshev coefficients c_n of a profile sampled at the NZAA Gauss points
    (profile(x_z) = sum_n c_n T_n(x_z), x_z = cos(pi (2z+1) / (2 NZAA)))."""
    N = profile.shape[-1]
    z = np.arange(N)
    n = np.arange(N)[:, None]
    C = np.cos(np.pi * n * (2 * z + 1) / (2 * N)) * (2.0 / N)
    c = C @ profile
    c[0] *= 0.5
    return c


def wall_gradients(profile: np.ndarray, gap: float) -> Tuple[float, float]:
    """d<f>/dz at the bottom (z = center - gap/2) and top wall from the
    Chebyshev series of the profile: T_n'(1) = n^2, T_n'(-1) = (-1)^(n+1) n^2."""
    c = chebyshev_coefficients(profile)
    n = np.arange(c.size)
    top = float(np.sum(c * n * n)) * 2.0 / gap
    bot = float(np.sum(c * n * n * np.where(n % 2 == 0, -1.0, 1.0))) * 2.0 / gap
    return bot, top


def nusselt_wall_differences(profile: np.ndarray, gap: float) -> Tuple[float, float]:
    """The finite-difference wall fluxes of
    etc/benchmarks/02_RRBC_julien_JFM96/compute_nusselt.py: the horizontal average
    at the first (top) and last (bottom) Gauss point divided by their distance to
    the wall, (top, bottom) = (<T>_1 / dz, -<T>_N / dz)."""
    N = profile.shape[-1]
    dz = 0.5 * gap * (1.0 - math.cos(math.pi / (2 * N)))
    return float(profile[0] / dz), float(-profile[-1] / dz)


def nusselt_quarters(time: np.ndarray, t_top: np.ndarray, t_bot: np.ndarray, NZAA: int, gap: float = 1.0):
    """Nusselt number of a run from its timeseries, as
    etc/benchmarks/02_RRBC_julien_JFM96/compute_nusselt.py:16-57 forms it.

    ``time`` are the sample times and ``t_top`` / ``t_bot`` the horizontal
    averages of the temperature at the first (next to the top wall) and last
    Gauss point.  The series is cut into four quarters of
    q = (len(time) - 1) // 4 samples (the script takes len(time) // 4 with two
    time samples more than temperature samples; here every sample k < 4q has
    its weight dt_k); in quarter a the wall flux is the time-weighted mean
    sum(dt_k T_k) / sum(dt_k) over k in [a q, (a+1) q) with dt_k = time[k+1] -
    time[k], divided by the distance dz of the Gauss point to its wall (top:
    +<T>/dz, bottom: -<T>/dz).  Returns (quarters, mean, std): the (top,
    bottom) pair of each quarter, and the mean and population standard
    deviation of the six values of quarters 2 to 4 (the first quarter is the
    transient)."""
    time = np.asarray(time, np.float64)
    t_top = np.asarray(t_top, np.float64)
    t_bot = np.asarray(t_bot, np.float64)
    q = (time.shape[0] - 1) // 4
    if q < 1 or t_top.shape[0] < 4 * q or t_bot.shape[0] < 4 * q:
        raise ValueError("need at least four samples of time and both temperatures")
    dz = 0.5 * gap * (1.0 - math.cos(math.pi / (2 * NZAA)))
    d = np.diff(time)
    quarters = []
    for a in range(4):
        sl = slice(a * q, (a + 1) * q)
        w = np.sum(d[sl])
        quarters.append((float(np.sum(d[sl] * t_top[sl]) / w / dz), float(-np.sum(d[sl] * t_bot[sl]) / w / dz)))
    vals = np.array([v for a in (1, 2, 3) for v in quarters[a]])
    mean = float(np.mean(vals))
    std = float(np.sqrt(np.mean((vals - mean) ** 2)))
    return quarters, mean, std


def horizontal_spectrum(k_state: np.ndarray, positions, NX: int, NY: int, Lx: float, Ly: float) -> np.ndarray:
    """Horizontal spectrum of a kxky system's state (NX/2, NYAA, total): the
    squared moduli of the Galerkin coefficients at `positions` (e.g. the
    velocity variables), summed over them and per wavenumber shell
    |k| / (2 pi / Lx) rounded to the nearest integer; kx > 0 counted twice
    (the conjugate half of the real fields' spectrum); dealiased slots
    excluded.  Entry m is shell m."""
    from .ic import wavenumber_indices
    kx, ky, kept = wavenumber_indices(NX, NY)
    e = np.sum(np.abs(k_state[:, :, positions]) ** 2, axis=2)
    e = np.where(kept, e, 0.0) * np.where(np.asarray(kx)[:, None] > 0, 2.0, 1.0)
    kk = np.sqrt((np.asarray(kx)[:, None] * 2 * np.pi / Lx) ** 2 +
                 (np.asarray(ky)[None, :] * 2 * np.pi / Ly) ** 2) / (2 * np.pi / Lx)
    shell = np.floor(kk + 0.5).astype(int)
    return np.bincount(shell.ravel(), weights=e.ravel())
