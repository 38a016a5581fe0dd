"""Seeded initial conditions in Coral's spectral state layout.

Coral draws its noisy initial condition from an unseeded ``random_number``
(``src/plane_layer/PL_IMEX_timestepping.f90:425-486``), so runs cannot be
repeated.  This module builds a *seeded* perturbation instead, identical on
every machine, so that the GPU path and any checker start from the same state:

* generator: splitmix64 (Steele, Lea & Flood 2014) feeding Box-Muller normals;
* the perturbation is put on one variable of a kxky system (the temperature
  ``t`` for Rayleigh-Benard), on Galerkin modes ``0..nmodes-1`` and on integer
  wavevectors with ``kx_int**2 + ky_int**2 <= kmax**2``;
* the kx = 0 plane is made Hermitian in ky (``c(-ky) = conj(c(ky))``), as the
  real fields that Coral's forward transform produces are;
* the (0, 0) mode and every dealiased mode stay zero;
* the coefficients are scaled so that their l2 norm over the full Hermitian
  spectrum equals the requested amplitude (the Galerkin functions have O(1)
  norms, so the physical rms of the perturbation is of that order).

The state layout is Coral's per-system array ``(total, NYAA, NX/2)`` in Fortran
order, i.e. a C-contiguous ``(NX/2, NYAA, total)`` complex128 array; interleaved
Galerkin positions follow ``build_kxky_shuffle_matrices_for_system``
(``PL_IMEX_timestepping.f90:1344-1380``).
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np

_MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform in (0, 1): 53 random bits, never exactly 0."""
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / (1 << 53))

    def normal_pair(self):
        u1, u2 = self.uniform(), self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        return r * math.cos(2.0 * math.pi * u2), r * math.sin(2.0 * math.pi * u2)


def wavenumber_indices(NX: int, NY: int):
    """Integer wavenumbers of the spectral slots, Coral's slab layout
    (``PL_geometry.f90:179-222``): kx = 0..NX/2-1 and, for iy = 0..NYAA-1,
    ky = iy if iy+1 < NYAA/2+1 else -(NYAA - iy).  Returns (kx_int, ky_int,
    kept[ix, iy]) where kept excludes |k| >= N/2 (the 2/3 rule)."""
    NYAA = (3 * NY) // 2
    NXAA = (3 * NX) // 2
    ky = np.array([iy if iy + 1 < NYAA // 2 + 1 else -(NYAA - iy) for iy in range(NYAA)], np.int64)
    kx = np.array([ix if ix + 1 < NXAA // 2 + 1 else -(NXAA - ix) for ix in range(NX // 2)], np.int64)
    # kcut compare as Coral does it: |k*k_box| >= k_box*N/2 <=> |k| >= N/2 for even N
    kept = (np.abs(kx)[:, None] * 2 < NX) & (np.abs(ky)[None, :] * 2 < NY)
    return kx, ky, kept


def interleaved_positions(nvars: int, var_sizes: Sequence[int], var: int) -> np.ndarray:
    """Interleaved (degree-shuffled) positions of variable ``var``'s Galerkin
    coefficients (``PL_IMEX_timestepping.f90:1344-1380``)."""
    nmin = min(var_sizes)
    pos: List[int] = []
    offset = 0
    for v in range(nvars):
        for c in range(nmin):
            if v == var:
                pos.append(v + c * nvars)
        for _ in range(var_sizes[v] - nmin):
            if v == var:
                pos.append(nmin * nvars + offset)
            offset += 1
    return np.array(pos, np.int64)


def seeded_perturbation(NX: int, NY: int, NZ: int, bc_codes: Sequence[int], var: int,
                        amplitude: float, kmax: int, nmodes: int, seed: int) -> np.ndarray:
    """Seeded perturbation of one variable of a kxky system; returns the
    system's state as a C-contiguous complex128 array (NX/2, NYAA, total)."""
    nvars = len(bc_codes)
    sizes = [NZ - bc // 10 for bc in bc_codes]
    total = sum(sizes)
    kx, ky, kept = wavenumber_indices(NX, NY)
    NYAA = ky.size
    state = np.zeros((NX // 2, NYAA, total), np.complex128)
    pos = interleaved_positions(nvars, sizes, var)[: min(nmodes, sizes[var])]
    rng = SplitMix64(seed)
    order = np.argsort(ky, kind="stable")           # ky ascending: -.., 0, ..+
    for ix in range(NX // 2):
        for iy in order:
            if not kept[ix, iy] or kx[ix] ** 2 + ky[iy] ** 2 > kmax ** 2:
                continue
            if kx[ix] == 0 and ky[iy] <= 0:
                continue                              # filled by symmetry below / (0,0) stays 0
            for p in pos:
                a, b = rng.normal_pair()
                state[ix, iy, p] = complex(a, b)
    # Hermitian kx = 0 plane: c(0, -ky) = conj(c(0, ky))
    for iy in range(NYAA):
        if ky[iy] < 0 and kept[0, iy]:
            jy = int(np.nonzero(ky == -ky[iy])[0][0])
            state[0, iy, :] = np.conj(state[0, jy, :])
    # normalise: the l2 norm of the coefficients over the full (Hermitian) spectrum
    # equals `amplitude`, so the physical perturbation has an rms of that order
    norm2 = np.sum(np.abs(state[0]) ** 2) + 2.0 * np.sum(np.abs(state[1:]) ** 2)
    if norm2 > 0:
        state *= amplitude / math.sqrt(norm2)
    return state


def rbc_initial_state(recipe, NX: int, NY: int, NZ: int, amplitude: float, kmax: int,
                      nmodes: int, seed: int, var_name: str = "t"):
    """Per-system initial states for a recipe: the seeded perturbation on
    ``var_name`` and zeros elsewhere (mean fields zero: conductive background)."""
    kstates = []
    for s in recipe.kxky_systems:
        if var_name in s.var_names:
            kstates.append(seeded_perturbation(NX, NY, NZ, s.bc_codes, s.var_names.index(var_name),
                                               amplitude, kmax, nmodes, seed))
        else:
            total = sum(NZ - bc // 10 for bc in s.bc_codes)
            kstates.append(np.zeros((NX // 2, (3 * NY) // 2, total), np.complex128))
    zstates = [np.zeros(sum(NZ - bc // 10 for bc in s.bc_codes)) for s in recipe.zero_systems]
    return kstates, zstates


def noisy_initial_state(recipe, NX: int, NY: int, NZ: int, Lx: float, Ly: float, amp_kxky: float,
                        amp_zero: float, seed: int):
    """Coral's default noisy initial condition (``noisy_initial_conditions``,
    ``src/plane_layer/PL_IMEX_timestepping.f90:425-486``) with seeded phases.

    Every Galerkin coefficient of position ``n`` (counted inside its variable)
    of every kxky variable at wavevector |k| gets two random-phase bumps,
    ``exp(-(n-3.5)^2/8) exp(-(|k|-6 k_box)^2/8) amp`` and
    ``3 exp(-(n-1.5)^2/8) exp(-(|k|-3 k_box)^2/32) amp``, with k_box = 2 pi / Lx
    (the reference's ``kx_box``) and |k| from the physical wavenumbers; mean
    variables get ``exp(-(n-1.5)^2/8) cos(p) amp_zero`` at their storage
    position p (1-based).  The phases come from SplitMix64 uniforms instead of
    the unseeded ``random_number``; the kx = 0 plane is made Hermitian in ky
    and the (0, 0) and dealiased modes are zero (the step keeps them so)."""
    kx, ky, kept = wavenumber_indices(NX, NY)
    NYAA = ky.size
    kx_box, ky_box = 2.0 * math.pi / Lx, 2.0 * math.pi / Ly
    kabs = np.sqrt((kx[:, None] * kx_box) ** 2 + (ky[None, :] * ky_box) ** 2)
    rng = SplitMix64(seed)
    kstates = []
    for s in recipe.kxky_systems:
        sizes = [NZ - bc // 10 for bc in s.bc_codes]
        state = np.zeros((NX // 2, NYAA, sum(sizes)), np.complex128)
        for v in range(len(sizes)):
            pos = interleaved_positions(len(sizes), sizes, v)
            for n, p in enumerate(pos):
                ph1 = np.array([[rng.uniform() for _ in range(NYAA)] for _ in range(NX // 2)])
                ph2 = np.array([[rng.uniform() for _ in range(NYAA)] for _ in range(NX // 2)])
                b1 = math.exp(-(n - 3.5) ** 2 / 8.0) * np.exp(-(kabs - 6.0 * kx_box) ** 2 / 8.0)
                b2 = 3.0 * math.exp(-(n - 1.5) ** 2 / 8.0) * np.exp(-(kabs - 3.0 * kx_box) ** 2 / 32.0)
                state[:, :, p] = amp_kxky * (b1 * np.exp(2j * math.pi * ph1) + b2 * np.exp(2j * math.pi * ph2))
        state[~kept] = 0.0
        state[0, 0] = 0.0
        for iy in range(NYAA):
            if ky[iy] < 0 and kept[0, iy]:
                jy = int(np.nonzero(ky == -ky[iy])[0][0])
                state[0, iy, :] = np.conj(state[0, jy, :])
        kstates.append(state)
    zstates = []
    for s in recipe.zero_systems:
        sizes = [NZ - bc // 10 for bc in s.bc_codes]
        z = np.zeros(sum(sizes))
        first = 0
        for v, size in enumerate(sizes):
            for n in range(size):
                z[first + n] = math.exp(-(n - 1.5) ** 2 / 8.0) * math.cos(float(first + n + 1)) * amp_zero
            first += size
        zstates.append(z)
    return kstates, zstates
