"""Ensembles of independent plane-layer runs stepped together on one GPU.

``Ensemble`` wraps ``coral_amd_ensemble_*``: every run is a ``Solver`` of its
own (its recipe, grid, operators, dt and state), and one ensemble step
advances every run by one ARS step: the runs are spread over branches, each
branch's step one captured HIP graph replayed on a stream of its own.
``Ensemble.tune`` runs the library's tuner (per-run launch shapes, then the
branch count).

``draw_runs`` draws the per-run parameters of an uneven Rayleigh-Benard
ensemble from a seed: NX = NY in {64, 96, 128}, NZ in {33, 49, 65}, Ra
log-uniform in [1e5, 1e7], Pr in {0.7, 1, 7}, each run with the reference's
primitive-variable recipe (``coral.equations.RBC_primitiveVars_Vanilla``
form) and its own parameters.  The reference runs one problem per
executable; an ensemble there is that many separate jobs.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import check, lib
from .equations import Recipe, parse_equations
from .ic import SplitMix64

HERE = os.path.dirname(os.path.abspath(__file__))
TEMPLATE = os.path.join(HERE, "inputs", "coral.equations.rbc_ra1e5")


@dataclass
class RunSpec:
    NX: int
    NZ: int
    Ra: float
    Pr: float
    dt: float
    seed: int

    def label(self) -> str:
        return f"{self.NX}^2x{self.NZ} Ra={self.Ra:.3g} Pr={self.Pr:g}"


def rbc_recipe(Ra: float, Pr: float) -> Recipe:
    """The non-rotating Rayleigh-Benard recipe with the given Rayleigh and
    Prandtl numbers (the template's ``Prandtl`` and ``RayleighPr`` = Ra Pr
    parameter lines replaced)."""
    lines = []
    for ln in open(TEMPLATE).read().splitlines():
        if ln.strip().startswith(">>Param :: Prandtl"):
            ln = f">>Param :: Prandtl := {Pr!r} <<"
        elif ln.strip().startswith(">>Param :: RayleighPr"):
            ln = f">>Param :: RayleighPr := {Ra * Pr!r} <<"
        lines.append(ln)
    return parse_equations("\n".join(lines) + "\n")


def run_dt(Ra: float, Pr: float) -> float:
    """Fixed time step of an ensemble run, 5e-5 (Ra Pr / 1e5)^(-1/2): the
    free-fall velocity in thermal-diffusion units grows like (Ra Pr)^(1/2), so
    the step shrinks with it (5e-6 at Ra Pr = 1e7)."""
    return 5e-5 * math.sqrt(1e5 / (Ra * Pr))


def draw_runs(nrun: int, seed: int, sizes: Sequence[int] = (64, 96, 128), depths: Sequence[int] = (33, 49, 65),
              ra_range: Tuple[float, float] = (1e5, 1e7), prandtl: Sequence[float] = (0.7, 1.0, 7.0)) -> List[RunSpec]:
    """Per-run parameters drawn from SplitMix64(seed); dt from ``run_dt``."""
    rng = SplitMix64(seed)
    runs = []
    for i in range(nrun):
        NX = sizes[rng.next_u64() % len(sizes)]
        NZ = depths[rng.next_u64() % len(depths)]
        lo, hi = math.log(ra_range[0]), math.log(ra_range[1])
        Ra = math.exp(lo + (hi - lo) * rng.uniform())
        Pr = prandtl[rng.next_u64() % len(prandtl)]
        dt = run_dt(Ra, Pr)
        runs.append(RunSpec(NX, NZ, Ra, Pr, dt, seed * 1000 + i))
    return runs


class Ensemble:
    def __init__(self, solvers: Sequence):
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.handle.value for s in self.solvers])
        h = ctypes.c_void_p()
        check(lib.coral_amd_ensemble_create(arr, len(self.solvers), ctypes.byref(h)), "coral_amd_ensemble_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib.coral_amd_ensemble_destroy(self._h), "coral_amd_ensemble_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def tune(self, trial_steps: int = 10) -> Tuple[int, float]:
        """Run the library's tuner; returns (branches, ms per ensemble step);
        a negative branch count is the lockstep schedule over that many
        branches."""
        nb, ms = ctypes.c_int32(), ctypes.c_double()
        check(lib.coral_amd_ensemble_tune(self._h, int(trial_steps), ctypes.byref(nb), ctypes.byref(ms)),
              "coral_amd_ensemble_tune")
        return nb.value, ms.value

    def set_schedule(self, nbranch: int, assign: Optional[Sequence[int]] = None):
        """nbranch > 0: parallel branches of whole run steps; nbranch <= 0:
        the lockstep schedule (stage fronts over max(1, -nbranch) branches,
        each stage's solves of all runs batched)."""
        a = None if assign is None else (ctypes.c_int32 * len(assign))(*assign)
        check(lib.coral_amd_ensemble_set_schedule(self._h, int(nbranch), a), "coral_amd_ensemble_set_schedule")

    def step(self, nsteps: int = 1):
        """Advance every run by nsteps steps of its own factorized dt."""
        check(lib.coral_amd_ensemble_step(self._h, int(nsteps)), "coral_amd_ensemble_step")

    def synchronize(self):
        check(lib.coral_amd_ensemble_synchronize(self._h), "coral_amd_ensemble_synchronize")
