"""Adaptive time step: Coral's ``timings_T`` (``src/misc/timeKeeping.f90``)
driving the GPU solver.

Per step, as in ``plane_layer_main.f90:136-200``:

1. ``check_for_update`` - adopt the CFL-based target dt when it differs from
   the dt in use by more than ``tolerance * dt_current``; the operators are then
   refactorized on the device (``coral_amd_set_dt``);
2. ``march_forward(dt_current)``;
3. ``apply_cfl`` - target dt = CFL_C * min(dx/u, dy/v, dz/w) of the last
   stage (the maxima are reduced on the device inside the x pass), capped by
   dt_max; a target below dt_min (1e-9) stops the run;
4. ``increment`` - time += dt_current.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional

from . import state_io


class MinimalTimestepReached(RuntimeError):
    """The CFL-based time step fell below dt_min (the reference stops there)."""


@dataclass
class Timings:
    dt_max: float
    tolerance_factor: float
    CFL_C: float
    dt: float = field(init=False)
    dt_current: float = field(init=False)
    dt_min: float = 1e-9
    CFL_flag: bool = True
    absolute_time: float = 0.0
    i_timestep: int = 1
    refactorization_is_due: bool = True

    def __post_init__(self):
        # initialize_timings (timeKeeping.f90:57-71)
        self.dt = self.dt_max
        self.dt_current = self.dt_max

    def apply_cfl(self, dx_over_velocity_min: float) -> None:
        """compare_dt_against_CFL (timeKeeping.f90:78-89)."""
        if self.CFL_flag:
            self.dt = dx_over_velocity_min * self.CFL_C
        if self.dt > self.dt_max:
            self.dt = self.dt_max
        if self.dt < self.dt_min:
            raise MinimalTimestepReached("minimal timestep reached")

    def check_for_update(self) -> None:
        """update_dt_if_needed (timeKeeping.f90:171-183), single process."""
        if abs(self.dt - self.dt_current) > self.tolerance_factor * self.dt_current:
            self.refactorization_is_due = True
            self.dt_current = self.dt

    def increment(self) -> None:
        self.i_timestep += 1
        self.absolute_time += self.dt_current

    def fromfile(self, path: str) -> None:
        """read_timings_from_file (timeKeeping.f90:92-134)."""
        dt, dtc, t, step = state_io.read_dt_sav(path, self.dt_max)
        self.dt, self.dt_current, self.absolute_time = dt, dtc, t
        if step is not None:
            self.i_timestep = step
        self.refactorization_is_due = True

    def tofile(self, path: str) -> None:
        """write_timings_to_file (timeKeeping.f90:137-147)."""
        state_io.write_dt_sav(path, self.dt_current, self.absolute_time, self.i_timestep)


def march_adaptive(solver, timings: Timings, nsteps: int,
                   after_step: Optional[Callable[[object, Timings], None]] = None) -> List[float]:
    """Run `nsteps` steps with the reference's adaptive-dt rules; returns the dt
    used at each step.  `after_step(solver, timings)` runs after each step
    (outputs, checkpoints)."""
    used = []
    for _ in range(nsteps):
        timings.check_for_update()
        if timings.refactorization_is_due:
            solver.factorize_operators(timings.dt_current)
            timings.refactorization_is_due = False
        solver.march_forward(timings.dt_current, 1)
        used.append(timings.dt_current)
        timings.apply_cfl(solver.cfl_dt())
        timings.increment()
        if after_step is not None:
            after_step(solver, timings)
    return used
