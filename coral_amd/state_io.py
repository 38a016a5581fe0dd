"""Coral's state files: QuickSave volumes and ``dt.sav``.

``export_checkpoints`` / ``import_quicksave`` follow
``src/plane_layer/PL_IMEX_timestepping_output.f90:309-456``: one raw FP64 file
per coupled variable, ``<dir>/QuickSave.kxky.sysNNN.varNNN.rollingR`` holding
the physical values of the variable (z fastest, then y, then x, no header) and
``QuickSave.zero.sysNNN.varNNN.rollingR`` holding the NZAA-point profile of a
mean variable.  Export writes into ``CheckPoints/``, import reads from
``Restart/`` and takes rolling index 1 when its first file exists, else 2.

``dt.sav`` follows ``src/misc/timeKeeping.f90:92-147``: ``dt, time`` (16
bytes) or ``dt, time, i_timestep`` (20 bytes, int32 last); on reading, the
stored dt is divided by 50 and capped by dt_max.
"""
from __future__ import annotations

import os
import struct
from typing import List, Optional, Tuple

import numpy as np


def quicksave_name(kind: str, isys: int, ivar: int, rolling: int) -> str:
    """File name of coupled variable `ivar` of system `isys` (0-based)."""
    if kind not in ("kxky", "zero"):
        raise ValueError("kind must be 'kxky' or 'zero'")
    return f"QuickSave.{kind}.sys{isys + 1:03d}.var{ivar + 1:03d}.rolling{rolling:d}"


def _systems(solver):
    r = solver.recipe
    return [("kxky", i, s.nvars) for i, s in enumerate(r.kxky_systems)] + \
           [("zero", i, s.nvars) for i, s in enumerate(r.zero_systems)]


def export_checkpoints(solver, output_dir: str, rolling: int = 1, subdir: str = "CheckPoints") -> List[str]:
    """Write every coupled variable of the solver's current state; returns the paths."""
    d = os.path.join(output_dir, subdir)
    os.makedirs(d, exist_ok=True)
    paths = []
    for kind, isys, nvars in _systems(solver):
        for ivar in range(nvars):
            path = os.path.join(d, quicksave_name(kind, isys, ivar, rolling))
            solver.variable_physical(kind, isys, ivar).tofile(path)
            paths.append(path)
    return paths


def find_rolling(restart_dir: str) -> int:
    for rolling in (1, 2):
        if os.path.exists(os.path.join(restart_dir, quicksave_name("kxky", 0, 0, rolling))):
            return rolling
    raise FileNotFoundError(f"No QuickSave was found in {restart_dir}")


def import_quicksave(solver, output_dir: str, subdir: str = "Restart") -> int:
    """Set the solver's state from the QuickSave files; returns the rolling index used."""
    d = os.path.join(output_dir, subdir)
    rolling = find_rolling(d)
    for kind, isys, nvars in _systems(solver):
        n = solver.NXAA * solver.NYAA * solver.NZAA if kind == "kxky" else solver.NZAA
        for ivar in range(nvars):
            path = os.path.join(d, quicksave_name(kind, isys, ivar, rolling))
            data = np.fromfile(path, dtype=np.float64)
            if data.size != n:
                raise ValueError(f"{path}: {data.size} values, expected {n}")
            solver.set_variable_physical(kind, isys, ivar, data)
    return rolling


def write_dt_sav(path: str, dt: float, time: float, i_timestep: Optional[int] = None) -> None:
    with open(path, "wb") as f:
        f.write(struct.pack("<dd", dt, time))
        if i_timestep is not None:
            f.write(struct.pack("<i", i_timestep))


def read_dt_sav(path: str, dt_max: float) -> Tuple[float, float, float, Optional[int]]:
    """Returns (dt, dt_current, time, i_timestep) as the reference sets them on restart."""
    raw = open(path, "rb").read()
    if len(raw) == 16:
        dt, time = struct.unpack("<dd", raw)
        step = None
    elif len(raw) == 20:
        dt, time, step = struct.unpack("<ddi", raw)
        tfile = os.path.join(os.path.dirname(os.path.abspath(path)), "..", "Timeseries", "time.dat")
        if not os.path.exists(tfile):
            step = 1
    else:
        raise ValueError(f"{path}: {len(raw)} bytes, expected 16 or 20")
    dt = dt / 50.0
    if dt > dt_max:
        dt = dt_max
    return dt, dt, time, step
