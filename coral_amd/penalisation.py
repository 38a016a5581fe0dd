"""Penalisation masks of full variables.

A ``coral.equations`` block ``>>enable_penalisation :: GAMMA <<``,
``>>penalisation_strength :: s <<``, ``>>penalisation_width :: w <<`` makes
``GAMMA`` a full variable whose physical values are fixed at setup
(``init_penalisation``, ``src/plane_layer/PL_IMEX_timestepping_marching.f90:778-822``):

    GAMMA(x, y, z) = s (1 + tanh(w max_i m_i(x, y, z))) / 2

where the m_i are the masks that user code defines
(``src/plane_layer/PL_user_penalisation_mask.f90``).  Here a mask definition
is a callable ``m(x, y, z, Lx, Ly)`` of broadcastable coordinate arrays that
returns a list of mask arrays (or one array); ``Solver.set_penalisation_mask``
evaluates it on the reference's coordinates and hands the field to the
library (``coral_amd_set_fixed_physical``).

The reference evaluates the masks at ``x = (ix + 1) Lx / NXAA`` and
``y = (iy + 1) Ly / NYAA`` for the 1-based point indices ``ix``, ``iy``
(``PL_geometry.f90:136-145`` with ``phys_iStart = 1`` on one process), i.e.
two grid spacings beyond the FFT point the value is stored at; ``z`` is the
Gauss-Chebyshev grid.  ``reference_coordinates`` returns those arrays.

``SHIPPED`` restates the two example masks of the shipped user file.
"""
from __future__ import annotations

import math
from typing import Callable, List, Sequence, Union

import numpy as np

from .diagnostics import gauss_chebyshev_grid

MaskDefinition = Callable[..., Union[np.ndarray, Sequence[np.ndarray]]]


def reference_coordinates(NXAA: int, NYAA: int, NZAA: int, Lx: float, Ly: float, center: float, gap: float):
    """x [NXAA, 1, 1], y [1, NYAA, 1], z [1, 1, NZAA] as the reference's
    mask loop sees them (0-based point i stores the value at (i + 2) L / N)."""
    x = (np.arange(NXAA) + 2.0) * Lx / NXAA
    y = (np.arange(NYAA) + 2.0) * Ly / NYAA
    z = gauss_chebyshev_grid(NZAA, center, gap)
    return x[:, None, None], y[None, :, None], z[None, None, :]


def mask_field(definition: MaskDefinition, strength: float, width: float, NXAA: int, NYAA: int, NZAA: int,
               Lx: float, Ly: float, center: float, gap: float) -> np.ndarray:
    """strength (1 + tanh(width max_i m_i)) / 2 on the padded grid, [x, y, z]."""
    x, y, z = reference_coordinates(NXAA, NYAA, NZAA, Lx, Ly, center, gap)
    masks = definition(x, y, z, Lx, Ly)
    if isinstance(masks, np.ndarray):
        masks = [masks]
    shape = (NXAA, NYAA, NZAA)
    pm = np.full(shape, -np.inf)
    for m in masks:
        pm = np.maximum(pm, np.broadcast_to(np.asarray(m, np.float64), shape))
    pm = np.tanh(width * pm)
    pm = (1.0 + pm) / 2.0
    return np.ascontiguousarray(strength * pm)


def _shipped(x, y, z, lx, ly) -> List[np.ndarray]:
    pi = 3.14159265358979311599796346854
    above = z - (0.9 + 0.1 * np.cos(5 * (x + y) * 2 * pi / lx))
    below = (0.4 * np.exp(-((6.0 + np.cos(2 * pi * x / lx)) ** 2)) / math.exp(-25.0)) - z
    return [above, below]


# src/plane_layer/PL_user_penalisation_mask.f90:1-29
SHIPPED: MaskDefinition = _shipped

__all__ = ["reference_coordinates", "mask_field", "SHIPPED"]
