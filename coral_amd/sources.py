"""Source profiles of the mean systems.

In Coral a source is a vertical profile defined in user code
(``src/plane_layer/PL_user_sources_definitions.f90``, included into
``init_all_sources``, ``PL_IMEX_timestepping_marching.f90:829-843``): the
array ``sourceDef`` on the NZAA Gauss-Chebyshev points ``z`` (top wall first)
becomes source NN through ``add_source``, and a ``coral.equations`` line
``>>add_rhs_source :: - Iz.Iz.source02 <<`` adds it to a mean equation.  Here
a definition is a Python callable ``f(z, p)`` returning the profile at the
points ``z``, with ``p`` the ``sourceParams`` lookup of the recipe's
``sourceParam`` lines, or an array of the NZAA values; ``Solver.set_sources``
hands them to the library, which takes their Chebyshev coefficients.

``SHIPPED`` restates the definitions file shipped in ``src/plane_layer`` and
``BAROCLINIC_TURBULENCE`` the one of ``etc/baroclinic_turbulence`` (the case
whose readme asks for it to replace the shipped one).
"""
from __future__ import annotations

from typing import Callable, Dict, Union

import numpy as np


class SourceParams:
    """``sourceParams(name)`` (``PL_IMEX_timestepping_marching.f90:711-734``):
    the value of the ``sourceParam`` line whose name's first six characters
    match.  With no ``sourceParam`` line at all the reference returns 0; with
    some but none matching it stops, as this raises."""

    def __init__(self, values: Dict[str, float]):
        self.values = {k[:6]: float(v) for k, v in values.items()}

    def __call__(self, name: str) -> float:
        if not self.values:
            return 0.0
        key = name[:6]
        if key not in self.values:
            raise KeyError(f"source parameter {name!r} matches no sourceParam line of coral.equations")
        return self.values[key]


Definition = Union[Callable[[np.ndarray, SourceParams], np.ndarray], np.ndarray, float]


def _bottom_heating(z, p):
    h = p("HovEll")
    return h * np.exp(-z * h) / (1.0 - np.exp(-h)) - 1.0


def _uniform_cooling(z, p):
    return np.full_like(z, -1.0)


# src/plane_layer/PL_user_sources_definitions.f90:4-24
SHIPPED: Dict[int, Definition] = {1: _bottom_heating, 2: _uniform_cooling}

# etc/baroclinic_turbulence/PL_user_sources_definitions.f90:4-47
BAROCLINIC_TURBULENCE: Dict[int, Definition] = {
    1: _bottom_heating,
    2: _uniform_cooling,
    3: lambda z, p: np.exp(-(1.0 - z) / p("shearL")),
    4: lambda z, p: np.exp(-(1.0 - z) / p("shearL")) / p("shearL"),
    5: lambda z, p: np.exp(-(1.0 - z) / p("shearL")) / p("shearL") ** 2,
}


def evaluate(definition: Definition, z: np.ndarray, params: SourceParams) -> np.ndarray:
    """A definition's NZAA values at the points z (a constant is broadcast)."""
    vals = definition(z, params) if callable(definition) else definition
    vals = np.asarray(vals, np.float64)
    if vals.ndim == 0:
        vals = np.full(z.shape, float(vals))
    if vals.shape != z.shape:
        raise ValueError(f"a source profile has {z.size} values on the Gauss-Chebyshev grid, got shape {vals.shape}")
    return np.ascontiguousarray(vals)


__all__ = ["SourceParams", "SHIPPED", "BAROCLINIC_TURBULENCE", "evaluate"]
