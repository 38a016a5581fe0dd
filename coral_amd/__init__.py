"""coral_amd -- MI355X-native time-stepper for Coral's plane-layer IMEX hot path.

Host-side pieces (no GPU needed): ``equations`` (coral.equations reader),
``params`` (coral.parameters.in reader), ``schemes`` (ARS tableaux), ``ic``
(seeded initial conditions).  ``solver`` drives the HIP library through the
C ABI of ``include/coral_amd.h`` and is imported lazily, because it needs the
built ``lib/libcoral_amd.so``.
"""
from . import equations, ic, params, schemes  # noqa: F401

__all__ = ["equations", "params", "schemes", "ic"]


def load_solver():
    """Import and return the GPU solver module (raises if the library is missing)."""
    from . import solver
    return solver
