"""ARS97 IMEX Runge-Kutta tableaux, as Coral defines them.

``src/timesteppers/IMEX_schemes.f90:21-402``: ``A`` is the implicit tableau
(s x s), ``A_hat`` the explicit one ((s+1) x (s+1)).  The default scheme of every
shipped parameter file is ARS_222 (``IMEX_schemes.f90:210-254``).  ARS_343 is
written in the reference with default-real (single precision) literals for
gamma, 1-gamma, b1 and b2 (``:296-301``); they are rounded to float32 here too,
so the two codes use the same numbers.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

SCHEMES = ("ARS_111", "ARS_122", "ARS_233", "ARS_232", "ARS_222", "ARS_343", "ARS_443")


def tableau(name: str) -> Tuple[np.ndarray, np.ndarray]:
    if name == "ARS_111":
        A = np.array([[1.0]])
        Ah = np.zeros((2, 2))
        Ah[1, 0] = 1.0
    elif name == "ARS_122":
        A = np.array([[0.5]])
        Ah = np.zeros((2, 2))
        Ah[1, 0] = 0.5
    elif name == "ARS_233":
        g = (3.0 + math.sqrt(3.0)) / 6.0
        A = np.array([[g, 0.0], [1.0 - 2.0 * g, g]])
        Ah = np.zeros((3, 3))
        Ah[1, 0], Ah[2, 0], Ah[2, 1] = g, g - 1.0, 2.0 * (1.0 - g)
    elif name == "ARS_232":
        g = 1.0 - 1.0 / math.sqrt(2.0)
        d = -2.0 * math.sqrt(2.0) / 3.0
        A = np.array([[g, 0.0], [1.0 - g, g]])
        Ah = np.zeros((3, 3))
        Ah[1, 0], Ah[2, 0], Ah[2, 1] = g, d, 1.0 - d
    elif name == "ARS_222":
        g = 1.0 - 1.0 / math.sqrt(2.0)
        d = 1.0 - 1.0 / (2.0 * g)
        A = np.array([[g, 0.0], [1.0 - g, g]])
        Ah = np.zeros((3, 3))
        Ah[1, 0], Ah[2, 0], Ah[2, 1] = g, d, 1.0 - d
    elif name == "ARS_343":
        f = lambda v: float(np.float32(v))
        g, umg, b1, b2 = f(0.43586652150845895), f(0.28206673924577053), f(1.2084966491760099), f(-0.64436317068446880)
        a42 = 0.5529291479
        A = np.array([[g, 0.0, 0.0], [umg, g, 0.0], [b1, b2, g]])
        Ah = np.zeros((4, 4))
        Ah[1, 0] = g
        p = 15.0 / 4.0 - 15.0 * g + 21.0 / 4.0 * g ** 2
        Ah[2, 0] = p * a42 - 7.0 / 2.0 + 13.0 * g - 4.5 * g ** 2
        Ah[2, 1] = -p * a42 + 4.0 - 12.5 * g + 4.5 * g ** 2
        Ah[3, 0], Ah[3, 1], Ah[3, 2] = 1.0 - a42 - a42, a42, a42
    elif name == "ARS_443":
        A = np.array([[0.5, 0, 0, 0], [1.0 / 6.0, 0.5, 0, 0], [-0.5, 0.5, 0.5, 0], [1.5, -1.5, 0.5, 0.5]])
        Ah = np.zeros((5, 5))
        Ah[1, 0] = 0.5
        Ah[2, 0], Ah[2, 1] = 11.0 / 18.0, 1.0 / 18.0
        Ah[3, 0], Ah[3, 1], Ah[3, 2] = 5.0 / 6.0, -5.0 / 6.0, 0.5
        Ah[4, 0], Ah[4, 1], Ah[4, 2], Ah[4, 3] = 0.25, 7.0 / 4.0, 0.75, -7.0 / 4.0
    else:
        raise ValueError(f"Invalid time-stepping scheme code {name!r}; valid: {', '.join(SCHEMES)}")
    return A.astype(np.float64), Ah.astype(np.float64)
