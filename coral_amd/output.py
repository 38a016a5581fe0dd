"""Coral's user outputs in the reference's file layout: the ``coral.usrOutput``
objects (volumes, x / y / z slices, vertically summed slices and horizontal
profiles of full and quadratic variables in physical space) and the
``Geometry`` files.

Objects and their schedule follow ``add_outputs``
(``src/plane_layer/PL_equations.f90:325-440``) and
``initialise_bookkeeping_counters`` / ``output_slices_volumes_and_profiles``
(``src/plane_layer/PL_IMEX_timestepping_marching.f90:620-705``): every time
step, at its first stage, the time integer advances by one and each object's
counter by one; an object whose counter reaches its period writes its file
and restarts from zero.  ``UserOutput.before_step`` is that call: make it
right before each step, on the state the step starts from.

Files (``src/plane_layer/PL_IMEX_timestepping_output.f90:250-306,474-845``,
and the single-slab writers of ``src/slabs.fftw3mpi/transforms.f90:53-70,
777-858``): raw native (little-endian) FP64, no header, Fortran order of the
reference's (z, y, x) array, i.e. z fastest:

* ``Volumes/{linear,quadra}_var%02d_time%08d_full.dat``: NZAA x NYAA x NXAA;
* ``Slices/..._z%05d_...``: the Gauss level k (1-based from the top wall),
  NYAA x NXAA (y fastest); ``_y%05d``: NZAA x NXAA; ``_x%05d``: NZAA x NYAA;
* ``Slices/..._zAvged_time%08d_full.dat``: sum_z w_z f with the
  Gauss-Chebyshev weights, NYAA x NXAA;
* ``Profiles/..._time%08d_full.dat``: the mean over x and y, NZAA values.

Variables are numbered from 1 among the full (``linear``) and the quadratic
(``quadra``) variables.  A quadratic variable's physical values are those the
reference writes: the pointwise product of its full variables on the padded
grid (``PL_IMEX_timestepping_marching.f90:278-290``), or, for a
``zFiltered`` one, the physical values of its spectrum after the removal of
the vertical mean (``:293-299``).
"""
from __future__ import annotations

import os
import sys
from dataclasses import dataclass
from typing import List

import numpy as np

from .diagnostics import gauss_chebyshev_weights

_KINDS = {">>Volume ::": "volume", ">>XSlice ::": "xSlice", ">>YSlice ::": "ySlice", ">>ZSlice ::": "zSlice",
          ">>Zsummed::": "zAvged", ">>Profile::": "profil"}


@dataclass
class OutputObject:
    kind: str          # volume, xSlice, ySlice, zSlice, zAvged, profil
    name: str          # full or quadratic variable
    period: int
    position: int = 0  # 1-based slice index
    quadratic: bool = False
    index: int = 0     # 0-based among the full or the quadratic variables
    counter: int = 0


def _value(line: str, key: str, end: int) -> int:
    i = line.index(key) + len(key)
    return int(line[i:end].strip().split()[0])


def parse_usr_output(text: str, recipe, NXAA: int, NYAA: int, NZAA: int) -> List[OutputObject]:
    """Objects of a ``coral.usrOutput`` file (lines ``>>Volume :: uu ;;
    Period= 100 <<``, ``>>ZSlice :: tFull ;; Period= 100 ;; Position= 14 <<``,
    ..., up to ``>>EOF<<``), following ``add_outputs``
    (``PL_equations.f90:325-440``): the kind from the line's first eleven
    characters, the name between ``::`` and ``;; Period=``, slices beyond the
    padded grid ignored with a message, a name that is neither a full nor a
    quadratic variable ignored.  An unknown kind raises, as the reference
    stops; so does a ``Zsummed`` object of a quadratic variable, on which the
    reference's output routine stops (``PL_IMEX_timestepping_marching.f90:
    699-701``)."""
    full = [f.name for f in recipe.full_vars]
    quad = [q.name for q in recipe.quad_vars]
    limits = {"xSlice": NXAA, "ySlice": NYAA, "zSlice": NZAA}
    out: List[OutputObject] = []
    for line in text.splitlines():
        if line[2:5] == "EOF":
            break
        if not line.strip():
            continue
        kind = _KINDS.get(line[:11])
        if kind is None:
            raise ValueError("coral.usrOutput: each line starts with >>Volume ::, >>XSlice ::, >>YSlice ::, "
                             f">>ZSlice ::, >>Profile:: or >>Zsummed:: (got {line.strip()!r})")
        end = line.index("<<")
        pi = line.index(";; Period=")
        name = line[line.index("::") + 2:pi].strip()
        if kind in limits:
            qi = line.index(";; Position=")
            period = _value(line[:qi], ";; Period=", qi)
            position = _value(line, ";; Position=", end)
            if position > limits[kind]:
                print(f"coral.usrOutput: {kind} of {name} at index {position} beyond the padded grid is ignored",
                      file=sys.stderr)
                continue
        else:
            period, position = _value(line, ";; Period=", end), 0
        if period < 1:
            raise ValueError(f"coral.usrOutput: period must be positive ({line.strip()!r})")
        if name in full:
            out.append(OutputObject(kind, name, period, position, False, full.index(name)))
        if name in quad:
            if kind == "zAvged":
                raise ValueError(f"coral.usrOutput: Zsummed of the quadratic variable {name} is not an output "
                                 "the reference writes")
            out.append(OutputObject(kind, name, period, position, True, quad.index(name)))
    return out


def file_name(obj: OutputObject, time_integer: int) -> str:
    """Path of an object's file under the output directory
    (``PL_IMEX_timestepping_output.f90:258-259,282-287,496-497,691-692``)."""
    lq = "quadra" if obj.quadratic else "linear"
    v = obj.index + 1
    if obj.kind == "volume":
        return f"Volumes/{lq}_var{v:02d}_time{time_integer:08d}_full.dat"
    if obj.kind in ("zSlice", "ySlice", "xSlice"):
        return f"Slices/{lq}_var{v:02d}_{obj.kind[0]}{obj.position:05d}_time{time_integer:08d}_full.dat"
    if obj.kind == "zAvged":
        return f"Slices/{lq}_var{v:02d}_zAvged_time{time_integer:08d}_full.dat"
    return f"Profiles/{lq}_var{v:02d}_time{time_integer:08d}_full.dat"


class UserOutput:
    """Writes a solver's ``coral.usrOutput`` objects on their schedule."""

    def __init__(self, solver, objects: List[OutputObject], directory: str, time_index: int = 0):
        self.solver = solver
        self.objects = objects
        self.directory = directory
        # initialise_bookkeeping_counters (PL_IMEX_timestepping_marching.f90:620-637)
        self.time_integer = time_index
        for o in objects:
            o.counter = time_index % o.period
        self.weights = gauss_chebyshev_weights(solver.NZAA, solver.gap)

    def field(self, obj: OutputObject) -> np.ndarray:
        """Physical values [NXAA, NYAA, NZAA] of an object's variable."""
        g = self.solver
        if not obj.quadratic:
            return g.physical(obj.index)
        q = g.recipe.quad_vars[obj.index]
        if q.remove_z_integral:
            return g.bwd_transform(g.products()[obj.index])
        if q.ivar1 == -1:
            return g.physical(q.ivar2)
        if q.ivar2 == -1:
            return g.physical(q.ivar1)
        return g.physical(q.ivar1) * g.physical(q.ivar2)

    def data(self, obj: OutputObject, f: np.ndarray) -> np.ndarray:
        """An object's array in the file's order (C order of the reverse of
        the reference's Fortran shape), from f[x, y, z]."""
        if obj.kind == "volume":
            return f                                   # z fastest, then y, then x
        if obj.kind == "zSlice":
            return f[:, :, obj.position - 1]           # (x, y): y fastest
        if obj.kind == "ySlice":
            return f[:, obj.position - 1, :]           # (x, z): z fastest
        if obj.kind == "xSlice":
            return f[obj.position - 1, :, :]           # (y, z): z fastest
        if obj.kind == "zAvged":
            return f @ self.weights                    # (x, y)
        return f.mean(axis=(0, 1))                     # profile over z

    def before_step(self) -> List[str]:
        """output_slices_volumes_and_profiles: advance the time integer and the
        counters; write the objects that are due.  Returns the files written."""
        self.time_integer += 1
        written = []
        cache = {}
        # the full variables' objects, then the quadratic ones, each variable's
        # in the order of the file (the reference's loops over iVar, iObj)
        for o in sorted(self.objects, key=lambda o: (o.quadratic, o.index)):
            o.counter += 1
            if o.counter != o.period:
                continue
            o.counter = 0
            key = (o.quadratic, o.index)
            if key not in cache:
                cache[key] = self.field(o)
            path = os.path.join(self.directory, file_name(o, self.time_integer))
            os.makedirs(os.path.dirname(path), exist_ok=True)
            np.ascontiguousarray(self.data(o, cache[key]), np.float64).astype("<f8").tofile(path)
            written.append(path)
        return written


def write_geometry(solver, directory: str):
    """``export_geometry_to_disk`` (``src/plane_layer/PL_geometry.f90:242-267``)
    for one process holding the whole domain, in the single-slab decomposition
    of ``src/slabs.fftw3mpi/domain_decomposition.f90:245-276``.
    ``Geometry/domDecmp.core0000``: int32 NX, NY, NZ, NXAA, NYAA, NZAA, the
    spectral local NX (NX/2) and NY (NYAA), Nvar (never assigned by the
    plane-layer code; written as 0), then phys start / size / end and spec
    start / size / end, three each, in the (z, x, y) and (z, y, x) axis
    orders of those arrays.  ``Geometry/geometry.core0000``: float64 Lx, Ly,
    gap, center, kx_box = 2 pi / Lx, ky_box = 2 pi / Ly, then kx (NX/2
    values, k kx_box) and ky (NYAA values: k ky_box for k < NYAA/2, else
    -(NYAA - k) ky_box) (``PL_geometry.f90:179-221``)."""
    g = solver
    d = os.path.join(directory, "Geometry")
    os.makedirs(d, exist_ok=True)
    phys = [1, 1, 1, g.NZAA, g.NXAA, g.NYAA, g.NZAA, g.NXAA, g.NYAA]
    spec = [1, 1, 1, g.NZ, g.NYAA, g.NX // 2, g.NZ, g.NYAA, g.NX // 2]
    ints = [g.NX, g.NY, g.NZ, g.NXAA, g.NYAA, g.NZAA, g.NX // 2, g.NYAA, 0] + phys + spec
    np.asarray(ints, "<i4").tofile(os.path.join(d, "domDecmp.core0000"))
    kxb, kyb = 2 * np.pi / g.Lx, 2 * np.pi / g.Ly
    kx = np.arange(g.NX // 2) * kxb
    k = np.arange(g.NYAA)
    ky = np.where(k >= g.NYAA // 2, -(g.NYAA - k) * kyb, k * kyb)
    vals = np.concatenate(([g.Lx, g.Ly, g.gap, g.center, kxb, kyb], kx, ky))
    np.asarray(vals, "<f8").tofile(os.path.join(d, "geometry.core0000"))


def read_usr_output_file(path: str, recipe, NXAA: int, NYAA: int, NZAA: int) -> List[OutputObject]:
    with open(path) as f:
        return parse_usr_output(f.read(), recipe, NXAA, NYAA, NZAA)


__all__ = ["OutputObject", "parse_usr_output", "read_usr_output_file", "file_name", "UserOutput", "write_geometry"]
