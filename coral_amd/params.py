"""Reader/writer for Coral's fixed-layout ``coral.parameters.in``.

Layout (18 lines, one value at the start of each line; the ``0`` lines are
section separators): ``src/plane_layer/PL_geometry.f90:269-305``.  NX, NY, NZ
are the dealiased mode counts; the padded grid is NXAA = 3NX/2, NYAA = 3NY/2,
NZAA = 3NZ/2 with integer division (``PL_geometry.f90:99-101``).
"""
from __future__ import annotations

import dataclasses


@dataclasses.dataclass
class Parameters:
    NX: int
    NY: int
    NZ: int
    Lx: float
    Ly: float
    gap: float = 1.0
    center: float = 0.5
    dt_max: float = 1e-3
    tolerance: float = 0.3
    cfl_c: float = 0.2
    scheme: str = "ARS_222"
    restart: bool = False
    time_limit: float = 3600.0

    @property
    def NXAA(self) -> int:
        return (3 * self.NX) // 2

    @property
    def NYAA(self) -> int:
        return (3 * self.NY) // 2

    @property
    def NZAA(self) -> int:
        return (3 * self.NZ) // 2

    def validate(self) -> None:
        """The slab layout keeps kx = 0..NX/2-1, which needs NX even; the padded
        grids must be representable (PL_geometry.f90:99-101)."""
        if self.NX < 2 or self.NY < 2 or self.NZ < 4:
            raise ValueError("resolution too small")
        if self.NX % 2 or self.NY % 2:
            raise ValueError("NX and NY must be even")
        if not (self.Lx > 0 and self.Ly > 0 and self.gap > 0):
            raise ValueError("box sizes must be positive")


def _value(line: str) -> str:
    return line.strip().split()[0]


def _fortran_float(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def parse_parameters(text: str) -> Parameters:
    """Values are read positionally from the first token of lines 0..17, as the
    Fortran list-directed reads at PL_geometry.f90:276-295 do."""
    lines = [ln for ln in text.splitlines() if ln.strip()]
    if len(lines) < 18:
        raise ValueError("coral.parameters.in must have 18 lines")
    v = [_value(ln) for ln in lines[:18]]
    p = Parameters(NX=int(v[1]), NY=int(v[2]), NZ=int(v[3]),
                   Lx=_fortran_float(v[5]), Ly=_fortran_float(v[6]),
                   gap=_fortran_float(v[7]), center=_fortran_float(v[8]),
                   dt_max=_fortran_float(v[10]), tolerance=_fortran_float(v[11]),
                   cfl_c=_fortran_float(v[12]), scheme=v[13].strip("'\""),
                   restart=v[15].strip(".").lower() in ("true", "t"),
                   time_limit=_fortran_float(v[17]))
    p.validate()
    return p


def read_parameters(path: str) -> Parameters:
    with open(path, "r") as f:
        return parse_parameters(f.read())


def format_parameters(p: Parameters) -> str:
    """Write a parameters file in Coral's 17-line layout."""
    def d(x: float) -> str:
        r = repr(float(x))
        return r.replace("e", "d") if "e" in r else r + "d0"
    rows = [
        ("0", "........... RESOLUTION ............"),
        (str(p.NX), "01 ! (INT) ... along X"),
        (str(p.NY), "02 ! (INT) ... along Y"),
        (str(p.NZ), "03 ! (INT) ... along Z"),
        ("0", "............ GEOMETRY ............."),
        (d(p.Lx), "05 ! (DBL) x/z aspect ratio"),
        (d(p.Ly), "06 ! (DBL) y/z aspect ratio"),
        (d(p.gap), "07 ! (DBL) gap"),
        (d(p.center), "08 ! (DBL) center"),
        ("0", ".......... TIMESTEPPING ..........."),
        (d(p.dt_max), "10 ! (DBL) maximum timestep"),
        (d(p.tolerance), "11 ! (DBL) timestep tolerance"),
        (d(p.cfl_c), "12 ! (DBL) CFL Factor"),
        (f"'{p.scheme}'", "13 ! scheme"),
        ("0", "......... QUICKSAVE RESTART ......."),
        (".True." if p.restart else ".False.", "15 ! (BOO) quicksave restart"),
        ("0", "............ TIMER ................"),
        (d(p.time_limit), "17 ! (DBL) Time limit (in sec.)"),
    ]
    return "\n".join(f"{a:<12}! {b}" for a, b in rows) + "\n"
