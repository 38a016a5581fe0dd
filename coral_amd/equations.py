"""Reader for Coral's ``coral.equations`` text interface.

This is the host-side half of the drop-in boundary: it turns the equation text
into the recipe tables that the C-ABI (``coral_amd_create``) consumes.  The
grammar, the meaning of each token and the bookkeeping follow Coral's parser:

* line layout ``>>keyword :: payload <<`` and the keyword state machine:
  ``src/plane_layer/PL_equations.f90:1758-2043`` (``determine_kind_of_information``);
* atom splitting on ``.`` with the operator tokens ``dx``, ``dy``, ``Iz`` (``dz``
  inside full-variable definitions) and ``mz``:
  ``src/plane_layer/PL_string_to_data.f90:53-118``;
* a signed term's scalar is its sign times the product of the named
  parameters: ``PL_equations.f90:1510-1535`` (stiffness), ``:1089-1154`` (NL);
* variables of a coupled set are numbered in the order of
  ``linearly_coupled_var`` lines and equations in the order of ``new_equation``
  lines (``PL_equations.f90:706-811``);
* a quadratic variable is the product of two full variables, ``ONE`` standing
  for 1 (``PL_equations.f90:890-954``); ``zFiltered_quadratic_variable`` sets
  ``remove_z_integral`` (``:918-925``).

* ``sourceParam`` values and ``add_rhs_source`` terms of the mean systems
  (``PL_equations.f90:985-996,1156-1230``); the sources themselves are
  profiles that the caller defines (``Solver.set_sources``), as the
  reference's user include ``PL_user_sources_definitions.f90`` does;
* penalisation variables (``enable_penalisation``, ``penalisation_strength``,
  ``penalisation_width``, ``PL_equations.f90:509-517``), whose masks the
  caller defines (``Solver.set_penalisation_mask``), as the user include
  ``PL_user_penalisation_mask.f90`` does, and the Smagorinsky variable
  ``nuSmago`` of a run with ``--withHorizontalSmagorinsky``.
"""
from __future__ import annotations

import dataclasses
import re
from typing import Dict, List, Optional, Tuple


class EquationsError(ValueError):
    """Malformed or inconsistent ``coral.equations`` input (Coral: ``stop``)."""


@dataclasses.dataclass
class LinearPiece:
    """One piece ``dsca * (dx)^kx (dy)^ky Iz^iz mz^mz`` acting on variable ``var``
    inside equation ``eqn`` (both 0-based).  Mirrors ``linear_operator_recipe_T``
    (``PL_equations.f90:59-68``)."""
    dsca: float
    kx: int
    ky: int
    iz: int
    mz: int
    var: int
    eqn: int


@dataclasses.dataclass
class NLTerm:
    """``dsca * (dx)^kx (dy)^ky Iz^iz mz^mz Q[qvar]`` added to equation ``eqn``.
    Mirrors ``atom_NL_term_T`` (``PL_equations.f90:22-30``)."""
    dsca: float
    kx: int
    ky: int
    iz: int
    mz: int
    qvar: int
    eqn: int


@dataclasses.dataclass
class SourceTerm:
    """``dsca * Iz^iz sourceNN`` added to equation ``eqn`` of mean system
    ``system`` (0-based); ``index`` = NN.  Mirrors ``source_term_T``
    (``PL_equations.f90:37-45``)."""
    dsca: float
    iz: int
    index: int
    eqn: int
    system: int


@dataclasses.dataclass
class CoupledSystem:
    """A set of linearly coupled variables with one equation per variable."""
    var_names: List[str]
    bc_codes: List[int]
    eqn_orders: List[int]
    mass: List[LinearPiece]
    stif: List[LinearPiece]
    nl: List[NLTerm]

    @property
    def nvars(self) -> int:
        return len(self.var_names)


@dataclasses.dataclass
class FullTerm:
    """``dsca * (dx)^kx (dy)^ky (dz)^dz mz^mz var`` in a full-variable build.
    ``kind`` is ``'kxky'`` or ``'zero'``; ``system``/``position`` locate the
    variable inside its coupled set (0-based).  Mirrors
    ``a_linear_contribution_T`` (``PL_equations.f90:71-79``)."""
    dsca: float
    kx: int
    ky: int
    dz: int
    mz: int
    kind: str
    system: int
    position: int


@dataclasses.dataclass
class FullVariable:
    """A full variable; extract: "nowhere", or "BotSurf" / "TopSurf" (the
    z-constant field of its value at the bottom / top wall,
    ``linear_variable_{bottom,top}_boundary_value``); zdecomp: "", or
    "barotropic" (only its vertical mean kept) / "baroclinic" (its vertical
    mean removed), ``vertical_decomposition`` (``PL_equations.f90:1052-1068``)."""
    name: str
    terms: List[FullTerm]
    extract: str = "nowhere"
    zdecomp: str = ""
    # penalisation variable (enable_penalisation, PL_equations.f90:1035-1050):
    # its physical values are fixed, strength (1 + tanh(width m)) / 2 of a mask m
    penalisation: bool = False
    penalisation_strength: float = 0.0
    penalisation_width: float = 0.0
    # nuSmago, the horizontal Smagorinsky viscosity (PL_equations.f90:519-525)
    smagorinsky: bool = False


@dataclasses.dataclass
class QuadraticVariable:
    """``Q = F1 * F2``; a factor index of ``-1`` stands for the constant ``ONE``."""
    name: str
    ivar1: int
    ivar2: int
    remove_z_integral: bool = False


@dataclasses.dataclass
class Recipe:
    """Parsed equations: mirrors ``full_problem_recipe_T``
    (``PL_equations.f90:143-219``), restricted to what the time step uses."""
    params: Dict[str, float]
    kxky_vars: List[str]
    zero_vars: List[str]
    full_vars: List[FullVariable]
    quad_vars: List[QuadraticVariable]
    kxky_systems: List[CoupledSystem]
    zero_systems: List[CoupledSystem]
    # sourceParam values by name, the name cut to its first six characters as
    # the reference's sourceParams_T%str (character(len=6), PL_equations.f90:47-50)
    source_params: Dict[str, float] = dataclasses.field(default_factory=dict)
    sources: List[SourceTerm] = dataclasses.field(default_factory=list)
    smagorinsky: bool = False

    def var_location(self, name: str) -> Tuple[str, int, int]:
        for kind, systems in (("kxky", self.kxky_systems), ("zero", self.zero_systems)):
            for isys, s in enumerate(systems):
                if name in s.var_names:
                    return kind, isys, s.var_names.index(name)
        raise EquationsError(f"variable {name!r} belongs to no coupled set")


_LINE = re.compile(r"^>>(.*)<<\s*$")


def _fortran_float(text: str) -> float:
    """Fortran real literal (``1.d0``, ``2.81D5``, ``707.1067811``) to float."""
    t = text.strip().replace("d", "e").replace("D", "e")
    try:
        return float(t)
    except ValueError as exc:
        raise EquationsError(f"cannot read a number from {text!r}") from exc


def _split_signed_atoms(expr: str) -> List[Tuple[float, str]]:
    """``"+ a.b - c"`` -> [(+1, 'a.b'), (-1, 'c')].  Coral's
    ``break_linear_operators_in_pieces`` (``PL_string_to_data.f90:330-466``)
    splits on the signs that separate atoms; a leading atom without a sign is
    taken as positive."""
    expr = expr.strip()
    out: List[Tuple[float, str]] = []
    sign = 1.0
    cur = ""
    for ch in expr:
        if ch in "+-":
            if cur.strip():
                out.append((sign, cur.strip()))
            cur = ""
            sign = 1.0 if ch == "+" else -1.0
        else:
            cur += ch
    if cur.strip():
        out.append((sign, cur.strip()))
    if not out:
        raise EquationsError(f"empty linear expression {expr!r}")
    return out


def _atom_tokens(atom: str, dz_token: str):
    """Split one atom on '.' and count operator tokens (PL_string_to_data.f90:94-110)."""
    kx = ky = iz = mz = 0
    names: List[str] = []
    for tok in atom.split("."):
        tok = tok.strip()
        if tok == "dx":
            kx += 1
        elif tok == "dy":
            ky += 1
        elif tok == dz_token:
            iz += 1
        elif tok == "mz":
            mz += 1
        elif tok:
            names.append(tok)
    if len(names) > 5:
        raise EquationsError(f"too many unrecognized tokens in {atom!r}")
    return kx, ky, iz, mz, names


def _source_term(atom: str, params: Dict[str, float], sgn: float, eqn: int, system: int) -> SourceTerm:
    """One ``add_rhs_source`` atom (``build_zero_sources_in_full_recipe``,
    ``PL_equations.f90:1156-1230``): ``Iz`` tokens count the exponent, an
    eight-character ``sourceNN`` names the source, any other name is a
    parameter multiplying the term.  As in the reference, the names after
    ``sourceNN`` in an atom are not read, and an ``mz`` token has no effect
    (the source enters through trunc(I^iz) alone, ``PL_IMEX_timestepping_marching.f90:407-409``)."""
    kx, ky, iz, _mz, names = _atom_tokens(atom, "Iz")
    if kx or ky:
        raise EquationsError(f"horizontal derivative in a mean-field source: {atom!r}")
    dsca, index = sgn, 0
    for nm in names:
        if len(nm) == 8 and nm[:6] == "source":
            try:
                index = int(nm[6:8])
            except ValueError as exc:
                raise EquationsError("Source terms must be of the form sourceNN, where NN are 2 integers") from exc
            break
        if nm in params:
            dsca = params[nm] * dsca
        else:
            raise EquationsError(f"unknown string in equations: {nm!r}")
    if not 1 <= index <= 99:
        raise EquationsError(f"invalid source_index in {atom!r}")
    return SourceTerm(dsca, iz, index, eqn, system)


# The keywords each keyword may be followed by (determine_kind_of_information,
# PL_equations.f90:1758-2043): a file that breaks the order stops the
# reference ("A different prefix was expected"), and raises here.  Keywords
# inside a coupled set are told apart by the set's kind; "term" stands for
# the add_d/dt_term / add_rhs_* lines of an equation.
_START = ("Param", "sourceParam", "linear_variable_kxky")
_QUADS = ("add_set_of_coupled_kxky_equations", "quadratic_variable", "zFiltered_quadratic_variable")
_SET_VAR = ("linearly_coupled_var", "new_equation", "add_BC_for_this_var")
_KXKY_TERMS = ("add_d/dt_term", "add_rhs_linear", "add_rhs_NL")
_ZERO_TERMS = _KXKY_TERMS + ("add_rhs_source",)
_GRAMMAR = {
    "start": _START, "Param": _START, "sourceParam": _START,
    "linear_variable_kxky": ("linear_variable_kxky", "linear_variable_mean"),
    "linear_variable_mean": ("linear_variable_mean", "linear_variable_full"),
    "linear_variable_full": ("linear_variable_full_build", "vertical_decomposition"),
    "linear_variable_bottom_boundary_value": ("linear_variable_full_build",),
    "linear_variable_top_boundary_value": ("linear_variable_full_build",),
    "vertical_decomposition": ("linear_variable_full_build",),
    "linear_variable_full_build": ("linear_variable_full_build", "linear_variable_full",
                                   "linear_variable_bottom_boundary_value", "linear_variable_top_boundary_value",
                                   "enable_penalisation", "quadratic_variable"),
    "enable_penalisation": ("penalisation_strength",),
    "penalisation_strength": ("penalisation_width",),
    "penalisation_width": ("quadratic_variable",),
    "quadratic_variable": _QUADS, "zFiltered_quadratic_variable": _QUADS,
    "add_set_of_coupled_kxky_equations": ("linearly_coupled_var",),
    "kxky:linearly_coupled_var": _SET_VAR,
    "kxky:add_BC_for_this_var": ("linearly_coupled_var", "new_equation"),
    "kxky:new_equation": _KXKY_TERMS,
    "kxky:term": _KXKY_TERMS + ("new_equation", "add_set_of_coupled_kxky_equations",
                                "add_set_of_coupled_zero_equations"),
    "add_set_of_coupled_zero_equations": ("linearly_coupled_var",),
    "zero:linearly_coupled_var": _SET_VAR,
    "zero:add_BC_for_this_var": ("linearly_coupled_var", "new_equation"),
    "zero:new_equation": _ZERO_TERMS,
    "zero:term": _ZERO_TERMS + ("new_equation", "add_set_of_coupled_zero_equations", "EOF"),
}


def _grammar_state(key: str, kind: str) -> str:
    if key in _ZERO_TERMS:
        return f"{kind}:term"
    if key in ("linearly_coupled_var", "add_BC_for_this_var", "new_equation"):
        return f"{kind}:{key}"
    return key


def _one_float(payload: str) -> float:
    """extract_one_float: the number after ``::``."""
    fields = payload.split()
    if not fields:
        raise EquationsError("missing number")
    return _fortran_float(fields[0])


def parse_equations(text: str, smagorinsky: bool = False) -> Recipe:
    """Parse the text of a ``coral.equations`` file into a :class:`Recipe`.
    ``smagorinsky`` is the run's ``--withHorizontalSmagorinsky`` flag
    (``plane_layer_main.f90:82``): a full variable ``nuSmago`` is then added
    where the first ``quadratic_variable`` line starts the quadratic
    variables (``PL_equations.f90:519-525``)."""
    params: Dict[str, float] = {}
    source_params: Dict[str, float] = {}
    sources: List[SourceTerm] = []
    kxky_vars: List[str] = []
    zero_vars: List[str] = []
    full_vars: List[FullVariable] = []
    quad_specs: List[Tuple[str, str, str, bool]] = []
    kxky_systems: List[CoupledSystem] = []
    zero_systems: List[CoupledSystem] = []
    current: Optional[CoupledSystem] = None
    current_kind = ""
    full_build_pending: List[Tuple[int, str]] = []  # (full var index, expression)
    saw_eof = False
    smagorinsky_pending = True

    state = "start"
    for lineno, raw in enumerate(text.splitlines(), 1):
        line = raw.rstrip("\r\n")
        if line.startswith(">>="):
            continue
        m = _LINE.match(line.strip())
        if not m:
            raise EquationsError(f"line {lineno}: {line!r} is not a '>>keyword :: ... <<' line")
        body = m.group(1)
        if "::" in body:
            key, payload = body.split("::", 1)
        else:
            key, payload = body, ""
        key = key.strip()
        payload = payload.strip()
        if key not in _GRAMMAR[state]:
            raise EquationsError(f"line {lineno}: {key!r} where one of {', '.join(_GRAMMAR[state])} was expected "
                                 "(a different prefix was expected)")
        if key == "EOF":
            saw_eof = True
            break
        state = _grammar_state(key, "zero" if key == "add_set_of_coupled_zero_equations" else
                               "kxky" if key == "add_set_of_coupled_kxky_equations" else current_kind)
        if key == "Param":
            name, val = payload.split(":=", 1)
            params[name.strip()] = _fortran_float(val)
        elif key == "sourceParam":
            # add_source_parameter_to_full_recipe (PL_equations.f90:985-996)
            name, val = payload.split(":=", 1)
            source_params[name.strip()[:6]] = _fortran_float(val)
        elif key == "add_rhs_source":
            if current is None or not current.eqn_orders:
                raise EquationsError(f"{key} outside an equation")
            if current_kind != "zero":
                raise EquationsError("add_rhs_source in a kxky coupled set: sources drive the mean systems only")
            eqn = len(current.eqn_orders) - 1
            for sgn, atom in _split_signed_atoms(payload):
                sources.append(_source_term(atom, params, sgn, eqn, len(zero_systems) - 1))
        elif key == "linear_variable_kxky":
            kxky_vars.append(payload)
        elif key == "linear_variable_mean":
            zero_vars.append(payload)
        elif key == "linear_variable_full":
            full_vars.append(FullVariable(payload, []))
        elif key in ("linear_variable_bottom_boundary_value", "linear_variable_top_boundary_value"):
            # PL_equations.f90:496-500, 1850-1853
            full_vars.append(FullVariable(payload, [], "BotSurf" if "bottom" in key else "TopSurf"))
        elif key == "vertical_decomposition":
            # PL_equations.f90:501-503 (get_linear_variable_zDecomposition,
            # PL_string_to_data.f90:322-342: exactly 'baroclinic' or 'barotropic')
            if not full_vars:
                raise EquationsError("vertical_decomposition before linear_variable_full")
            if payload not in ("baroclinic", "barotropic"):
                raise EquationsError(f"vertical_decomposition must be baroclinic or barotropic, got {payload!r}")
            full_vars[-1].zdecomp = payload
        elif key == "linear_variable_full_build":
            if not full_vars:
                raise EquationsError("linear_variable_full_build before linear_variable_full")
            full_build_pending.append((len(full_vars) - 1, payload))
        elif key == "enable_penalisation":
            full_vars.append(FullVariable(payload, [], penalisation=True))
        elif key in ("penalisation_strength", "penalisation_width"):
            if not full_vars or not full_vars[-1].penalisation:
                raise EquationsError(f"{key} does not follow enable_penalisation")
            setattr(full_vars[-1], key, _one_float(payload))
        elif key in ("quadratic_variable", "zFiltered_quadratic_variable"):
            if key == "quadratic_variable" and smagorinsky_pending:
                if smagorinsky:
                    full_vars.append(FullVariable("nuSmago", [], smagorinsky=True))
                smagorinsky_pending = False
            qv, rhs = payload.split(":=", 1)
            parts = rhs.strip().split(".")
            if len(parts) != 2:
                raise EquationsError(f"bad quadratic variable definition {payload!r}")
            quad_specs.append((qv.strip(), parts[0].strip(), parts[1].strip(),
                               key.startswith("zFiltered")))
        elif key in ("add_set_of_coupled_kxky_equations", "add_set_of_coupled_zero_equations"):
            current = CoupledSystem([], [], [], [], [], [])
            current_kind = "kxky" if "kxky" in key else "zero"
            (kxky_systems if current_kind == "kxky" else zero_systems).append(current)
        elif key == "linearly_coupled_var":
            if current is None:
                raise EquationsError("linearly_coupled_var outside a coupled set")
            pool = kxky_vars if current_kind == "kxky" else zero_vars
            if payload not in pool:
                raise EquationsError(f"unknown {current_kind} variable {payload!r}")
            current.var_names.append(payload)
            current.bc_codes.append(0)
        elif key == "add_BC_for_this_var":
            current.bc_codes[-1] = int(payload)
        elif key == "new_equation":
            current.eqn_orders.append(int(payload))
        elif key in ("add_d/dt_term", "add_rhs_linear", "add_rhs_NL"):
            if current is None or not current.eqn_orders:
                raise EquationsError(f"{key} outside an equation")
            eqn = len(current.eqn_orders) - 1
            for sgn, atom in _split_signed_atoms(payload):
                kx, ky, iz, mz, names = _atom_tokens(atom, "Iz")
                if current_kind == "zero" and (kx or ky):
                    raise EquationsError(f"horizontal derivative in a mean-field equation: {atom!r}")
                dsca = sgn
                target = None
                for nm in names:
                    if key == "add_rhs_NL":
                        hit = [i for i, q in enumerate(quad_specs) if q[0] == nm]
                    else:
                        hit = [i for i, v in enumerate(current.var_names) if v == nm]
                    if hit:
                        target = hit[0]
                    elif nm in params:
                        dsca = params[nm] * dsca
                    else:
                        raise EquationsError(f"unknown string in equations: {nm!r}")
                if target is None:
                    raise EquationsError(f"term {atom!r} names no variable")
                if key == "add_rhs_NL":
                    current.nl.append(NLTerm(dsca, kx, ky, iz, mz, target, eqn))
                else:
                    piece = LinearPiece(dsca, kx, ky, iz, mz, target, eqn)
                    (current.mass if key == "add_d/dt_term" else current.stif).append(piece)
        else:
            raise EquationsError(f"unknown coral.equations keyword {key!r}")

    if not saw_eof:
        raise EquationsError("missing >>EOF<< line")

    recipe = Recipe(params, kxky_vars, zero_vars, full_vars, [], kxky_systems, zero_systems, source_params, sources,
                    smagorinsky)

    # full variables are resolved once all coupled sets are known
    for ifull, expr in full_build_pending:
        for sgn, atom in _split_signed_atoms(expr):
            kx, ky, dz, mz, names = _atom_tokens(atom, "dz")
            dsca = sgn
            loc = None
            for nm in names:
                if nm in zero_vars or nm in kxky_vars:
                    loc = recipe.var_location(nm)
                elif nm in params:
                    dsca = dsca * params[nm]
                else:
                    raise EquationsError(f"unknown string in equations: {nm!r}")
            if loc is None:
                raise EquationsError(f"invalid full variable definition {atom!r}")
            kind, isys, pos = loc
            full_vars[ifull].terms.append(FullTerm(dsca, kx, ky, dz, mz, kind, isys, pos))

    names = [f.name for f in full_vars]
    for qv, a, b, filt in quad_specs:
        i1 = -1 if a == "ONE" else (names.index(a) if a in names else None)
        i2 = -1 if b == "ONE" else (names.index(b) if b in names else None)
        if i1 is None or i2 is None:
            raise EquationsError("bad NL var definition (wrong reference to linear var.)")
        if i1 == -1 and i2 == -1:
            raise EquationsError("a quadratic variable cannot be ONE.ONE")
        recipe.quad_vars.append(QuadraticVariable(qv, i1, i2, filt))

    for kind, systems in (("kxky", kxky_systems), ("zero", zero_systems)):
        for isys, s in enumerate(systems):
            if len(s.eqn_orders) != s.nvars:
                raise EquationsError(
                    f"{kind} system #{isys + 1}: {len(s.eqn_orders)} equations for {s.nvars} variables")
            for p in s.mass + s.stif:
                if p.iz > s.eqn_orders[p.eqn]:
                    raise EquationsError(
                        f"{kind} system #{isys + 1}: Iz^{p.iz} exceeds the order of equation {p.eqn + 1}")
    return recipe


def read_equations(path: str, smagorinsky: bool = False) -> Recipe:
    with open(path, "r") as f:
        return parse_equations(f.read(), smagorinsky)
