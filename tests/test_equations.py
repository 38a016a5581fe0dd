"""Host-side readers of the drop-in input files (coral.equations,
coral.parameters.in), checked term by term against the reference's fixtures."""
import pytest

from coral_amd.equations import EquationsError, parse_equations
from coral_amd.params import format_parameters, parse_parameters, read_parameters
from conftest import input_path, recipe


def test_rbc_vanilla_recipe():
    r = recipe("rbc_vanilla.coral.equations")
    assert r.params == {"Prandtl": 1.0, "iEkman": 1000.0, "RayleighPr": 1e5}
    assert r.kxky_vars == ["u", "v", "w", "t", "p", "chi"]
    assert r.zero_vars == ["uMean", "vMean", "tMean"]
    s = r.kxky_systems[0]
    # order of linearly_coupled_var lines, not of the declarations
    assert s.var_names == ["u", "v", "w", "p", "t", "chi"]
    assert s.bc_codes == [20, 20, 20, 0, 20, 0]
    assert s.eqn_orders == [2, 2, 1, 1, 0, 2]
    assert len(s.mass) == 4 and len(s.stif) == 24 and len(s.nl) == 12
    # + iEkman.Iz.Iz.v in the u equation
    p = s.stif[0]
    assert (p.dsca, p.kx, p.ky, p.iz, p.var, p.eqn) == (1000.0, 0, 0, 2, 1, 0)
    # - Iz.Iz.dx.p
    assert any(q.dsca == -1.0 and q.kx == 1 and q.iz == 2 and q.var == 3 and q.eqn == 0 for q in s.stif)
    # + RayleighPr.Iz.t in the w equation
    assert any(q.dsca == 1e5 and q.iz == 1 and q.var == 4 and q.eqn == 2 for q in s.stif)
    # continuity: + chi + dx.u + dy.v
    cont = [q for q in s.stif if q.eqn == 4]
    assert sorted((q.var, q.kx, q.ky, q.iz) for q in cont) == [(0, 1, 0, 0), (1, 0, 1, 0), (5, 0, 0, 0)]
    # NL: - Iz.Iz.dx.uu - Iz.Iz.dy.uv - Iz.uw in the u equation
    nl0 = [t for t in s.nl if t.eqn == 0]
    assert [(t.qvar, t.kx, t.ky, t.iz, t.dsca) for t in nl0] == [(0, 1, 0, 2, -1.0), (1, 0, 1, 2, -1.0), (2, 0, 0, 1, -1.0)]
    assert [q.name for q in r.quad_vars] == ["uu", "uv", "uw", "vv", "vw", "ww", "tu", "tv", "tw"]
    assert (r.quad_vars[6].ivar1, r.quad_vars[6].ivar2) == (3, 0)
    # uFull = uMean + u: mean term then kxky term
    uf = r.full_vars[0]
    assert [(t.kind, t.system, t.position) for t in uf.terms] == [("zero", 0, 0), ("kxky", 0, 0)]
    # mean systems
    z0, z1 = r.zero_systems
    assert z0.var_names == ["uMean", "vMean"] and z1.var_names == ["tMean"]
    assert any(q.dsca == 1000.0 * 1.0 and q.var == 1 and q.eqn == 0 for q in z0.stif)


def test_parameter_product_is_left_to_right():
    r = recipe("rrbc_julien.coral.equations.velocityPressure")
    s = r.kxky_systems[0]
    # + Prandtl.iEkman.Iz.Iz.v with Prandtl=1, iEkman=707.1067811
    assert any(abs(q.dsca - 707.1067811) < 1e-12 and q.var == 1 and q.eqn == 0 for q in s.stif)
    assert any(q.dsca == 2.81e5 and q.var == 4 and q.eqn == 2 for q in s.stif)


def test_heat_equation_recipe():
    r = recipe("heat_equation.coral.equations")
    assert [s.bc_codes for s in r.kxky_systems] == [[22], [23], [20]]
    assert [s.bc_codes for s in r.zero_systems] == [[22], [23], [20]]
    assert r.params["Prandtl"] == 0.33
    assert [(t.kind, t.system) for t in r.full_vars[2].terms] == [("kxky", 2), ("zero", 2)]


def skeleton(kxky=">>add_d/dt_term :: + Iz.Iz.u <<\n>>add_rhs_linear :: + a.u <<\n",
             zero=">>add_d/dt_term :: + Iz.Iz.m <<\n>>add_rhs_linear :: + a.m <<\n", head="", full=None):
    """A minimal recipe in the reference's keyword order, with the given
    equation lines for its kxky and mean systems."""
    full = full if full is not None else ">>linear_variable_full :: f <<\n>>linear_variable_full_build :: + u <<\n"
    return (">>Param :: a := 1.d0 <<\n" + head + ">>linear_variable_kxky :: u <<\n>>linear_variable_mean :: m <<\n"
            + full + ">>quadratic_variable :: ff := f.f <<\n"
            ">>add_set_of_coupled_kxky_equations <<\n>>linearly_coupled_var :: u <<\n>>add_BC_for_this_var :: 20 <<\n"
            ">>new_equation :: 2 <<\n" + kxky +
            ">>add_set_of_coupled_zero_equations <<\n>>linearly_coupled_var :: m <<\n>>add_BC_for_this_var :: 20 <<\n"
            ">>new_equation :: 2 <<\n" + zero + ">>EOF<<\n")


def test_parse_errors():
    r = parse_equations(skeleton())
    assert r.kxky_systems[0].var_names == ["u"] and r.zero_systems[0].var_names == ["m"]
    with pytest.raises(EquationsError):
        parse_equations(skeleton()[: -len(">>EOF<<\n")])  # no EOF
    with pytest.raises(EquationsError, match="unknown string"):
        parse_equations(skeleton(kxky=">>add_d/dt_term :: + Iz.Iz.u <<\n>>add_rhs_linear :: + b.u <<\n"))
    with pytest.raises(EquationsError):   # penalisation_strength only follows enable_penalisation
        parse_equations(skeleton(full=">>linear_variable_full :: f <<\n>>linear_variable_full_build :: + u <<\n"
                                      ">>penalisation_strength :: 1.e3 <<\n"))
    with pytest.raises(EquationsError, match="horizontal derivative"):
        parse_equations(skeleton(zero=">>add_d/dt_term :: + Iz.Iz.m <<\n>>add_rhs_linear :: + dx.m <<\n"))


@pytest.mark.parametrize("change", [
    ("blank line", lambda t: t.replace(">>linear_variable_mean", "\n>>linear_variable_mean", 1)),
    ("no mean variable", lambda t: t.replace(">>linear_variable_mean :: m <<\n", "", 1)),
    ("parameter after the variables", lambda t: t.replace(">>linear_variable_mean :: m <<\n",
                                                          ">>linear_variable_mean :: m <<\n>>Param :: b := 2.d0 <<\n")),
    ("vertical decomposition after a build line", lambda t: t.replace(
        ">>quadratic_variable", ">>vertical_decomposition :: baroclinic <<\n>>quadratic_variable", 1)),
    ("zFiltered first", lambda t: t.replace(">>quadratic_variable :: ff", ">>zFiltered_quadratic_variable :: ff", 1)),
    ("kxky set after a mean set", lambda t: t.replace(
        ">>EOF<<", ">>add_set_of_coupled_kxky_equations <<\n>>linearly_coupled_var :: u <<\n"
                   ">>new_equation :: 0 <<\n>>add_rhs_linear :: + u <<\n>>EOF<<")),
    ("new equation with no term", lambda t: t.replace(">>new_equation :: 2 <<\n>>add_d/dt_term :: + Iz.Iz.m",
                                                      ">>new_equation :: 2 <<\n>>new_equation :: 2 <<\n"
                                                      ">>add_d/dt_term :: + Iz.Iz.m")),
    ("source in a kxky set", lambda t: t.replace(">>add_rhs_linear :: + a.u <<",
                                                 ">>add_rhs_linear :: + a.u <<\n>>add_rhs_source :: + source01 <<")),
    ("no closing marks", lambda t: t.replace(">>linear_variable_mean :: m <<", ">>linear_variable_mean :: m")),
], ids=lambda c: c[0] if isinstance(c, tuple) else None)
def test_keyword_order_of_the_reference_grammar(change):
    """determine_kind_of_information (PL_equations.f90:1758-2043): each
    keyword may only follow certain keywords; the reference stops on any other
    line ('A different prefix was expected'), blank lines included, and only
    reads up to an EOF line that closes a mean system's equation."""
    parse_equations(skeleton())
    with pytest.raises(EquationsError):
        parse_equations(change[1](skeleton()))
    # lines starting with '>>=' are skipped anywhere
    parse_equations(skeleton().replace(">>linear_variable_mean", ">>=== a comment <<\n>>linear_variable_mean", 1))


def test_parameters_file():
    p = read_parameters(input_path("rrbc_julien.coral.parameters.in"))
    assert (p.NX, p.NY, p.NZ) == (96, 96, 48)
    assert (p.NXAA, p.NYAA, p.NZAA) == (144, 144, 72)
    assert (p.Lx, p.Ly, p.gap, p.center) == (2.0, 2.0, 1.0, 0.5)
    assert p.scheme == "ARS_222" and p.dt_max == 1e-3 and p.restart is False
    q = parse_parameters(format_parameters(p))
    assert q == p
    h = read_parameters(input_path("heat_equation.coral.parameters.in"))
    assert h.scheme == "ARS_443" and h.restart is True and h.NZ == 128
