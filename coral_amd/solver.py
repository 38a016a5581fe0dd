"""Host-side mirror of Coral's solver object on top of the C ABI.

``Solver`` plays the role of ``full_problem_data_structure_T``
(``src/plane_layer/PL_IMEX_timestepping.f90:143-219``):

==========================  =============================================
Coral                        coral_amd
==========================  =============================================
``main%init(scheme_id)``      ``Solver(recipe, params, scheme)``
``factorize_operators(dt)``   ``Solver.factorize_operators(dt)``
``march_forward(dt)``         ``Solver.march_forward(dt, nsteps=1)``
``coupled_kxky_set(i)%field`` ``Solver.get_field / set_field``
==========================  =============================================

States use Coral's layout: a kxky system's field is a complex128 array of C
shape ``(NX/2, NYAA, total)`` (Fortran ``(total, NYAA, NX/2)``), a mean
system's field a float64 array ``(total,)``.  Every compute call runs on the
GPU through ``libcoral_amd.so``.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib, sources
from ._lib import check, lib
from .diagnostics import gauss_chebyshev_grid
from .equations import Recipe
from .schemes import tableau


def _arr(ctype, values):
    values = list(values)
    return (ctype * max(1, len(values)))(*values)


def build_problem(recipe: Recipe, NX: int, NY: int, NZ: int, Lx: float, Ly: float, gap: float = 1.0,
                  center: float = 0.5, scheme: str = "ARS_222"):
    """ctypes ``coral_amd_problem`` for a recipe; returns (problem, keepalive)."""
    A, Ah = tableau(scheme)
    keep = []

    def system(s):
        bc = _arr(ctypes.c_int32, s.bc_codes)
        eo = _arr(ctypes.c_int32, s.eqn_orders)
        mass = _arr(_lib.Piece, [_lib.Piece(p.dsca, p.kx, p.ky, p.iz, p.mz, p.var, p.eqn) for p in s.mass])
        stif = _arr(_lib.Piece, [_lib.Piece(p.dsca, p.kx, p.ky, p.iz, p.mz, p.var, p.eqn) for p in s.stif])
        nl = _arr(_lib.NLTerm, [_lib.NLTerm(t.dsca, t.kx, t.ky, t.iz, t.mz, t.qvar, t.eqn) for t in s.nl])
        keep.extend([bc, eo, mass, stif, nl])
        return _lib.System(s.nvars, bc, eo, len(s.mass), mass, len(s.stif), stif, len(s.nl), nl)

    ks = _arr(_lib.System, [system(s) for s in recipe.kxky_systems])
    zs = _arr(_lib.System, [system(s) for s in recipe.zero_systems])
    fvs = []
    for f in recipe.full_vars:
        terms = _arr(_lib.FullTerm, [_lib.FullTerm(t.dsca, t.kx, t.ky, t.dz, t.mz, 0 if t.kind == "kxky" else 1,
                                                   t.system, t.position) for t in f.terms])
        keep.append(terms)
        fvs.append(_lib.FullVar(len(f.terms), terms))
    fv = _arr(_lib.FullVar, fvs)
    qv = _arr(_lib.QuadVar, [_lib.QuadVar(q.ivar1, q.ivar2, int(q.remove_z_integral)) for q in recipe.quad_vars])
    a_arr = _arr(ctypes.c_double, A.ravel())
    a_hat = _arr(ctypes.c_double, Ah.ravel())
    keep.extend([ks, zs, fv, qv, a_arr, a_hat])
    prob = _lib.Problem(NX, NY, NZ, Lx, Ly, gap, center, A.shape[0], a_arr, a_hat,
                        len(recipe.kxky_systems), ks, len(recipe.zero_systems), zs,
                        len(recipe.full_vars), fv, len(recipe.quad_vars), qv)
    return prob, keep


def assembled_operator(recipe: Recipe, NX: int, NY: int, NZ: int, Lx: float, Ly: float, isys: int, which: str,
                       kx_int: int, ky_int: int, gap: float = 1.0, center: float = 0.5):
    """Dense mass (``which='mass'``) or stiffness operator of kxky system ``isys``
    at integer wavenumbers, as assembled by the library (no GPU needed).
    Returns (matrix complex (total, total), nl, nu)."""
    prob, keep = build_problem(recipe, NX, NY, NZ, Lx, Ly, gap, center)
    t, nl, nu = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    w = 0 if which == "mass" else 1
    check(lib.coral_amd_assembled_operator(ctypes.byref(prob), isys, w, kx_int, ky_int, None, ctypes.byref(t),
                                           ctypes.byref(nl), ctypes.byref(nu)), "coral_amd_assembled_operator")
    out = np.zeros((t.value, t.value), np.complex128)
    check(lib.coral_amd_assembled_operator(ctypes.byref(prob), isys, w, kx_int, ky_int, out.ctypes.data, None,
                                           None, None), "coral_amd_assembled_operator")
    return out, nl.value, nu.value


class Solver:
    def __init__(self, recipe: Recipe, NX: int, NY: int, NZ: int, Lx: float, Ly: float,
                 gap: float = 1.0, center: float = 0.5, scheme: str = "ARS_222", device: int = 0,
                 smagorinsky_prefactor: float = 1.0):
        self.recipe = recipe
        self.NX, self.NY, self.NZ = NX, NY, NZ
        self.NXAA, self.NYAA, self.NZAA = 3 * NX // 2, 3 * NY // 2, 3 * NZ // 2
        self.Lx, self.Ly, self.gap, self.center = Lx, Ly, gap, center
        self.scheme = scheme
        self.A, self.Ah = tableau(scheme)
        prob, keep = build_problem(recipe, NX, NY, NZ, Lx, Ly, gap, center, scheme)
        ctx = ctypes.c_void_p()
        check(lib.coral_amd_create(ctypes.byref(prob), device, ctypes.byref(ctx)), "coral_amd_create")
        self._ctx = ctx
        # boundary values and vertical decompositions of the full variables
        for i, fv in enumerate(recipe.full_vars):
            ex = {"nowhere": 0, "BotSurf": 1, "TopSurf": 2}[fv.extract]
            zd = {"": 0, "barotropic": 1, "baroclinic": 2}[fv.zdecomp]
            if ex or zd:
                check(lib.coral_amd_set_full_variable_filter(self._ctx, i, ex, zd),
                      "coral_amd_set_full_variable_filter")
        # nuSmago (a recipe read with smagorinsky=True): its profile from the
        # squares of the full variables the reference names
        # (PL_IMEX_timestepping_marching.f90:519-534); --smagorinsky-prefactor
        nu = [i for i, f in enumerate(recipe.full_vars) if f.smagorinsky]
        if nu:
            grads = [i for i, f in enumerate(recipe.full_vars)
                     if f.name in ("dxUfull", "dyUfull", "dxVfull", "dyVfull")]
            g = _arr(ctypes.c_int32, grads)
            check(lib.coral_amd_set_smagorinsky(self._ctx, nu[0], g, len(grads), float(smagorinsky_prefactor)),
                  "coral_amd_set_smagorinsky")
        # the mean systems' source terms; the sources themselves: set_sources
        for t in recipe.sources:
            check(lib.coral_amd_add_source_term(self._ctx, t.system, t.eqn, t.iz, t.dsca, t.index),
                  "coral_amd_add_source_term")
        self.dt: Optional[float] = None
        self.totals = [self.system_info(i)[0] for i in range(len(recipe.kxky_systems))]
        self.zero_totals = [sum(NZ - bc // 10 for bc in s.bc_codes) for s in recipe.zero_systems]

    # --- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            check(lib.coral_amd_destroy(self._ctx), "coral_amd_destroy")
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._ctx

    # --- Coral's verbs ---------------------------------------------------------
    def factorize_operators(self, dt: float):
        check(lib.coral_amd_set_dt(self._ctx, float(dt)), "coral_amd_set_dt")
        self.dt = float(dt)

    def march_forward(self, dt: float, nsteps: int = 1):
        if self.dt is None or dt != self.dt:
            self.factorize_operators(dt)
        check(lib.coral_amd_step(self._ctx, int(nsteps)), "coral_amd_step")

    def synchronize(self):
        check(lib.coral_amd_synchronize(self._ctx), "coral_amd_synchronize")

    # --- state -----------------------------------------------------------------
    def set_field(self, kind: str, isys: int, data: np.ndarray):
        if kind == "kxky":
            a = np.ascontiguousarray(data, np.complex128)
            exp = (self.NX // 2, self.NYAA, self.totals[isys])
            if a.shape != exp:
                raise ValueError(f"kxky field must have shape {exp}, got {a.shape}")
            check(lib.coral_amd_upload_state(self._ctx, 0, isys, a.ctypes.data), "coral_amd_upload_state")
        else:
            a = np.ascontiguousarray(data, np.float64)
            if a.shape != (self.zero_totals[isys],):
                raise ValueError(f"mean field must have shape ({self.zero_totals[isys]},)")
            check(lib.coral_amd_upload_state(self._ctx, 1, isys, a.ctypes.data), "coral_amd_upload_state")

    def get_field(self, kind: str, isys: int) -> np.ndarray:
        if kind == "kxky":
            a = np.empty((self.NX // 2, self.NYAA, self.totals[isys]), np.complex128)
            check(lib.coral_amd_download_state(self._ctx, 0, isys, a.ctypes.data), "coral_amd_download_state")
        else:
            a = np.empty(self.zero_totals[isys], np.float64)
            check(lib.coral_amd_download_state(self._ctx, 1, isys, a.ctypes.data), "coral_amd_download_state")
        return a

    def set_state(self, kstates: List[np.ndarray], zstates: List[np.ndarray]):
        for i, a in enumerate(kstates):
            self.set_field("kxky", i, a)
        for i, a in enumerate(zstates):
            self.set_field("zero", i, a)

    def get_state(self) -> Tuple[List[np.ndarray], List[np.ndarray]]:
        return ([self.get_field("kxky", i) for i in range(len(self.totals))],
                [self.get_field("zero", i) for i in range(len(self.zero_totals))])

    # --- diagnostics -------------------------------------------------------------
    def cfl_dt(self) -> float:
        out = ctypes.c_double()
        check(lib.coral_amd_get_cfl_dt(self._ctx, ctypes.byref(out)), "coral_amd_get_cfl_dt")
        return out.value

    def physical(self, ifull: int) -> np.ndarray:
        out = np.empty((self.NXAA, self.NYAA, self.NZAA), np.float64)
        check(lib.coral_amd_get_physical(self._ctx, ifull, out.ctypes.data), "coral_amd_get_physical")
        return out

    def wavenumbers(self):
        kx = np.empty(self.NX // 2, np.int32)
        ky = np.empty(self.NYAA, np.int32)
        dx = np.empty(self.NX // 2, np.uint8)
        dy = np.empty(self.NYAA, np.uint8)
        check(lib.coral_amd_wavenumbers(self._ctx, kx.ctypes.data, ky.ctypes.data, dx.ctypes.data, dy.ctypes.data),
              "coral_amd_wavenumbers")
        return kx, ky, dx.astype(bool), dy.astype(bool)

    def system_info(self, isys: int):
        v = [ctypes.c_int32() for _ in range(4)]
        check(lib.coral_amd_system_info(self._ctx, isys, *[ctypes.byref(x) for x in v]), "coral_amd_system_info")
        return tuple(x.value for x in v)  # total, nl, nu, nkept

    def variable_physical(self, kind: str, isys: int, ivar: int) -> np.ndarray:
        """Physical values of coupled variable `ivar` of system `isys` (kind
        "kxky": [NXAA, NYAA, NZAA]; "zero": [NZAA] profile)."""
        k = 0 if kind == "kxky" else 1
        shape = (self.NXAA, self.NYAA, self.NZAA) if k == 0 else (self.NZAA,)
        out = np.empty(shape, np.float64)
        check(lib.coral_amd_get_variable_physical(self._ctx, k, isys, ivar, out.ctypes.data_as(ctypes.c_void_p)),
              "coral_amd_get_variable_physical")
        return out

    def set_variable_physical(self, kind: str, isys: int, ivar: int, values: np.ndarray):
        """Replace the Galerkin coefficients of one coupled variable from its
        physical values (the inverse of `variable_physical`)."""
        k = 0 if kind == "kxky" else 1
        shape = (self.NXAA, self.NYAA, self.NZAA) if k == 0 else (self.NZAA,)
        v = np.ascontiguousarray(values, np.float64).reshape(shape)
        check(lib.coral_amd_set_variable_physical(self._ctx, k, isys, ivar, v.ctypes.data_as(ctypes.c_void_p)),
              "coral_amd_set_variable_physical")

    def horizontal_averages(self):
        """(full [nfull, NZAA], quadratic [nquad, NZAA]) horizontal averages of the
        current state at the Gauss points."""
        nf, nq = len(self.recipe.full_vars), len(self.recipe.quad_vars)
        full = np.zeros((max(nf, 1), self.NZAA))
        quad = np.zeros((max(nq, 1), self.NZAA))
        check(lib.coral_amd_horizontal_averages(self._ctx, full.ctypes.data_as(ctypes.c_void_p),
                                                quad.ctypes.data_as(ctypes.c_void_p)),
              "coral_amd_horizontal_averages")
        return full[:nf], quad[:nq]

    def factor_is_real(self, isys: int = 0) -> bool:
        """True when kxky system `isys` is factored in real form."""
        v = ctypes.c_int32()
        check(lib.coral_amd_factor_kind(self._ctx, isys, ctypes.byref(v)), "coral_amd_factor_kind")
        return bool(v.value)

    def factor_blocks(self, isys: int = 0, kind: int = 0):
        """(nblk, blk_n, blk_nl, blk_nu): structural blocks factored per system
        (kind 0 = kxky, 1 = mean); nblk = 1 when whole systems are factored."""
        v = [ctypes.c_int32() for _ in range(4)]
        check(lib.coral_amd_factor_blocks(self._ctx, kind, isys, *[ctypes.byref(x) for x in v], None),
              "coral_amd_factor_blocks")
        return tuple(x.value for x in v)

    def u_entries_per_column(self, isys: int = 0, kind: int = 0) -> float:
        """Mean number of U entries per column the solve reads (records cut to
        their longest nonzero prefix over all wavenumbers; after factorization)."""
        v = [ctypes.c_int32() for _ in range(4)]
        u = ctypes.c_double()
        check(lib.coral_amd_factor_blocks(self._ctx, kind, isys, *[ctypes.byref(x) for x in v], ctypes.byref(u)),
              "coral_amd_factor_blocks")
        return u.value

    def bwd_transform(self, spec: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(spec, np.complex128)
        if s.shape != (self.NX // 2, self.NYAA, self.NZ):
            raise ValueError("spectral field must have shape (NX/2, NYAA, NZ)")
        out = np.empty((self.NXAA, self.NYAA, self.NZAA), np.float64)
        check(lib.coral_amd_bwd_transform(self._ctx, s.ctypes.data, out.ctypes.data), "coral_amd_bwd_transform")
        return out

    def fwd_transform(self, phys: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(phys, np.float64)
        if p.shape != (self.NXAA, self.NYAA, self.NZAA):
            raise ValueError("physical field must have shape (NXAA, NYAA, NZAA)")
        out = np.empty((self.NX // 2, self.NYAA, self.NZ), np.complex128)
        check(lib.coral_amd_fwd_transform(self._ctx, p.ctypes.data, out.ctypes.data), "coral_amd_fwd_transform")
        return out

    def solve_system(self, isys: int, rhs: np.ndarray) -> np.ndarray:
        """(M - dt a11 L) x = rhs for every kept wavenumber of kxky system
        `isys` with the factors of the last factorization and the solve kernel
        of the time step.  rhs, x: complex (NX/2, NYAA, total)."""
        r = np.ascontiguousarray(rhs, np.complex128)
        exp = (self.NX // 2, self.NYAA, self.totals[isys])
        if r.shape != exp:
            raise ValueError(f"rhs must have shape {exp}, got {r.shape}")
        x = np.empty_like(r)
        check(lib.coral_amd_system_solve(self._ctx, isys, r.ctypes.data, x.ctypes.data), "coral_amd_system_solve")
        return x

    # --- timing ---------------------------------------------------------------------
    def products(self) -> np.ndarray:
        """Spectra of the quadratic variables of the current state,
        complex (nquad, NX/2, NYAA, NZ); dealiased ky slots are zero."""
        nq = len(self.recipe.quad_vars)
        out = np.zeros((nq, self.NX // 2, self.NYAA, self.NZ), np.complex128)
        check(lib.coral_amd_products(self._ctx, out.ctypes.data), "coral_amd_products")
        return out

    def set_sources(self, definitions: Dict[int, "sources.Definition"]):
        """Define sources by index NN (the ``sourceNN`` of ``add_rhs_source``
        lines): a callable ``f(z, params)`` of the NZAA Gauss-Chebyshev points
        (top wall first) and the recipe's ``sourceParams`` lookup, an array of
        the NZAA values, or a constant (``init_all_sources``,
        ``PL_IMEX_timestepping_marching.f90:829-843``).  Only the sources that
        the recipe's terms name are evaluated; ``coral_amd.sources.SHIPPED``
        is the reference's shipped definitions file.  Stepping fails while a
        named source has no definition."""
        z = gauss_chebyshev_grid(self.NZAA, self.center, self.gap)
        params = sources.SourceParams(self.recipe.source_params)
        used = {t.index for t in self.recipe.sources}
        for nn, d in definitions.items():
            if nn not in used:
                continue
            vals = sources.evaluate(d, z, params)
            check(lib.coral_amd_set_source(self._ctx, int(nn), vals.ctypes.data), "coral_amd_set_source")

    def set_penalisation_mask(self, definition=None, name: Optional[str] = None):
        """Fix the physical values of a penalisation variable (the one named,
        or every one of the recipe) to strength (1 + tanh(width max_i m_i)) / 2
        of the masks m_i of ``definition`` (``coral_amd.penalisation``; the
        shipped example masks by default), as ``init_penalisation`` does
        (``PL_IMEX_timestepping_marching.f90:778-822``)."""
        from . import penalisation
        definition = definition or penalisation.SHIPPED
        for i, f in enumerate(self.recipe.full_vars):
            if not f.penalisation or (name is not None and f.name != name):
                continue
            field = penalisation.mask_field(definition, f.penalisation_strength, f.penalisation_width, self.NXAA,
                                            self.NYAA, self.NZAA, self.Lx, self.Ly, self.center, self.gap)
            self.set_fixed_physical(i, field)

    def set_fixed_physical(self, ifull: int, values: Optional[np.ndarray]):
        """Fixed physical values [x, y, z] of full variable ``ifull``
        (``coral_amd_set_fixed_physical``); None returns it to its spectral
        definition."""
        if values is None:
            check(lib.coral_amd_set_fixed_physical(self._ctx, ifull, None), "coral_amd_set_fixed_physical")
            return
        v = np.ascontiguousarray(values, np.float64)
        if v.shape != (self.NXAA, self.NYAA, self.NZAA):
            raise ValueError(f"fixed physical values have shape {(self.NXAA, self.NYAA, self.NZAA)}, got {v.shape}")
        check(lib.coral_amd_set_fixed_physical(self._ctx, ifull, v.ctypes.data), "coral_amd_set_fixed_physical")

    def set_precision(self, mode: str = "fp64", refine: int = 0):
        """``"fp64"``: the FP64 step; ``"mixed"``: FP32 transforms and products
        (compensated FP32 MFMA z products), FP64 NL sums, right-hand sides,
        factors and solves.  ``refine``: FP64 residual-refinement sweeps after
        the solves (0 none, 1 after the last stage of each step, 2 after every
        stage)."""
        m = {"fp64": 0, "mixed": 1}[mode]
        check(lib.coral_amd_set_precision(self._ctx, m, int(refine)), "coral_amd_set_precision")

    def set_kx0_projection(self, on: bool = True):
        """Replace the kx = 0 plane by its Hermitian part after every step
        (``coral_amd_set_kx0_projection``): removes the component the physical
        fields cannot see, which otherwise grows without bound in long
        supercritical runs."""
        check(lib.coral_amd_set_kx0_projection(self._ctx, int(bool(on))), "coral_amd_set_kx0_projection")

    def set_closure(self, model=None):
        """Attach a trained closure (``coral_amd.closure.Model``: layout, flat
        weights and normalisation) or remove it (None): every stage then adds
        the network's correction to the quadratic variables
        (``coral_amd_set_closure``)."""
        if model is None:
            check(lib.coral_amd_set_closure(self._ctx, None), "coral_amd_set_closure")
            return
        f32 = lambda a: np.ascontiguousarray(a, np.float32)
        arrs = [f32(model.params), f32(model.in_mean), f32(model.in_std), f32(model.out_mean), f32(model.out_std)]
        P = ctypes.POINTER(ctypes.c_float)
        d = _lib.ClosureDesc(model.layout.width, model.layout.hidden, model.layout.layers, model.modes,
                             arrs[0].ctypes.data_as(P), arrs[0].size, arrs[1].ctypes.data_as(P),
                             arrs[2].ctypes.data_as(P), arrs[3].ctypes.data_as(P), arrs[4].ctypes.data_as(P))
        check(lib.coral_amd_set_closure(self._ctx, ctypes.byref(d)), "coral_amd_set_closure")

    def set_launch_shape(self, xpass_wpb: int = 0, z_plain: bool = False):
        """Launch shapes (``coral_amd_set_launch_shape``): x pass wavefronts per
        workgroup (0 = most that fit) and vector-GEMM z transforms."""
        check(lib.coral_amd_set_launch_shape(self._ctx, int(xpass_wpb), int(bool(z_plain))),
              "coral_amd_set_launch_shape")

    def launch_shape(self) -> Tuple[int, bool]:
        w, z = ctypes.c_int32(), ctypes.c_int32()
        check(lib.coral_amd_get_launch_shape(self._ctx, ctypes.byref(w), ctypes.byref(z)), "coral_amd_get_launch_shape")
        return w.value, bool(z.value)

    def use_graphs(self, on: bool = True):
        """Replay steps as HIP graphs (captured at the next step); results are
        the eager ones bit for bit."""
        check(lib.coral_amd_use_graphs(self._ctx, 1 if on else 0), "coral_amd_use_graphs")

    def timers_enable(self, on: bool = True):
        check(lib.coral_amd_timers_enable(self._ctx, 1 if on else 0), "coral_amd_timers_enable")

    def timers(self) -> Dict[str, Tuple[float, int]]:
        name = ctypes.create_string_buffer(64)
        ms = ctypes.c_double()
        cnt = ctypes.c_int32()
        n = check(lib.coral_amd_timer_read(self._ctx, 0, name, 64, ctypes.byref(ms), ctypes.byref(cnt)))
        out = {}
        for i in range(n):
            check(lib.coral_amd_timer_read(self._ctx, i, name, 64, ctypes.byref(ms), ctypes.byref(cnt)))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out


def banded_solve(ab: np.ndarray, kl: int, ku: int, b: np.ndarray, device: int = 0):
    """Batched LAPACK-band LU + solve on the GPU.  ``ab`` complex (nsys, n, 2kl+ku+1)
    (column j of system s contiguous), ``b`` complex (nsys, n).  Returns (x, info)."""
    ab = np.ascontiguousarray(ab, np.complex128)
    x = np.ascontiguousarray(b, np.complex128).copy()
    nsys, n, ldab = ab.shape
    if ldab != 2 * kl + ku + 1:
        raise ValueError("ab must have 2kl+ku+1 rows per column")
    info = np.zeros(nsys, np.int32)
    check(lib.coral_amd_banded_solve(device, n, kl, ku, nsys, ab.ctypes.data, x.ctypes.data, info.ctypes.data),
          "coral_amd_banded_solve")
    return x, info
