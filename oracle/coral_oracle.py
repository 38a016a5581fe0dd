"""ORACLE -- test infrastructure only.

CPU restatement of Coral's plane-layer IMEX time step, used by ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` as the
checker.  The product (``coral_amd``) never imports this module.

Each function cites the reference file:line it restates (paths relative to the
reference checkout).  Conventions are the reference's single-rank slab layout
(``src/slabs.fftw3mpi``): spectral arrays hold kx = 0..NX/2-1 and all NYAA ky
slots, systems are stored ``[ix, iy, :]`` (C order), i.e. Fortran
``(total, NYAA, NX/2)`` as in ``PL_IMEX_timestepping.f90:602-623``.

Third-party arithmetic that the reference takes from FFTW 3.3.10 and Intel MKL
(``etc/set_env_examples/set_env_template.sh:20,29``) is restated with
numpy/scipy (pocketfft for the FFTs, DCT-II/III = FFTW REDFT10/REDFT01
unnormalised) and with ``oracle/banded.c`` (LAPACK xGBTF2/xGBTRS).

Pins (see tests/test_oracle_*.py): closed-form wavenumbers and masks, the
Gauss-Chebyshev DCT normalisation, transform round trip, brute-force dealiased
products on tiny grids, Chebyshev integration/multiplication recurrences,
Galerkin stencils annihilated by their boundary functionals, the structural
band of the RBC system, the stress-free Rayleigh-Benard growth rate
(24.137487) and the heat-equation error table of
``etc/benchmarks/00_heat_equation/error_report.py:4-15``; the banded LU is
checked against scipy's LAPACK zgbtrf/zgbtrs and the 5x5 known answer of
``src/plane_layer/driver_algebra.f90:24-47``.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
# threads for scipy's FFTs (the C parts use OpenMP with OMP_NUM_THREADS)
WORKERS = int(os.environ.get("ORACLE_THREADS", os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle_banded.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle C library not built: run `make -C {_HERE}`")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int
        lib.oracle_zgbtrf_batch.argtypes = [I, I, I, I, P, I, P, P]
        lib.oracle_zgbtrs_batch.argtypes = [I, I, I, I, P, I, P, P, I]
        lib.oracle_band_matvec_batch.argtypes = [I, I, I, I, P, I, P, P, I, P, I]
        _LIB = lib
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------------------
# Banded LU (restates LAPACK as called at PL_algebra_z_3D.f90:189-276)
# ----------------------------------------------------------------------------

def zgbtrf_batch(ab: np.ndarray, kl: int, ku: int):
    """Factor in place.  ``ab``: complex128, shape (nsys, n, ldab) with
    ldab = 2kl+ku+1 (column j of LAPACK band storage contiguous).
    Returns (ipiv[int32, (nsys, n)], info[int32, nsys])."""
    assert ab.dtype == np.complex128 and ab.flags.c_contiguous
    nsys, n, ldab = ab.shape
    assert ldab == 2 * kl + ku + 1
    ipiv = np.zeros((nsys, n), np.int32)
    info = np.zeros(nsys, np.int32)
    _lib().oracle_zgbtrf_batch(n, kl, ku, nsys, _ptr(ab), ldab, _ptr(ipiv), _ptr(info))
    return ipiv, info


def zgbtrs_batch(ab: np.ndarray, kl: int, ku: int, ipiv: np.ndarray, b: np.ndarray):
    """Solve in place; ``b`` complex128 (nsys, n)."""
    nsys, n, ldab = ab.shape
    assert b.dtype == np.complex128 and b.flags.c_contiguous and b.shape == (nsys, n)
    _lib().oracle_zgbtrs_batch(n, kl, ku, nsys, _ptr(ab), ldab, _ptr(ipiv), _ptr(b), n)


def band_matvec_batch(bands: np.ndarray, nl: int, nu: int, coef: np.ndarray, x: np.ndarray):
    """y[s] = sum_p coef[s,p] * B_p x[s] for banded real pieces ``bands``
    (npieces, nl+nu+1, n) with B_p(i, i+d) = bands[p, d+nl, i]."""
    npieces, w, n = bands.shape
    nsys = x.shape[0]
    bands = np.ascontiguousarray(bands, np.float64)
    coef = np.ascontiguousarray(coef, np.complex128)
    x = np.ascontiguousarray(x, np.complex128)
    y = np.empty_like(x)
    _lib().oracle_band_matvec_batch(n, nl, nu, npieces, _ptr(bands), nsys, _ptr(coef),
                                    _ptr(x), n, _ptr(y), n)
    return y


def dense_to_lapack_band(A: np.ndarray, kl: int, ku: int) -> np.ndarray:
    """(n, n) -> (n, 2kl+ku+1) LAPACK band with the band copied into rows
    kl..2kl+ku (0-based), as PL_algebra_z_3D.f90:258 does from the DIA form."""
    n = A.shape[0]
    ldab = 2 * kl + ku + 1
    ab = np.zeros((n, ldab), A.dtype)
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[j, kl + ku + i - j] = A[i, j]
    return ab


# ----------------------------------------------------------------------------
# Geometry (PL_geometry.f90:80-236, single rank, slab layout)
# ----------------------------------------------------------------------------

class Geometry:
    """Wavenumbers, dealiasing masks and grids of the single-rank slab layout.

    * NXAA = 3NX/2, NYAA = 3NY/2, NZAA = 3NZ/2 (PL_geometry.f90:99-101);
    * spectral extents: z NZ, y NYAA, x NX/2 (slabs.fftw3mpi/domain_decomposition.f90:215-233);
    * ky(iy) = (k-1)*ky_box if k < NYAA/2+1 else -(NYAA+1-k)*ky_box, k = iy+1
      (PL_geometry.f90:179-198); kx likewise (:200-222);
    * deAliase = |k| >= k_box*N/2 (:193-197, :217-221).
    """

    def __init__(self, NX, NY, NZ, Lx, Ly, gap=1.0, center=0.5):
        self.NX, self.NY, self.NZ = int(NX), int(NY), int(NZ)
        self.NXAA, self.NYAA, self.NZAA = (3 * NX) // 2, (3 * NY) // 2, (3 * NZ) // 2
        self.Lx, self.Ly, self.gap, self.center = float(Lx), float(Ly), float(gap), float(center)
        self.nkx = self.NX // 2
        self.nky = self.NYAA
        pi = 4.0 * math.atan(1.0)
        self.kx_box = 2.0 * pi / self.Lx
        self.ky_box = 2.0 * pi / self.Ly
        self.ky_int = np.empty(self.nky, np.int64)
        for iy in range(self.nky):
            k = iy + 1
            self.ky_int[iy] = -(self.NYAA + 1 - k) if k >= self.NYAA // 2 + 1 else k - 1
        self.kx_int = np.empty(self.nkx, np.int64)
        for ix in range(self.nkx):
            k = ix + 1
            self.kx_int[ix] = -(self.NXAA + 1 - k) if k >= self.NXAA // 2 + 1 else k - 1
        self.ky = self.ky_int.astype(np.float64) * self.ky_box
        self.kx = self.kx_int.astype(np.float64) * self.kx_box
        kcut_y = self.ky_box * self.NY / 2.0
        kcut_x = self.kx_box * self.NX / 2.0
        self.dealias_y = np.abs(self.ky) >= kcut_y
        self.dealias_x = np.abs(self.kx) >= kcut_x
        self.px = 1j * self.kx
        self.py = 1j * self.ky
        # mask[ix, iy]: True where the reference zeroes the mode
        self.dealias = self.dealias_x[:, None] | self.dealias_y[None, :]

    def gauss_chebyshev_grid(self, n: Optional[int] = None) -> np.ndarray:
        """z_i = center + gap/2 cos(pi (2i-1)/(2N)), i=1..N (PL_cheby_misc.f90:8-27)."""
        n = self.NZAA if n is None else n
        pi = 4.0 * math.atan(1.0)
        i = np.arange(1, n + 1)
        return self.center + np.cos(pi * (2 * i - 1) / (2.0 * n)) * 0.5 * self.gap

    def gauss_chebyshev_weight(self, n: Optional[int] = None) -> np.ndarray:
        """Midpoint widths of the Gauss-Chebyshev grid (PL_cheby_misc.f90:29-55)."""
        n = self.NZAA if n is None else n
        pi = 4.0 * math.atan(1.0)
        g = np.cos(pi * (2 * np.arange(1, n + 1) - 1) / (2.0 * n))
        mid = np.empty(n + 1)
        mid[0], mid[n] = 1.0, -1.0
        mid[1:n] = 0.5 * (g[1:] + g[:-1])
        return (mid[:-1] - mid[1:]) / 2.0 * self.gap


# ----------------------------------------------------------------------------
# Chebyshev elementary matrices and Galerkin stencils
# ----------------------------------------------------------------------------

def chebyshev_elementary_matrices(N: int, center: float, gap: float, padding: int):
    """CEM_I (integration) and CEM_M (multiplication by z), size N+padding
    (chebyshev_elementary.f90:28-151).  Returns dense (I, M)."""
    Np = N + padding
    M = np.zeros((Np, Np))
    M[1, 0] = gap / 2.0                      # :66-70
    for i in range(Np - 1):                  # super-diagonal gap/4 (:72-77)
        M[i, i + 1] = gap / 4.0
    for i in range(1, Np - 1):               # sub-diagonal gap/4 (:80-85)
        M[i + 1, i] = gap / 4.0
    M[np.arange(Np), np.arange(Np)] = center  # :88-93
    I = np.zeros((Np, Np))
    I[1, 0] = 2.0                            # :118-123
    # sub-diagonal, 1-based (row i+1, col i) = 1/i for i = 2..Np-1 (:125-130)
    I[np.arange(2, Np), np.arange(1, Np - 1)] = 1.0 / np.arange(2, Np, dtype=np.float64)
    I[np.arange(1, Np - 1), np.arange(2, Np)] = -1.0 / np.arange(1, Np - 1, dtype=np.float64)  # :132-137
    I *= gap / 4.0                           # :139
    return I, M


def cheb_IM(NZ: int, center: float, gap: float, iz: int, mz: int, padding: int = 15) -> np.ndarray:
    """trunc_{NZ x NZ}(I^iz M^mz) built with padded factors
    (PL_IMEX_timestepping.f90:726-791, sparse_formats_d.f90:152-220)."""
    I, M = chebyshev_elementary_matrices(NZ, center, gap, padding)
    A = np.linalg.matrix_power(I, iz) if iz > 0 else np.eye(I.shape[0])
    if mz > 0:
        A = A @ np.linalg.matrix_power(M, mz)
    return A[:NZ, :NZ].copy()


def galerkin_stencil(N: int, bc: int) -> np.ndarray:
    """Dense N x (N - bc//10) Galerkin stencil (chebyshev_galerkin_2.f90:7-440)."""
    if bc == 0:
        return np.eye(N)
    ncol = N - bc // 10
    S = np.zeros((N, ncol))
    for j0 in range(ncol):
        j = j0 + 1  # 1-based column index of the reference loops
        if bc == 10:
            S[j0, j0], S[j0 + 1, j0] = -1.0, 1.0
        elif bc == 11:
            S[j0, j0], S[j0 + 1, j0] = 1.0, 1.0
        elif bc == 12:
            S[j0, j0], S[j0 + 1, j0] = -(j + 1.0) ** 2, float(j) ** 2
        elif bc == 13:
            S[j0, j0], S[j0 + 1, j0] = (j + 1.0) ** 2, float(j) ** 2
        elif bc == 20:
            S[j0, j0], S[j0 + 2, j0] = -1.0, 1.0
        elif bc == 21:
            n = float(j + 1)
            S[j0, j0], S[j0 + 2, j0] = -1.0, (n - 2.0) ** 2 / n ** 2
        elif bc in (22, 23):
            n = float(j - 1)
            sgn = -1.0 if bc == 22 else 1.0
            S[j0, j0] = 1.0
            S[j0 + 1, j0] = sgn * (4.0 * n + 4.0) / (2.0 * n ** 2 + 6.0 * n + 5.0)
            S[j0 + 2, j0] = -(n ** 2 + (n + 1.0) ** 2) / ((n + 1.0) ** 2 + (n + 2.0) ** 2)
        elif bc == 24:
            if j <= 2:
                S[j0, j0] = 1.0
            else:
                n = float(j + 1)
                S[j0 + 2, j0] = 1.0
                S[j0, j0] = -n ** 2 * (n ** 2 + 1.0) / ((n ** 2 - 3.0 * n + 2.0) * (n ** 2 - 5.0 * n + 6.0))
        elif bc == 40:
            n = float(j + 3)
            S[j0, j0] = -1.0
            S[j0 + 2, j0] = (2.0 * n - 4.0) / (n - 1.0)
            S[j0 + 4, j0] = -(n - 3.0) / (n - 1.0)
        elif bc == 41:
            n = float(j + 3)
            m = 24.0 * n / (2.0 * n ** 2 - 4.0 * n + 3.0) - 18.0 / (n - 1.0)
            S[j0, j0], S[j0 + 2, j0], S[j0 + 4, j0] = -1.0, m + 2.0, -m - 1.0
        elif bc in (42, 43):
            n = float(j - 1)
            s = -1.0 if bc == 42 else 1.0
            S[j0, j0] = 1.0
            S[j0 + 1, j0] = s * (2.0 * n + 2.0) / (n ** 2 + 5.0 * n + 7.0)
            S[j0 + 2, j0] = -(2.0 * n ** 3 + 12.0 * n ** 2 + 28.0 * n + 24.0) / (n ** 3 + 8.0 * n ** 2 + 22.0 * n + 21.0)
            S[j0 + 3, j0] = -s * (2.0 * n + 2.0) / (n ** 2 + 5.0 * n + 7.0)
            S[j0 + 4, j0] = (n ** 3 + 4.0 * n ** 2 + 6.0 * n + 3.0) / (n ** 3 + 8.0 * n ** 2 + 22.0 * n + 21.0)
        else:
            raise ValueError(f"unknown boundary-condition code {bc}")
    return S


def chebyshev_derivative_matrix(NZ: int, center: float, gap: float) -> np.ndarray:
    """d/dz on NZ Chebyshev coefficients as the reference computes it: shift
    the coefficients down by one and solve with trunc(CEM_I)[2:NZ, 1:NZ-1]; the
    last coefficient is zero (PL_IMEX_timestepping.f90:276-319, :488-506)."""
    I, _ = chebyshev_elementary_matrices(NZ, center, gap, 10)
    It = I[1:NZ, 0:NZ - 1]
    D = np.zeros((NZ, NZ))
    shift = np.zeros((NZ - 1, NZ))
    shift[np.arange(NZ - 1), np.arange(1, NZ)] = 1.0
    D[:NZ - 1, :] = np.linalg.solve(It, shift)
    return D


# ----------------------------------------------------------------------------
# IMEX tableaux (IMEX_schemes.f90:21-402)
# ----------------------------------------------------------------------------

def imex_scheme(name: str):
    """Returns (A (s x s), A_hat ((s+1) x (s+1))) of the ARS97 scheme."""
    if name == "ARS_111":
        A = np.array([[1.0]])
        Ah = np.zeros((2, 2)); Ah[1, 0] = 1.0
    elif name == "ARS_122":
        A = np.array([[0.5]])
        Ah = np.zeros((2, 2)); Ah[1, 0] = 0.5
    elif name == "ARS_233":
        g = (3.0 + math.sqrt(3.0)) / 6.0
        A = np.array([[g, 0.0], [1.0 - 2.0 * g, g]])
        Ah = np.zeros((3, 3)); Ah[1, 0] = g; Ah[2, 0] = g - 1.0; Ah[2, 1] = 2.0 * (1.0 - g)
    elif name == "ARS_232":
        d = -2.0 * math.sqrt(2.0) / 3.0
        g = 1.0 - 1.0 / math.sqrt(2.0)
        A = np.array([[g, 0.0], [1.0 - g, g]])
        Ah = np.zeros((3, 3)); Ah[1, 0] = g; Ah[2, 0] = d; Ah[2, 1] = 1.0 - d
    elif name == "ARS_222":
        g = 1.0 - 1.0 / math.sqrt(2.0)
        d = 1.0 - 1.0 / (2.0 * g)
        A = np.array([[g, 0.0], [1.0 - g, g]])
        Ah = np.zeros((3, 3)); Ah[1, 0] = g; Ah[2, 0] = d; Ah[2, 1] = 1.0 - d
    elif name == "ARS_343":
        f32 = lambda v: float(np.float32(v))  # default-real literals (IMEX_schemes.f90:296-301)
        g, umg = f32(0.43586652150845895), f32(0.28206673924577053)
        b1, b2 = f32(1.2084966491760099), f32(-0.64436317068446880)
        a42 = 0.5529291479
        A = np.array([[g, 0, 0], [umg, g, 0], [b1, b2, g]], dtype=np.float64)
        Ah = np.zeros((4, 4))
        Ah[1, 0] = g
        Ah[2, 0] = (15.0 / 4.0 - 15.0 * g + 21.0 / 4.0 * g ** 2) * a42 - 7.0 / 2.0 + 13.0 * g - 4.5 * g ** 2
        Ah[2, 1] = -(15.0 / 4.0 - 15.0 * g + 21.0 / 4.0 * g ** 2) * a42 + 4.0 - 12.5 * g + 4.5 * g ** 2
        Ah[3, 0] = 1.0 - a42 - a42
        Ah[3, 1] = a42
        Ah[3, 2] = a42
    elif name == "ARS_443":
        A = np.array([[0.5, 0, 0, 0], [1.0 / 6.0, 0.5, 0, 0], [-0.5, 0.5, 0.5, 0], [1.5, -1.5, 0.5, 0.5]])
        Ah = np.zeros((5, 5))
        Ah[1, 0] = 0.5
        Ah[2, 0], Ah[2, 1] = 11.0 / 18.0, 1.0 / 18.0
        Ah[3, 0], Ah[3, 1], Ah[3, 2] = 5.0 / 6.0, -5.0 / 6.0, 0.5
        Ah[4, 0], Ah[4, 1], Ah[4, 2], Ah[4, 3] = 0.25, 7.0 / 4.0, 0.75, -7.0 / 4.0
    else:
        raise ValueError(f"unknown IMEX scheme {name!r}")
    return A, Ah


# ----------------------------------------------------------------------------
# Operators of one coupled system (PL_IMEX_timestepping.f90:870-1574)
# ----------------------------------------------------------------------------

class SystemOps:
    """Shapes, shuffles and real operator pieces of one coupled system.

    * sizes: var_size = NZ - bc//10, eqn_size = NZ - order, total = sum
      (compute_systems_shapes, PL_IMEX_timestepping.f90:870-936);
    * shuffles: the first nmin coefficients of every variable (equation) are
      interleaved by degree, the remainders appended in order
      (build_*_shuffle_matrices_for_system, :1344-1414);
    * a piece's block is dsca * trunc_rows(order)(I^iz M^mz) * stencil(var)
      placed at (eqn_first, var_first) (add_*_times_MIblock, :793-856), then the
      whole matrix is permuted Q A P (change_of_basis_zCSR_dMat,
      PL_algebra_z_3D.f90:40-53).
    """

    def __init__(self, system, NZ: int, center: float, gap: float):
        self.sys = system
        self.NZ = NZ
        nv = system.nvars
        self.nvars = nv
        self.var_size = [NZ - bc // 10 for bc in system.bc_codes]
        self.eqn_size = [NZ - o for o in system.eqn_orders]
        self.total = sum(self.var_size)
        if self.total != sum(self.eqn_size):
            raise ValueError("incompatible number of boundary conditions and equation orders")
        self.var_first = np.concatenate([[0], np.cumsum(self.var_size)[:-1]]).astype(int)
        self.eqn_first = np.concatenate([[0], np.cumsum(self.eqn_size)[:-1]]).astype(int)
        self.posP = self._shuffle(self.var_size)   # var-order index -> interleaved
        self.posQ = self._shuffle(self.eqn_size)   # eqn-order index -> interleaved
        self.stencils = [galerkin_stencil(NZ, bc) for bc in system.bc_codes]
        self._im_cache: Dict[Tuple[int, int], np.ndarray] = {}
        self.center, self.gap = center, gap

    def _shuffle(self, sizes: Sequence[int]) -> np.ndarray:
        nmin = min(sizes)
        nv = len(sizes)
        pos = []
        offset = 0
        for v in range(nv):
            for c in range(nmin):
                pos.append(v + c * nv)
            for _ in range(sizes[v] - nmin):
                pos.append(nmin * nv + offset)
                offset += 1
        return np.array(pos, dtype=np.int64)

    def IM(self, iz: int, mz: int) -> np.ndarray:
        key = (iz, mz)
        if key not in self._im_cache:
            self._im_cache[key] = cheb_IM(self.NZ, self.center, self.gap, iz, mz)
        return self._im_cache[key]

    def piece_matrix(self, p) -> np.ndarray:
        """Real total x total matrix of one piece (dsca folded in), shuffled."""
        order = self.sys.eqn_orders[p.eqn]
        block = self.IM(p.iz, p.mz)[order:, :] @ self.stencils[p.var]
        A = np.zeros((self.total, self.total))
        r0, c0 = self.eqn_first[p.eqn], self.var_first[p.var]
        A[r0:r0 + block.shape[0], c0:c0 + block.shape[1]] += p.dsca * block
        S = np.zeros_like(A)
        S[np.ix_(self.posQ, self.posP)] = A
        return S

    def grouped_pieces(self, pieces) -> List[Tuple[Tuple[int, int], np.ndarray]]:
        """Sum pieces sharing the same (kx, ky) exponents."""
        groups: Dict[Tuple[int, int], np.ndarray] = {}
        for p in pieces:
            key = (p.kx, p.ky)
            m = self.piece_matrix(p)
            groups[key] = groups.get(key, 0) + m
        return sorted(groups.items())

    def extract(self, var: int, aux: np.ndarray) -> np.ndarray:
        """padStencilExtractShuffle(var) . aux: Chebyshev coefficients (NZ) of
        variable ``var`` from interleaved Galerkin coefficients
        (PL_IMEX_timestepping.f90:1199-1209).  ``aux`` (..., total)."""
        idx = self.posP[self.var_first[var]: self.var_first[var] + self.var_size[var]]
        g = aux[..., idx]
        return g @ self.stencils[var].T

    def insert_rows(self, eqn: int) -> np.ndarray:
        """Interleaved positions of equation ``eqn``'s rows (shuffleInsert)."""
        return self.posQ[self.eqn_first[eqn]: self.eqn_first[eqn] + self.eqn_size[eqn]]


def bandwidth(mats: Sequence[np.ndarray]) -> Tuple[int, int]:
    """nl, nu of the union of the structures (z_csr2dia, sparse_formats_z.f90:812-829)."""
    nl = nu = 0
    for m in mats:
        r, c = np.nonzero(m)
        if r.size:
            nl = max(nl, int(np.max(r - c)))
            nu = max(nu, int(np.max(c - r)))
    return nl, nu


def to_diagonals(m: np.ndarray, nl: int, nu: int) -> np.ndarray:
    """(nl+nu+1, n) with D[d+nl, i] = m[i, i+d]."""
    n = m.shape[0]
    D = np.zeros((nl + nu + 1, n))
    for d in range(-nl, nu + 1):
        i = np.arange(max(0, -d), min(n, n - d))
        D[d + nl, i] = m[i, i + d]
    return D


# ----------------------------------------------------------------------------
# Transforms (slabs.fftw3mpi/transforms.f90:155-301)
# ----------------------------------------------------------------------------

def spec_to_phys(S: np.ndarray, geo: Geometry) -> np.ndarray:
    """c2r_transform (slabs.fftw3mpi/transforms.f90:155-225).
    S: complex (nkx=NX/2, NYAA, NZ) -> real (NXAA, NYAA, NZAA), axes (x, y, z).
    A complex64 S runs the whole chain in single precision (scipy's FFTs keep
    the input precision); complex128 in double."""
    A = sfft.ifft(S, axis=1, workers=WORKERS) * geo.NYAA                  # y c2c FFTW_BACKWARD (:172-178)
    A[:, :, 0] *= 2.0                                    # Chebyshev k=0 row x2 (:184)
    B = np.zeros((geo.NXAA // 2 + 1, geo.NYAA, geo.NZ), S.dtype)
    B[: geo.NX // 2] = A * 0.5                           # pad x and halve (:197-198)
    C = sfft.irfft(B, n=geo.NXAA, axis=0, workers=WORKERS) * geo.NXAA     # x c2r, unnormalised (:204-210)
    D = np.zeros((geo.NXAA, geo.NYAA, geo.NZAA), C.dtype)
    D[:, :, : geo.NZ] = C                                # pad z (:216-217)
    return sfft.dct(D, type=3, axis=2, workers=WORKERS)                   # REDFT01 (:223)


def phys_to_spec(P: np.ndarray, geo: Geometry) -> np.ndarray:
    """r2c_transform (slabs.fftw3mpi/transforms.f90:227-301), inverse of the above
    (in the precision of P: float32 or float64)."""
    B5 = sfft.dct(P, type=2, axis=2, workers=WORKERS)                     # REDFT10 (:243)
    B4 = B5[:, :, : geo.NZ]                              # shrink z (:250)
    B3 = sfft.rfft(B4, axis=0, workers=WORKERS)                           # x r2c (:256-262)
    B2 = B3[: geo.NX // 2].copy()                        # shrink x (:268)
    B2[:, :, 0] *= 0.5                                   # Chebyshev k=0 row x1/2 (:279)
    S = sfft.fft(B2, axis=1, workers=WORKERS)                             # y c2c FFTW_FORWARD (:285-292)
    return S / S.real.dtype.type(geo.NXAA * geo.NYAA * geo.NZAA)   # normalisation (:293-296)


# ----------------------------------------------------------------------------
# The solver (PL_IMEX_timestepping.f90 + PL_IMEX_timestepping_marching.f90)
# ----------------------------------------------------------------------------

def remove_z_integral(s, NZ):
    """In place: the Chebyshev row 0 of spectra s[..., n] set so that the
    z-integral of the series vanishes, summing the even rows 2q for
    q = 1 .. NZ/2 - 1 (remove_vertical_mean_of_quadratic_var,
    PL_IMEX_timestepping.f90:223-235; for odd NZ the loop bound leaves row
    NZ - 1 out)."""
    s[..., 0] = 0.0
    for k in range(1, NZ // 2):
        s[..., 0] -= s[..., 2 * k] * (0.5 / (2.0 * k + 1.0) - 0.5 / (2.0 * k - 1.0))
    return s


def keep_z_integral_only(s, NZ):
    """In place: only the z-integral of the series kept, on row 0: row 0 plus
    the even rows 2q times their integral weights, rows 2q - 1 and 2q zeroed,
    q = 1 .. NZ/2 - 1 (remove_vertical_fluctuations_of_linear_var,
    PL_IMEX_timestepping.f90:252-274; rows past 2 (NZ/2 - 1) are left as they
    are, as there)."""
    for k in range(1, NZ // 2):
        s[..., 0] += s[..., 2 * k] * (0.5 / (2.0 * k + 1.0) - 0.5 / (2.0 * k - 1.0))
        s[..., 2 * k - 1] = 0.0
        s[..., 2 * k] = 0.0
    return s


def boundary_value(s, NZ, where):
    """In place: the series replaced by the constant of its value at the
    bottom (T_n(-1) = (-1)^n) or top (T_n(1) = 1) wall, on row 0, rows
    1 .. NZ - 1 zeroed, summed in the order of n
    (compute_full_variables_in_physical_space, PL_IMEX_timestepping_marching.f90:544-562)."""
    for n in range(1, NZ):
        if where == "TopSurf" or n % 2 == 0:
            s[..., 0] += s[..., n]
        else:
            s[..., 0] -= s[..., n]
        s[..., n] = 0.0
    return s


def project_kx0_plane(f, geo):
    """In place: the kx = 0 plane of a kxky state (nkx, NYAA, total) replaced by
    its Hermitian part, c(0, ky) = (c(0, ky) + conj c(0, -ky)) / 2 over the
    kept ky (no counterpart in the reference; see OracleSolver.project_kx0)."""
    ky = geo.ky_int
    for iy in range(geo.nky):
        m = int(ky[iy])
        if m <= 0 or geo.dealias[0, iy]:
            continue
        jy = int(np.nonzero(ky == -m)[0][0])
        h = 0.5 * (f[0, iy] + np.conj(f[0, jy]))
        f[0, iy] = h
        f[0, jy] = np.conj(h)
    return f


class OracleSolver:
    """State and one_step_beyond of ``full_problem_data_structure_T``.

    ``field[isys]``: complex (nkx, NYAA, total) for kxky systems;
    ``zfield[isys]``: real (total,) for mean (zero) systems."""

    def __init__(self, recipe, geo: Geometry, scheme: str = "ARS_222"):
        self.recipe, self.geo = recipe, geo
        self.A, self.Ah = imex_scheme(scheme)
        self.source_cheb: Dict[int, np.ndarray] = {}
        # penalisation variables: fixed physical values (init_penalisation,
        # marching.f90:778-822); nuSmago: its z-profile as last computed
        # (it starts at zero) and prefactor (--smagorinsky-prefactor)
        self.fixed_phys: Dict[int, np.ndarray] = {}
        self.nu_index = next((i for i, f in enumerate(recipe.full_vars) if getattr(f, "smagorinsky", False)), None)
        self.nu_prof = np.zeros(geo.NZAA)
        self.smagorinsky_prefactor = 1.0
        self.s = self.A.shape[0]
        NZ = geo.NZ
        self.kops = [SystemOps(s, NZ, geo.center, geo.gap) for s in recipe.kxky_systems]
        self.zops = [SystemOps(s, NZ, geo.center, geo.gap) for s in recipe.zero_systems]
        nkx, nky = geo.nkx, geo.nky
        self.keep = ~geo.dealias
        self.keep00 = self.keep.copy()
        self.keep00[0, 0] = False      # the (0,0) mode of kxky systems is pinned to zero
        # kxky operator pieces grouped by (kx, ky) exponent, as banded diagonals
        self.kmass, self.kstif, self.kband = [], [], []
        for ops in self.kops:
            mg = ops.grouped_pieces(ops.sys.mass)
            sg = ops.grouped_pieces(ops.sys.stif)
            nl, nu = bandwidth([m for _, m in mg] + [m for _, m in sg])
            self.kband.append((nl, nu))
            self.kmass.append(([k for k, _ in mg], np.stack([to_diagonals(m, nl, nu) for _, m in mg])
                               if mg else np.zeros((0, nl + nu + 1, ops.total))))
            self.kstif.append(([k for k, _ in sg], np.stack([to_diagonals(m, nl, nu) for _, m in sg])
                               if sg else np.zeros((0, nl + nu + 1, ops.total))))
        # zero systems: dense real matrices
        self.zmass, self.zstif = [], []
        for ops in self.zops:
            self.zmass.append(sum((ops.piece_matrix(p) for p in ops.sys.mass), np.zeros((ops.total, ops.total))))
            self.zstif.append(sum((ops.piece_matrix(p) for p in ops.sys.stif), np.zeros((ops.total, ops.total))))
        self.field = [np.zeros((nkx, nky, o.total), np.complex128) for o in self.kops]
        self.zfield = [np.zeros(o.total) for o in self.zops]
        self.dt = None
        self.D = chebyshev_derivative_matrix(NZ, geo.center, geo.gap)
        self.cfl_dt = None
        # precision of the transform chain and the products (marching.f90:279-300):
        # np.float64 as the reference, or np.float32 (scipy's single-precision
        # FFTs, products rounded to FP32), the yardstick of a mixed-precision step
        self.transform_dtype = np.float64
        # after each step, the kx = 0 plane replaced by its Hermitian part (the
        # HIP path's coral_amd_set_kx0_projection); False restates the reference
        self.project_kx0 = False

    # --- per-mode monomials (recipe_to_prefactors_array, :1558-1574) -------
    def monomials(self, keys) -> np.ndarray:
        """coef[ix, iy, g] = px^a py^b for each (a, b) in keys."""
        g = self.geo
        out = np.empty((g.nkx, g.nky, len(keys)), np.complex128)
        for k, (a, b) in enumerate(keys):
            out[:, :, k] = (g.px[:, None] ** a) * (g.py[None, :] ** b)
        return out

    # --- factorize_operators (PL_IMEX_timestepping.f90:351-422) ------------
    def factorize(self, dt: float):
        self.dt = dt
        a11 = self.A[0, 0]
        self.klu = []
        for isys, ops in enumerate(self.kops):
            nl, nu = self.kband[isys]
            mk, mb = self.kmass[isys]
            sk, sb = self.kstif[isys]
            keys = list(mk) + list(sk)
            bands = np.concatenate([mb, sb], axis=0)
            coef = self.monomials(keys)
            coef[:, :, len(mk):] *= -dt * a11
            kept = np.argwhere(self.keep00)
            n = ops.total
            ldab = 2 * nl + nu + 1
            ab = np.zeros((len(kept), n, ldab), np.complex128)
            ipiv = np.zeros((len(kept), n), np.int32)
            info = np.zeros(len(kept), np.int32)
            cvec_all = coef[kept[:, 0], kept[:, 1], :]      # (nk, ng)
            chunk = 2048
            for c0 in range(0, len(kept), chunk):
                cvec = cvec_all[c0:c0 + chunk]
                # A(i, i+d) = sum_g coef_g bands[g, d+nl, i] -> LAPACK column j=i+d, row kl+ku+i-j
                full = np.einsum("kg,gdi->kdi", cvec, bands)    # (nk, nl+nu+1, n)
                sub = ab[c0:c0 + chunk]
                for d in range(-nl, nu + 1):
                    i = np.arange(max(0, -d), min(n, n - d))
                    j = i + d
                    sub[:, j, nl + nu + i - j] = full[:, d + nl, i]
                p, inf = zgbtrf_batch(sub, nl, nu)
                ipiv[c0:c0 + chunk], info[c0:c0 + chunk] = p, inf
            if np.any(info):
                raise RuntimeError("singular evolution operator in kxky system %d" % (isys + 1))
            self.klu.append((kept, ab, ipiv, nl, nu))
        self.zlu = []
        for isys, ops in enumerate(self.zops):
            Az = self.zmass[isys] - dt * a11 * self.zstif[isys]
            nl, nu = bandwidth([Az])
            ab = dense_to_lapack_band(Az.astype(np.complex128), nl, nu)[None].copy()
            ipiv, info = zgbtrf_batch(ab, nl, nu)
            if np.any(info):
                raise RuntimeError("singular evolution operator in zero system %d" % (isys + 1))
            self.zlu.append((ab, ipiv, nl, nu))

    # --- operator application ----------------------------------------------
    def _apply(self, isys: int, which: str, x: np.ndarray) -> np.ndarray:
        """y = Op(kx,ky) x on every mode (zcsrmv loop, PL_algebra_z_3D.f90:127-157).
        Dealiased modes use the same operator (the reference applies it to all
        modes); the (0,0) kxky mode uses M = 0, L = identity (:1052-1058)."""
        keys, bands = (self.kmass if which == "mass" else self.kstif)[isys]
        nl, nu = self.kband[isys]
        g = self.geo
        n = x.shape[-1]
        coef = self.monomials(keys).reshape(g.nkx * g.nky, len(keys))
        y = band_matvec_batch(bands, nl, nu, coef, x.reshape(-1, n)).reshape(x.shape)
        y[0, 0, :] = 0.0 if which == "mass" else x[0, 0, :]
        return y

    # --- full variables in physical space (marching.f90:420-588) -----------
    def full_variables_spec(self, aux, zaux) -> List[np.ndarray]:
        g = self.geo
        out = []
        for fv in self.recipe.full_vars:
            spec = np.zeros((g.nkx, g.nky, g.NZ), np.complex128)
            for t in fv.terms:
                if t.kind == "kxky":
                    ops = self.kops[t.system]
                    buf = t.dsca * ops.extract(t.position, aux[t.system])
                    if t.kx == 0 and t.ky == 0 and t.dz == 0:
                        spec += buf
                    else:
                        for _ in range(t.dz):
                            buf = buf @ self.D.T
                        spec += buf * ((g.px[:, None] ** t.kx) * (g.py[None, :] ** t.ky))[:, :, None]
                else:
                    ops = self.zops[t.system]
                    buf = ops.extract(t.position, zaux[t.system])
                    for _ in range(t.dz):
                        buf = self.D @ buf
                    spec[0, 0, :] += buf * t.dsca
            # boundary-value extraction, then the vertical decomposition
            # (PL_IMEX_timestepping_marching.f90:544-577)
            where = getattr(fv, "extract", "nowhere")
            if where in ("BotSurf", "TopSurf"):
                boundary_value(spec, g.NZ, where)
            zd = getattr(fv, "zdecomp", "")
            if zd == "barotropic":
                keep_z_integral_only(spec, g.NZ)
            elif zd == "baroclinic":
                remove_z_integral(spec, g.NZ)
            out.append(spec)
        return out

    def cfl(self, phys: List[np.ndarray]) -> float:
        """compute_cfl_based_timestep (marching.f90:200-254), cflFactor_along_z = 1."""
        g = self.geo
        u = np.max(np.abs(phys[0])) * g.NXAA / g.Lx
        v = np.max(np.abs(phys[1])) * g.NYAA / g.Ly
        w8 = g.gauss_chebyshev_weight()
        w = max(v, float(np.max(np.abs(phys[2]) / w8[None, None, :])))
        m = max(u, v, w)
        return 1.0 / m if m > 0 else math.inf

    # --- NL terms (marching.f90:257-418) -----------------------------------
    def set_fixed_physical(self, ifull: int, phys: np.ndarray):
        """Fixed physical values [x, y, z] of a full variable (a penalisation
        variable's mask)."""
        self.fixed_phys[ifull] = np.asarray(phys, np.float64).copy()

    def _physical(self, specs, advance: bool):
        """compute_full_variables_in_physical_space (marching.f90:420-588):
        spec_to_phys of every full variable but the penalisation variables
        (their fixed values) and nuSmago, whose values are
        prefactor sqrt(<nuSmago + sum of the squares of dxUfull, dyUfull,
        dxVfull, dyVfull>_xy) (:514-541) with nuSmago its previous values
        (the reference's phys array is not reset between calls); ``advance``
        keeps the new profile."""
        g = self.geo
        cdt = np.complex64 if self.transform_dtype == np.float32 else np.complex128
        phys = []
        for i, s in enumerate(specs):
            if i in self.fixed_phys:
                phys.append(self.fixed_phys[i])
            elif i == self.nu_index:
                phys.append(None)
            else:
                phys.append(spec_to_phys(s.astype(cdt), g))
        if self.nu_index is not None:
            arr = np.broadcast_to(self.nu_prof, (g.NXAA, g.NYAA, g.NZAA)).copy()
            for j, fv in enumerate(self.recipe.full_vars):
                if fv.name in ("dxUfull", "dyUfull", "dxVfull", "dyVfull"):
                    arr = arr + phys[j] ** 2
            avg = arr.sum(axis=(0, 1)) / g.NXAA / g.NYAA
            prof = self.smagorinsky_prefactor * np.sqrt(avg)
            if advance:
                self.nu_prof = prof
            phys[self.nu_index] = np.broadcast_to(prof, (g.NXAA, g.NYAA, g.NZAA)).copy()
        return phys

    def quad_spectra(self, aux, zaux, advance: bool = True):
        """Spectra of the quadratic variables (marching.f90:279-300 and the
        z-integral removal of PL_IMEX_timestepping.f90:223-235), complex
        (nkx, NYAA, NZ) each; sets cfl_dt from the full variables, with the
        Smagorinsky bound (marching.f90:241-253) when nuSmago is present."""
        g = self.geo
        specs = self.full_variables_spec(aux, zaux)
        phys = self._physical(specs, advance)
        if len(phys) >= 3:
            self.cfl_dt = self.cfl(phys)
            if self.nu_index is not None:
                self.cfl_dt = min(self.cfl_dt, 0.95 / ((2 * 3.141592) ** 2 * float(np.max(phys[-1]))
                                                       * ((g.NXAA / g.Lx) ** 2 + (g.NYAA / g.Ly) ** 2)))
        qhat = []
        for q in self.recipe.quad_vars:
            if q.ivar1 == -1:
                prod = phys[q.ivar2]
            elif q.ivar2 == -1:
                prod = phys[q.ivar1]
            else:
                prod = phys[q.ivar1] * phys[q.ivar2]
            s = phys_to_spec(prod, g).astype(np.complex128)
            if q.remove_z_integral:
                remove_z_integral(s, g.NZ)
            qhat.append(s)
        return qhat

    def nl_terms(self, aux, zaux):
        g = self.geo
        qhat = self.quad_spectra(aux, zaux)
        khat = []
        for isys, ops in enumerate(self.kops):
            K = np.zeros((g.nkx, g.nky, ops.total), np.complex128)
            for t in ops.sys.nl:
                order = ops.sys.eqn_orders[t.eqn]
                coef = t.dsca * (g.px[:, None] ** t.kx) * (g.py[None, :] ** t.ky)
                W = ops.IM(t.iz, t.mz)[order:, :]
                vals = (qhat[t.qvar][:, :, : g.NZ] * coef[:, :, None]) @ W.T
                K[:, :, ops.insert_rows(t.eqn)] += vals
            K[~self.keep] = 0.0
            K[0, 0, :] = 0.0
            khat.append(K)
        zkhat = []
        for isys, ops in enumerate(self.zops):
            K = np.zeros(ops.total)
            for t in ops.sys.nl:
                order = ops.sys.eqn_orders[t.eqn]
                W = ops.IM(t.iz, t.mz)[order:, :]
                K[ops.insert_rows(t.eqn)] += W @ (t.dsca * qhat[t.qvar][0, 0, : g.NZ].real)
            # sources (marching.f90:399-416): dsca times the source's first NZ
            # Chebyshev coefficients through trunc(I^iz)[order:, :]
            for t in getattr(self.recipe, "sources", []):
                if t.system != isys:
                    continue
                if t.index not in self.source_cheb:
                    raise ValueError(f"source{t.index:02d} has no definition")
                order = ops.sys.eqn_orders[t.eqn]
                W = ops.IM(t.iz, 0)[order:, :]
                K[ops.insert_rows(t.eqn)] += W @ (self.source_cheb[t.index][: g.NZ] * t.dsca)
            zkhat.append(K)
        return khat, zkhat

    def set_source(self, index: int, physical: np.ndarray):
        """Source ``index`` from its NZAA values on the Gauss-Chebyshev grid
        (top wall first): Compute_1dCheby_transform (marching.f90:845-866),
        REDFT10 / NZAA with the first coefficient halved."""
        c = sfft.dct(np.asarray(physical, np.float64), type=2) / self.geo.NZAA
        c[0] *= 0.5
        self.source_cheb[index] = c

    # --- one_step_beyond (marching.f90:3-48) -------------------------------
    def step(self, dt: float):
        if self.dt is None or dt != self.dt:
            self.factorize(dt)
        A, Ah = self.A, self.Ah
        aux = [f.copy() for f in self.field]
        zaux = [f.copy() for f in self.zfield]
        Khat, Kstd, zKhat, zKstd = [], [], [], []
        for i in range(self.s):
            kh, zkh = self.nl_terms(aux, zaux)
            Khat.append(kh)
            zKhat.append(zkh)
            new_aux, new_zaux, ks, zks = [], [], [], []
            for isys, ops in enumerate(self.kops):
                rhs = self._apply(isys, "mass", self.field[isys])
                for j in range(i):
                    rhs = rhs + dt * A[i, j] * Kstd[j][isys]
                    rhs = rhs + dt * Ah[i + 1, j] * Khat[j][isys]
                rhs = rhs + dt * Ah[i + 1, i] * Khat[i][isys]
                kept, ab, ipiv, nl, nu = self.klu[isys]
                x = rhs.copy()   # identity systems (dealiased, (0,0)) keep x = rhs
                b = np.ascontiguousarray(rhs[kept[:, 0], kept[:, 1], :])
                zgbtrs_batch(ab, nl, nu, ipiv, b)
                x[kept[:, 0], kept[:, 1], :] = b
                new_aux.append(x)
                k = self._apply(isys, "stif", x)
                k[~self.keep] = 0.0
                k[0, 0, :] = 0.0
                ks.append(k)
            for isys, ops in enumerate(self.zops):
                rhs = self.zmass[isys] @ self.zfield[isys]
                for j in range(i):
                    rhs = rhs + dt * A[i, j] * zKstd[j][isys]
                    rhs = rhs + dt * Ah[i + 1, j] * zKhat[j][isys]
                rhs = rhs + dt * Ah[i + 1, i] * zKhat[i][isys]
                ab, ipiv, nl, nu = self.zlu[isys]
                b = rhs.astype(np.complex128)[None].copy()
                zgbtrs_batch(ab, nl, nu, ipiv, b)
                xz = b[0].real.copy()
                new_zaux.append(xz)
                zks.append(self.zstif[isys] @ xz)
            aux, zaux = new_aux, new_zaux
            Kstd.append(ks)
            zKstd.append(zks)
        for isys in range(len(self.kops)):
            f = aux[isys]
            f[~self.keep] = 0.0
            f[0, 0, :] = 0.0
            if self.project_kx0:
                project_kx0_plane(f, self.geo)
            self.field[isys] = f
        self.zfield = zaux

    def run(self, dt: float, nsteps: int):
        for _ in range(nsteps):
            self.step(dt)

    # --- diagnostics ---------------------------------------------------------
    def physical_full_variables(self) -> List[np.ndarray]:
        """The full variables of the current state [x, y, z]; nuSmago as last
        computed by a step."""
        specs = self.full_variables_spec(self.field, self.zfield)
        phys = self._physical(specs, advance=False)
        if self.nu_index is not None:
            phys[self.nu_index] = np.broadcast_to(self.nu_prof, phys[self.nu_index].shape).copy()
        return phys
