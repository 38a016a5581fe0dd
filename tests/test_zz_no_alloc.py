"""coral_amd_banded_solve_noalloc: the allocation-free batched banded LU and
solve, called through the C ABI on device buffers, against the allocating
entry point coral_amd_banded_solve on the same systems.

Device buffers come from the HIP runtime itself (libamdhip64 through ctypes),
so the test does not depend on another allocator.  Two calls in a row, the
second smaller, check that the work space the calls share needs no clearing;
integer outputs (LAPACK info) must be identical, solutions equal to rounding
(relative L2 <= 1e-12: the two entry points use different solve kernels)."""
import ctypes

import numpy as np
import pytest

from oracle import coral_oracle as O

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs, straight from libamdhip64."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.h.hipSetDevice.argtypes = [ctypes.c_int]
        self.bufs = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def to_device(self, a: np.ndarray) -> int:
        p = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)), "hipMalloc")
        self.bufs.append(p.value)
        self.ok(self.h.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def to_host(self, p: int, a: np.ndarray) -> np.ndarray:
        self.ok(self.h.hipMemcpy(a.ctypes.data, p, a.nbytes, 2), "hipMemcpy D2H")
        return a

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.fixture(scope="module")
def lib(native_lib):
    return native_lib


def random_band(rng, ns, n, kl, ku):
    abs_ = []
    for _ in range(ns):
        A = np.zeros((n, n), np.complex128)
        for i in range(n):
            for j in range(max(0, i - kl), min(n, i + ku + 1)):
                A[i, j] = rng.normal() + 1j * rng.normal()
        abs_.append(O.dense_to_lapack_band(A, kl, ku))
    return np.ascontiguousarray(np.stack(abs_)), rng.normal(size=(ns, n)) + 1j * rng.normal(size=(ns, n))


def solve_noalloc(lib, hip, ab, kl, ku, b):
    ns, n, _ = ab.shape
    dab = hip.to_device(np.ascontiguousarray(ab))
    dx = hip.to_device(np.ascontiguousarray(b, np.complex128))
    dinfo = hip.to_device(np.full(ns, -7, np.int32))
    lib.check(lib.lib.coral_amd_banded_solve_noalloc(0, n, kl, ku, ns, dab, dx, dinfo, None),
              "coral_amd_banded_solve_noalloc")
    hip.ok(hip.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
    x = hip.to_host(dx, np.empty((ns, n), np.complex128))
    info = hip.to_host(dinfo, np.empty(ns, np.int32))
    return x, info


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_noalloc_matches_allocating_entry_point(lib):
    from coral_amd.solver import banded_solve
    hip = Hip()
    hip.ok(hip.h.hipSetDevice(0), "hipSetDevice")
    rng = np.random.default_rng(21)
    try:
        # first call: Coral's band (kl = 16, ku = 27); second: a smaller batch
        # and band, with a singular system among them (LAPACK info = 3)
        for ns, n, kl, ku, singular in ((6, 190, 16, 27, False), (3, 40, 2, 5, True)):
            print("noalloc call", ns, n, kl, ku, flush=True)
            ab, b = random_band(rng, ns, n, kl, ku)
            if singular:
                A = np.zeros((n, n), np.complex128)
                for j in range(n):
                    for i in range(max(0, j - ku), min(n, j + kl + 1)):
                        A[i, j] = ab[1, j, kl + ku + i - j]
                A[:, 2] = 0.0
                ab[1] = O.dense_to_lapack_band(A, kl, ku)
            assert lib.lib.coral_amd_banded_solve_noalloc_bytes(n, kl, ku, ns) <= 64 * 1024 * 1024
            x_ref, info_ref = banded_solve(ab, kl, ku, b)
            x, info = solve_noalloc(lib, hip, ab, kl, ku, b)
            assert np.array_equal(info, info_ref)
            for s in range(ns):
                if info_ref[s] == 0:
                    assert rel(x[s], x_ref[s]) <= 1e-12
            if singular:
                assert info[1] == 3
    finally:
        hip.free()


def test_noalloc_rejects_batches_beyond_its_work_space(lib):
    """A batch whose work space exceeds the fixed array is refused with an
    error before any device work (nothing is read from the null buffers)."""
    from coral_amd._lib import CoralError
    n, kl, ku, ns = 766, 16, 27, 1000
    need = lib.lib.coral_amd_banded_solve_noalloc_bytes(n, kl, ku, ns)
    assert need > 64 * 1024 * 1024
    fake = ctypes.c_void_p(1)
    with pytest.raises(CoralError, match="work space"):
        lib.check(lib.lib.coral_amd_banded_solve_noalloc(0, n, kl, ku, ns, fake, fake, fake, None),
                  "coral_amd_banded_solve_noalloc")
