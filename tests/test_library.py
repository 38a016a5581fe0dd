"""The C ABI library: it loads, exports every symbol the header declares, and
reports bad input from the host-side assembly without touching a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT, recipe


def header_symbols():
    text = open(os.path.join(ROOT, "include", "coral_amd.h")).read()
    return sorted(set(re.findall(r"^(?:int|size_t|int64_t|int32_t)\s+(coral_amd_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_header_symbol(native_lib):
    syms = header_symbols()
    assert len(syms) >= 18
    so = native_lib.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", so], stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(so)
    for s in syms:
        assert getattr(lib, s) is not None
    assert set(syms) == set(native_lib.EXPORTS)


def test_library_is_built_for_gfx950(native_lib):
    out = subprocess.run(["strings", native_lib.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    assert "gfx950" in out


def test_assembly_errors_are_reported_without_a_device(native_lib):
    from coral_amd._lib import CoralError
    from coral_amd.solver import assembled_operator
    r = recipe("rbc_vanilla.coral.equations")
    r.kxky_systems[0].bc_codes[0] = 25   # unknown Galerkin code with a valid size
    with pytest.raises(CoralError, match="boundary-condition code 25"):
        assembled_operator(r, 16, 16, 17, 1.0, 1.0, 0, "mass", 0, 0)
    r = recipe("rbc_vanilla.coral.equations")
    r.kxky_systems[0].eqn_orders[0] = 1  # orders no longer match the boundary conditions
    with pytest.raises(CoralError, match="incompatible number of boundary conditions"):
        assembled_operator(r, 16, 16, 17, 1.0, 1.0, 0, "mass", 0, 0)
    with pytest.raises(CoralError, match="even NX"):
        assembled_operator(recipe("rbc_vanilla.coral.equations"), 15, 16, 17, 1.0, 1.0, 0, "mass", 0, 0)


def test_solver_without_device_fails_loudly(native_lib):
    """On a machine without a GPU the product raises; it never falls back to the CPU."""
    import importlib
    torch = None
    try:
        torch = importlib.import_module("torch")
    except Exception:
        pass
    if torch is not None and torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from coral_amd._lib import CoralError
    from coral_amd.solver import Solver
    with pytest.raises(CoralError):
        Solver(recipe("rbc_vanilla.coral.equations"), 16, 16, 17, 1.0, 1.0)
