"""The Fortran host side of the boundary: ``include/coral_amd_binding.f90``
(iso_c_binding interface to coral_amd.h), compiled with amdflang by
``__graft_entry__.build()`` together with the programs in ``tests/fortran``.

- The bind(C) types have the C compiler's layout of the structs in
  coral_amd.h (sizes and the offsets of fields after padding).
- A Fortran program runs the reference's 5x5 algebra known-answer test
  (src/plane_layer/driver_algebra.f90:24-47) on the GPU through the C-ABI.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "fortran", "_build")


def _prog(name):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built: run __graft_entry__.build() (it compiles the Fortran host programs)")
    return path


def _run(name, timeout=60):
    return subprocess.run([_prog(name)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=timeout)


def test_fortran_binding_layout_matches_c_header():
    f = _run("binding_layout")
    c = _run("c_layout")
    assert f.returncode == 0, f.stdout
    assert c.returncode == 0, c.stdout
    fl, cl = f.stdout.split(), c.stdout.split()
    assert len(fl) == len(cl) == 2 * 22
    assert fl == cl


@pytest.mark.gpu
def test_fortran_host_banded_kat():
    print("fortran banded_kat", flush=True)
    r = _run("banded_kat", timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    assert "info 0" in r.stdout and "banded_kat: ok" in r.stdout
