"""The Fortran host side of the boundary: ``include/coral_amd_binding.f90``
(iso_c_binding interface to coral_amd.h), compiled with amdflang by
``__graft_entry__.build()`` together with the programs in ``tests/fortran``.

- The bind(C) types have the C compiler's layout of the structs in
  coral_amd.h (sizes and the offsets of fields after padding).
- A Fortran program runs the reference's 5x5 algebra known-answer test
  (src/plane_layer/driver_algebra.f90:24-47) on the GPU through the C-ABI.
- A Fortran host (tests/fortran/step_driver.f90) drives whole time steps of
  rotating RBC through create / set_dt / upload / step / download, as the
  reference's main loop would (plane_layer_main.f90:147,157); its result is
  the Python host's bit for bit and matches the oracle.
"""
import math
import os
import struct
import subprocess

import numpy as np
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


def _write_problem(path, recipe, NX, NY, NZ, Lx, Ly, scheme, gap=1.0, center=0.5):
    """The parsed recipe in the binary layout step_driver.f90 reads."""
    from coral_amd.schemes import tableau
    A, Ah = tableau(scheme)
    with open(path, "wb") as f:
        f.write(struct.pack("=3i4di", NX, NY, NZ, Lx, Ly, gap, center, A.shape[0]))
        f.write(np.ascontiguousarray(A, "=f8").tobytes())
        f.write(np.ascontiguousarray(Ah, "=f8").tobytes())
        for systems in (recipe.kxky_systems, recipe.zero_systems):
            f.write(struct.pack("=i", len(systems)))
            for sy in systems:
                f.write(struct.pack("=i", sy.nvars))
                f.write(np.asarray(sy.bc_codes, "=i4").tobytes())
                f.write(np.asarray(sy.eqn_orders, "=i4").tobytes())
                for lst in (sy.mass, sy.stif):
                    f.write(struct.pack("=i", len(lst)))
                    for p in lst:
                        f.write(struct.pack("=d6i", p.dsca, p.kx, p.ky, p.iz, p.mz, p.var, p.eqn))
                f.write(struct.pack("=i", len(sy.nl)))
                for t in sy.nl:
                    f.write(struct.pack("=d6i", t.dsca, t.kx, t.ky, t.iz, t.mz, t.qvar, t.eqn))
        f.write(struct.pack("=i", len(recipe.full_vars)))
        for fv in recipe.full_vars:
            f.write(struct.pack("=i", len(fv.terms)))
            for t in fv.terms:
                f.write(struct.pack("=d7i", t.dsca, t.kx, t.ky, t.dz, t.mz, 0 if t.kind == "kxky" else 1,
                                    t.system, t.position))
        f.write(struct.pack("=i", len(recipe.quad_vars)))
        for q in recipe.quad_vars:
            f.write(struct.pack("=3i", q.ivar1, q.ivar2, int(q.remove_z_integral)))


def _write_state(path, dt, nsteps, kst, zst):
    with open(path, "wb") as f:
        f.write(struct.pack("=di", dt, nsteps))
        for arrs, dtype in ((kst, np.complex128), (zst, np.float64)):
            f.write(struct.pack("=i", len(arrs)))
            for a in arrs:
                a = np.ascontiguousarray(a, dtype)
                f.write(struct.pack("=q", a.size))
                f.write(a.tobytes())


def _read_state(path, kshapes, zshapes):
    buf = open(path, "rb").read()
    off = struct.calcsize("=di")
    out = []
    for shapes, dtype in ((kshapes, np.complex128), (zshapes, np.float64)):
        (n,) = struct.unpack_from("=i", buf, off)
        off += 4
        assert n == len(shapes)
        arrs = []
        for shp in shapes:
            (cnt,) = struct.unpack_from("=q", buf, off)
            off += 8
            a = np.frombuffer(buf, dtype, cnt, off).reshape(shp)
            off += a.nbytes
            arrs.append(a)
        out.append(arrs)
    (cfl,) = struct.unpack_from("=d", buf, off)
    return out[0], out[1], cfl


@pytest.mark.gpu
def test_fortran_host_drives_time_steps(tmp_path):
    from coral_amd import ic
    from coral_amd.equations import read_equations
    from coral_amd.solver import Solver
    from oracle import coral_oracle as O

    NX, NY, NZ, L, dt, nsteps = 16, 16, 17, 2 * math.pi, 2e-5, 3
    r = read_equations(os.path.join(ROOT, "coral_amd", "inputs", "coral.equations.rrbc_ra5e7_ek1e-4"))
    k0, z0 = ic.rbc_initial_state(r, NX, NY, NZ, 1e-2, 6, 6, 3)
    _write_problem(tmp_path / "problem.bin", r, NX, NY, NZ, L, L, "ARS_222")
    _write_state(tmp_path / "state.bin", dt, nsteps, k0, z0)
    print("fortran step_driver", NX, NY, NZ, nsteps, flush=True)
    res = subprocess.run([_prog("step_driver"), str(tmp_path / "problem.bin"), str(tmp_path / "state.bin"),
                          str(tmp_path / "out.bin")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=120)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stdout
    kf, zf, cfl_f = _read_state(tmp_path / "out.bin", [a.shape for a in k0], [a.shape for a in z0])

    with Solver(r, NX, NY, NZ, L, L, scheme="ARS_222") as g:
        g.set_state(k0, z0)
        g.march_forward(dt, nsteps)
        kp, zp = g.get_state()
        cfl_p = g.cfl_dt()
    for a, b in zip(kf, kp):
        assert np.array_equal(a, b)
    for a, b in zip(zf, zp):
        assert np.array_equal(a, b)
    assert cfl_f == cfl_p

    ora = O.OracleSolver(r, O.Geometry(NX, NY, NZ, L, L), "ARS_222")
    ora.field, ora.zfield = [a.copy() for a in k0], [a.copy() for a in z0]
    ora.run(dt, nsteps)
    assert np.linalg.norm(kf[0] - ora.field[0]) <= 1e-10 * np.linalg.norm(ora.field[0])
    for a, b in zip(zf, ora.zfield):
        assert np.linalg.norm(a - b) <= 1e-10 * max(np.linalg.norm(b), 1e-300)
