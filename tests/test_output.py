"""User outputs and Geometry files (coral_amd/output.py) on the CPU, with a
stand-in solver whose physical fields are known arrays: the parser on the
reference's own coral.usrOutput files, the file names and layouts against the
reference's writers (the arrays the Fortran code passes to its writers,
restated index by index), the schedule of the counters, and the Geometry
records."""
import os

import numpy as np
import pytest

from coral_amd import output as O
from coral_amd.equations import read_equations
from conftest import ROOT

INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")


def recipe(name="rrbc_julien.coral.equations.velocityPressure"):
    return read_equations(os.path.join(INPUTS, name))


class StubSolver:
    """Physical field f[x, y, z] = 1000 ifull + 100 x + 10 y + z / 10 (distinct
    everywhere), the solver's geometry attributes, and for a filtered
    quadratic variable a stand-in spectrum whose back transform is marked."""

    def __init__(self, r, NX=16, NY=12, NZ=9, Lx=2.0, Ly=3.0, gap=1.0, center=0.5):
        self.recipe = r
        self.NX, self.NY, self.NZ = NX, NY, NZ
        self.NXAA, self.NYAA, self.NZAA = 3 * NX // 2, 3 * NY // 2, 3 * NZ // 2
        self.Lx, self.Ly, self.gap, self.center = Lx, Ly, gap, center

    def physical(self, i):
        x, y, z = np.meshgrid(np.arange(self.NXAA), np.arange(self.NYAA), np.arange(self.NZAA), indexing="ij")
        return 1000.0 * i + 100.0 * x + 10.0 * y + z / 10.0

    def products(self):
        return np.full((len(self.recipe.quad_vars), 1), 7.0)

    def bwd_transform(self, spec):
        return np.full((self.NXAA, self.NYAA, self.NZAA), float(spec[0]) + 0.25)


def test_parse_reference_usr_output_files():
    r = recipe()
    objs = O.read_usr_output_file(os.path.join(INPUTS, "rrbc_julien.coral.usrOutput"), r, 144, 144, 72)
    assert [(o.kind, o.name, o.period, o.position) for o in objs[:5]] == [
        ("volume", "uFull", 10000, 0), ("volume", "vFull", 10000, 0), ("volume", "wFull", 10000, 0),
        ("volume", "tFull", 10000, 0), ("zSlice", "tFull", 1000, 1)]
    assert len(objs) == 17
    assert objs[-1].kind == "profil" and objs[-1].period == 50 and objs[-1].index == 3
    assert all(not o.quadratic for o in objs)
    # the same file on a grid with NZAA = 48 drops the slices at 72 (PL_equations.f90:383-389)
    assert len(O.read_usr_output_file(os.path.join(INPUTS, "rrbc_julien.coral.usrOutput"), r, 144, 144, 48)) == 16
    ex = O.read_usr_output_file(os.path.join(INPUTS, "example.coral.usrOutput"), recipe("rbc_vanilla.coral.equations"),
                                96, 96, 48)
    assert [(o.kind, o.name, o.quadratic, o.index, o.period, o.position) for o in ex] == [
        ("volume", "uu", True, 0, 10, 0), ("volume", "uFull", False, 0, 10, 0), ("zSlice", "uFull", False, 0, 10, 32),
        ("profil", "tFull", False, 3, 5, 0)]
    with pytest.raises(ValueError):
        O.parse_usr_output(">>Profile :: tFull ;; Period= 5 <<\n>>EOF<<\n", r, 8, 8, 8)   # space before ::
    with pytest.raises(ValueError):
        O.parse_usr_output(">>Zsummed:: uu ;; Period= 5 <<\n>>EOF<<\n", r, 8, 8, 8)


def test_file_names_follow_the_reference_formats():
    o = O.OutputObject("zSlice", "tFull", 10, 24, False, 3)
    assert O.file_name(o, 1234) == "Slices/linear_var04_z00024_time00001234_full.dat"
    assert O.file_name(O.OutputObject("xSlice", "uu", 1, 7, True, 0), 5) == "Slices/quadra_var01_x00007_time00000005_full.dat"
    assert O.file_name(O.OutputObject("volume", "uu", 1, 0, True, 8), 99999999) == \
        "Volumes/quadra_var09_time99999999_full.dat"
    assert O.file_name(O.OutputObject("zAvged", "tFull", 1, 0, False, 3), 2) == \
        "Slices/linear_var04_zAvged_time00000002_full.dat"
    assert O.file_name(O.OutputObject("profil", "tFull", 1, 0, False, 3), 2) == "Profiles/linear_var04_time00000002_full.dat"


def test_files_hold_the_reference_arrays(tmp_path):
    """Each file read back as the Fortran array the reference writes:
    transposed_phys(iz, iy, ix) for volumes; aux2d(iy, ix) = phys(k, iy, ix),
    aux2d(iz, ix) = phys(iz, j, ix), aux2d(iz, iy) = phys(iz, iy, i) for
    slices; sum_z w_z phys(iz, iy, ix) for the summed slice; the x-y mean
    for profiles (src/slabs.fftw3mpi/transforms.f90:53-70,777-858,
    PL_IMEX_timestepping_output.f90:514-529)."""
    r = recipe()
    g = StubSolver(r)
    text = (">>Volume :: tFull ;; Period= 1 <<\n>>ZSlice :: tFull ;; Period= 1 ;; Position= 4 <<\n"
            ">>YSlice :: uFull ;; Period= 1 ;; Position= 5 <<\n>>XSlice :: vFull ;; Period= 1 ;; Position= 6 <<\n"
            ">>Zsummed:: wFull ;; Period= 1 <<\n>>Profile:: tFull ;; Period= 1 <<\n>>Volume :: tw ;; Period= 1 <<\n"
            ">>EOF<<\n")
    objs = O.parse_usr_output(text, r, g.NXAA, g.NYAA, g.NZAA)
    out = O.UserOutput(g, objs, str(tmp_path))
    files = out.before_step()
    assert len(files) == 7
    NX, NY, NZ = g.NXAA, g.NYAA, g.NZAA

    def read(name, shape):
        return np.fromfile(os.path.join(tmp_path, name), "<f8").reshape(shape, order="F")
    phys = {i: g.physical(i) for i in range(4)}         # [x, y, z]
    vol = read("Volumes/linear_var04_time00000001_full.dat", (NZ, NY, NX))
    for iz, iy, ix in [(0, 0, 0), (3, 2, 7), (NZ - 1, NY - 1, NX - 1)]:
        assert vol[iz, iy, ix] == phys[3][ix, iy, iz]
    zs = read("Slices/linear_var04_z00004_time00000001_full.dat", (NY, NX))
    assert all(zs[iy, ix] == phys[3][ix, iy, 3] for iy in range(NY) for ix in range(NX))
    ys = read("Slices/linear_var01_y00005_time00000001_full.dat", (NZ, NX))
    assert all(ys[iz, ix] == phys[0][ix, 4, iz] for iz in range(NZ) for ix in range(NX))
    xs = read("Slices/linear_var02_x00006_time00000001_full.dat", (NZ, NY))
    assert all(xs[iz, iy] == phys[1][5, iy, iz] for iz in range(NZ) for iy in range(NY))
    w = O.gauss_chebyshev_weights(NZ, g.gap)
    za = read("Slices/linear_var03_zAvged_time00000001_full.dat", (NY, NX))
    for iy, ix in [(0, 0), (5, 9)]:
        assert za[iy, ix] == pytest.approx(sum(phys[2][ix, iy, iz] * w[iz] for iz in range(NZ)), rel=1e-14)
    prof = np.fromfile(os.path.join(tmp_path, "Profiles/linear_var04_time00000001_full.dat"), "<f8")
    assert prof.shape == (NZ,)
    np.testing.assert_allclose(prof, [phys[3][:, :, iz].sum() / NX / NY for iz in range(NZ)], rtol=1e-14)
    tw = read("Volumes/quadra_var09_time00000001_full.dat", (NZ, NY, NX))
    assert tw[2, 3, 4] == phys[3][4, 3, 2] * phys[2][4, 3, 2]


def test_schedule_and_filtered_quadratics(tmp_path):
    """Counters start at time_index mod period and an object writes when its
    counter reaches the period (PL_IMEX_timestepping_marching.f90:620-650); a
    zFiltered quadratic variable is written from its filtered spectrum."""
    r = recipe()
    r.quad_vars[8].remove_z_integral = True
    g = StubSolver(r)
    objs = O.parse_usr_output(">>Profile:: tFull ;; Period= 3 <<\n>>Volume :: tw ;; Period= 2 <<\n>>EOF<<\n", r,
                              g.NXAA, g.NYAA, g.NZAA)
    out = O.UserOutput(g, objs, str(tmp_path), time_index=5)
    times = []
    for _ in range(7):
        times += [os.path.basename(f) for f in out.before_step()]
    assert times == ["linear_var04_time00000006_full.dat", "quadra_var09_time00000006_full.dat",
                     "quadra_var09_time00000008_full.dat", "linear_var04_time00000009_full.dat",
                     "quadra_var09_time00000010_full.dat", "linear_var04_time00000012_full.dat",
                     "quadra_var09_time00000012_full.dat"]
    v = np.fromfile(os.path.join(tmp_path, "Volumes/quadra_var09_time00000006_full.dat"), "<f8")
    assert np.all(v == 7.25)


def test_geometry_records(tmp_path):
    g = StubSolver(recipe(), NX=16, NY=12, NZ=9, Lx=2.0, Ly=3.0, gap=1.0, center=0.5)
    O.write_geometry(g, str(tmp_path))
    ints = np.fromfile(os.path.join(tmp_path, "Geometry", "domDecmp.core0000"), "<i4")
    assert list(ints[:9]) == [16, 12, 9, 24, 18, 13, 8, 18, 0]
    assert list(ints[9:18]) == [1, 1, 1, 13, 24, 18, 13, 24, 18]
    assert list(ints[18:]) == [1, 1, 1, 9, 18, 8, 9, 18, 8]
    vals = np.fromfile(os.path.join(tmp_path, "Geometry", "geometry.core0000"), "<f8")
    assert vals.size == 6 + 8 + 18
    assert list(vals[:4]) == [2.0, 3.0, 1.0, 0.5]
    assert vals[4] == pytest.approx(np.pi) and vals[5] == pytest.approx(2 * np.pi / 3)
    # kx = (k_index - 1) kx_box; ky per PL_geometry.f90:183-191, restated from k_index
    kyb = 2 * np.pi / 3
    ky = [((k - 1) if k < 18 // 2 + 1 else -(18 + 1 - k)) * kyb for k in range(1, 19)]
    np.testing.assert_allclose(vals[6:14], np.arange(8) * np.pi, rtol=1e-15)
    np.testing.assert_allclose(vals[14:], ky, rtol=1e-15)
