"""Adaptive time-step rules (timeKeeping.f90) on a scripted solver (no GPU)."""
import pytest

from coral_amd.timekeeping import MinimalTimestepReached, Timings, march_adaptive


class Scripted:
    """Stands in for a solver: returns a scripted CFL time step after every step."""

    def __init__(self, cfl):
        self.cfl = list(cfl)
        self.calls = []

    def factorize_operators(self, dt):
        self.calls.append(("factor", dt))

    def march_forward(self, dt, n):
        self.calls.append(("step", dt))

    def cfl_dt(self):
        return self.cfl.pop(0)


def test_refactorization_only_beyond_tolerance():
    t = Timings(dt_max=1e-2, tolerance_factor=0.3, CFL_C=0.5)
    # CFL dt * 0.5: 1 (capped to 1e-2), 1.8e-2 (capped), 1.6e-2 -> 8e-3 (20 % change: kept),
    # 1.2e-2 -> 6e-3 (40 % change from 1e-2: refactorize), 1.1e-2 -> 5.5e-3 (8 %: kept)
    s = Scripted([2.0, 3.6e-2, 1.6e-2, 1.2e-2, 1.1e-2])
    used = march_adaptive(s, t, 5)
    assert used == [1e-2, 1e-2, 1e-2, 1e-2, 6e-3]
    assert [c for c in s.calls if c[0] == "factor"] == [("factor", 1e-2), ("factor", 6e-3)]
    assert abs(t.absolute_time - (4e-2 + 6e-3)) < 1e-15
    assert t.i_timestep == 6


def test_minimal_timestep_stops_the_run():
    t = Timings(dt_max=1e-2, tolerance_factor=0.3, CFL_C=0.5)
    with pytest.raises(MinimalTimestepReached):
        march_adaptive(Scripted([1e-9]), t, 1)


def test_cfl_flag_off_keeps_dt_max():
    t = Timings(dt_max=1e-3, tolerance_factor=0.1, CFL_C=0.5)
    t.CFL_flag = False
    used = march_adaptive(Scripted([1e-6, 1e-6]), t, 2)
    assert used == [1e-3, 1e-3]


def test_timings_file_round_trip(tmp_path):
    t = Timings(dt_max=1.0, tolerance_factor=0.3, CFL_C=0.5)
    t.dt_current, t.absolute_time, t.i_timestep = 5e-3, 1.25, 40
    p = str(tmp_path / "dt.sav")
    t.tofile(p)
    u = Timings(dt_max=1.0, tolerance_factor=0.3, CFL_C=0.5)
    u.fromfile(p)
    # the reference restarts with dt / 50 (timeKeeping.f90:106-117)
    assert (u.dt, u.dt_current, u.absolute_time) == (1e-4, 1e-4, 1.25)
