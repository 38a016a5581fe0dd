"""State file formats on the host (no GPU): names and dt.sav."""
import os
import struct

import pytest

from coral_amd import state_io


def test_quicksave_names_follow_the_fortran_formats():
    # PL_IMEX_timestepping_output.f90:336-337, 406-407: i3.3 system and variable, i1.1 rolling index
    assert state_io.quicksave_name("kxky", 0, 4, 1) == "QuickSave.kxky.sys001.var005.rolling1"
    assert state_io.quicksave_name("zero", 11, 0, 2) == "QuickSave.zero.sys012.var001.rolling2"
    with pytest.raises(ValueError):
        state_io.quicksave_name("mean", 0, 0, 1)


def test_rolling_index_prefers_one(tmp_path):
    with pytest.raises(FileNotFoundError):
        state_io.find_rolling(str(tmp_path))
    (tmp_path / state_io.quicksave_name("kxky", 0, 0, 2)).write_bytes(b"")
    assert state_io.find_rolling(str(tmp_path)) == 2
    (tmp_path / state_io.quicksave_name("kxky", 0, 0, 1)).write_bytes(b"")
    assert state_io.find_rolling(str(tmp_path)) == 1


def test_dt_sav_round_trip(tmp_path):
    p = tmp_path / "dt.sav"
    # 16-byte form, as generate_initial_condition.py writes it: dt, time
    state_io.write_dt_sav(str(p), 5e-3, 0.25)
    assert os.path.getsize(p) == 16
    dt, dtc, t, step = state_io.read_dt_sav(str(p), dt_max=1.0)
    assert (dt, dtc, t, step) == (5e-3 / 50, 5e-3 / 50, 0.25, None)
    # capped by dt_max
    assert state_io.read_dt_sav(str(p), dt_max=1e-5)[:2] == (1e-5, 1e-5)
    # 20-byte form: dt, time, i_timestep; without Timeseries/time.dat the step restarts at 1
    state_io.write_dt_sav(str(p), 1e-3, 2.0, 77)
    assert os.path.getsize(p) == 20
    assert state_io.read_dt_sav(str(p), dt_max=1.0) == (1e-3 / 50, 1e-3 / 50, 2.0, 1)
    (tmp_path / ".." / "Timeseries").mkdir(exist_ok=True)
    (tmp_path / ".." / "Timeseries" / "time.dat").write_bytes(b"")
    assert state_io.read_dt_sav(str(p), dt_max=1.0)[3] == 77
    assert struct.unpack("<ddi", p.read_bytes()) == (1e-3, 2.0, 77)
