"""The output layer's host side on the CPU: the deck's outputs list, the
field names, NetCDF name sanitising and the budget series file (written and
read back with scipy)."""

import numpy as np
import pytest

from snapy_amd import output


def test_outputs_list_is_read_with_positional_ids():
    outs = output.read_outputs([
        {"type": "restart", "dt": 300.0},
        {"type": "pnetcdf", "variables": ["prim", "uov", "scalar_prim"], "dt": 300.0},
        {"type": "budget", "dt": 10.0},
    ])
    assert [o.file_id() for o in outs] == ["out0", "out1", "out2"]
    assert outs[1].variables == ["prim", "uov", "scalar_prim"] and not outs[1].double_precision
    with pytest.raises(ValueError, match="output file type must be specified"):
        output.read_outputs([{"dt": 1.0}])
    with pytest.raises(ValueError, match="x2_slice"):
        output.read_outputs([{"type": "netcdf", "x2_slice": [0.0]}])
    with pytest.raises(ValueError, match="output_sumx1"):
        output.read_outputs([{"type": "netcdf", "output_sumx1": True}])
    with pytest.raises(ValueError, match="'hdf5'"):
        output.read_outputs([{"type": "hdf5"}])


def test_field_names_and_sanitising():
    assert output._names("vel", 3) == ["vel1", "vel2", "vel3"]
    assert output._names("rho", 1) == ["rho"]
    assert output._names("a;b", 2) == ["a", "b"]
    assert output._names("q?_bar", 2) == ["q1_bar", "q2_bar"]
    assert output.sanitize("species/1") == "species_1"
    assert output.sanitize("1abc") == "_1abc"
    assert output.row_names(5) == ["mass", "momentum1", "momentum2", "momentum3", "energy"]


def test_budget_series_appends_records(tmp_path):
    from scipy.io import netcdf_file

    p = str(tmp_path / "run.budget.nc")
    rows = output.row_names(5)
    for n in range(3):
        rec = output.BudgetRecord(rows=rows, total=np.arange(5.0) + n, change=np.full(5, 0.5 * n),
                                  boundary=np.full(5, -0.5 * n), time=10.0 * n, interval=10.0, dt=0.1, cycle=7 * n)
        output.append_budget(p, rec, new=(n == 0))
    with netcdf_file(p, "r", mmap=False) as f:
        assert f.variables["time"][:].tolist() == [0.0, 10.0, 20.0]
        assert f.variables["cycle"][:].tolist() == [0.0, 7.0, 14.0]
        assert f.variables["total"][:].shape == (3, 5)
        np.testing.assert_array_equal(f.variables["change"][2], np.full(5, 1.0))
        names = [b"".join(r).rstrip(b"\0").decode() for r in f.variables["row_name"][:]]
        assert names == rows
        assert int(f.snapy_budget_version) == 1
