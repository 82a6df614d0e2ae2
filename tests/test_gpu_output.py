"""The output layer on the GPU: a NetCDF snapshot of a tracer block read
back (dimensions, coordinates, names, values at NC_FLOAT), the budget
records of an open block (every row's change plus its outward boundary
transfer is zero for a deck without sources, through outflow faces, and the
boundary transfer is zero through periodic faces), the output schedule and
the cycle line."""

import os

import numpy as np
import pytest
import torch

from oracle import fv

pytestmark = pytest.mark.gpu


def _open_block(precision="float64", bcs=("outflow", "outflow", "periodic", "periodic", "periodic", "periodic"),
                outputs=None):
    from snapy_amd import MeshBlock
    from snapy_amd.options import CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions

    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", 0.0, 1.0, 0.0, 0.5, -0.5, 0.5, 32, 16, 1, 3)
    op.hydro = HydroOptions(riemann="hllc", precision=precision)
    op.hydro.eos.gammad = 1.4
    op.intg = IntegratorOptions(cfl=0.8, execution="eager")
    op.bcnames = list(bcs)
    op.outputs = outputs or []
    blk = MeshBlock(op)
    blk.to("cuda")
    g = fv.Grid(32, 16, 1, 0.0, 1.0, 0.0, 0.5, -0.5, 0.5)
    x1 = g.centers(1)[None, None, :] + np.zeros(g.shape)
    x2 = g.centers(2)[None, :, None] + np.zeros(g.shape)
    w = np.zeros((5,) + g.shape)
    w[fv.IDN] = 1.0 + 0.5 * np.exp(-((x1 - 0.7) ** 2 + (x2 - 0.25) ** 2) / 0.01)
    w[fv.IVX] = 0.6
    w[fv.IVY] = 0.1 * np.sin(2 * np.pi * x2 / 0.5)
    w[fv.IPR] = 1.0
    vars, _ = blk.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    return blk, vars


def _run(blk, vars, outs, steps, t=0.0):
    for _ in range(steps):
        dt = blk.max_time_step(vars)
        for s in range(3):
            blk.forward(vars, dt, s)
        assert blk.check_redo(vars) == 0
        blk.inc_cycle()
        t += dt
        outs.observe(vars, dt)
        outs.make_outputs(vars, t)
    return t


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_budget_change_is_the_boundary_transfer(tmp_path, precision):
    from scipy.io import netcdf_file

    from snapy_amd.output import BlockOutputs, OutputOptions

    print(f"output: budget of an open block, {precision}", flush=True)
    blk, vars = _open_block(precision)
    outs = BlockOutputs(blk, [OutputOptions(fid=0, file_type="budget", dt=0.05)], basename="open",
                        output_dir=str(tmp_path))
    t = _run(blk, vars, outs, 30)
    outs.finalize(vars, t)
    with netcdf_file(os.path.join(str(tmp_path), "open.budget.nc"), "r", mmap=False) as f:
        total, change, boundary = (f.variables[k][:].copy() for k in ("total", "change", "boundary"))
        times = f.variables["time"][:].copy()
    assert len(times) >= 3
    # mass leaves through the x1 outflow faces: the records see it, and every
    # row's change is minus its outward transfer (no sources in this deck)
    assert boundary[1:, 0].sum() > 1e-4 * total[0, 0]
    # each row's magnitude: sum |u| V over the interior (the x2 momentum's
    # total is near zero by symmetry, so its total is no scale)
    c = blk.options.coord
    scale = (blk.total_u(vars)[blk.part(exterior=False)].abs().sum(dim=(1, 2, 3))
             * (c.dx(3) * (c.dx(2) * c.dx(1)))).cpu().numpy()
    # (the x3 momentum is zero throughout: its residual is taken absolute)
    resid = np.abs(change + boundary) / np.where(scale > 0, scale, 1.0)
    print("budget residual", resid.max(axis=0), flush=True)
    assert resid[1:].max() < 1e-12


def test_periodic_faces_carry_no_boundary_transfer(tmp_path):
    from snapy_amd.output import BudgetConsumer

    print("output: budget of a periodic block", flush=True)
    blk, vars = _open_block(bcs=("periodic",) * 6)
    bc = BudgetConsumer(blk)
    bc.observe(vars, 0.0)
    for _ in range(5):
        dt = blk.max_time_step(vars)
        for s in range(3):
            blk.forward(vars, dt, s)
        bc.observe(vars, dt)
    rec = bc.flush(vars, 1.0)
    assert np.array_equal(rec.boundary, np.zeros(5))
    assert abs(rec.change[0]) <= 1e-14 * rec.total[0]


def test_netcdf_snapshot_reads_back(tmp_path):
    from scipy.io import netcdf_file

    from snapy_amd.output import BlockOutputs, OutputOptions, cycle_line
    from tests.test_gpu_scalar import GRID2, _pair

    print("output: netcdf snapshot of a tracer block", flush=True)
    g = GRID2
    _, blk, vars = _pair(g, ("reflecting",) * 6, 9, True)
    opts = [OutputOptions(fid=0, file_type="netcdf", dt=1e9, variables=["prim", "cons", "scalar_prim", "uov"]),
            # "d" selects the density alone (shouldOutputPrimitive({"d"}))
            OutputOptions(fid=1, file_type="netcdf", dt=1e9, variables=["d"], double_precision=True)]
    outs = BlockOutputs(blk, opts, basename="dc", output_dir=str(tmp_path))
    outs.make_outputs(vars, 0.0)
    t = _run(blk, vars, outs, 2)
    outs.finalize(vars, t)
    names = sorted(os.listdir(str(tmp_path)))
    assert names == ["dc.block0.out0.00000.nc", "dc.block0.out0.00001.nc", "dc.block0.out1.00000.nc",
                     "dc.block0.out1.00001.nc"]
    w = vars["hydro_w"].cpu().numpy()
    ng = 3
    inner = w[:, 0, ng : ng + g.nx2, ng : ng + g.nx1]  # (5, x2, x1)
    with netcdf_file(os.path.join(str(tmp_path), "dc.block0.out0.00001.nc"), "r", mmap=False) as f:
        assert f.dimensions["x1"] == g.nx1 and f.dimensions["x2"] == g.nx2 and f.dimensions["x3"] == 1
        assert "x1f" in f.dimensions and "x3f" not in f.dimensions
        assert f.variables["x1"].axis == b"Z" and f.variables["x2"].long_name == b"X-coordinate at cell center"
        np.testing.assert_allclose(f.variables["x1"][:], g.centers(1)[ng : ng + g.nx1], rtol=1e-6)
        np.testing.assert_allclose(f.variables["x1f"][:], g.faces(1)[ng : ng + g.nx1 + 1], rtol=1e-6, atol=1e-3)
        keys = set(f.variables)
        for k in ("rho", "press", "vel1", "vel2", "vel3", "dens", "Etot", "mom1", "mom2", "mom3", "r_t0", "r_t1"):
            assert k in keys, k
        rho = f.variables["rho"][:]
        # NetCDF stores big-endian: NC_FLOAT reads back as '>f4'
        assert rho.dtype.kind == "f" and rho.dtype.itemsize == 4 and rho.shape == (1, g.nx1, 1, g.nx2)
        # (time, x1, x3, x2) against the block's (x3, x2, x1)
        np.testing.assert_array_equal(rho[0, :, 0, :], inner[0].T.astype(np.float32))
        assert f.variables["rho"].units == b"kg/m^3"
        assert float(f.variables["time"][0]) == pytest.approx(t, rel=1e-6)
    with netcdf_file(os.path.join(str(tmp_path), "dc.block0.out1.00001.nc"), "r", mmap=False) as f:
        rho = f.variables["rho"][:]
        assert rho.dtype.itemsize == 8 and set(f.variables) == {"time", "x1", "x1f", "x2", "x2f", "x3", "rho"}
        np.testing.assert_array_equal(rho[0, :, 0, :], inner[0].T)
    line = cycle_line(blk, vars, t, 0.5)
    assert line.startswith(f"cycle={blk.cycle()} redo=0 time=")
    mass = float(blk.budget(vars)[0])
    assert f"mass0={mass:.13e}" in line and " ie=" in line and " pe=" in line


def test_restart_output_resumes_schedule_and_budget(tmp_path):
    """A run with budget and restart outputs, and the same run resumed from
    one of its combined restart files (the reference's TorchScript archive)
    in another directory: the resumed block's state is bit for bit the
    uninterrupted one's, its restart files carry on the same numbers, and its
    budget records after the resume (the open interval carried) are the
    uninterrupted run's, bit for bit."""
    from scipy.io import netcdf_file

    from snapy_amd import restart
    from snapy_amd.output import BlockOutputs, OutputOptions

    print("output: restart output resumed with its schedule and budget carry", flush=True)
    opts = [OutputOptions(fid=0, file_type="budget", dt=0.05), OutputOptions(fid=1, file_type="restart", dt=0.05)]
    da, db = tmp_path / "a", tmp_path / "b"
    blk, vars = _open_block(outputs=opts)
    outs = BlockOutputs(blk, opts, basename="open", output_dir=str(da))
    t = _run(blk, vars, outs, 30)
    names_a = sorted(p.name for p in da.glob("*.restart"))
    assert len(names_a) >= 4 and not list(da.glob("*.part"))
    path = str(da / names_a[2])
    data = restart.load_restart(path)
    t0, c0 = float(data["last_time"][0]), int(data["last_cycle"][0])
    b2, _ = _open_block(outputs=opts)
    vb, tb = b2.initialize_from_restart(path)
    assert tb == t0 and b2.cycle() == c0
    outs_b = BlockOutputs(b2, opts, basename="open", output_dir=str(db))
    outs_b.resume(data, tb)
    assert outs_b.file_number[1] == 3
    tb = _run(b2, vb, outs_b, 30 - c0, tb)
    assert tb == t
    assert torch.equal(vb["hydro_u"], vars["hydro_u"])
    assert sorted(p.name for p in db.glob("*.restart")) == names_a[3:]

    def records(d):
        with netcdf_file(os.path.join(str(d), "open.budget.nc"), "r", mmap=False) as f:
            return {k: f.variables[k][:].copy() for k in ("time", "total", "change", "boundary", "interval")}

    ra, rb = records(da), records(db)
    later = ra["time"] > t0
    assert later.sum() == len(rb["time"]) >= 2
    for k in ra:
        np.testing.assert_array_equal(ra[k][later], rb[k], err_msg=k)
