"""The vertical implicit correction (implicit-schemes 9 and 1) on the GPU
against the CPU FP64 oracle: one kernel call on random columns, then whole 2D/3D
blocks with gravity and the implicit x1 over several RK3 steps."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import fv, implicit
from snapy_amd import _capi
from snapy_amd.geometry import vic_table, x1_gravity_table
from tests._gpu_util import dev, grid_desc, host, phys_desc, rel_l2
from tests.test_gpu_gravity import GAMMA, stratified

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("entry", ["snapy_vic_full", "snapy_vic_full_serial", "snapy_vic_partial"])
@pytest.mark.parametrize("shape", [(16, 6, 1), (12, 4, 3), (40, 37, 3), (400, 6, 1), (5, 9, 1), (16, 70, 60)])
def test_vic_kernel_matches_oracle(shape, entry):
    """One call on random columns against the oracle; the 400-layer columns
    take the finish kernel's fallback to fewer columns per workgroup (its
    LDS holds 3 (n + 1) doubles per column); the 5-layer columns (the
    fewest scheme 1 takes) put both Woodbury rows on layer 2.  Scheme 9
    solves grids of at most 1536 columns by cyclic reduction and larger ones
    with the forward sweep's lane kernels (the 70 x 60 grid)."""
    if (shape[0] > 100 or shape[0] < 8) and entry != "snapy_vic_partial":
        pytest.skip("the tall and the 5-layer columns exercise the scheme-1 kernels' edge cases")
    if shape[1] * shape[2] > 1536 and entry != "snapy_vic_full":
        pytest.skip("the 4200-column grid exercises scheme 9's lane kernels")
    lib = _capi.load()
    g = fv.Grid(*shape, 0.0, 50.0 * shape[0], -400.0, 400.0, 0.0, 300.0)
    w = stratified(g, seed=5, noise=0.5)
    rng = np.random.default_rng(9)
    du = np.zeros_like(w)
    sub = (slice(None),) + g.interior()
    du[sub] = 1e-3 * (rng.random(du[sub].shape) - 0.5) * np.array([1, 1, 1, 1, 1e3])[:, None, None, None]
    dt = 0.7
    pd = phys_desc(gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=("reflecting",) * 6,
                   grav1=-9.81)
    scheme = 1 if entry == "snapy_vic_partial" else 9
    pd.implicit_scheme = scheme
    gd = grid_desc(g)
    x1f, x1v = torch.from_numpy(g.faces(1)), torch.from_numpy(g.centers(1))
    a1, vol = g.dx(3) * g.dx(2), g.dx(3) * (g.dx(2) * g.dx(1))
    tab, t0, t1 = vic_table(x1f, x1v, 3, g.nx1, -9.81, a1, vol)
    xt = x1_gravity_table(x1f, x1v, 3, g.nx1, -9.81)
    wsize = lib.snapy_vic_partial_workspace if scheme == 1 else lib.snapy_vic_workspace
    ws = torch.empty(int(wsize(ctypes.byref(gd))), dtype=torch.float64, device="cuda")
    st = torch.zeros(g.nx2 * g.nx3, dtype=torch.int32, device="cuda")
    u0 = torch.zeros(w.shape, dtype=torch.float64, device="cuda")
    u = torch.zeros_like(u0)
    tabd, xtd, wd, dud = tab.cuda(), xt.cuda(), dev(w), dev(du)
    print(f"vic kernel {entry} {shape}", flush=True)
    moved = torch.full(g.shape, -7.0, dtype=torch.float64, device="cuda")
    rc = getattr(lib, entry)(ctypes.byref(gd), ctypes.byref(pd), wd.data_ptr(), dud.data_ptr(), u0.data_ptr(),
                            u.data_ptr(), tabd.data_ptr(), xtd.data_ptr(), t0, t1, dt, None, 1.0, 0.0, 0.0, 1.0,
                            ws.data_ptr(), st.data_ptr(), moved.data_ptr(), None)
    assert rc == 0, lib.snapy_last_error()
    assert int(host(st).max()) == 0
    got = host(u)
    # oracle on the same columns
    k, j, i = g.interior()
    n = g.nx1
    wc = w[:, k, j, 2 : 3 + n + 1]
    ncol = wc.shape[1] * wc.shape[2]
    wcol = np.moveaxis(wc, 0, -1).reshape(ncol, n + 2, 5)
    dcol = np.moveaxis(du[:, k, j, 3 : 3 + n], 0, -1).reshape(ncol, n, 5)
    solve = implicit.vic_partial if scheme == 1 else implicit.vic_full
    ref, ref_moved = solve(dcol, wcol, GAMMA, -9.81, dt, g.dx(1), g.faces(1)[3 : 3 + n + 1], g.centers(1)[3 : 3 + n],
                           a1, vol, return_moved=True)
    gc = np.moveaxis(got[:, k, j, 3 : 3 + n], 0, -1).reshape(ncol, n, 5)
    for v in range(5):
        err = rel_l2(gc[..., v], ref[..., v])
        print("row", v, err, flush=True)
        assert err < 1e-11
    # the dry-gas transfer through the face below each cell, the passive
    # scalars' input (0 through the column's bottom face)
    mv = host(moved)[k, j, 3 : 3 + n].reshape(ncol, n)
    assert np.array_equal(mv[:, 0], np.zeros(ncol))
    err = rel_l2(mv, ref_moved[:, :n])
    print("moved", err, flush=True)
    # the transfer is a running sum over the column of the implicit minus
    # the explicit mass increment, nearly equal numbers, so it carries the
    # rows' rounding (1e-13 here) magnified by their ratio to the
    # difference; the cyclic reduction rounds in another order than the
    # oracle's sweep
    assert err < (1e-11 if scheme == 1 else 3e-11)


CASES = [
    (fv.Grid(24, 16, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0), ("reflecting",) * 6, False, 0.0),
    (fv.Grid(12, 8, 6, 0.0, 3200.0, -3200.0, 3200.0, -1600.0, 1600.0), ("reflecting",) * 6, False, 0.0),
    # the 3D convection physics: periodic x2/x3, covariance on, bottom heating
    (fv.Grid(16, 8, 8, 0.0, 800.0, 0.0, 800.0, 0.0, 800.0),
     ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic"), True, 100.0),
]


@pytest.mark.parametrize("scheme", [9, 1])
@pytest.mark.parametrize("g,bcs,cov,heat", CASES)
def test_whole_block_implicit(g, bcs, cov, heat, scheme):
    from snapy_amd import MeshBlock
    from snapy_amd.options import CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions

    d = fv.Deck(grid=g, gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=bcs,
                grav1=-9.81, cfl=0.9, implicit_scheme=scheme, flux_covariance=cov, bot_heat_flux=heat)
    w = stratified(g, seed=4, noise=0.5)
    ref = fv.Block(d)
    ref.initialize(w)
    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", g.x1min, g.x1max, g.x2min, g.x2max, g.x3min, g.x3max,
                                 g.nx1, g.nx2, g.nx3, 3)
    op.hydro = HydroOptions(riemann="lmars", grav1=-9.81, flux_covariance=cov, bot_heat_flux=heat)
    op.hydro.eos.gammad = GAMMA
    op.hydro.recon1.shock = False
    op.hydro.recon23.shock = False
    op.intg = IntegratorOptions(cfl=0.9, implicit_scheme=scheme)
    op.bcnames = list(bcs)
    block = MeshBlock(op)
    print(f"whole block scheme {scheme} {g.nx1}x{g.nx2}x{g.nx3}", flush=True)
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    for n in range(3):
        dt = block.max_time_step(vars)
        dt_ref = ref.max_time_step()
        assert abs(dt - dt_ref) <= 1e-12 * dt_ref
        print(f"step {n} dt {dt}", flush=True)
        for s in range(3):
            block.forward(vars, dt, s)
            ref.forward(dt, s)
        assert block.check_redo(vars) == 0
    u = host(vars["hydro_u"])
    sub = (slice(None),) + g.interior()
    for rows in ([0], [1, 2, 3], [4]):
        err = rel_l2(u[sub][rows], ref.u[sub][rows])
        print("rows", rows, err, flush=True)
        assert err < 1e-12
