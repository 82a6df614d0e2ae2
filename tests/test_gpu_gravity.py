"""The gravity path on the GPU against the CPU FP64 oracle: the hydrostatic
x1 reference, the well-balanced x1 sweep, the gravity update, and a whole 2D
stratified block over several RK3 steps (1e-12 relative L2)."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import fv, gravity
from snapy_amd import _capi
from snapy_amd.geometry import x1_gravity_table
from tests._gpu_util import dev, grid_desc, host, phys_desc, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12
RD, CP = 287.0, 1004.0
GAMMA = CP / (CP - RD)


def stratified(g, seed=0, bubble=True, noise=1e-3):
    """A dry isentropic column (theta = 300 K) with an optional cold bubble
    and a little velocity noise."""
    rng = np.random.default_rng(seed)
    z = g.centers(1)[None, None, :]
    x = g.centers(2)[None, :, None]
    t = 300.0 - 9.81 * z / CP + 0.0 * x
    p = 1.0e5 * (t / 300.0) ** (CP / RD)
    if bubble:
        L = np.sqrt((x / 4000.0) ** 2 + ((z - 3000.0) / 2000.0) ** 2)
        t = t - np.where(L <= 1.0, 15.0 * (1 + np.cos(np.pi * L)) / 2, 0.0)
    w = np.zeros((5,) + g.shape)
    w[fv.IPR] = p
    w[fv.IDN] = p / (RD * t)
    for v in (fv.IVX, fv.IVY):
        w[v] = noise * (rng.random(g.shape) - 0.5)
    return w


GRIDS = [fv.Grid(24, 20, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0),
         fv.Grid(16, 10, 8, 0.0, 3200.0, -3200.0, 3200.0, -1600.0, 1600.0)]


@pytest.mark.parametrize("g", GRIDS)
def test_reference_and_wb_sweep(g):
    lib = _capi.load()
    w = stratified(g)
    bcs = ("reflecting",) * 6
    pd = phys_desc(gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=bcs, grav1=-9.81)
    gd = grid_desc(g)
    wd = dev(w)
    ref = torch.zeros((5,) + g.shape, dtype=torch.float64, device="cuda")
    print(f"hydro_ref_x1 {g.nx1}x{g.nx2}x{g.nx3}", flush=True)
    assert lib.snapy_hydro_ref_x1(ctypes.byref(gd), ctypes.byref(pd), wd.data_ptr(), ref.data_ptr(), None) == 0
    iu = g.nc(1) - 4
    r = gravity.hydro_ref_x1(w, np.full(g.nc(1), g.dx(1)), iu, 9.81)
    got = host(ref)
    for n, name in enumerate(("psf_lo", "psf_hi", "pref", "dsf", "dref")):
        assert rel_l2(got[n], r[n]) < TOL, name
    flx = torch.zeros_like(wd)
    fp = torch.zeros_like(wd[0])
    assert lib.snapy_sweep_x1_wb(ctypes.byref(gd), ctypes.byref(pd), 0, wd.data_ptr(), ref.data_ptr(),
                                 flx.data_ptr(), fp.data_ptr(), None) == 0
    wlr = gravity.wb_reconstruct_x1(w, g, -9.81, False)
    fo, po = fv.lmars(wlr[0], wlr[1], GAMMA, 3)
    fv.prescribe_walls(fo, 3, bcs, 1, g)
    k, j, i = g.interior()
    sl = (k, j, slice(i.start, i.stop + 1))
    f = host(flx)
    for v in range(5):
        assert rel_l2(f[v][sl], fo[v][sl]) < TOL, v
    assert rel_l2(host(fp)[sl], po[sl]) < TOL


@pytest.mark.parametrize("g", GRIDS)
def test_whole_block_with_gravity(g):
    bcs = ("reflecting",) * 6
    d = fv.Deck(grid=g, gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=bcs,
                grav1=-9.81, cfl=0.9)
    w = stratified(g, seed=3)
    ref = fv.Block(d)
    ref.initialize(w)
    from snapy_amd import MeshBlock
    from snapy_amd.options import MeshBlockOptions, CoordinateOptions, HydroOptions, IntegratorOptions

    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", g.x1min, g.x1max, g.x2min, g.x2max, g.x3min, g.x3max,
                                 g.nx1, g.nx2, g.nx3, 3)
    op.hydro = HydroOptions(riemann="lmars", grav1=-9.81, flux_covariance=False)
    op.hydro.eos.gammad = GAMMA
    op.hydro.recon1.shock = False
    op.hydro.recon23.shock = False
    op.intg = IntegratorOptions(cfl=0.9)
    op.bcnames = list(bcs)
    block = MeshBlock(op)
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    for n in range(4):
        dt = block.max_time_step(vars)
        dt_ref = ref.max_time_step()
        assert abs(dt - dt_ref) <= 1e-12 * dt_ref
        print(f"step {n} dt {dt}", flush=True)
        for s in range(3):
            block.forward(vars, dt, s)
            ref.forward(dt, s)
    u = host(vars["hydro_u"])
    sub = (slice(None),) + g.interior()
    for rows in ([0], [1, 2, 3], [4]):
        err = rel_l2(u[sub][rows], ref.u[sub][rows])
        print("rows", rows, err, flush=True)
        assert err < TOL


def test_gravity_table_matches_oracle_weights():
    g = fv.Grid(20, 1, 1, 0.0, 2000.0)
    x1f, x1v = torch.from_numpy(g.faces(1)), torch.from_numpy(g.centers(1))
    tab = x1_gravity_table(x1f, x1v, 3, 20, -9.81).numpy()
    idx, wt = gravity.centroid_slope_weights(g.centers(1)[3:23])
    for r in range(3):
        assert np.array_equal(tab[4 + r, 3:23], wt[r])


@pytest.mark.parametrize("g", GRIDS + [fv.Grid(20, 7, 5, 0.0, 4000.0, -700.0, 700.0, 0.0, 500.0),
                                       # a tall column: fewer columns per workgroup
                                       fv.Grid(640, 3, 1, 0.0, 16000.0, -300.0, 300.0, 0.0, 100.0)])
@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("bcs", [("reflecting",) * 6, ("outflow", "outflow") + ("periodic",) * 4])
def test_fused_reference_sweep_is_the_two_launches(g, dtype, bcs):
    """snapy_sweep_x1_wb_ref gives the same bits as snapy_hydro_ref_x1 then
    snapy_sweep_x1_wb: the fluxes, the face pressure and the psf_lo plane of
    every interior column (a partial last workgroup included)."""
    lib = _capi.load()
    w = stratified(g, seed=11, noise=0.3)
    pd = phys_desc(gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=bcs, grav1=-9.81)
    gd = grid_desc(g)
    wd = dev(w)
    fdt = torch.float64 if dtype == 0 else torch.float32
    ref_a = torch.zeros((5,) + g.shape, dtype=torch.float64, device="cuda")
    ref_b = torch.zeros_like(ref_a)
    fa, fb = torch.zeros(w.shape, dtype=fdt, device="cuda"), torch.zeros(w.shape, dtype=fdt, device="cuda")
    pa, pb = torch.zeros_like(wd[0]), torch.zeros_like(wd[0])
    print(f"fused x1 reference sweep {g.nx1}x{g.nx2}x{g.nx3} dtype {dtype} {bcs[0]}", flush=True)
    assert lib.snapy_hydro_ref_x1(ctypes.byref(gd), ctypes.byref(pd), wd.data_ptr(), ref_a.data_ptr(), None) == 0
    assert lib.snapy_sweep_x1_wb(ctypes.byref(gd), ctypes.byref(pd), dtype, wd.data_ptr(), ref_a.data_ptr(),
                                 fa.data_ptr(), pa.data_ptr(), None) == 0
    assert lib.snapy_sweep_x1_wb_ref(ctypes.byref(gd), ctypes.byref(pd), dtype, wd.data_ptr(), ref_b.data_ptr(),
                                     fb.data_ptr(), pb.data_ptr(), None) == 0
    k, j, i = g.interior()
    cols = (slice(None), k, j)
    assert torch.equal(fa[cols], fb[cols])
    assert torch.equal(pa[k, j], pb[k, j])
    assert torch.equal(ref_a[0][k, j], ref_b[0][k, j])
