"""The Cartesian shallow-water stage on the GPU against the CPU FP64 oracle:
one shallow-roe sweep per direction, then whole periodic and walled blocks
over several RK3 steps."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import fv, swe
from snapy_amd import _capi
from tests._gpu_util import dev, grid_desc, host, phys_desc, rel_l2

pytestmark = pytest.mark.gpu


def hump(g, seed=0):
    rng = np.random.default_rng(seed)
    x = g.centers(1)[None, None, :]
    y = g.centers(2)[None, :, None]
    w = np.zeros((4,) + g.shape)
    w[0] = 10.0 + 2.0 * np.exp(-(x**2 + y**2)) + 0.01 * rng.random(g.shape)
    w[1] = 0.5 + 0.1 * (rng.random(g.shape) - 0.5)
    w[2] = -0.3 + 0.1 * (rng.random(g.shape) - 0.5)
    return w


def swe_phys(bcs):
    pd = phys_desc(riemann="shallow-roe", shock_x1=False, shock_x23=False, bcs=bcs)
    pd.eos = 1
    return pd


@pytest.mark.parametrize("dim", [1, 2])
def test_swe_sweep_matches_oracle(dim):
    lib = _capi.load()
    g = fv.Grid(20, 12, 1, -10.0, 10.0, -5.0, 5.0, -0.5, 0.5)
    bcs = ("reflecting",) * 2 + ("periodic",) * 4
    w = hump(g, 1)
    fv.apply_boundaries(w, bcs, 1.4, primitive=True)
    gd = grid_desc(g, nvar=4)
    pd = swe_phys(bcs)
    flux = torch.zeros((4,) + g.shape, dtype=torch.float64, device="cuda")
    print(f"swe sweep x{dim}", flush=True)
    wd = dev(w)  # referenced until the results are read back
    rc = lib.snapy_sweep(dim, ctypes.byref(gd), ctypes.byref(pd), 0, wd.data_ptr(), flux.data_ptr(), None, None)
    assert rc == 0, lib.snapy_last_error()
    tdim = 4 - dim
    wlr = fv.reconstruct(w, tdim, False)
    ref = swe.shallow_roe(wlr[fv.ILT], wlr[fv.IRT], tdim)
    swe._prescribe_walls(ref, tdim, bcs, dim, g)
    k, j, i = g.interior()
    if dim == 1:
        i = slice(i.start, i.stop + 1)
    else:
        j = slice(j.start, j.stop + 1)
    got = host(flux)[:, k, j, i]
    assert np.array_equal(got, ref[:, k, j, i]), np.max(np.abs(got - ref[:, k, j, i]))


BLOCKS = [
    (fv.Grid(32, 24, 1, -10.0, 10.0, -5.0, 5.0, -0.5, 0.5), ("periodic",) * 6),
    (fv.Grid(24, 16, 1, -10.0, 10.0, -5.0, 5.0, -0.5, 0.5), ("reflecting",) * 4 + ("periodic",) * 2),
    (fv.Grid(64, 1, 1, -10.0, 10.0, -0.5, 0.5, -0.5, 0.5), ("periodic",) * 6),
]


@pytest.mark.parametrize("g,bcs", BLOCKS)
def test_swe_block_matches_oracle(g, bcs):
    from snapy_amd import MeshBlock, MeshBlockOptions
    from snapy_amd.options import CoordinateOptions, HydroOptions, IntegratorOptions

    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", g.x1min, g.x1max, g.x2min, g.x2max, g.x3min, g.x3max,
                                 g.nx1, g.nx2, g.nx3, 3)
    op.hydro = HydroOptions(riemann="shallow-roe")
    op.hydro.eos.type = "shallow-water"
    op.hydro.recon1.shock = False
    op.hydro.recon23.shock = False
    op.intg = IntegratorOptions(cfl=0.9)
    op.bcnames = list(bcs)
    block = MeshBlock(op)
    block.to("cuda")
    w = hump(g, 2)
    ref = swe.Block(swe.Deck(grid=g, bcs=bcs))
    ref.initialize(w)
    vars, _ = block.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    print(f"swe block {g.nx1}x{g.nx2} {bcs[0]}", flush=True)
    for n in range(5):
        dt = block.max_time_step(vars)
        dt_ref = ref.max_time_step()
        assert abs(dt - dt_ref) <= 1e-12 * dt_ref, (dt, dt_ref)
        for s in range(3):
            block.forward(vars, dt, s)
            ref.forward(dt, s)
        assert block.check_redo(vars) == 0
    u = host(vars["hydro_u"])
    sub = (slice(None),) + g.interior()
    for rows in ([0], [1, 2]):
        err = rel_l2(u[sub][rows], ref.u[sub][rows])
        print("rows", rows, err, flush=True)
        assert err < 1e-12
