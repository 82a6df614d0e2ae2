"""Isotropic diffusion on the GPU against the CPU FP64 oracle: the face
fluxes with the energy posts and the momentum increment of one call (2D
walls, 3D periodic), then whole blocks: a viscous shear layer without
gravity and the density current with gravity and the vertical implicit
correction, over several RK3 steps."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import diffusion, fv
from snapy_amd import _capi
from tests._gpu_util import dev, grid_desc, host, phys_desc, rel_l2
from tests.test_gpu_gravity import GAMMA, stratified

pytestmark = pytest.mark.gpu
WEIGHT = 0.02897025302492418
RD = diffusion.RGAS / WEIGHT

CASES = [
    (fv.Grid(12, 10, 1, 0.0, 1200.0, -500.0, 500.0, 0.0, 100.0),
     ("reflecting", "reflecting", "reflecting", "reflecting", "periodic", "periodic")),
    (fv.Grid(8, 6, 7, 0.0, 800.0, 0.0, 600.0, 0.0, 700.0), ("periodic",) * 6),
    (fv.Grid(10, 8, 6, 0.0, 1000.0, 0.0, 800.0, 0.0, 600.0),
     ("reflecting", "reflecting") + ("periodic",) * 4),
]


def smooth_state(g, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros((5,) + g.shape)
    w[fv.IDN] = 1.1 + 0.2 * rng.random(g.shape)
    w[fv.IPR] = 1.0e5 * (1.0 + 0.05 * rng.random(g.shape))
    for v in (fv.IVX, fv.IVY, fv.IVZ):
        w[v] = 2.0 * (rng.random(g.shape) - 0.5)
    return w


@pytest.mark.parametrize("g,bcs", CASES)
def test_diffusion_kernel_matches_oracle(g, bcs):
    lib = _capi.load()
    w = smooth_state(g, 3)
    fv.apply_boundaries(w, bcs, GAMMA, None, primitive=True)
    nu, kappa, dt = 75.0, 60.0, 0.25
    rng = np.random.default_rng(5)
    axes = [a for a in (1, 2, 3) if g.nc(a) > 1]
    totals = {a: rng.random((5,) + g.shape) for a in axes}
    pd = phys_desc(gammad=GAMMA, bcs=bcs)
    pd.nu_iso, pd.kappa_iso, pd.gas_r = nu, kappa, RD
    gd = grid_desc(g)
    td = {a: dev(totals[a]) for a in axes}
    tp = (ctypes.c_void_p * 3)(*[td[a].data_ptr() if a in td else None for a in (1, 2, 3)])
    work = torch.zeros(int(lib.snapy_diffusion_workspace(ctypes.byref(gd))), dtype=torch.float64, device="cuda")
    dsrc = torch.full((5,) + g.shape, 7.0, dtype=torch.float64, device="cuda")
    wd = dev(w)
    print(f"diffusion kernel {g.nx1}x{g.nx2}x{g.nx3} {bcs[0]}", flush=True)
    rc = lib.snapy_diffusion(ctypes.byref(gd), ctypes.byref(pd), 0, wd.data_ptr(), tp, work.data_ptr(),
                             dsrc.data_ptr(), dt, None, None)
    assert rc == 0, lib.snapy_last_error()
    fd = diffusion.cartesian_fluxes(w, g, nu, kappa, GAMMA, WEIGHT, bcs)
    ref_tot = {a: totals[a].copy() for a in axes}
    diffusion.post_fluxes(ref_tot, fd, g, bcs)
    # the hydro totals of a wall face were prescribed by the sweep: the
    # post leaves them; the oracle zeroes them, so compare off the walls
    for a in axes:
        got = host(td[a])
        ref = ref_tot[a]
        tdim = {1: 3, 2: 2, 3: 1}[a]
        fv.prescribe_walls(got, tdim, bcs, a, g)
        assert np.array_equal(got[fv.IPR], ref[fv.IPR]), (a, np.max(np.abs(got[fv.IPR] - ref[fv.IPR])))
        # the other rows are untouched
        assert np.array_equal(got[fv.IVX : fv.IPR], totals[a][fv.IVX : fv.IPR])
    du = np.zeros_like(w)
    diffusion.add_momentum(du, fd, g, dt)
    sub = (slice(None),) + g.interior()
    got = host(dsrc)[sub]
    assert np.array_equal(got, du[sub]), np.max(np.abs(got - du[sub]))


def _block(g, bcs, grav1, scheme, nu, kappa, riemann):
    from snapy_amd import MeshBlock
    from snapy_amd.options import CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions

    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", g.x1min, g.x1max, g.x2min, g.x2max, g.x3min, g.x3max,
                                 g.nx1, g.nx2, g.nx3, 3)
    op.hydro = HydroOptions(riemann=riemann, grav1=grav1, nu_iso=nu, kappa_iso=kappa)
    op.hydro.eos.gammad = GAMMA
    op.hydro.eos.weight = WEIGHT
    op.hydro.recon1.shock = False
    op.hydro.recon23.shock = False
    op.intg = IntegratorOptions(cfl=0.9, implicit_scheme=scheme)
    op.bcnames = list(bcs)
    return MeshBlock(op)


BLOCKS = [
    # a viscous, conducting shear layer without gravity (the explicit update)
    (fv.Grid(16, 24, 1, 0.0, 1600.0, -1200.0, 1200.0, 0.0, 100.0),
     ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic"), 0.0, 0, 50.0),
    # the density current: gravity, reflecting walls, the implicit x1
    (fv.Grid(24, 32, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0), ("reflecting",) * 6, -9.81, 9, 75.0),
    # 3D with gravity, explicit
    (fv.Grid(12, 8, 6, 0.0, 2400.0, -800.0, 800.0, -600.0, 600.0),
     ("reflecting", "reflecting") + ("periodic",) * 4, -9.81, 0, 75.0),
]


@pytest.mark.parametrize("g,bcs,grav1,scheme,nu", BLOCKS)
def test_whole_block_with_diffusion(g, bcs, grav1, scheme, nu):
    if grav1 != 0.0:
        w = stratified(g, seed=6, noise=0.5)
    else:
        w = smooth_state(g, 8)
        w[fv.IVY] = 20.0 * np.tanh(g.centers(1)[None, None, :] / 200.0 - 4.0) + 0.0 * w[fv.IDN]
    d = fv.Deck(grid=g, gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=bcs, grav1=grav1,
                cfl=0.9, implicit_scheme=scheme, nu_iso=nu, kappa_iso=nu, weight=WEIGHT, flux_covariance=True)
    ref = fv.Block(d)
    ref.initialize(w)
    block = _block(g, bcs, grav1, scheme, nu, nu, "lmars")
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    for n in range(3):
        dt = block.max_time_step(vars)
        dt_ref = ref.max_time_step()
        assert abs(dt - dt_ref) <= 1e-12 * dt_ref, (dt, dt_ref)
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
