"""Passive scalars (tracers) on the GPU against the CPU FP64 oracle
(oracle/scalar.py): whole blocks with gravity, diffusion, the covariance and
its positivity limiter, and both implicit schemes with the donor-capped
tracer transfer, over several RK3 steps at 1e-12 (FP64 form) and 1e-5 (FP32
form); r = 1 follows the gas density; a captured stage replays the eager
bits; a redo restores the tracers bit for bit; a resume continues them."""

import numpy as np
import pytest
import torch

from oracle import fv
from tests._gpu_util import host, rel_l2
from tests.test_gpu_gravity import GAMMA, stratified

pytestmark = pytest.mark.gpu

GRID2 = fv.Grid(16, 12, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0)
GRID3 = fv.Grid(12, 8, 6, 0.0, 3200.0, -3200.0, 3200.0, -1600.0, 1600.0)
PERIODIC = ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic")


def _tracers(g, ns=2):
    """Row 0 a sharp slab (the limiter's case), row 1 linear in height."""
    x2 = g.centers(2)[None, :, None] + np.zeros(g.shape)
    z = g.centers(1)[None, None, :] + np.zeros(g.shape)
    r = np.zeros((ns,) + g.shape)
    r[0] = np.where(np.abs(x2) < 2000.0, 1.0, 0.0)
    if ns > 1:
        r[1] = 0.2 + z / (g.x1max - g.x1min)
    return r


def _pair(g, bcs, scheme, cov, precision="float64", execution="eager", ns=2, r0=None, nu=25.0):
    from snapy_amd import MeshBlock
    from snapy_amd.options import (CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions,
                                   ScalarOptions)

    d = fv.Deck(grid=g, gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=bcs, grav1=-9.81,
                cfl=0.9, implicit_scheme=scheme, flux_covariance=cov, nu_iso=nu, kappa_iso=nu, nscalar=ns)
    w = stratified(g, seed=6, noise=0.5)
    r = _tracers(g, ns) if r0 is None else r0
    ref = fv.Block(d)
    ref.initialize(w, r=r)
    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", g.x1min, g.x1max, g.x2min, g.x2max, g.x3min, g.x3max,
                                 g.nx1, g.nx2, g.nx3, 3)
    op.hydro = HydroOptions(riemann="lmars", grav1=-9.81, flux_covariance=cov, nu_iso=nu, kappa_iso=nu,
                            precision=precision)
    op.hydro.eos.gammad = GAMMA
    op.hydro.eos.weight = fv.Deck.weight
    op.hydro.recon1.shock = False
    op.hydro.recon23.shock = False
    op.intg = IntegratorOptions(cfl=0.9, implicit_scheme=scheme, execution=execution)
    op.bcnames = list(bcs)
    op.scalar = ScalarOptions(nvar=ns, names=[f"t{n}" for n in range(ns)])
    block = MeshBlock(op)
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": torch.from_numpy(w).cuda(), "scalar_r": torch.from_numpy(r).cuda()})
    return ref, block, vars


def _steps(ref, block, vars, n, check_dt=True):
    for _ in range(n):
        dt = block.max_time_step(vars)
        if ref is not None:
            dt_ref = ref.max_time_step()
            if check_dt:
                assert abs(dt - dt_ref) <= 1e-12 * dt_ref
        for s in range(3):
            block.forward(vars, dt, s)
            if ref is not None:
                ref.forward(dt, s)
        assert block.check_redo(vars) == 0


CASES = [
    (GRID2, ("reflecting",) * 6, 0, True),
    (GRID2, ("reflecting",) * 6, 9, True),
    (GRID2, ("reflecting",) * 6, 1, True),
    (GRID2, ("reflecting",) * 6, 9, False),
    (GRID3, PERIODIC, 1, True),
    (GRID3, PERIODIC, 0, False),
]


@pytest.mark.parametrize("g,bcs,scheme,cov", CASES)
def test_tracer_block_matches_oracle(g, bcs, scheme, cov):
    print(f"scalar block {g.nx1}x{g.nx2}x{g.nx3} scheme {scheme} cov {cov}", flush=True)
    ref, block, vars = _pair(g, bcs, scheme, cov)
    _steps(ref, block, vars, 3)
    sub = (slice(None),) + g.interior()
    u = host(vars["hydro_u"])
    for rows in ([0], [1, 2, 3], [4]):
        assert rel_l2(u[sub][rows], ref.u[sub][rows]) < 1e-12
    s, r = host(vars["scalar_s"]), host(vars["scalar_r"])
    for n in range(2):
        es, er = rel_l2(s[n][sub[1:]], ref.s[n][sub[1:]]), rel_l2(r[n][sub[1:]], ref.r[n][sub[1:]])
        print("tracer", n, es, er, flush=True)
        assert es < 1e-12 and er < 1e-12
    # the ghosts too (reflecting mirror / periodic copy of s, r = s / rho)
    assert rel_l2(s, ref.s) < 1e-12
    # the limiter keeps the slab non-negative
    if cov:
        assert s[0][sub[1:]].min() >= 0.0


def test_fp32_form_tracers_within_policy():
    g = GRID2
    print("scalar block FP32 form, scheme 9", flush=True)
    ref, block, vars = _pair(g, ("reflecting",) * 6, 9, True, precision="float32")
    _steps(ref, block, vars, 3, check_dt=False)
    sub = g.interior()
    s = host(vars["scalar_s"])
    for n in range(2):
        err = rel_l2(s[n][sub], ref.s[n][sub])
        print("tracer", n, err, flush=True)
        assert err < 1e-5


@pytest.mark.parametrize("scheme", [9, 1])
def test_unit_ratio_follows_the_gas_density(scheme):
    g = GRID2
    print(f"scalar r = 1, scheme {scheme}", flush=True)
    _, block, vars = _pair(g, ("reflecting",) * 6, scheme, True, ns=1, r0=np.ones((1,) + g.shape))
    _steps(None, block, vars, 4)
    sub = (slice(None),) + g.interior()
    r = host(vars["scalar_r"])
    assert np.abs(r[sub] - 1.0).max() < 1e-13


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_captured_stage_replays_the_eager_bits(precision):
    g = GRID3
    print(f"scalar graph == eager {precision}", flush=True)
    _, b0, v0 = _pair(g, PERIODIC, 1, True, precision=precision)
    _, b1, v1 = _pair(g, PERIODIC, 1, True, precision=precision, execution="hip-graph")
    for _ in range(4):
        for b, v in ((b0, v0), (b1, v1)):
            dt = b.device_time_step(v)
            for s in range(3):
                b.forward(v, dt, s)
        torch.cuda.synchronize()
        assert torch.equal(v0["scalar_s"], v1["scalar_s"])
        assert torch.equal(v0["scalar_r"], v1["scalar_r"])
    assert b1.execution_report()["replays"] > 0


def test_redo_restores_the_tracers():
    g = GRID2
    print("scalar redo", flush=True)
    _, block, vars = _pair(g, ("reflecting",) * 6, 9, True)
    _steps(None, block, vars, 1)
    s_before, r_before = vars["scalar_s"].clone(), vars["scalar_r"].clone()
    dt = block.max_time_step(vars)
    for s in range(3):
        block.forward(vars, dt, s)
    vars["hydro_u"][0, 0, 6, 6] = float("nan")
    assert block.check_redo(vars) > 0
    torch.cuda.synchronize()
    assert torch.equal(vars["scalar_s"], s_before)
    assert torch.equal(vars["scalar_r"], r_before)


def test_resume_continues_the_tracers(tmp_path):
    g = GRID2
    print("scalar resume", flush=True)
    _, a, va = _pair(g, ("reflecting",) * 6, 1, True)
    _steps(None, a, va, 2)
    path = a.write_restart(va, str(tmp_path / "t.block0.00001.part"), 1.0)
    _, b, vb = _pair(g, ("reflecting",) * 6, 1, True)
    vb, t = b.initialize_from_restart(path)
    assert t == 1.0
    _steps(None, a, va, 2)
    _steps(None, b, vb, 2)
    torch.cuda.synchronize()
    assert torch.equal(va["scalar_s"], vb["scalar_s"])
    assert torch.equal(va["hydro_u"], vb["hydro_u"])
