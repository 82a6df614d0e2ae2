"""The FP32 state form (dynamics/precision: float32) on the GPU against the
CPU FP64 oracle: FP32 face fluxes, the compensated FP32 state pair, FP64
primitives, x1 reference, implicit solve and budgets.  The bars are the
FP32 policy's: relative L2 of density, momentum and energy within 1e-5 of
the FP64 oracle, mass conserved to 1e-13 relative, and a hydrostatic rest
state keeping |m1| below 2e-8 (the reference's invariant,
tests/test_fp32_policy.py:54-65)."""

import numpy as np
import pytest
import torch

from oracle import fv
from tests._gpu_util import host, rel_l2
from tests.test_gpu_gravity import GAMMA, stratified

pytestmark = pytest.mark.gpu
WEIGHT = 0.02897025302492418
TOL = 1e-5


def _block(g, bcs, precision, grav1=0.0, scheme=0, riemann="lmars", shock=False, nu=0.0, cov=True,
           heat=0.0, gamma=GAMMA, cfl=0.9):
    from snapy_amd import MeshBlock
    from snapy_amd.options import CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions

    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", g.x1min, g.x1max, g.x2min, g.x2max, g.x3min, g.x3max,
                                 g.nx1, g.nx2, g.nx3, 3)
    op.hydro = HydroOptions(riemann=riemann, grav1=grav1, nu_iso=nu, kappa_iso=nu, precision=precision,
                            flux_covariance=cov, bot_heat_flux=heat)
    op.hydro.eos.gammad = gamma
    op.hydro.eos.weight = WEIGHT
    op.hydro.recon1.shock = shock
    op.hydro.recon23.shock = shock
    op.intg = IntegratorOptions(cfl=cfl, implicit_scheme=scheme)
    op.bcnames = list(bcs)
    return MeshBlock(op)


def _run_pair(g, bcs, w, steps, **kw):
    """Advance the FP32 block and the FP64 oracle from the same state with
    the oracle's dt; return (FP32 total state, oracle state, mass drift)."""
    deck = fv.Deck(grid=g, gammad=kw.get("gamma", GAMMA), riemann=kw.get("riemann", "lmars"),
                   shock_x1=kw.get("shock", False), shock_x23=kw.get("shock", False), bcs=bcs,
                   grav1=kw.get("grav1", 0.0), cfl=kw.get("cfl", 0.9), implicit_scheme=kw.get("scheme", 0),
                   nu_iso=kw.get("nu", 0.0), kappa_iso=kw.get("nu", 0.0), weight=WEIGHT,
                   flux_covariance=kw.get("cov", True), bot_heat_flux=kw.get("heat", 0.0))
    ref = fv.Block(deck)
    ref.initialize(w)
    block = _block(g, bcs, "float32", **kw)
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    assert vars["hydro_u"].dtype == torch.float32
    assert vars["hydro_w"].dtype == torch.float64
    assert block.flux_register().total(1).dtype == torch.float32
    m0 = float(block.budget(vars)[0].item())
    for n in range(steps):
        dt = ref.max_time_step()
        for s in range(3):
            block.forward(vars, dt, s)
            ref.forward(dt, s)
        assert block.check_redo(vars) == 0
    m1 = float(block.budget(vars)[0].item())
    u = host(block.total_u(vars))
    return u, ref.u, abs(m1 - m0) / abs(m0)


def _errors(u, uref, g):
    sub = (slice(None),) + g.interior()
    return [rel_l2(u[sub][rows], uref[sub][rows]) for rows in ([0], [1, 2, 3], [4])]


def test_fp32_sod_within_policy_tolerance():
    g = fv.Grid(256, 1, 1, 0.0, 1.0, -0.5, 0.5, -0.5, 0.5)
    w = fv.sod_initial(g)
    print("fp32 sod 256", flush=True)
    u, uref, drift = _run_pair(g, ("outflow",) * 6, w, 20, riemann="hllc", shock=True, gamma=1.4)
    errs = _errors(u, uref, g)
    print("rel L2", errs, "mass drift", drift, flush=True)
    assert max(errs) < TOL
    assert drift < 1e-13


CASES = [
    # the 3D convection physics: periodic x2/x3, covariance, bottom heating,
    # the implicit x1
    (fv.Grid(16, 8, 8, 0.0, 800.0, 0.0, 800.0, 0.0, 800.0),
     ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic"),
     dict(grav1=-9.81, scheme=9, heat=100.0)),
    # the density current physics: walls, diffusion, the implicit x1
    (fv.Grid(24, 32, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0), ("reflecting",) * 6,
     dict(grav1=-9.81, scheme=9, nu=75.0)),
    # gravity, explicit
    (fv.Grid(24, 20, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0), ("reflecting",) * 6, dict(grav1=-9.81)),
    # the 3D convection physics with the partial implicit matrix (scheme 1)
    (fv.Grid(16, 8, 8, 0.0, 800.0, 0.0, 800.0, 0.0, 800.0),
     ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic"),
     dict(grav1=-9.81, scheme=1, heat=100.0)),
]


@pytest.mark.parametrize("g,bcs,kw", CASES)
def test_fp32_stratified_within_policy_tolerance(g, bcs, kw):
    w = stratified(g, seed=4, noise=0.5)
    print(f"fp32 stratified {g.nx1}x{g.nx2}x{g.nx3} {kw}", flush=True)
    u, uref, drift = _run_pair(g, bcs, w, 3, **kw)
    errs = _errors(u, uref, g)
    print("rel L2", errs, "mass drift", drift, flush=True)
    assert max(errs) < TOL
    assert drift < 1e-13


def test_fp32_hydrostatic_rest_stays_at_rest():
    g = fv.Grid(32, 16, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0)
    w = stratified(g, seed=0, bubble=False, noise=0.0)
    bcs = ("reflecting",) * 6
    block = _block(g, bcs, "float32", grav1=-9.81, scheme=9)
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    print("fp32 hydrostatic rest", flush=True)
    for n in range(5):
        dt = block.max_time_step(vars)
        for s in range(3):
            block.forward(vars, dt, s)
        assert block.check_redo(vars) == 0
    u = host(block.total_u(vars))
    sub = g.interior()
    m1 = np.abs(u[fv.IVX][sub]).max()
    print("max |m1|", m1, flush=True)
    assert m1 < 2e-8
