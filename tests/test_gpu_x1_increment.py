"""The FP32 x1 reference sweep with the stage increment fused
(snapy_sweep_x1_wb_inc_f32): on the same primitives, its x1 fluxes, face
pressures, reference face pressure plane and FP64 stage increment are the
bits of snapy_sweep_x1_wb_ref followed by snapy_update_gravity_f32 with
du_out (walls, outflow, bot-heat, a 2D block without x3 fluxes, a member
axis), and the block using it stays within the FP32 bar of the oracle."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import fv
from snapy_amd import _capi
from tests._gpu_util import host
from tests.test_gpu_fp32 import _block
from tests.test_gpu_gravity import stratified

pytestmark = pytest.mark.gpu

CASES = [
    # (grid, bcs, heat)
    (fv.Grid(16, 8, 8, 0.0, 800.0, 0.0, 800.0, 0.0, 800.0),
     ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic"), 100.0),
    (fv.Grid(13, 7, 5, 0.0, 1300.0, -700.0, 700.0, 0.0, 500.0), ("reflecting",) * 6, 0.0),
    (fv.Grid(40, 9, 1, 0.0, 4000.0, -900.0, 900.0, 0.0, 100.0),
     ("reflecting", "outflow", "periodic", "periodic", "periodic", "periodic"), 50.0),
    (fv.Grid(70, 6, 4, 0.0, 7000.0, 0.0, 600.0, 0.0, 400.0),
     ("reflecting", "reflecting", "reflecting", "reflecting", "periodic", "periodic"), 0.0),
]


def _prepare(g, bcs, heat, members=1):
    from snapy_amd import MeshBlock

    blk = _block(g, bcs, "float32", grav1=-9.81, scheme=1, heat=heat)
    if members > 1:
        blk = MeshBlock(blk.options, members=members)
    blk.to("cuda")
    w = stratified(g, seed=5, noise=0.5)
    if members > 1:
        # the members stacked along x3, each with its own velocities
        wm = [torch.from_numpy(w).cuda() for _ in range(members)]
        for m, x in enumerate(wm):
            x[1:4] += 0.1 * (m + 1)
        vars, _ = blk.initialize({"hydro_w": torch.cat(wm, dim=1)})
    else:
        vars, _ = blk.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    assert blk._increment_in_x1_sweep()
    # one step, then stage 1's primitives with their ghosts
    dt = blk.device_time_step(vars)
    for s in range(3):
        blk.forward(vars, dt, s)
    dt = blk.device_time_step(vars)
    blk.forward(vars, dt, 0)
    calls = dict(blk._stage_calls(vars, 1))
    _capi.check(calls["apply_boundaries"](), "apply_boundaries")
    for d in (2, 3):
        if f"sweep_x{d}" in calls:
            _capi.check(calls[f"sweep_x{d}"](), f"sweep_x{d}")
    torch.cuda.synchronize()
    return blk, vars


def _both(blk, vars):
    lib = _capi.load()
    g, p, P = ctypes.byref(blk._grid), ctypes.byref(blk._phys), _capi.ptr
    freg = blk._freg
    w = vars["hydro_w"]
    out = []
    for fused in (False, True):
        f1 = torch.full_like(freg.total(1), float("nan"))
        fp = torch.full_like(freg.face_pressure1, float("nan"))
        ref = torch.full_like(blk._ref, float("nan"))
        du = torch.full_like(blk._vic["du"], float("nan"))
        f2, f3 = freg.total(2), freg.total(3)
        if fused:
            _capi.check(
                lib.snapy_sweep_x1_wb_inc_f32(g, p, P(w), P(ref), P(f1), P(fp), P(f2), P(f3), P(blk._x1tab), P(du),
                                              0.0, P(blk._dt_dev), None),
                "sweep_x1_wb_inc_f32",
            )
        else:
            _capi.check(lib.snapy_sweep_x1_wb_ref(g, p, 1, P(w), P(ref), P(f1), P(fp), None), "sweep_x1_wb_ref")
            fl = (ctypes.c_void_p * 3)(P(f1), P(f2), P(f3))
            _capi.check(
                lib.snapy_update_gravity_f32(g, p, fl, P(ref), None, None, None, None, P(du), P(w), P(blk._x1tab),
                                             None, 0.0, P(blk._dt_dev), 1.0, 0.0, 1.0, None),
                "update_gravity_f32",
            )
        torch.cuda.synchronize()
        out.append((f1, fp, ref[0], du))
    return out


def _same(a, b):
    return torch.equal(torch.nan_to_num(a, nan=7.0), torch.nan_to_num(b, nan=7.0))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fused_increment_is_the_separate_launches_bits(case):
    g, bcs, heat = CASES[case]
    print(f"x1 increment case {case}: {g.nx1}x{g.nx2}x{g.nx3} {bcs[:2]} heat {heat}", flush=True)
    blk, vars = _prepare(g, bcs, heat)
    (a, b) = _both(blk, vars)
    inner = blk.part(exterior=False)
    for name, x, y in zip(("F1", "face pressure", "psf_lo", "du"), a, b):
        if name == "du":
            x, y = x[inner], y[inner]
        assert _same(x, y), name


def test_fused_increment_member_axis():
    g, bcs, heat = CASES[2]
    print("x1 increment member axis", flush=True)
    blk, vars = _prepare(g, bcs, heat, members=3)
    (a, b) = _both(blk, vars)
    inner = blk.part(exterior=False)
    for name, x, y in zip(("F1", "face pressure", "psf_lo", "du"), a, b):
        if name == "du":
            x, y = x[inner], y[inner]
        assert _same(x, y), name
