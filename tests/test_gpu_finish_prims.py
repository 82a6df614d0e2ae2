"""The FP32 state form with the partial implicit solve (scheme 1): the
solve's finish writes the primitives of the combined pair into the interior
of hydro_w (snapy_vic_partial_prim_f32), so stages 1 and 2 start with the
ghost fill alone and the refresh between steps fills the ghosts.  The
primitives must be those cons2prim_f32 forms from the pair, bit for bit, at
the entry (against snapy_vic_partial_f32 + snapy_cons2prim_f32 on the same
inputs) and after whole steps; a dt after a step is that of a full refresh;
the block's state is that of the oracle within the FP32 bar."""

import ctypes

import pytest
import torch

from oracle import fv
from snapy_amd import _capi
from tests._gpu_util import host, rel_l2
from tests.test_gpu_fp32 import _block, _errors
from tests.test_gpu_gravity import GAMMA, stratified

pytestmark = pytest.mark.gpu

CONV = fv.Grid(16, 8, 8, 0.0, 800.0, 0.0, 800.0, 0.0, 800.0)
CONV_BCS = ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic")


def _conv_block(execution="eager"):
    from snapy_amd.meshblock import StageExecutor

    blk = _block(CONV, CONV_BCS, "float32", grav1=-9.81, scheme=1, heat=100.0)
    blk.options.intg.execution = execution
    blk._executor = StageExecutor(blk, execution)
    blk.to("cuda")
    w = stratified(CONV, seed=9, noise=0.5)
    vars, _ = blk.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    return blk, vars, w


def _prims_of_pair(blk, vars):
    lib = _capi.load()
    w2 = torch.full_like(vars["hydro_w"], float("nan"))
    _capi.check(
        lib.snapy_cons2prim_f32(ctypes.byref(blk._grid), ctypes.byref(blk._phys), _capi.ptr(vars["hydro_u"]),
                                _capi.ptr(blk._state.comp()), _capi.ptr(w2), None),
        "cons2prim_f32",
    )
    return w2


@pytest.mark.parametrize("execution", ["eager", "hip-graph"])
def test_interior_primitives_after_steps_are_cons2prim_of_the_pair(execution):
    blk, vars, _ = _conv_block(execution)
    assert blk._prims_from_finish()
    inner = blk.part(exterior=False)
    for n in range(3):
        dt = blk.device_time_step(vars)
        for s in range(3):
            blk.forward(vars, dt, s)
            torch.cuda.synchronize()
            w_ref = _prims_of_pair(blk, vars)
            print(f"{execution} step {n} stage {s}", flush=True)
            assert torch.equal(vars["hydro_w"][inner], w_ref[inner]), (n, s)
    # the dt of the ghost-only refresh equals that of a full refresh
    dt_fast = float(blk.max_time_step(vars))
    blk._w_stale, blk._w_interior_fresh = True, False
    dt_full = float(blk.max_time_step(vars))
    assert dt_fast == dt_full


def test_entry_equals_the_plain_finish_and_cons2prim():
    blk, vars, _ = _conv_block()
    dt = blk.device_time_step(vars)
    blk.forward(vars, dt, 0)
    # stage 1 up to its vertical solve, then the solve twice from the same
    # inputs: the plain FP32 finish + cons2prim_f32, and the prim entry
    calls = blk._stage_calls(vars, 1)
    names = [n for n, _ in calls]
    vi = names.index(blk._vic["name"])
    for name, fn in calls[:vi]:
        _capi.check(fn(), name)
    lib = _capi.load()
    g, p, P = ctypes.byref(blk._grid), ctypes.byref(blk._phys), _capi.ptr
    u, c, w = vars["hydro_u"], blk._state.comp(), vars["hydro_w"]
    u0, c0 = blk._state.u0(), blk._state.comp0()
    vic, st = blk._vic, blk.intg.stages[1]
    u_in, c_in, w_in = u.clone(), c.clone(), w.clone()
    outs = []
    for entry in (lib.snapy_vic_partial_f32, lib.snapy_vic_partial_prim_f32):
        u.copy_(u_in)
        c.copy_(c_in)
        w.copy_(w_in)
        _capi.check(
            entry(g, p, P(w), P(vic["du"]), P(u0), P(c0), P(u), P(c), P(vic["tab"]), P(blk._x1tab), vic["t0"],
                  vic["t1"], 0.0, P(blk._dt_dev), st.wght2(), st.wght0(), st.wght1(), st.wght2(), P(vic["ws"]),
                  P(vic["status"]), P(vic["moved"]), None),
            "vic",
        )
        torch.cuda.synchronize()
        outs.append((u.clone(), c.clone(), w.clone()))
    (u1, c1, w1), (u2, c2, w2) = outs
    assert torch.equal(u1, u2) and torch.equal(c1, c2)
    inner = blk.part(exterior=False)
    w_ref = _prims_of_pair(blk, {"hydro_u": u2, "hydro_w": w2})
    assert torch.equal(w2[inner], w_ref[inner])
    # the ghosts are left as they were
    mask = torch.ones_like(w2, dtype=torch.bool)
    mask[inner] = False
    assert torch.equal(w2[mask], w_in[mask])


def test_block_against_oracle_within_fp32_bar():
    blk, vars, w = _conv_block()
    d = fv.Deck(grid=CONV, gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=CONV_BCS,
                grav1=-9.81, cfl=0.9, implicit_scheme=1, weight=0.02897025302492418, flux_covariance=True,
                bot_heat_flux=100.0)
    ref = fv.Block(d)
    ref.initialize(w)
    for n in range(3):
        dt = ref.max_time_step()
        for s in range(3):
            blk.forward(vars, dt, s)
            ref.forward(dt, s)
        assert blk.check_redo(vars) == 0
    u = host(blk.total_u(vars))
    errs = _errors(u, ref.u, CONV)
    print("errors", errs, flush=True)
    assert max(errs) < 1e-5
