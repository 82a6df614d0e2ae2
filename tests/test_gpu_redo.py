"""The failure and retry paths and the step-level contracts of the block on
the GPU:

* the vertical implicit solve's column status words (bit 0: this call
  failed; bit 1: a call since the last clear failed), for both schemes;
* check_redo (src/mesh/meshblock_redo.cpp:32-89,170-206): a NaN injected
  mid-step asks for a redo, the state comes back bit for bit from the step's
  backup, the primitives are those of the restored state and the next dt
  is half the step's; a failed implicit column alone (the step latch) asks
  for a redo as well; past max_redo a floor/NaN cause stops at once;
* the FluxRegister step fields, S_d <- w1 S_d + w2 dt T_d per stage, against
  the weighted stage sum of the reference's own test
  (tests/test_flux_register.cpp:211-248), and the kernel alone against the
  oracle's fold;
* graph == eager bit for bit for the Cartesian block (the reference's
  tests/test_stage_executor_capture.py:67-105), both precisions and both
  implicit schemes."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import fv, implicit
from snapy_amd import _capi
from snapy_amd.geometry import vic_table, x1_gravity_table
from tests._gpu_util import dev, grid_desc, host, phys_desc, rel_l2
from tests.test_gpu_fp32 import _block
from tests.test_gpu_gravity import GAMMA, stratified

pytestmark = pytest.mark.gpu

CONV = fv.Grid(16, 8, 8, 0.0, 800.0, 0.0, 800.0, 0.0, 800.0)
CONV_BCS = ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic")


# ---------------------------------------------------------------------------
# the implicit solve's status words
# ---------------------------------------------------------------------------
def _vic_call(lib, scheme, g, w, du, st, dt=0.7):
    pd = phys_desc(gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=("reflecting",) * 6,
                   grav1=-9.81)
    pd.implicit_scheme = scheme
    gd = grid_desc(g)
    x1f, x1v = torch.from_numpy(g.faces(1)), torch.from_numpy(g.centers(1))
    a1, vol = g.dx(3) * g.dx(2), g.dx(3) * (g.dx(2) * g.dx(1))
    tab, t0, t1 = vic_table(x1f, x1v, 3, g.nx1, -9.81, a1, vol)
    xt = x1_gravity_table(x1f, x1v, 3, g.nx1, -9.81)
    wsize = lib.snapy_vic_partial_workspace if scheme == 1 else lib.snapy_vic_workspace
    entry = lib.snapy_vic_partial if scheme == 1 else lib.snapy_vic_full
    ws = torch.empty(int(wsize(ctypes.byref(gd))), dtype=torch.float64, device="cuda")
    u0 = torch.zeros(w.shape, dtype=torch.float64, device="cuda")
    u = torch.zeros_like(u0)
    # every operand stays referenced until the results are read back (a
    # temporary's memory would go back to the allocator at once)
    tabd, xtd, wd, dud = tab.cuda(), xt.cuda(), dev(w), dev(du)
    rc = entry(ctypes.byref(gd), ctypes.byref(pd), wd.data_ptr(), dud.data_ptr(), u0.data_ptr(),
               u.data_ptr(), tabd.data_ptr(), xtd.data_ptr(), t0, t1, dt, None, 1.0, 0.0, 0.0, 1.0, ws.data_ptr(),
               st.data_ptr(), None, None)
    assert rc == 0, lib.snapy_last_error()
    out = host(u)
    del wd, dud, tabd, xtd, ws
    return out


@pytest.mark.parametrize("scheme", [9, 1])
def test_vic_status_marks_the_failed_column_and_latches_it(scheme):
    """A NaN in one column's explicit increment fails that column's solve:
    bits 0 and 1 of its word are set, the column keeps its explicit
    increment, the other columns are solved as the oracle solves them.  A
    second call with a clean increment clears bit 0 and keeps bit 1."""
    lib = _capi.load()
    g = fv.Grid(16, 6, 1, 0.0, 800.0, -400.0, 400.0, 0.0, 300.0)
    w = stratified(g, seed=5, noise=0.5)
    rng = np.random.default_rng(9)
    du = np.zeros_like(w)
    sub = (slice(None),) + g.interior()
    du[sub] = 1e-3 * (rng.random(du[sub].shape) - 0.5) * np.array([1, 1, 1, 1, 1e3])[:, None, None, None]
    bad = du.copy()
    jbad = 3 + 2  # interior column 2
    bad[fv.IPR, 0, jbad, 3 + 7] = np.nan
    st = torch.zeros(g.nx2 * g.nx3, dtype=torch.int32, device="cuda")
    print(f"vic status scheme {scheme}: NaN in column 2", flush=True)
    got = _vic_call(lib, scheme, g, w, bad, st)
    words = host(st).tolist()
    assert words == [0, 0, 3, 0, 0, 0], words
    # the failed column: its explicit increment (u = w2 du with u0 = u = 0)
    np.testing.assert_array_equal(got[:, 0, jbad, 3:-3], bad[:, 0, jbad, 3:-3])
    # the others: the oracle's solve
    n = g.nx1
    wc = w[:, 0:1, 3:-3, 2 : 3 + n + 1]
    wcol = np.moveaxis(wc, 0, -1).reshape(g.nx2, n + 2, 5)
    dcol = np.moveaxis(du[:, 0:1, 3:-3, 3 : 3 + n], 0, -1).reshape(g.nx2, n, 5)
    a1, vol = g.dx(3) * g.dx(2), g.dx(3) * (g.dx(2) * g.dx(1))
    solve = implicit.vic_partial if scheme == 1 else implicit.vic_full
    ref = solve(dcol, wcol, GAMMA, -9.81, 0.7, g.dx(1), g.faces(1)[3 : 3 + n + 1], g.centers(1)[3 : 3 + n], a1, vol)
    gc = np.moveaxis(got[:, 0, 3:-3, 3 : 3 + n], 0, -1)
    keep = [c for c in range(g.nx2) if c != 2]
    for v in range(5):
        assert rel_l2(gc[keep][..., v], ref[keep][..., v]) < 1e-11
    # a clean call: bit 0 clears, the step latch (bit 1) stays
    print(f"vic status scheme {scheme}: clean call keeps the latch", flush=True)
    _vic_call(lib, scheme, g, w, du, st)
    assert host(st).tolist() == [0, 0, 2, 0, 0, 0]


# ---------------------------------------------------------------------------
# check_redo
# ---------------------------------------------------------------------------
def _conv_block(precision, scheme, execution="eager"):
    blk = _block(CONV, CONV_BCS, precision, grav1=-9.81, scheme=scheme, heat=100.0)
    blk.options.intg.execution = execution
    from snapy_amd.meshblock import StageExecutor

    blk._executor = StageExecutor(blk, execution)
    blk.to("cuda")
    w = stratified(CONV, seed=4, noise=0.5)
    vars, _ = blk.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    return blk, vars


def _snapshot(blk, vars):
    torch.cuda.synchronize()
    c = blk._state.comp()
    return (vars["hydro_u"].clone(), None if c is None else c.clone(), vars["hydro_w"].clone())


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("execution", ["eager", "hip-graph"])
def test_nan_mid_step_redoes_from_the_backup(precision, execution):
    blk, vars = _conv_block(precision, 1, execution)
    print(f"redo: NaN mid-step {precision} {execution}", flush=True)
    # two accepted steps first (graph mode captures its stages)
    for _ in range(2):
        dt = blk.max_time_step(vars)
        for s in range(3):
            blk.forward(vars, dt, s)
        assert blk.check_redo(vars) == 0
    dt = blk.max_time_step(vars)
    u_start, c_start, w_start = _snapshot(blk, vars)
    blk.forward(vars, dt, 0)
    vars["hydro_u"][fv.IDN, 5, 6, 9] = float("nan")
    for s in (1, 2):
        blk.forward(vars, dt, s)
    rc = blk.check_redo(vars)
    causes = blk.last_redo_causes
    print("check_redo", rc, "causes", causes, flush=True)
    assert rc == 1
    assert causes & blk.REDO_NAN and causes & blk.REDO_VIC
    assert blk.intg.current_redo == 1
    torch.cuda.synchronize()
    assert torch.equal(vars["hydro_u"], u_start)
    if c_start is not None:
        assert torch.equal(blk._state.comp(), c_start)
    # the primitives of the restored state: those the step started from
    assert torch.equal(vars["hydro_w"], w_start)
    dt2 = blk.max_time_step(vars)
    assert dt2 == 0.5 * dt
    for s in range(3):
        blk.forward(vars, dt2, s)
    assert blk.check_redo(vars) == 0
    assert blk.intg.current_redo == 0
    assert bool(torch.isfinite(vars["hydro_u"]).all())


@pytest.mark.parametrize("scheme", [9, 1])
def test_failed_implicit_column_alone_asks_for_a_redo(scheme):
    """The step latch of a column that failed in an earlier stage of the
    step (a later stage's success does not clear it) is a retry cause by
    itself; past max_redo the block restores the state and then stops."""
    blk, vars = _conv_block("float64", scheme)
    print(f"redo: latched implicit failure scheme {scheme}", flush=True)
    # one accepted step: hydro_w is prim(u) from then on (after initialize
    # it is the caller's primitives, which prim(prim2cons(w)) rounds)
    dt = blk.max_time_step(vars)
    for s in range(3):
        blk.forward(vars, dt, s)
    assert blk.check_redo(vars) == 0
    dt = blk.max_time_step(vars)
    u_start, _, w_start = _snapshot(blk, vars)
    for s in range(3):
        blk.forward(vars, dt, s)
    # as a column that failed at stage 0 and succeeded at stages 1 and 2
    # leaves it: bit 1 kept, bit 0 clear
    blk._vic["status"][7] = 2
    assert blk.check_redo(vars) == 1
    assert blk.last_redo_causes == blk.REDO_VIC
    torch.cuda.synchronize()
    assert torch.equal(vars["hydro_u"], u_start)
    assert torch.equal(vars["hydro_w"], w_start)
    # the next stage 0 clears the latch: the redone step is accepted
    dt2 = blk.max_time_step(vars)
    assert dt2 == 0.5 * dt
    for s in range(3):
        blk.forward(vars, dt2, s)
    assert int(blk._vic["status"].max().item()) == 0
    assert blk.check_redo(vars) == 0
    # past max_redo: an implicit failure restores first, then stops
    for s in range(3):
        blk.forward(vars, dt2, s)
    u_step, _, _ = _snapshot(blk, vars)
    blk.intg.current_redo = blk.intg.max_redo
    blk._vic["status"][0] = 2
    assert blk.check_redo(vars) == -1
    assert not torch.equal(vars["hydro_u"], u_step)  # restored to the step's start


def test_floor_or_nan_past_max_redo_stops_without_restoring():
    blk, vars = _conv_block("float64", 1)
    print("redo: NaN past max_redo", flush=True)
    dt = blk.max_time_step(vars)
    for s in range(3):
        blk.forward(vars, dt, s)
    vars["hydro_u"][fv.IPR, 4, 4, 8] = float("nan")
    u_bad = vars["hydro_u"].clone()
    blk.intg.current_redo = blk.intg.max_redo
    assert blk.check_redo(vars) == -1
    assert blk.last_redo_causes & blk.REDO_NAN and not blk.last_redo_causes & blk.REDO_VIC
    # the same bits, the NaN included: nothing was restored
    assert torch.equal(vars["hydro_u"].view(torch.int64), u_bad.view(torch.int64))


# ---------------------------------------------------------------------------
# step fields
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_step_fields_are_the_weighted_stage_sum(precision):
    g = fv.Grid(24, 16, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0)
    blk = _block(g, ("reflecting",) * 6, precision, grav1=-9.81, scheme=9)
    blk.to("cuda")
    w = stratified(g, seed=3, noise=0.5)
    vars, _ = blk.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    reg = blk.flux_register()
    reg.request_step_field(1)
    reg.request_step_field(2)
    stages = blk.intg.stages
    print(f"step fields {precision}", flush=True)
    worst = 0.0
    for step in range(6):
        dt = blk.max_time_step(vars)
        T = []
        for s in range(3):
            blk.forward(vars, dt, s)
            T.append([reg.total(d).double().clone() for d in (1, 2)])
        for d in range(2):
            ref = torch.zeros_like(T[0][d])
            for s in range(3):
                c = stages[s].wght2()
                for t in range(s + 1, 3):
                    c *= stages[t].wght1()
                ref += (c * dt) * T[s][d]
            S = reg.step(d + 1)
            assert S.dtype == torch.float64
            rel = float((S - ref).abs().max() / ref.abs().max())
            worst = max(worst, rel)
            assert rel <= 1e-15, (step, d, rel)
        assert blk.check_redo(vars) == 0
    print("worst", worst, flush=True)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_step_field_kernel_is_the_fold(dtype):
    """snapy_step_field alone, with the dt from the device, against the
    oracle's fold (FluxRegister::end_stage), bit for bit."""
    lib = _capi.load()
    g = fv.Grid(12, 10, 6)
    rng = np.random.default_rng(5)
    T = rng.standard_normal((5,) + g.shape).astype(np.float32 if dtype == torch.float32 else np.float64)
    S = rng.standard_normal((5,) + g.shape)
    dt = 0.37
    Td, Sd = dev(T, dtype), dev(S)
    dtd = torch.tensor(dt, dtype=torch.float64, device="cuda")
    print(f"step field kernel {dtype}", flush=True)
    for w1, w2 in ((0.0, 1.0), (0.25, 0.25), (2.0 / 3.0, 2.0 / 3.0)):
        rc = lib.snapy_step_field(ctypes.byref(grid_desc(g)), _capi.dtype_code(dtype), Td.data_ptr(),
                                  Sd.data_ptr(), w1, w2, 0.0, dtd.data_ptr(), None)
        assert rc == 0, lib.snapy_last_error()
        S = fv.step_field_fold(S, T, w1, w2, dt)
        np.testing.assert_array_equal(host(Sd), S)


# ---------------------------------------------------------------------------
# graph == eager
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("scheme", [9, 1])
def test_graph_replay_is_eager_bit_for_bit(precision, scheme):
    outs = []
    for execution in ("eager", "hip-graph"):
        blk, vars = _conv_block(precision, scheme, execution)
        print(f"graph == eager {precision} scheme {scheme}: {execution}", flush=True)
        for _ in range(4):
            dt = blk.device_time_step(vars)
            for s in range(3):
                blk.forward(vars, dt, s)
        assert blk.check_redo(vars) == 0
        rep = blk.execution_report()
        if execution == "hip-graph":
            assert rep["replays"] >= 6 and rep["captures"] == 3
        outs.append(_snapshot(blk, vars))
    (ue, ce, we), (ug, cg, wg) = outs
    assert torch.equal(ue, ug)
    assert torch.equal(we, wg)
    if ce is not None:
        assert torch.equal(ce, cg)
