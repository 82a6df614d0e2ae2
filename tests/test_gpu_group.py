"""Grouped launches on the GPU: density-current blocks of different widths
and heights (one deck) stepped by one launch per kernel give each block the
same bits as its own steps, in both state forms and both implicit schemes;
plans that do not match are refused."""

import ctypes

import pytest
import torch

from snapy_amd import _capi, problems, serving
from snapy_amd.group import BlockGroup, group_key
from snapy_amd.meshblock import MeshBlock

pytestmark = pytest.mark.gpu

SIZES = [(48, 16, -15.0), (96, 32, -9.0), (64, 24, -19.0)]


def _block(nx, nz, dT, precision, scheme):
    op = serving._options("density_current", {"nx": nx, "nz": nz, "dT": dT, "precision": precision,
                                             "execution": "eager"})
    op.intg.implicit_scheme = scheme
    blk = MeshBlock(op)
    blk.to("cuda")
    vars, _ = blk.initialize({"hydro_w": problems.density_current(blk, dT=dT, grav=serving.GRAV).cuda()})
    return blk, vars


def _state(blk, vars):
    out = [vars["hydro_u"].clone()]
    if blk._state.comp() is not None:
        out.append(blk._state.comp().clone())
    return out


@pytest.mark.parametrize("precision,scheme", [("float64", 9), ("float32", 9), ("float64", 1), ("float32", 1)])
def test_grouped_steps_equal_each_blocks_own(precision, scheme):
    print(f"group of {SIZES}, {precision}, scheme {scheme}", flush=True)
    nsteps = 4
    solo = []
    for nx, nz, dT in SIZES:
        blk, vars = _block(nx, nz, dT, precision, scheme)
        for _ in range(nsteps):
            dt = blk.device_time_step(vars)
            for s in range(3):
                blk.forward(vars, dt, s)
        torch.cuda.synchronize()
        solo.append(_state(blk, vars))
        blk.release()
    members = [_block(nx, nz, dT, precision, scheme) for nx, nz, dT in SIZES]
    grp = BlockGroup(members)
    assert grp.launches(0) > 5
    for _ in range(nsteps):
        grp.step()
    torch.cuda.synchronize()
    for (blk, vars), ref in zip(members, solo):
        got = _state(blk, vars)
        assert torch.isfinite(got[0]).all()
        for a, b in zip(got, ref):
            assert torch.equal(a, b), (blk.options.coord.nx2, precision, scheme)
    grp.release()


def test_groups_refuse_blocks_that_launch_differently():
    a = _block(48, 16, -15.0, "float64", 9)
    b = _block(48, 16, -15.0, "float64", 1)  # the other implicit scheme: other kernels
    assert group_key(a[0]) != group_key(b[0])
    with pytest.raises(ValueError, match="share a deck"):
        BlockGroup([a, b])
    # the library refuses the plans themselves as well
    lib = _capi.load()
    plans = []
    for blk, vars in (a, b):
        assert lib.snapy_record_begin() == 0
        for name, fn in blk._stage_calls(vars, 0):
            if name != "vic_status_clear":
                _capi.check(fn(), name)
        p = ctypes.c_void_p()
        assert lib.snapy_record_end(ctypes.byref(p)) == 0
        assert lib.snapy_plan_launches(p) > 5
        plans.append(p)
    arr = (ctypes.c_void_p * 2)(*[p.value for p in plans])
    g = ctypes.c_void_p()
    rc = lib.snapy_group_build(arr, 2, ctypes.byref(g))
    for p in plans:
        lib.snapy_plan_free(p)
    assert rc == -21
    assert b"differ" in lib.snapy_last_error()
