"""The autotuner on the GPU: it times every candidate (workgroup size,
columns per workgroup) of every kernel family on a block, keeps a valid
choice, and a block run with the tuned shapes gives the same bits as one
run with the default shapes (a shape never changes a cell's arithmetic)."""

import os

import pytest
import torch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _block(precision):
    from snapy_amd import MeshBlock, MeshBlockOptions, problems

    op = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "convection3d.yaml"))
    op.coord.nx1, op.coord.nx2, op.coord.nx3 = 32, 24, 16
    op.coord.x1max, op.coord.x2max, op.coord.x3max = 1600.0, 2400.0, 1600.0
    op.intg.execution = "eager"
    op.hydro.precision = precision
    blk = MeshBlock(op)
    blk.to("cuda")
    vars, _ = blk.initialize({"hydro_w": problems.convection3d(blk).cuda()})
    return blk, vars


def _steps(blk, vars, n=2):
    for _ in range(n):
        dt = blk.max_time_step(vars)
        for s in range(3):
            blk.forward(vars, dt, s)
    torch.cuda.synchronize()
    return blk.total_u(vars).clone()


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_tuned_shapes_give_the_same_bits(precision, tmp_path):
    from snapy_amd import autotune

    print(f"autotune {precision}", flush=True)
    ref_blk, ref_vars = _block(precision)
    u_ref = _steps(ref_blk, ref_vars)
    blk, vars = _block(precision)
    res = autotune.tune_block(blk, vars, reps=3, cache=tmp_path / "tune.json")
    # the FP32 form forms the stage increment inside the x1 sweep's launch,
    # so it has no launch of the update family
    fams = {"eos", "sweep", "ref_cols"} | ({"update"} if precision == "float64" else set())
    assert set(res) == fams
    for fam, r in res.items():
        assert r["choice"] in autotune.FAMILIES[fam][2]
        assert autotune.get_choice(blk, fam) == r["choice"]
    # force a non-default shape on every family (the smallest candidate),
    # then step from a fresh state
    for fam in res:
        autotune.set_choice(blk, fam, autotune.FAMILIES[fam][2][0])
    # and the scheme-1 finish's column count, set by hand
    autotune.set_choice(blk, "vic_cols", 1)
    blk2, vars2 = _block(precision)
    u = _steps(blk2, vars2)
    assert torch.equal(u, u_ref)
    autotune.set_choice(blk, "vic_cols", 0)
    assert autotune.apply_cache(blk2, tmp_path / "tune.json") == {f: r["choice"] for f, r in res.items()}
    for fam in res:
        autotune.set_choice(blk, fam, 0)


def test_work_list_grid():
    from snapy_amd import autotune, serving
    from snapy_amd.meshblock import MeshBlock
    from snapy_amd.worklist import BlockBatch

    members = []
    for i, n in enumerate((100, 3000, 257)):
        op = serving._options("sod", {"n": n})
        blk = MeshBlock(op)
        blk.to("cuda")
        vars, _ = blk.initialize({"hydro_w": serving._initial_state("sod", blk, {"n": n}, i).cuda()})
        members.append((blk, vars))
    batch = BlockBatch(members)
    r = autotune.tune_work_list(batch, reps=3)
    assert r["choice"] in autotune.WL_GRIDS
    _capi_reset()


def _capi_reset():
    from snapy_amd import _capi

    _capi.check(_capi.load().snapy_tune_set(3, None, 0, 0), "tune_set")
