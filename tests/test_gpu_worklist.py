"""The work list (snapy_amd/worklist.py, csrc/worklist.hip) on the GPU:
shock tubes of uneven size and every x1 boundary kind, advanced together by
one launch per phase, are bit-identical to the same tubes stepped alone
through MeshBlock.device_time_step / forward; a tube can leave and another
join the batch between steps."""

import pytest
import torch

pytestmark = pytest.mark.gpu

TUBES = [(37, "outflow"), (128, "reflecting"), (300, "periodic"), (1000, "extrapolation"),
         (4099, "outflow"), (64, "outflow")]


def _tube(n, seed, bc):
    from snapy_amd import serving
    from snapy_amd.meshblock import MeshBlock
    from snapy_amd.options import IntegratorOptions

    op = serving._options("sod", {"n": n})
    op.bcnames = [bc] * 6
    op.intg = IntegratorOptions(cfl=0.9, execution="eager")
    blk = MeshBlock(op)
    blk.to("cuda")
    w = serving._initial_state("sod", blk, {"n": n}, seed)
    vars, _ = blk.initialize({"hydro_w": w.cuda()})
    return blk, vars


def _solo_step(blk, vars):
    dt = blk.device_time_step(vars)
    for s in range(3):
        blk.forward(vars, dt, s)


def _assert_same(solo, batched):
    for (b0, v0), (b1, v1) in zip(solo, batched):
        assert torch.equal(v0["hydro_u"], v1["hydro_u"]), b0.options.coord.nx1
        assert torch.equal(b0._dt_dev, b1._dt_dev)


def test_batch_is_bitwise_the_solo_blocks():
    from snapy_amd.worklist import BlockBatch

    print("work list: 6 tubes, 8 steps", flush=True)
    solo = [_tube(n, 100 + i, bc) for i, (n, bc) in enumerate(TUBES)]
    batched = [_tube(n, 100 + i, bc) for i, (n, bc) in enumerate(TUBES)]
    batch = BlockBatch(batched)
    assert batch.launches_per_step() == 15
    for _ in range(8):
        for blk, vars in solo:
            _solo_step(blk, vars)
        batch.step()
    torch.cuda.synchronize()
    _assert_same(solo, batched)
    for _, vars in batched:
        assert torch.isfinite(vars["hydro_u"]).all()


def test_members_leave_and_join_between_steps():
    from snapy_amd.worklist import BlockBatch

    print("work list: membership change", flush=True)
    solo = [_tube(n, 200 + i, bc) for i, (n, bc) in enumerate(TUBES[:4])]
    batched = [_tube(n, 200 + i, bc) for i, (n, bc) in enumerate(TUBES[:4])]
    batch = BlockBatch(batched)
    for _ in range(3):
        for blk, vars in solo:
            _solo_step(blk, vars)
        batch.step()
    # the second tube leaves, a new one joins
    solo.pop(1)
    batched.pop(1)
    solo.append(_tube(777, 300, "outflow"))
    batched.append(_tube(777, 300, "outflow"))
    batch = BlockBatch(batched)
    for _ in range(4):
        for blk, vars in solo:
            _solo_step(blk, vars)
        batch.step()
    torch.cuda.synchronize()
    _assert_same(solo, batched)


def test_ineligible_blocks_are_refused():
    from snapy_amd import serving
    from snapy_amd.meshblock import MeshBlock
    from snapy_amd.worklist import BlockBatch, eligible

    op = serving._options("density_current", {"nx": 16, "nz": 8, "dT": -10.0})
    blk = MeshBlock(op)
    blk.to("cuda")
    assert not eligible(blk)
    with pytest.raises(ValueError):
        BlockBatch([(blk, {"hydro_u": None})])
