"""A block of members (MeshBlock(options, members=M)) on the GPU: M
same-shape 2D density currents (different cold bubbles) stacked along x3
and advanced by one launch per kernel, each with its own dt, are bit for bit
the same M blocks stepped alone -- both state forms, both implicit schemes,
eager and captured -- and so are their time steps; shock tubes likewise."""

import pytest
import torch

pytestmark = pytest.mark.gpu

DTS = (-15.0, -5.0, -25.0, -10.0)


def _dc(dT, members=1, precision="float64", scheme=9, execution="eager", nx=64, nz=32):
    from snapy_amd import problems, serving
    from snapy_amd.meshblock import MeshBlock

    p = {"nx": nx, "nz": nz, "dT": dT, "precision": precision, "execution": execution}
    op = serving._options("density_current", p)
    op.intg.implicit_scheme = scheme
    blk = MeshBlock(op, members=members)
    blk.to("cuda")
    if members == 1:
        w = problems.density_current(blk, dT=dT, grav=serving.GRAV)
    else:
        w = torch.cat([problems.density_current(blk, dT=d, grav=serving.GRAV)[:, i : i + 1]
                       for i, d in enumerate(DTS[:members])], dim=1)
    vars, _ = blk.initialize({"hydro_w": w.cuda()})
    return blk, vars


def _step(blk, vars, steps):
    for _ in range(steps):
        dt = blk.device_time_step(vars)
        for s in range(3):
            blk.forward(vars, dt, s)


def _compare(solo, batch, m):
    (b0, v0), (b1, v1) = solo, batch
    assert torch.equal(v0["hydro_u"][:, 0], v1["hydro_u"][:, m]), m
    c0, c1 = b0._state.comp(), b1._state.comp()
    if c0 is not None:
        assert torch.equal(c0[:, 0], c1[:, m]), m
    assert torch.equal(b0._dt_dev.reshape(()), b1._dt_dev[m]), m


@pytest.mark.parametrize("precision,scheme,execution", [
    ("float64", 9, "eager"),
    ("float64", 1, "eager"),
    ("float32", 9, "eager"),
    ("float32", 1, "hip-graph"),
    ("float64", 9, "hip-graph"),
])
def test_members_are_bitwise_the_solo_blocks(precision, scheme, execution):
    M, steps = len(DTS), 4
    print(f"members: density current x{M}, {precision}, scheme {scheme}, {execution}", flush=True)
    batch = _dc(0.0, members=M, precision=precision, scheme=scheme, execution=execution)
    solo = [_dc(d, precision=precision, scheme=scheme, execution=execution) for d in DTS]
    assert batch[0].shape() == (5, M, batch[0].options.coord.nc2(), batch[0].options.coord.nc1())
    # the initial time steps differ by member (different bubbles) and match
    dts = batch[0].max_time_step(batch[1])
    assert dts == [b.max_time_step(v) for b, v in solo]
    _step(*batch, steps)
    for b, v in solo:
        _step(b, v, steps)
    torch.cuda.synchronize()
    for m in range(M):
        _compare(solo[m], batch, m)
    assert batch[0].max_time_step(batch[1]) == [b.max_time_step(v) for b, v in solo]
    assert batch[0].check_redo(batch[1]) == 0
    assert torch.isfinite(batch[1]["hydro_u"]).all()
    # the batch budget is the members' budgets together, to rounding of each
    # row's magnitude (the momentum totals cancel to near zero)
    tot = sum(b.budget(v) for b, v in solo)
    c = batch[0].options.coord
    inner = batch[0].part(exterior=False)
    scale = batch[0].total_u(batch[1])[inner].abs().sum(dim=(1, 2, 3)) * (c.dx(1) * c.dx(2) * c.dx(3))
    assert (batch[0].budget(batch[1]) - tot).abs().le(1e-13 * scale).all()


def test_member_tubes_are_bitwise_the_solo_tubes():
    """1D shock tubes (x2 and the deck's x3 single cells) with HLLC: the
    member axis of a block with no x2 either."""
    from snapy_amd import problems, serving
    from snapy_amd.meshblock import MeshBlock

    print("members: 3 Sod tubes, 300 cells", flush=True)
    states = [((1.0, 0.0, 1.0), (0.125, 0.0, 0.1)), ((0.5, 0.2, 2.0), (0.2, 0.0, 0.3)),
              ((0.9, -0.1, 0.4), (1.0, 0.0, 1.5))]

    def tube(members):
        op = serving._options("sod", {"n": 300, "execution": "eager"})
        blk = MeshBlock(op, members=members)
        blk.to("cuda")
        return blk

    solo = []
    for l, r in states:
        b = tube(1)
        v, _ = b.initialize({"hydro_w": problems.sod(b, left=l, right=r).cuda()})
        solo.append((b, v))
    bb = tube(len(states))
    w = torch.cat([problems.sod(bb, left=l, right=r)[:, i : i + 1] for i, (l, r) in enumerate(states)], dim=1)
    bv, _ = bb.initialize({"hydro_w": w.cuda()})
    for b, v in solo:
        _step(b, v, 10)
    _step(bb, bv, 10)
    torch.cuda.synchronize()
    for m in range(len(states)):
        _compare(solo[m], (bb, bv), m)


def test_member_dt_from_host_values():
    """forward takes one host dt per member (a list) or one for all."""
    print("members: host dt forms", flush=True)
    b1, v1 = _dc(0.0, members=2)
    b2, v2 = _dc(0.0, members=2)
    dts = b1.max_time_step(v1)
    for s in range(3):
        b1.forward(v1, dts, s)
        b2.forward(v2, torch.tensor(dts, dtype=torch.float64, device="cuda"), s)
    torch.cuda.synchronize()
    assert torch.equal(v1["hydro_u"], v2["hydro_u"])
    b3, v3 = _dc(0.0, members=2)
    b4, v4 = _dc(0.0, members=2)
    for s in range(3):
        b3.forward(v3, min(dts), s)
        b4.forward(v4, [min(dts)] * 2, s)
    torch.cuda.synchronize()
    assert torch.equal(v3["hydro_u"], v4["hydro_u"])
