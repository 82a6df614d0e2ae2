"""The host side of a block of members (no GPU): what the constructor
accepts and the shapes and descriptor it sets."""

import pytest

from snapy_amd import serving
from snapy_amd.meshblock import MeshBlock


def _op(kind="density_current", **p):
    base = {"nx": 64, "nz": 32, "dT": -15.0, "n": 100}
    base.update(p)
    return serving._options(kind, base)


def test_member_descriptor_and_shapes():
    op = _op()
    blk = MeshBlock(op, members=6)
    c = op.coord
    assert blk.shape() == (5, 6, c.nc2(), c.nc1())
    g = blk._grid
    assert (g.nx3, g.nc3, g.members3) == (6, 6, 6)
    # the interior slice covers every member along x3
    inner = blk.part(exterior=False)
    assert inner[1] == slice(0, 6)
    assert inner[2] == slice(c.nghost, c.nghost + c.nx2)
    solo = MeshBlock(op)
    assert solo.shape() == (5, 1, c.nc2(), c.nc1())
    assert (solo._grid.nx3, solo._grid.members3) == (1, 0)


def test_member_refusals():
    with pytest.raises(ValueError, match="members >= 1"):
        MeshBlock(_op(), members=0)
    op = _op()
    op.coord.nx3 = 4
    with pytest.raises(ValueError, match="x3 must be a single cell"):
        MeshBlock(op, members=2)
    swe = _op("swe", n=16).block
    swe.coord.nx3 = 1
    with pytest.raises(ValueError, match="ideal-gas"):
        MeshBlock(swe, members=2)


def test_part_offsets_follow_the_reference():
    """MeshBlockImpl::part (meshblock.cpp:242-284): offset 0 is the interior
    (widened by extend_xd on both sides) whatever exterior says; -1 / +1 are
    the ghost layers (exterior) or the interior layers next to them."""
    from snapy_amd.options import CoordinateOptions, HydroOptions, MeshBlockOptions

    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 10, 6, 1, 3)
    op.hydro = HydroOptions(riemann="hllc")
    op.bcnames = ["outflow"] * 6
    b = MeshBlock(op)
    ng = 3
    assert b.part((0, 0, 0), True) == (slice(None), slice(0, 1), slice(ng, ng + 6), slice(ng, ng + 10))
    assert b.part((0, 0, 0), False) == b.part((0, 0, 0), True)
    assert b.part((0, 0, 0), False, extend_x1=1)[3] == slice(ng - 1, ng + 10 + 1)
    assert b.part((0, 0, 0), False, extend_x2=2)[2] == slice(ng - 2, ng + 6 + 2)
    # x1 outer ghosts, x1 inner ghosts, and the interior layers beside them
    assert b.part((0, 0, 1), True)[3] == slice(ng + 10, ng + 10 + ng)
    assert b.part((0, 0, -1), True)[3] == slice(0, ng)
    assert b.part((0, 0, 1), False)[3] == slice(10, 10 + ng)
    assert b.part((0, 0, -1), False)[3] == slice(ng, 2 * ng)
    assert b.part((0, 1, 0), True)[2] == slice(ng + 6, ng + 6 + ng)
    # the x3 axis has one cell and no ghosts: its +-1 parts are empty
    assert b.part((1, 0, 0), True)[1] == slice(1, 1)
    assert b.part((-1, 0, 0), False)[1] == slice(0, 0)
    with pytest.raises(ValueError, match="offset"):
        b.part((0, 0, 2))
    # a member axis: every member at offset 0, nothing beyond
    m = MeshBlock(_op(), members=4)
    assert m.part((0, 0, 0))[1] == slice(0, 4)
    assert m.part((1, 0, 0))[1] == slice(0, 0)
