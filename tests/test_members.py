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
