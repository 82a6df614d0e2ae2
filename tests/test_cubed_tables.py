"""The cubed-sphere mesh's host tables (snapy_amd/cubed.py) against the
oracle's restatement (oracle/cubed.py), bit for bit, on the CPU: the
gnomonic planes, the ghost source positions, the frame matrices, and the
ghost index maps applied on the host to a random state against the
oracle's literal serialize / flip / transpose / interpolate exchange."""

import numpy as np
import pytest

from oracle import cubed as oc
from snapy_amd import cubed as pc

IDN = 0


@pytest.mark.parametrize("n", [8, 13, 192])
def test_planes_are_the_oracles(n):
    g = oc.Grid(n, radius=6371000.0, depth=1.0)
    p = pc.PanelGrid(n, 6371000.0, 6371001.0)
    np.testing.assert_array_equal(p.x2f, g.x2f)
    assert p.x1v == g.x1v[0]
    for name_p, name_o in (("cos_c", "cos_c"), ("sin_c", "sin_c"), ("cos_f2", "cos_f2"), ("sin_f2", "sin_f2"),
                           ("cos_f3", "cos_f3"), ("sin_f3", "sin_f3"), ("area2", "area2"), ("area3", "area3"),
                           ("vol", "vol"), ("x_ov_rD", "x_ov_rD"), ("y_ov_rC", "y_ov_rC"),
                           ("width2", "width2"), ("width3", "width3")):
        np.testing.assert_array_equal(getattr(p, name_p), getattr(g, name_o)[..., 0], err_msg=name_p)


def test_ghost_source_positions_are_the_oracles():
    for n in (8, 16, 192):
        np.testing.assert_array_equal(pc.ghost_source_positions(n, 3), oc.ghost_usrc(n, 3))


@pytest.mark.parametrize("face", range(6))
def test_frame_matrices_are_the_oracles(face):
    g = pc.PanelGrid(10, 1.0, 2.0)
    X = np.broadcast_to(g.tx[None, :], (g.nc, g.nc))
    Y = np.broadcast_to(g.ty[:, None], (g.nc, g.nc))
    al = np.broadcast_to(g.x2v[None, :], (g.nc, g.nc))
    be = np.broadcast_to(g.x3v[:, None], (g.nc, g.nc))
    for to in (True, False):
        for cos in (None, g.cos_c):
            np.testing.assert_array_equal(pc.frame_matrix(face, X, Y, to, cos),
                                          oc.rotation_matrix(face, al, be, to, cos))


def _apply_ghost_tables(grid, tables, us, conserved):
    """The ghost kernel's arithmetic on the host, in the same order."""
    idx, w, mc, mp = tables
    mats = mc if conserved else mp
    nc = grid.nc
    flat = [u.reshape(4, nc * nc).copy() for u in us]
    src_all = [u.reshape(4, nc * nc) for u in us]

    def matvec(m, v):
        m = m.reshape(3, 3)
        return [m[i, 0] * v[0] + m[i, 1] * v[1] + m[i, 2] * v[2] for i in range(3)]

    for e in range(idx.shape[0]):
        p, dst, q, s0, s1, s2, s3, npts = (int(v) for v in idx[e])
        src = src_all[q]
        if npts == 1:
            flat[p][:, dst] = src[:, s0]
            continue
        b = []
        for t, s in enumerate((s0, s1, s2, s3)[:npts]):
            v = src[:, s]
            b.append([v[0]] + matvec(mats[e][9 * t : 9 * t + 9], v[1:]))
        if npts == 2:
            x = w[e, 0]
            g = [b[0][v] + x * (b[1][v] - b[0][v]) for v in range(4)]
        else:
            c0, c2, c3 = w[e]
            g = [b[1][v] + c0 * (b[0][v] - b[1][v]) + c2 * (b[2][v] - b[1][v]) + c3 * (b[3][v] - b[1][v])
                 for v in range(4)]
        o = matvec(mats[e][36:45], g[1:])
        flat[p][:, dst] = [g[0]] + o
    return [f.reshape(4, nc, nc, 1) for f in flat]


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("conserved", [False, True])
def test_ghost_index_maps_reproduce_the_exchange(order, conserved):
    n = 8
    grid = pc.PanelGrid(n, 6371000.0, 6371001.0)
    og = oc.Grid(n, interp_order=order)
    rng = np.random.default_rng(order + 10 * conserved)
    us = []
    for f in range(6):
        u = rng.normal(size=(4, grid.nc, grid.nc, 1))
        u[IDN] = 10.0 + u[IDN]
        us.append(u)
    ref = [u.copy() for u in us]
    oc.exchange_ghosts(og, ref, conserved)
    out = _apply_ghost_tables(grid, pc.build_ghost_tables(grid, order), us, conserved)
    mask = np.ones((grid.nc, grid.nc), bool)
    ng = grid.ng
    for a in (slice(0, ng), slice(ng + n, grid.nc)):
        for b in (slice(0, ng), slice(ng + n, grid.nc)):
            mask[a, b] = False  # corners are not exchanged
    for f in range(6):
        np.testing.assert_array_equal(out[f][:, mask], ref[f][:, mask])


def test_seam_table_pairs_every_edge_once():
    n = 6
    idx, dat = pc.build_seam_tables(pc.PanelGrid(n, 1.0, 2.0))
    assert idx.shape == (12 * n, 8) and dat.shape == (12 * n, 30)
    faces = set()
    for row in idx:
        p, s, i, q, qs, ib, sign, _ = (int(v) for v in row)
        assert p < q
        faces.add((p, s, i))
        faces.add((q, qs, ib))
    assert len(faces) == 24 * n  # every panel side's every face exactly once
