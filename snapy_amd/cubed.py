"""The cubed-sphere shallow-water mesh on one device.

``CubedSphereMesh`` is the reference's ``Mesh`` (src/mesh/mesh.hpp:73;
``Mesh::forward(MeshVariables&, dt, stage)``) for a cubed-sphere layout with
one block per panel, all six panels resident on the device.  Several
independent meshes (ensemble members) can share it: the state is one tensor
of ``members * 6`` blocks, block ``b = member * 6 + panel``, and every stage
kernel is one launch over all of them.

The grid-only tables are built here on the host, once, in FP64, with the
reference's arithmetic (GnomonicEquiangleImpl::reset,
src/coord/gnomonic_equiangle.cpp:37-250; PanelRotation::matrix,
src/coord/coord_rotation.cpp:108-150; cs_build_ghost_usrc,
src/coord/cubed_sphere_utils.cpp:82-113), and the exchanges are expressed as
index maps over them:

* per edge ghost cell: the source cells on the neighbour panel (the layer
  at the same depth, at the along-edge positions the ghost centre maps to),
  the interpolation weights, the sender's frame matrix at each source cell
  and the receiver's at the ghost cell;
* per seam face: the owner (the lower panel), the other panel's face, the
  sign and along-edge order of the connectivity, and the frame operands at
  both faces.

Transcendental functions are evaluated with the C library (``math``) on the
1D coordinates; everything else is IEEE arithmetic in numpy, so the tables
are reproducible.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _capi
from .meshblock import StageExecutor
from .options import MeshOptions

IDN, IVX, IVY, IVZ = 0, 1, 2, 3
NROWS = 4
SIDE_L, SIDE_R, SIDE_B, SIDE_T = 0, 1, 2, 3

# CS_FACE_EDGES (src/layout/cubed_sphere_layout.cpp:352-383): leaving panel f
# by side s lands on (panel, side, along-edge order reversed)
CS_EDGES = (
    ((4, SIDE_R, 0), (1, SIDE_L, 0), (5, SIDE_T, 0), (3, SIDE_B, 0)),
    ((0, SIDE_R, 0), (2, SIDE_L, 0), (5, SIDE_R, 1), (3, SIDE_R, 0)),
    ((1, SIDE_R, 0), (4, SIDE_L, 0), (5, SIDE_B, 1), (3, SIDE_T, 1)),
    ((4, SIDE_T, 1), (1, SIDE_T, 0), (0, SIDE_T, 0), (2, SIDE_T, 1)),
    ((2, SIDE_R, 0), (0, SIDE_L, 0), (5, SIDE_L, 0), (3, SIDE_L, 1)),
    ((4, SIDE_B, 0), (1, SIDE_B, 1), (2, SIDE_B, 1), (0, SIDE_B, 0)),
)
# cs_local_to_cart_vel / cs_cart_to_local_vel
# (src/layout/cubed_sphere_constants.h:22-37): (index, sign) per component
_L2C = (
    ((1, 1), (2, 1), (0, 1)),
    ((2, 1), (1, -1), (0, 1)),
    ((1, -1), (2, -1), (0, 1)),
    ((0, 1), (2, 1), (1, -1)),
    ((2, -1), (1, 1), (0, 1)),
    ((0, -1), (2, 1), (1, 1)),
)
_C2L = (
    ((2, 1), (0, 1), (1, 1)),
    ((2, 1), (1, -1), (0, 1)),
    ((2, 1), (0, -1), (1, -1)),
    ((0, 1), (2, -1), (1, 1)),
    ((2, 1), (1, 1), (0, -1)),
    ((0, -1), (2, 1), (1, 1)),
)


def _tan(a):
    a = np.asarray(a, dtype=np.float64)
    return np.array([math.tan(float(v)) for v in a.ravel()]).reshape(a.shape)


def _faces(lo, hi, n, ng):
    """block_faces_ (src/coord/coordinate.cpp:369-380): linspace over the
    ghosts, torch's scalar rule (counted up from the start for the first
    half, down from the end for the rest)."""
    dx = (hi - lo) / n
    start, end, steps = lo - ng * dx, hi + ng * dx, n + 2 * ng + 1
    step = (end - start) / (steps - 1)
    i = np.arange(steps)
    up = start + step * i
    down = end - step * (steps - i - 1)
    return np.where(i < steps // 2, up, down)


class PanelGrid:
    """The gnomonic-equiangle tables of a panel (the same on all six):
    metric cosines and sines at cells and faces, face areas, volumes, the
    metric source factors and the CFL widths."""

    def __init__(self, n: int, x1min: float, x1max: float, nghost: int = 3):
        self.n, self.ng = n, nghost
        self.nc = nc = n + 2 * nghost
        q = 0.25 * math.pi
        self.x2f = _faces(-q, q, n, nghost)
        self.x3f = self.x2f.copy()
        self.x2v = 0.5 * (self.x2f[:-1] + self.x2f[1:])
        self.x3v = 0.5 * (self.x3f[:-1] + self.x3f[1:])
        x1f = np.array([x1min, x1max])  # linspace of 2 points: the ends
        self.x1v = float(0.5 * (x1f[0] + x1f[1]))
        dx1 = (x1max - x1min) / 1
        # tangents of the 1D coordinates; the 2D planes broadcast them
        tx, txf = _tan(self.x2v), _tan(self.x2f)
        ty, tyf = _tan(self.x3v), _tan(self.x3f)
        self.tx, self.txf, self.ty, self.tyf = tx, txf, ty, tyf
        x, xf = tx[None, :], txf[None, :]
        y, yf = ty[:, None], tyf[:, None]
        C, Cf = np.sqrt(1.0 + x * x), np.sqrt(1.0 + xf * xf)
        D, Df = np.sqrt(1.0 + y * y), np.sqrt(1.0 + yf * yf)
        self.cos_c = -x * y / (C * D)
        self.sin_c = np.sqrt(1.0 + x * x + y * y) / (C * D)
        self.cos_f2 = -xf * y / (Cf * D)
        self.sin_f2 = np.sqrt(1.0 + xf * xf + y * y) / (Cf * D)
        self.cos_f3 = -x * yf / (C * Df)
        self.sin_f3 = np.sqrt(1.0 + x * x + yf * yf) / (C * Df)

        def arc(e, a, b):
            e2 = e * e
            v = (1.0 + e2 + a * b) / (np.sqrt(1.0 + e2 + a * a) * np.sqrt(1.0 + e2 + b * b))
            return np.array([math.acos(float(t)) for t in v.ravel()]).reshape(v.shape)

        lo2, hi2 = txf[None, :nc], txf[None, 1 : nc + 1]
        lo3, hi3 = tyf[:nc, None], tyf[1 : nc + 1, None]
        dx2f_ang = arc(y, lo2, hi2)
        dx2f_ang_face3 = arc(yf, lo2, hi2)
        dx3f_ang = arc(x, lo3, hi3)
        dx3f_ang_face2 = arc(xf, lo3, hi3)
        xc, yc = txf[None, :], tyf[:, None]
        t = xc * yc / np.sqrt(1.0 + xc * xc + yc * yc)
        corner = np.array([math.atan(float(v)) for v in t.ravel()]).reshape(t.shape)
        solid = corner[1:, 1:] - corner[1:, :-1] - corner[:-1, 1:] + corner[:-1, :-1]
        r1 = self.x1v * dx1
        self.area2 = r1 * dx3f_ang_face2
        self.area3 = r1 * dx2f_ang_face3
        rm, rp = x1f[0], x1f[1]
        self.vol = (dx1 * (rp * rp + rp * rm + rm * rm) / 3.0) * solid
        fx = self.area2 * self.sin_f2
        fy = self.area3 * self.sin_f3
        self.x_ov_rD = (fx[:, 1:] - fx[:, :-1]) / self.vol
        self.y_ov_rC = (fy[1:] - fy[:-1]) / self.vol
        self.width2 = self.x1v * dx2f_ang
        self.width3 = self.x1v * dx3f_ang
        self.solid_angle = solid

    def packed(self) -> np.ndarray:
        """The snapy_cs_mesh.geom layout."""
        parts = [self.cos_c, self.sin_c, self.vol, self.x_ov_rD, self.y_ov_rC, self.width2, self.width3,
                 self.cos_f2, self.sin_f2, self.area2, self.cos_f3, self.sin_f3, self.area3]
        return np.concatenate([np.ascontiguousarray(p, dtype=np.float64).ravel() for p in parts])

    def interior(self):
        return slice(self.ng, self.ng + self.n)


# ---------------------------------------------------------------------------
# frames (cs_metric_at, cs_contra_to_cart_at_, cs_cart_to_contra_at_,
# PanelRotation's local frame and matrix)
# ---------------------------------------------------------------------------
def _metric(x, y):
    return x, y, np.sqrt(x * x + y * y + 1), np.sqrt(1 + x * x), np.sqrt(1 + y * y)


def _contra_to_cart(v, m, face):
    x, y, delta, C, D = m
    vz, vx, vy = v[0], v[1], v[2]
    e = (
        (vz - vx * x / D - vy * y / C),
        (vz * x + vx * D - (vy * x * y) / C),
        (vz * y + vy * C - (vx * x * y) / D),
    )
    out = [None, None, None]
    for c in range(3):
        i, s = _L2C[face][c]
        out[i] = s * e[c] / delta
    return out


def _cart_to_contra(v, m, face):
    x, y, delta, C, D = m
    loc = [None, None, None]
    for c in range(3):
        i, s = _C2L[face][c]
        loc[i] = s * v[c]
    vz, vx, vy = loc
    return [(vz + x * vx + y * vy) / delta, D / delta * (vx - x * vz), C / delta * (vy - y * vz)]


def _raise(v, cth):
    sth2 = 1.0 - cth * cth
    return [v[0], v[1] / sth2 - v[2] * cth / sth2, -v[1] * cth / sth2 + v[2] / sth2]


def _lower(v, cth):
    return [v[0], v[1] + v[2] * cth, v[2] + v[1] * cth]


def frame_matrix(face, x, y, to_global, cosine=None):
    """[3, 3, ...]: the panel's components at tangents (x, y) to the local
    geographic frame (up, south, east) or back; cosine composes the metric
    raise before (to_global) or the lower after (conserved momenta)."""
    m = _metric(x, y)
    zero, one = np.zeros_like(x), np.ones_like(x)
    up = _contra_to_cart([one, zero, zero], m, face)
    fz, fx, fy = up
    s = np.sqrt(fx * fx + fy * fy)
    pole = s == 0.0
    safe = np.where(pole, 1.0, s)
    cphi = np.where(pole, 1.0, fx / safe)
    sphi = np.where(pole, 0.0, fy / safe)
    out = np.empty((3, 3) + np.shape(x))
    for inp in range(3):
        b = [one if c == inp else zero for c in range(3)]
        if to_global:
            if cosine is not None:
                b = _raise(b, cosine)
            vz, vx, vy = _contra_to_cart(b, m, face)
            horiz = vx * cphi + vy * sphi
            b = [horiz * s + vz * fz, horiz * fz - vz * s, -vx * sphi + vy * cphi]
        else:
            vr, vt, vp = b
            horiz = vr * s + vt * fz
            cart = [vr * fz - vt * s, horiz * cphi - vp * sphi, horiz * sphi + vp * cphi]
            b = _cart_to_contra(cart, m, face)
            if cosine is not None:
                b = _lower(b, cosine)
        for o in range(3):
            out[o, inp] = b[o]
    return out


# ---------------------------------------------------------------------------
# ghost and seam index maps
# ---------------------------------------------------------------------------
_FACE_XYZ = (
    lambda a, b: (1.0, a, b),
    lambda a, b: (-a, 1.0, b),
    lambda a, b: (-1.0, -a, b),
    lambda a, b: (-b, a, 1.0),
    lambda a, b: (a, -1.0, b),
    lambda a, b: (b, a, -1.0),
)
_FACE_AB = (
    lambda x, y, z: (math.atan2(y, x), math.atan2(z, x)),
    lambda x, y, z: (math.atan2(-x, y), math.atan2(z, y)),
    lambda x, y, z: (math.atan2(-y, -x), math.atan2(z, -x)),
    lambda x, y, z: (math.atan2(y, z), math.atan2(-x, z)),
    lambda x, y, z: (math.atan2(x, -y), math.atan2(z, -y)),
    lambda x, y, z: (math.atan2(y, -z), math.atan2(x, -z)),
)


def ghost_source_positions(n: int, nghost: int) -> np.ndarray:
    """usrc[depth - 1, j]: the fractional cell-centre index along the source
    panel's edge of the ghost centre (depth, j), built on panel 0's left
    edge and shared by every edge (cs_build_ghost_usrc with its default
    face and side)."""
    nface, nside, rev = CS_EDGES[0][SIDE_L]
    d = math.pi / (2.0 * n)
    out = np.empty((nghost, n))
    for depth in range(1, nghost + 1):
        perp = -1 * (math.pi / 4.0 + (depth - 0.5) * d)
        for j in range(n):
            along = -math.pi / 4.0 + (j + 0.5) * d
            x, y, z = _FACE_XYZ[0](math.tan(perp), math.tan(along))
            norm = math.sqrt(x * x + y * y + z * z)
            a_s, b_s = _FACE_AB[nface](x / norm, y / norm, z / norm)
            u = ((b_s if nside in (SIDE_L, SIDE_R) else a_s) + math.pi / 4.0) / d - 0.5
            out[depth - 1, j] = (n - 1) - u if rev else u
    return out


def _ghost_cell(side, depth, j, n, ng):
    """(row k, column j) of a panel's edge ghost at depth (1..ng), along j."""
    if side == SIDE_L:
        return ng + j, ng - depth
    if side == SIDE_R:
        return ng + j, ng + n - 1 + depth
    if side == SIDE_B:
        return ng - depth, ng + j
    return ng + n - 1 + depth, ng + j


def _layer_cell(side, depth, a, n, ng):
    """The interior cell at `depth` (1 = next to the edge) from `side`, along
    index a."""
    if side == SIDE_L:
        return ng + a, ng + depth - 1
    if side == SIDE_R:
        return ng + a, ng + n - depth
    if side == SIDE_B:
        return ng + depth - 1, ng + a
    return ng + n - depth, ng + a


def build_ghost_tables(grid: PanelGrid, interp_order: int):
    """(idx int32 [E, 8], weights [E, 3], matrices conserved [E, 45],
    matrices primitive [E, 45]) for the edge ghosts of the six panels, then
    the corner cells (copies of the nearest interior cell)."""
    n, ng, nc = grid.n, grid.ng, grid.nc
    usrc = ghost_source_positions(n, ng)
    # cell matrices of every panel: the sender's (to geographic) and the
    # receiver's (from geographic), conserved (with the cell cosine) and not
    X = np.broadcast_to(grid.tx[None, :], (nc, nc))
    Y = np.broadcast_to(grid.ty[:, None], (nc, nc))
    mats = {}
    for f in range(6):
        for to in (True, False):
            mats[(f, to, True)] = frame_matrix(f, X, Y, to, grid.cos_c)
            mats[(f, to, False)] = frame_matrix(f, X, Y, to, None)
    idx, wts, mc, mp = [], [], [], []
    shift = ng
    lo, hi = max(0, shift), min(nc - 1, shift + n - 1)
    for p in range(6):
        for side in range(4):
            q, qside, rev = CS_EDGES[p][side]
            for depth in range(1, ng + 1):
                for j in range(n):
                    u = usrc[depth - 1, j]
                    u0f = math.floor(u)
                    x = u - u0f
                    u0 = int(u0f) + shift  # index in the widened strip
                    if interp_order == 4 and hi - lo >= 3:
                        s = min(max(u0 - 1, lo), hi - 3)
                        t = float(u0 - s) + x
                        c0 = -(t - 1.0) * (t - 2.0) * (t - 3.0) / 6.0
                        c2 = -t * (t - 1.0) * (t - 3.0) / 2.0
                        c3 = t * (t - 1.0) * (t - 2.0) / 6.0
                        pts, w = [s, s + 1, s + 2, s + 3], (c0, c2, c3)
                    else:
                        pts, w = [u0, u0 + 1], (x, 0.0, 0.0)
                    if min(pts) < shift or max(pts) > shift + n - 1:
                        raise RuntimeError("a panel-edge ghost reads outside the source interior")
                    k, jj = _ghost_cell(side, depth, j, n, ng)
                    src = []
                    for i_buf in pts:
                        i = i_buf - shift
                        a = n - 1 - i if rev else i
                        src.append(_layer_cell(qside, depth, a, n, ng))
                    row = [p, k * nc + jj, q] + [r * nc + c for (r, c) in src] + [0] * (4 - len(src)) + [len(src)]
                    idx.append(row)
                    wts.append(w)
                    for store, cons in ((mc, True), (mp, False)):
                        m = np.zeros((5, 9))
                        for t_, (r, c) in enumerate(src):
                            m[t_] = mats[(q, True, cons)][:, :, r, c].ravel()
                        m[4] = mats[(p, False, cons)][:, :, k, jj].ravel()
                        store.append(m.ravel())
        # corners: the nearest interior cell of the same panel
        for k in list(range(ng)) + list(range(ng + n, nc)):
            for jj in list(range(ng)) + list(range(ng + n, nc)):
                kk = min(max(k, ng), ng + n - 1)
                jc = min(max(jj, ng), ng + n - 1)
                idx.append([p, k * nc + jj, p, kk * nc + jc, 0, 0, 0, 1])
                wts.append((0.0, 0.0, 0.0))
                mc.append(np.zeros(45))
                mp.append(np.zeros(45))
    return (np.array(idx, dtype=np.int32), np.array(wts, dtype=np.float64),
            np.array(mc, dtype=np.float64), np.array(mp, dtype=np.float64))


def build_seam_tables(grid: PanelGrid):
    """(idx int32 [12 n, 8], operands [12 n, 30]) for the twelve seams, each
    owned by its lower panel."""
    n, ng = grid.n, grid.ng
    plane = (ng, ng + n, ng, ng + n)

    def face_point(side, i):
        """(tan alpha, tan beta, face cosine) at a side's face i."""
        if side < 2:
            return grid.txf[plane[side]], grid.ty[ng + i], grid.cos_f2[ng + i, plane[side]]
        return grid.tx[ng + i], grid.tyf[plane[side]], grid.cos_f3[plane[side], ng + i]

    idx, dat = [], []
    for p in range(6):
        for side in range(4):
            q, qside, rev = CS_EDGES[p][side]
            if p > q:
                continue
            sign = -1 if (side % 2) == (qside % 2) else 1
            for i in range(n):
                ib = n - 1 - i if rev else i
                xa, ya, ca = face_point(side, i)
                xb, yb, cb = face_point(qside, ib)
                xa, ya, xb, yb = (np.float64(v) for v in (xa, ya, xb, yb))
                m_to_b = frame_matrix(q, xb, yb, True, None)
                m_from_a = frame_matrix(p, xa, ya, False, None)
                ma = _metric(xa, ya)
                mb = _metric(xb, yb)
                idx.append([p, side, i, q, qside, ib, sign, 0])
                dat.append(np.concatenate([m_to_b.ravel(), m_from_a.ravel(), [*ma, ca], [*mb, cb]]))
    return np.array(idx, dtype=np.int32), np.array(dat, dtype=np.float64)


def sphere_points(grid: PanelGrid, panel: int) -> np.ndarray:
    """(3, nc, nc) unit positions (x, y, z) of a panel's cell centres
    (cs_ab_to_xyz, src/coord/cubed_sphere_utils.cpp:47-52)."""
    a = np.broadcast_to(grid.tx[None, :], (grid.nc, grid.nc))
    b = np.broadcast_to(grid.ty[:, None], (grid.nc, grid.nc))
    one = np.ones_like(a)
    x, y, z = ((one, a, b), (-a, one, b), (-one, -a, b), (-b, a, one), (a, -one, b), (b, a, -one))[panel]
    norm = np.sqrt(x * x + y * y + z * z)
    return np.stack([x / norm, y / norm, z / norm])


# ---------------------------------------------------------------------------
# the mesh
# ---------------------------------------------------------------------------
class _MeshState:
    """The StateStore contract the executor keys on: a generation and the
    backup's address."""

    def __init__(self):
        self.generation = 0
        self._u0 = None

    def u0(self):
        return self._u0


class _MeshExecutor(StageExecutor):
    def _key(self, vars, stage):
        m = self.block
        return (stage, m._state.generation, m._u.data_ptr(), m._w.data_ptr(), m._u0.data_ptr())


class CubedSphereMesh:
    """Mesh::forward over six gnomonic panels (one block each) of `members`
    independent meshes, resident on one device.

    ``vars`` is the reference's MeshVariables: one dict per block, block
    ``member * 6 + panel``, each with ``hydro_w`` (and, after initialize,
    ``hydro_u``) of shape (4, nc, nc, 1); after initialize they are views of
    the mesh's resident state, which the stages update in place."""

    def __init__(self, options: MeshOptions, members: int = 1):
        if options.layout != "cubed-sphere":
            raise ValueError("CubedSphereMesh needs distribute/layout: cubed-sphere")
        self.options = options
        op = options.block
        c = op.coord
        self.n, self.ng, self.members = c.nx2, c.nghost, int(members)
        if self.members < 1:
            raise ValueError("members >= 1")
        self.nc = self.n + 2 * self.ng
        self.nblocks = 6 * self.members
        self.interp_order = c.interp_order
        self.shock = bool(op.hydro.recon23.shock)
        self.cfl = op.intg.cfl
        self.grid = PanelGrid(self.n, c.x1min, c.x1max, self.ng)
        self.device = torch.device("cpu")
        self._tables = None
        self._desc = None
        self._u = self._w = self._u0 = None
        self._state = _MeshState()
        self._dt_dev = None
        self._executor = _MeshExecutor(self, op.intg.execution)
        self._cycle = 0
        self.current_redo = 0
        self.max_redo = 5
        self.density_floor = op.hydro.eos.density_floor
        self._w_stale = False
        # dynamics/precision float32: the FP32 pair state, FP32 primitives,
        # fluxes and edge states, FP32 sweeps; FP64 EOS, gathers, seams,
        # update and budgets
        self.fp32 = op.hydro.precision == "float32"
        self._c = self._c0 = None

    # -- device set-up ------------------------------------------------------
    def to(self, device):
        device = torch.device(device)
        if device.type != "cuda":
            raise RuntimeError("CubedSphereMesh runs on a HIP device only (there is no CPU path)")
        _capi.load()
        self.device = device
        g = self.grid
        gi, gw, gmc, gmp = build_ghost_tables(g, self.interp_order)
        si, sd = build_seam_tables(g)

        def dev(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(device)

        self._tables = {
            "geom": dev(g.packed()),
            "ghost_idx": dev(gi),
            "ghost_w": dev(gw),
            "ghost_mat_c": dev(gmc),
            "ghost_mat_p": dev(gmp),
            "seam_idx": dev(si),
            "seam_d": dev(sd),
        }
        t = self._tables
        P = _capi.ptr
        self._desc = _capi.CsMesh(
            self.n, self.ng, self.members, self.interp_order, int(self.shock), 0, g.x1v,
            P(t["geom"]), P(t["ghost_idx"]), P(t["ghost_w"]), P(t["ghost_mat_c"]), P(t["ghost_mat_p"]),
            int(gi.shape[0]), P(t["seam_idx"]), P(t["seam_d"]),
        )
        lib = _capi.load()
        shape = (self.nblocks, NROWS, self.nc, self.nc, 1)
        z = dict(dtype=torch.float64, device=device)
        zf = dict(dtype=torch.float32 if self.fp32 else torch.float64, device=device)
        self._f2 = torch.zeros(shape, **zf)
        self._f3 = torch.zeros(shape, **zf)
        self._edge = torch.zeros(int(lib.snapy_cs_edge_size(ctypes.byref(self._desc))), **zf)
        self._dt_dev = torch.zeros(self.members, **z)
        return self

    def hbm_footprint(self) -> int:
        """Device bytes of the resident state: u, u0, w, two flux totals, the
        edge buffer and the tables (FP64; in the FP32 form the pair u, c and
        its backup, and FP32 primitives, fluxes and edges)."""
        cells = self.nblocks * NROWS * self.nc * self.nc
        tables = sum(t.numel() * t.element_size() for t in (self._tables or {}).values())
        if self.fp32:
            return 4 * cells * 4 + 3 * cells * 4 + self.nblocks * 4 * self.n * NROWS * 4 + tables
        return 5 * cells * 8 + self.nblocks * 4 * self.n * NROWS * 8 + tables

    def _stream(self):
        return _capi.stream_of(self.device)

    # -- the reference's Mesh surface ---------------------------------------
    def cycle(self) -> int:
        return self._cycle

    def inc_cycle(self) -> int:
        v = self._cycle
        self._cycle += 1
        return v

    def block_shape(self):
        return (NROWS, self.nc, self.nc, 1)

    def coordinates(self, panel: int):
        """(alpha, beta) of a panel's cell centres, each (nc, nc): x2v along
        the columns, x3v along the rows."""
        g = self.grid
        return (np.broadcast_to(g.x2v[None, :], (self.nc, self.nc)).copy(),
                np.broadcast_to(g.x3v[:, None], (self.nc, self.nc)).copy())

    def initialize(self, vars):
        """Mesh::initialize: each block's hydro_w (ghosts filled by the
        cubed-sphere exchange of the primitives), then hydro_u = prim2cons
        and its exchange (the first stage's entry fill).  Returns
        (vars, 0.0)."""
        if self._desc is None:
            raise RuntimeError("call to(device) before initialize")
        if len(vars) != self.nblocks:
            raise ValueError(f"initialize: {self.nblocks} blocks expected, got {len(vars)}")
        w = torch.stack([v["hydro_w"].to(self.device, torch.float64) for v in vars]).contiguous()
        if tuple(w.shape[1:]) != self.block_shape():
            raise ValueError(f"initialize: hydro_w has shape {tuple(w.shape[1:])}, expected {self.block_shape()}")
        lib, d, s = _capi.load(), ctypes.byref(self._desc), self._stream()
        P = _capi.ptr
        self._state = _MeshState()
        _capi.check(lib.snapy_cs_ghosts(d, 0, P(w), s), "cs_ghosts")
        if self.fp32:
            f32 = dict(dtype=torch.float32, device=self.device)
            self._u = torch.empty(w.shape, **f32)
            self._c = torch.empty(w.shape, **f32)
            self._u0 = torch.empty(w.shape, **f32)
            self._c0 = torch.empty(w.shape, **f32)
            self._w = torch.empty(w.shape, **f32)
            _capi.check(lib.snapy_cs_prim2cons_f32(d, P(w), P(self._u), P(self._c), s), "cs_prim2cons_f32")
            _capi.check(lib.snapy_cs_ghosts_f32(d, P(self._u), P(self._c), s), "cs_ghosts_f32")
            _capi.check(lib.snapy_cs_cons2prim_f32(d, P(self._u), P(self._c), P(self._w), s), "cs_cons2prim_f32")
        else:
            self._w = w
            self._u = torch.empty_like(w)
            self._u0 = torch.empty_like(w)
            _capi.check(lib.snapy_cs_prim2cons(d, P(self._w), P(self._u), s), "cs_prim2cons")
            _capi.check(lib.snapy_cs_ghosts(d, 1, P(self._u), s), "cs_ghosts")
        self._state._u0 = self._u0
        for b, v in enumerate(vars):
            v["hydro_w"] = self._w[b]
            v["hydro_u"] = self._u[b]
        self._w_stale = False
        return vars, 0.0

    def _stage_calls(self, stage: int):
        lib, d, s = _capi.load(), ctypes.byref(self._desc), self._stream()
        P = _capi.ptr
        w0, w1, w2 = ((1.0, 0.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0))[stage]
        u, w, u0, f2, f3, e = self._u, self._w, self._u0, self._f2, self._f3, self._edge
        if self.fp32:
            c, c0 = self._c, self._c0
            return [
                ("cs_cons2prim", lambda: lib.snapy_cs_cons2prim_f32(d, P(u), P(c), P(w), s)),
                ("cs_sweep_x2", lambda: lib.snapy_cs_sweep_f32(d, 2, P(w), P(f2), P(e), s)),
                ("cs_sweep_x3", lambda: lib.snapy_cs_sweep_f32(d, 3, P(w), P(f3), P(e), s)),
                ("cs_seams", lambda: lib.snapy_cs_seams_f32(d, P(f2), P(f3), P(e), s)),
                ("cs_update", lambda: lib.snapy_cs_update_f32(d, P(u0), P(c0), P(u), P(c), P(w), P(f2), P(f3),
                                                             0.0, P(self._dt_dev), w0, w1, w2, s)),
                ("cs_ghosts", lambda: lib.snapy_cs_ghosts_f32(d, P(u), P(c), s)),
            ]
        return [
            ("cs_cons2prim", lambda: lib.snapy_cs_cons2prim(d, P(u), P(w), s)),
            ("cs_sweep_x2", lambda: lib.snapy_cs_sweep(d, 2, P(w), P(f2), P(e), s)),
            ("cs_sweep_x3", lambda: lib.snapy_cs_sweep(d, 3, P(w), P(f3), P(e), s)),
            ("cs_seams", lambda: lib.snapy_cs_seams(d, P(f2), P(f3), P(e), s)),
            ("cs_update", lambda: lib.snapy_cs_update(d, P(u0), P(u), P(w), P(f2), P(f3), 0.0,
                                                     P(self._dt_dev), w0, w1, w2, s)),
            ("cs_ghosts", lambda: lib.snapy_cs_ghosts(d, 1, P(u), s)),
        ]

    def _launch_stage(self, vars, stage: int) -> None:
        if stage == 0:
            self._u0.copy_(self._u)
            if self._c is not None:
                self._c0.copy_(self._c)
        for name, fn in self._stage_calls(stage):
            _capi.check(fn(), name)

    def forward(self, vars, dt, stage: int) -> None:
        """Mesh::forward: one RK stage of every block.  dt: a host float
        (every member), a 0-d device tensor, or one device dt per member."""
        if isinstance(dt, torch.Tensor):
            self._dt_dev.copy_(dt.to(self._dt_dev.device, torch.float64).expand(self.members))
        else:
            self._dt_dev.fill_(float(dt))
        self._executor.run_stage(vars, stage)
        self._w_stale = True

    def _refresh_primitives(self) -> None:
        """w <- prim(u) on every cell of every panel (the reference's
        check_redo refreshes the primitives, meshblock_redo.cpp:107-112)."""
        lib = _capi.load()
        d, P = ctypes.byref(self._desc), _capi.ptr
        if self.fp32:
            _capi.check(lib.snapy_cs_cons2prim_f32(d, P(self._u), P(self._c), P(self._w), self._stream()),
                        "cs_cons2prim_f32")
        else:
            _capi.check(lib.snapy_cs_cons2prim(d, P(self._u), P(self._w), self._stream()), "cs_cons2prim")
        self._w_stale = False

    def device_time_step(self, vars=None) -> torch.Tensor:
        """The CFL dt of each member on the device (2^-redo * cfl * min over
        its six panels), with no host read; from the primitives of the
        current state."""
        if self._w_stale:
            self._refresh_primitives()
        lib = _capi.load()
        scale = math.pow(2.0, -self.current_redo) * self.cfl
        cfl = lib.snapy_cs_cfl_f32 if self.fp32 else lib.snapy_cs_cfl
        _capi.check(cfl(ctypes.byref(self._desc), _capi.ptr(self._w), scale,
                        _capi.ptr(self._dt_dev), self._stream()), "cs_cfl")
        return self._dt_dev

    def max_time_step(self, vars=None):
        """Mesh::max_time_step: one host read; a float for one member, else
        one per member."""
        dt = self.device_time_step(vars).clone().cpu().tolist()
        return dt[0] if self.members == 1 else dt

    def check_redo(self, vars=None) -> int:
        """>0 redo, 0 ok, <0 stop (meshblock_redo.cpp:32-89): the primitives
        are refreshed from u; a non-finite interior or a geopotential at or
        within 0.1% of the floor asks for a redo, which restores u0 (and
        its primitives) and halves dt; past max_redo the mesh stops."""
        self._refresh_primitives()
        i = slice(self.ng, self.ng + self.n)
        nan = ~torch.isfinite(self._u[:, :, i, i]).all()
        if self._c is not None:
            nan = nan | ~torch.isfinite(self._c[:, :, i, i]).all()
        floor = ~(self._w[:, IDN, i, i].min() > 1.001 * self.density_floor)
        flags = torch.stack([floor, nan]).cpu().tolist()
        if not any(flags):
            self.current_redo = 0
            return 0
        self.current_redo += 1
        if self.current_redo > self.max_redo:
            return -1
        self._u.copy_(self._u0)
        if self._c is not None:
            self._c.copy_(self._c0)
        self._refresh_primitives()
        return 1

    def mass(self) -> torch.Tensor:
        """Per member: the sum over its panels' interiors of phi times cell
        volume (FP64)."""
        i = slice(self.ng, self.ng + self.n)
        vol = torch.from_numpy(np.ascontiguousarray(self.grid.vol[i, i])).to(self.device)
        phi = self.total_u()[:, IDN, i, i, 0] * vol
        return phi.reshape(self.members, 6, -1).sum(dim=(1, 2))

    def total_u(self) -> torch.Tensor:
        """The conserved state of every block, FP64 (the pair's value
        u + c in the FP32 form)."""
        if self._c is not None:
            return self._u.double() + self._c.double()
        return self._u

    def execution_report(self) -> dict:
        return self._executor.report()

    def interior_state(self) -> torch.Tensor:
        """The conserved state of every block's interior, (blocks, 4, n, n)."""
        i = slice(self.ng, self.ng + self.n)
        return self.total_u()[:, :, i, i, 0]

    def release(self, sync: bool = True) -> None:
        """Drop the captured graphs (sync=False: the caller has synchronized
        the stream the stages ran on)."""
        if sync:
            torch.cuda.synchronize(self.device)
        self._executor.release()
