"""CPU FP64 restatement of the cubed-sphere shallow-water stage (test
infrastructure only: the product path never imports this module).

Six gnomonic-equiangle panels, one block per panel (the reference's
``distribute/layout: cubed-sphere`` with nb2 = nb3 = 1 and all six blocks in
one process), shallow water with the shallow-roe solver in the y-z plane.
Every function follows the reference's tensor operations in their order, so
that the HIP path can be held to it bit for bit:

* grid and metric   -- GnomonicEquiangleImpl::reset (src/coord/
  gnomonic_equiangle.cpp:37-250), face areas / volume (:262-284),
  block_faces_ and uniform_axis_metrics_ (src/coord/coordinate.cpp:369-425)
  with torch's linspace;
* frames            -- cs_cart_to_contra_at_ / cs_contra_to_cart_at_
  (src/coord/cubed_sphere_utils.cpp:240-290), the velocity tables
  (src/layout/cubed_sphere_constants.h:22-37), PanelRotation::matrix and
  its local geographic frame (src/coord/coord_rotation.cpp:12-150),
  coord_vec_raise/lower (src/coord/coord_utils_impl.h:9-25),
  call_cs_matrix_vec (src/coord/coord_dispatch.cpp:66-90);
* connectivity      -- CS_FACE_EDGES (src/layout/cubed_sphere_layout.cpp:
  352-383), cs_build_ghost_usrc (src/coord/cubed_sphere_utils.cpp:82-113);
* ghost exchange    -- CubedSphereLayoutImpl::serialize / deserialize
  (src/layout/cubed_sphere_layout.cpp:563-800) with interp_ghost
  (src/coord/gnomonic_equiangle.cpp:286-305,537-605);
* face-state sync   -- the cross-panel face-state exchange of
  HydroImpl::forward (src/hydro/hydro_forward.cpp:707-757);
* Riemann frames    -- prim2local2_/3_, flux2global2_/3_
  (src/coord/gnomonic_equiangle.cpp:342-470), shallow_roe_impl with
  shallow_axes dir yz (src/riemann/riemann_kernels.h:44-47,447-495);
* seams             -- reconcile_panel_seams (src/layout/seam_register.cpp:
  35-186): the lower rank owns a seam; its face flux, through the global
  Cartesian frame, replaces the other panel's;
* divergence        -- CoordinateImpl::divergence (src/coord/coordinate.cpp:
  618-662) and GnomonicEquiangleImpl::forward's shallow-water metric
  sources (src/coord/gnomonic_equiangle.cpp:472-521);
* EOS               -- ShallowWaterImpl::_cons2prim/_prim2cons
  (src/eos/shallow_water.cpp:31-55): contravariant velocities in w,
  covariant momenta in u;
* CFL               -- HydroImpl::max_time_step (src/hydro/hydro.cpp:372-397)
  with center_width2/3 (src/coord/gnomonic_equiangle.cpp:254-260).

* x1 mass source   -- GnomonicEquiangleImpl::x1_mass_source_
  (src/coord/gnomonic_equiangle.cpp:523-535) with m1 = rho v1, as
  hydro_forward.cpp:1224-1231 calls it for nc1 = 1.  The radial momentum is
  zero in exact arithmetic; the seam rotation leaves it at rounding level.

Transcendental functions (tan, atan, acos, atan2) are evaluated with the
C library through ``math``, element by element, so the tables do not depend
on numpy's vectorised kernels.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import fv, swe

IDN, IVX, IVY, IVZ = 0, 1, 2, 3
SIDE_L, SIDE_R, SIDE_B, SIDE_T = 0, 1, 2, 3
FACE_NAMES = ("+X", "+Y", "-X", "+Z", "-Y", "-Z")

# CS_FACE_EDGES: leaving face f by side s lands on (face, side, reversed)
EDGES = (
    ((4, SIDE_R, 0), (1, SIDE_L, 0), (5, SIDE_T, 0), (3, SIDE_B, 0)),
    ((0, SIDE_R, 0), (2, SIDE_L, 0), (5, SIDE_R, 1), (3, SIDE_R, 0)),
    ((1, SIDE_R, 0), (4, SIDE_L, 0), (5, SIDE_B, 1), (3, SIDE_T, 1)),
    ((4, SIDE_T, 1), (1, SIDE_T, 0), (0, SIDE_T, 0), (2, SIDE_T, 1)),
    ((2, SIDE_R, 0), (0, SIDE_L, 0), (5, SIDE_L, 0), (3, SIDE_L, 1)),
    ((4, SIDE_B, 0), (1, SIDE_B, 1), (2, SIDE_B, 1), (0, SIDE_B, 0)),
)

# cs_cart_to_local_vel / cs_local_to_cart_vel: (index, sign) per component
_C2L_IDX = ((2, 0, 1), (2, 1, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (0, 2, 1))
_C2L_SGN = ((1, 1, 1), (1, -1, 1), (1, -1, -1), (1, -1, 1), (1, 1, -1), (-1, 1, 1))
_L2C_IDX = ((1, 2, 0), (2, 1, 0), (1, 2, 0), (0, 2, 1), (2, 1, 0), (0, 2, 1))
_L2C_SGN = ((1, 1, 1), (1, -1, 1), (-1, -1, 1), (1, 1, -1), (-1, 1, 1), (-1, 1, 1))


def _map(fn, a):
    """A C-library function applied element by element."""
    a = np.asarray(a, dtype=np.float64)
    return np.array([fn(float(v)) for v in a.ravel()], dtype=np.float64).reshape(a.shape)


def torch_linspace(start, end, steps):
    """torch.linspace's scalar rule in FP64 (the reference's block_faces_):
    the first half counted up from start, the second half down from end.
    (torch's vectorised loop can round a middle element differently by an
    ulp; this build fixes the scalar rule on both sides of every parity
    test.)"""
    step = (end - start) / (steps - 1)
    half = steps // 2
    out = np.empty(steps)
    for i in range(steps):
        out[i] = start + step * i if i < half else end - step * (steps - i - 1)
    return out


# ---------------------------------------------------------------------------
# grid and metric
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    """One panel's gnomonic-equiangle grid: nx1 = 1 (the shell), nx2 = nx3 =
    n on [-pi/4, pi/4]; arrays are (nc3, nc2, nc1) with nc1 = 1."""

    n: int
    radius: float = 6371000.0
    depth: float = 1.0
    nghost: int = 3
    interp_order: int = 2

    def __post_init__(self):
        n, ng = self.n, self.nghost
        self.nc = n + 2 * ng
        lo, hi = -0.25 * math.pi, 0.25 * math.pi
        dx = (hi - lo) / n
        self.x2f = torch_linspace(lo - ng * dx, hi + ng * dx, n + 2 * ng + 1)
        self.x3f = self.x2f.copy()
        self.x2v = 0.5 * (self.x2f[:-1] + self.x2f[1:])
        self.x3v = 0.5 * (self.x3f[:-1] + self.x3f[1:])
        self.x1f = torch_linspace(self.radius, self.radius + self.depth, 2)
        self.x1v = 0.5 * (self.x1f[:1] + self.x1f[1:])
        self.dx1f = np.ones(1) * ((self.radius + self.depth - self.radius) / 1)
        nc = self.nc
        x = _map(math.tan, self.x2v)[None, :, None]
        xf = _map(math.tan, self.x2f)[None, :, None]
        y = _map(math.tan, self.x3v)[:, None, None]
        yf = _map(math.tan, self.x3f)[:, None, None]
        C = np.sqrt(1.0 + x * x)
        Cf = np.sqrt(1.0 + xf * xf)
        D = np.sqrt(1.0 + y * y)
        Df = np.sqrt(1.0 + yf * yf)
        self.cos_c = -x * y / (C * D)
        self.sin_c = np.sqrt(1.0 + x * x + y * y) / (C * D)
        self.cos_f2 = -xf * y / (Cf * D)
        self.sin_f2 = np.sqrt(1.0 + xf * xf + y * y) / (Cf * D)
        self.cos_f3 = -x * yf / (C * Df)
        self.sin_f3 = np.sqrt(1.0 + x * x + yf * yf) / (C * Df)

        def arc(e, a, b):
            e2 = e * e
            return _map(math.acos, (1.0 + e2 + a * b) / (np.sqrt(1.0 + e2 + a * a) * np.sqrt(1.0 + e2 + b * b)))

        x1 = _map(math.tan, self.x2f[:nc])[None, :, None]
        x2 = _map(math.tan, self.x2f[1 : nc + 1])[None, :, None]
        self.dx2f_ang = arc(y, x1, x2)
        self.dx2f_ang_face3 = arc(yf, x1, x2)
        y1 = _map(math.tan, self.x3f[:nc])[:, None, None]
        y2 = _map(math.tan, self.x3f[1 : nc + 1])[:, None, None]
        self.dx3f_ang = arc(x, y1, y2)
        self.dx3f_ang_face2 = arc(xf, y1, y2)
        xc = _map(math.tan, self.x2f)[None, :]
        yc = _map(math.tan, self.x3f)[:, None]
        corner = _map(math.atan, xc * yc / np.sqrt(1.0 + xc * xc + yc * yc))
        self.solid_angle = (corner[1:, 1:] - corner[1:, :-1] - corner[:-1, 1:] + corner[:-1, :-1])[..., None]
        r1 = (self.x1v * self.dx1f)[None, None, :]
        self.area2 = r1 * self.dx3f_ang_face2  # (nc3, nc2 + 1, 1)
        self.area3 = r1 * self.dx2f_ang_face3  # (nc3 + 1, nc2, 1)
        rm, rp = self.x1f[:1], self.x1f[1:]
        radial = (self.dx1f * (rp * rp + rp * rm + rm * rm) / 3.0)[None, None, :]
        self.vol = radial * self.solid_angle
        fx = self.area2 * self.sin_f2
        fy = self.area3 * self.sin_f3
        self.x_ov_rD = (fx[:, 1:] - fx[:, :-1]) / self.vol
        self.y_ov_rC = (fy[1:] - fy[:-1]) / self.vol
        self.width2 = self.x1v[None, None, :] * self.dx2f_ang
        self.width3 = self.x1v[None, None, :] * self.dx3f_ang

    @property
    def shape(self):
        return (self.nc, self.nc, 1)

    def interior(self):
        ng, n = self.nghost, self.n
        return (slice(ng, ng + n), slice(ng, ng + n), slice(0, 1))


# ---------------------------------------------------------------------------
# frames
# ---------------------------------------------------------------------------
def metric_at(alpha, beta):
    """cs_metric_at: x, y, delta, C, D at panel angles."""
    x = _map(math.tan, alpha)
    y = _map(math.tan, beta)
    return x, y, np.sqrt(x * x + y * y + 1), np.sqrt(1 + x * x), np.sqrt(1 + y * y)


def cart_to_contra(vel, m, face):
    """cs_cart_to_contra_at_: global Cartesian (z, x, y) -> panel
    contravariant, in place on vel[0:3]."""
    x, y, delta, C, D = m
    loc = [None, None, None]
    for c in range(3):
        loc[_C2L_IDX[face][c]] = _C2L_SGN[face][c] * vel[c]
    vz, vx, vy = loc
    vel[0] = (vz + x * vx + y * vy) / delta
    vel[1] = D / delta * (vx - x * vz)
    vel[2] = C / delta * (vy - y * vz)


def contra_to_cart(vel, m, face):
    """cs_contra_to_cart_at_: panel contravariant -> global Cartesian."""
    x, y, delta, C, D = m
    vz, vx, vy = vel[0].copy(), vel[1].copy(), vel[2].copy()
    i, s = _L2C_IDX[face], _L2C_SGN[face]
    vel[i[0]] = s[0] * (vz - vx * x / D - vy * y / C) / delta
    vel[i[1]] = s[1] * (vz * x + vx * D - (vy * x * y) / C) / delta
    vel[i[2]] = s[2] * (vz * y + vy * C - (vx * x * y) / D) / delta


def vec_raise(vel, cth):
    """coord_vec_raise_impl on rows 1, 2 of vel."""
    v, w = vel[1].copy(), vel[2].copy()
    sth2 = 1.0 - cth * cth
    vel[1] = v / sth2 - w * cth / sth2
    vel[2] = -v * cth / sth2 + w / sth2


def vec_lower(vel, cth):
    """coord_vec_lower_impl on rows 1, 2 of vel."""
    v, w = vel[1].copy(), vel[2].copy()
    vel[1] = v + w * cth
    vel[2] = w + v * cth


def local_frame(face, alpha, beta):
    """The local geographic frame from the unit position alone."""
    up = np.zeros((3,) + np.shape(alpha))
    up[0] = 1.0
    contra_to_cart(up, metric_at(alpha, beta), face)
    z, x, y = up[0], up[1], up[2]
    s = np.sqrt(x * x + y * y)
    pole = s == 0.0
    safe = np.where(pole, 1.0, s)
    cphi = np.where(pole, 1.0, x / safe)
    sphi = np.where(pole, 0.0, y / safe)
    return x, y, z, cphi, sphi, s


def cart_to_local(vel, f):
    x, y, z, cphi, sphi, s = f
    vz, vx, vy = vel[0].copy(), vel[1].copy(), vel[2].copy()
    horiz = vx * cphi + vy * sphi
    vel[0] = horiz * s + vz * z
    vel[1] = horiz * z - vz * s
    vel[2] = -vx * sphi + vy * cphi


def local_to_cart(vel, f):
    x, y, z, cphi, sphi, s = f
    vr, vt, vp = vel[0].copy(), vel[1].copy(), vel[2].copy()
    horiz = vr * s + vt * z
    vel[1] = horiz * cphi - vp * sphi
    vel[2] = horiz * sphi + vp * cphi
    vel[0] = vr * z - vt * s


def rotation_matrix(face, alpha, beta, to_global, cosine=None):
    """PanelRotation::matrix: [3, 3, ...] from the panel's components to the
    local geographic frame (to_global) or back; a defined cosine composes
    the metric raise before / the lower after (conserved momenta)."""
    alpha = np.asarray(alpha, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    shape = alpha.shape
    m = np.empty((3, 3) + shape)
    frame = local_frame(face, alpha, beta)
    for inp in range(3):
        basis = np.zeros((3,) + shape)
        basis[inp] = 1.0
        if to_global:
            if cosine is not None:
                vec_raise(basis, cosine)
            contra_to_cart(basis, metric_at(alpha, beta), face)
            cart_to_local(basis, frame)
        else:
            local_to_cart(basis, frame)
            cart_to_contra(basis, metric_at(alpha, beta), face)
            if cosine is not None:
                vec_lower(basis, cosine)
        m[:, inp] = basis
    return m


def apply_matrix(vel, m):
    """call_cs_matrix_vec: out_i = m_i0 v0 + m_i1 v1 + m_i2 v2."""
    v0, v1, v2 = vel[0].copy(), vel[1].copy(), vel[2].copy()
    for i in range(3):
        vel[i] = m[i, 0] * v0 + m[i, 1] * v1 + m[i, 2] * v2


# ---------------------------------------------------------------------------
# connectivity and the ghost source coordinates
# ---------------------------------------------------------------------------
def ab_to_xyz(face, alpha, beta):
    """cs_ab_to_xyz: the unit vector of panel angles."""
    a, b = math.tan(alpha), math.tan(beta)
    x, y, z = [(1.0, a, b), (-a, 1.0, b), (-1.0, -a, b), (-b, a, 1.0), (a, -1.0, b), (b, a, -1.0)][face]
    norm = math.sqrt(x * x + y * y + z * z)
    return x / norm, y / norm, z / norm


def xyz_to_ab(face, v):
    """cs_xyz_to_ab on a named face."""
    x, y, z = v
    return [
        (math.atan2(y, x), math.atan2(z, x)),
        (math.atan2(-x, y), math.atan2(z, y)),
        (math.atan2(-y, -x), math.atan2(z, -x)),
        (math.atan2(y, z), math.atan2(-x, z)),
        (math.atan2(x, -y), math.atan2(z, -y)),
        (math.atan2(y, -z), math.atan2(x, -z)),
    ][face]


def ghost_usrc(n, nghost, face_t=0, side_t=0):
    """cs_build_ghost_usrc: usrc[depth - 1, j], the fractional along-edge
    cell-centre index on the source panel of ghost (depth, j)."""
    nface, nside, rev = EDGES[face_t][side_t]
    src_ax_y = nside in (SIDE_L, SIDE_R)
    d = math.pi / (2.0 * n)
    sgn = 1 if side_t in (SIDE_R, SIDE_T) else -1
    out = np.empty((nghost, n))
    for depth in range(1, nghost + 1):
        for j in range(n):
            along = -math.pi / 4.0 + (j + 0.5) * d
            perp = sgn * (math.pi / 4.0 + (depth - 0.5) * d)
            if side_t in (SIDE_L, SIDE_R):
                alpha_t, beta_t = perp, along
            else:
                alpha_t, beta_t = along, perp
            v = ab_to_xyz(face_t, alpha_t, beta_t)
            alpha_s, beta_s = xyz_to_ab(nface, v)
            u = ((beta_s if src_ax_y else alpha_s) + math.pi / 4.0) / d - 0.5
            if rev:
                u = (n - 1) - u
            out[depth - 1, j] = u
    return out


def _side_of(dy, dx):
    return {(0, -1): SIDE_L, (0, 1): SIDE_R, (-1, 0): SIDE_B, (1, 0): SIDE_T}[(dy, dx)]


DIRECTIONS = ((0, -1), (0, 1), (-1, 0), (1, 0))  # (dy, dx): L, R, B, T


def _offset_of(side):
    return DIRECTIONS[side]


def _part(g, offset, exterior, depth=None, ext2=0, ext3=0):
    """MeshBlockImpl::part for (dy, dx) on a (nc3, nc2, 1) panel."""
    dy, dx = offset
    out = []
    for off, ext in ((dy, ext3), (dx, ext2)):
        nx, ng = g.n, g.nghost
        ln = ng if depth is None else min(ng, depth)
        if off == -1:
            s = ng * (0 if exterior else 1)
            out.append(slice(s, s + ln))
        elif off == 0:
            out.append(slice(ng - ext, ng - ext + nx + 2 * ext))
        else:
            s = nx + ng * (1 if exterior else 0)
            out.append(slice(s, s + ln))
    return (slice(None), out[0], out[1], slice(0, 1))


def _mesh_points(g, sub, kind):
    """(alpha, beta, cosine) at the points of sub: cells, x2 faces or x3
    faces (strip_coords, cubed_sphere_layout.cpp:123-137)."""
    s3, s2 = sub[1], sub[2]
    if kind == "cell":
        x3, x2, cs = g.x3v, g.x2v, g.cos_c[:, :, 0]
    elif kind == "face2":
        x3, x2, cs = g.x3v, g.x2f, g.cos_f2[:, :, 0]
    else:
        x3, x2, cs = g.x3f, g.x2v, g.cos_f3[:, :, 0]
    beta = np.broadcast_to(x3[s3][:, None], (len(x3[s3]), len(x2[s2])))
    alpha = np.broadcast_to(x2[s2][None, :], beta.shape)
    cosine = np.broadcast_to(cs, (len(x3), len(x2)))[s3, s2]
    return alpha[..., None], beta[..., None], cosine[..., None]


def _send_strip(g, face, var, offset, kind, conserved, exterior, depth, widen):
    """serialize for one direction: the strip, to the geographic frame, then
    the reverse / flip / transpose of the connectivity; returns the flat
    buffer."""
    dy, dx = offset
    ext2 = g.nghost if (widen and dy != 0) else 0
    ext3 = g.nghost if (widen and dx != 0) else 0
    sub = _part(g, offset, exterior, depth, ext2, ext3)
    buf = var[sub].copy()
    alpha, beta, cosine = _mesh_points(g, sub, kind)
    if conserved is not None:
        m = rotation_matrix(face, alpha, beta, True, cosine if conserved else None)
        apply_matrix(buf[IVX : IVX + 3], m)
    my_side = _side_of(dy, dx)
    nface, nb_side, rev = EDGES[face][my_side]
    trans = (my_side - 1.5) * (nb_side - 1.5) < 0
    flip = (my_side % 2) == (nb_side % 2)
    if rev:
        buf = np.flip(buf, axis=2 if dy != 0 else 1)
    if flip:
        buf = np.flip(buf, axis=1 if dy != 0 else 2)
    if trans:
        buf = np.ascontiguousarray(np.swapaxes(buf, 1, 2))
    return nface, nb_side, np.ascontiguousarray(buf).ravel()


def _interp_ghost(g, buf, usrc_t, shift, axis):
    """GnomonicEquiangleImpl::_interp_ghost along `axis` (1 = x3, 2 = x2) of
    a (nvar, rows, cols, 1) strip."""
    u0f = np.floor(usrc_t)
    x = usrc_t - u0f
    u0 = u0f.astype(np.int64) + shift
    ln = buf.shape[axis]
    lo = max(0, shift)
    hi = min(ln - 1, shift + g.n - 1)

    def gather(idx):
        if axis == 1:  # usrc_t (n, nghost): rows along x3, cols are depth
            return np.stack([buf[:, idx[:, c], c, :] for c in range(idx.shape[1])], axis=2)
        return np.stack([buf[:, r, idx[r, :], :] for r in range(idx.shape[0])], axis=1)

    xx = x[None, :, :, None]
    if g.interp_order == 4 and hi - lo >= 3:
        s = np.clip(u0 - 1, lo, hi - 3)
        t = (u0 - s).astype(np.float64)[None, :, :, None] + xx
        c0 = -(t - 1.0) * (t - 2.0) * (t - 3.0) / 6.0
        c2 = -t * (t - 1.0) * (t - 3.0) / 2.0
        c3 = t * (t - 1.0) * (t - 2.0) / 6.0
        b0, b1, b2, b3 = gather(s), gather(s + 1), gather(s + 2), gather(s + 3)
        return b1 + c0 * (b0 - b1) + c2 * (b2 - b1) + c3 * (b3 - b1)
    b0, b1 = gather(u0), gather(u0 + 1)
    return b0 + xx * (b1 - b0)


def exchange_ghosts(g, us, conserved):
    """exchange_state over the six panels (one block each): every panel's
    four edge strips, interpolated along the edge, in the receiver's
    components.  us: list of six (nvar, nc3, nc2, 1) arrays, in place."""
    usrc = ghost_usrc(g.n, g.nghost)
    usrc_bt = usrc  # (nghost, n)
    usrc_lr = usrc.T.copy()  # (n, nghost)
    shift = g.nghost
    # phase: every panel serializes from the same state, then deserializes
    sends = {}
    for p in range(6):
        for offset in DIRECTIONS:
            nface, nb_side, flat = _send_strip(g, p, us[p], offset, "cell", conserved, False, None, True)
            sends[(nface, nb_side)] = flat
    for p in range(6):
        for offset in DIRECTIONS:
            dy, dx = offset
            side = _side_of(dy, dx)
            flat = sends[(p, side)]
            ext2 = g.nghost if dy != 0 else 0
            ext3 = g.nghost if dx != 0 else 0
            sub_w = _part(g, offset, True, None, ext2, ext3)
            shape = us[p][sub_w].shape
            us[p][sub_w] = flat.reshape(shape)
            sub = _part(g, offset, True)
            wide = us[p][sub_w].copy()
            if dy != 0:
                ut = usrc_bt if dy > 0 else usrc_bt[::-1]
                us[p][sub] = _interp_ghost(g, wide, ut, shift, 2)
            else:
                ut = usrc_lr if dx > 0 else usrc_lr[:, ::-1]
                us[p][sub] = _interp_ghost(g, wide, ut, shift, 1)
            alpha, beta, cosine = _mesh_points(g, sub, "cell")
            m = rotation_matrix(p, alpha, beta, False, cosine if conserved else None)
            vel = us[p][sub][IVX : IVX + 3].copy()
            apply_matrix(vel, m)
            v = us[p][sub]
            v[IVX : IVX + 3] = vel
            us[p][sub] = v


def exchange_face_states(g, wlr2, wlr3):
    """The cross-panel face-state sync (kPrimitive, no interpolation): each
    panel's interior-side state at an edge face replaces the neighbour's
    exterior-side state there, through the geographic frame.
    wlr2/wlr3: lists of six (2, nvar, nc3, nc2, 1) arrays, in place."""
    sends = {}
    for p in range(6):
        for offset in DIRECTIONS:
            dy, dx = offset
            up = dy > 0 or dx > 0
            kind = "face2" if dx != 0 else "face3"
            src = wlr2[p] if dx != 0 else wlr3[p]
            # "hydro_wl:+" travels up, "hydro_wr:-" down
            var = src[0] if up else src[1]
            nface, nb_side, flat = _send_strip(g, p, var, offset, kind, False, up, 1, False)
            sends[(nface, nb_side)] = flat
    for p in range(6):
        for offset in DIRECTIONS:
            dy, dx = offset
            up = dy > 0 or dx > 0
            side = _side_of(dy, dx)
            kind = "face2" if dx != 0 else "face3"
            dst = wlr2[p] if dx != 0 else wlr3[p]
            var = dst[1] if up else dst[0]  # received: "-" from up, "+" from down
            sub = _part(g, offset, up, 1)
            var[sub] = sends[(p, side)].reshape(var[sub].shape)
            alpha, beta, _ = _mesh_points(g, sub, kind)
            m = rotation_matrix(p, alpha, beta, False, None)
            vel = var[sub][IVX : IVX + 3].copy()
            apply_matrix(vel, m)
            v = var[sub]
            v[IVX : IVX + 3] = vel
            var[sub] = v


# ---------------------------------------------------------------------------
# the Riemann solve in the face frame
# ---------------------------------------------------------------------------
def prim2local2(w, g):
    c = g.cos_f2[:, : g.nc]
    s = g.sin_f2[:, : g.nc]
    gi22 = 1.0 / (s * s)
    t22 = 1.0 / np.sqrt(gi22)
    t32 = c / np.sqrt(np.ones_like(c))
    t33 = np.sqrt(np.ones_like(c))
    uu2, uu3 = w[IVY].copy(), w[IVZ].copy()
    w[IVZ] = t32 * uu2 + t33 * uu3
    w[IVY] = t22 * uu2


def prim2local3(w, g):
    c = g.cos_f3[: g.nc]
    s = g.sin_f3[: g.nc]
    gi33 = 1.0 / (s * s)
    one = np.ones_like(c)
    t22 = np.sqrt(one)
    t23 = c / np.sqrt(one)
    t33 = 1.0 / np.sqrt(gi33)
    uu2, uu3 = w[IVY].copy(), w[IVZ].copy()
    w[IVY] = t22 * uu2 + t23 * uu3
    w[IVZ] = t33 * uu3


def flux2global2(f, g):
    c = g.cos_f2[:, : g.nc]
    s = g.sin_f2[:, : g.nc]
    gi22 = 1.0 / (s * s)
    one = np.ones_like(c)
    t22 = np.sqrt(gi22)
    t32 = -np.sqrt(gi22) * c / one
    t33 = 1.0 / np.sqrt(one)
    txy, txz = f[IVY].copy(), f[IVZ].copy()
    f[IVZ] = t32 * txy + t33 * txz
    f[IVY] = t22 * txy
    ty, tz = f[IVY].copy(), f[IVZ].copy()
    f[IVY] = ty + tz * c
    f[IVZ] = tz + ty * c


def flux2global3(f, g):
    c = g.cos_f3[: g.nc]
    s = g.sin_f3[: g.nc]
    gi33 = 1.0 / (s * s)
    one = np.ones_like(c)
    t22 = 1.0 / np.sqrt(one)
    t23 = -c / one * np.sqrt(gi33)
    t33 = np.sqrt(gi33)
    txy, txz = f[IVY].copy(), f[IVZ].copy()
    f[IVY] = t22 * txy + t23 * txz
    f[IVZ] = t33 * txz
    ty, tz = f[IVY].copy(), f[IVZ].copy()
    f[IVY] = ty + tz * c
    f[IVZ] = tz + ty * c


def solve(wlr, tdim, g):
    """RiemannSolverImpl::solve for shallow-roe, dir yz: to the face frame,
    the solver, back to covariant global rows."""
    wl, wr = wlr[0].copy(), wlr[1].copy()
    if tdim == 2:
        prim2local2(wl, g)
        prim2local2(wr, g)
    else:
        prim2local3(wl, g)
        prim2local3(wr, g)
    f = swe.shallow_roe(wl, wr, tdim, dir_yz=True)
    if tdim == 2:
        flux2global2(f, g)
    else:
        flux2global3(f, g)
    return f


# ---------------------------------------------------------------------------
# seams
# ---------------------------------------------------------------------------
def _seam_operands(g, side):
    """(metric, cosine) at a side's interior face points, each (n, 1)."""
    ng, n = g.nghost, g.n
    plane = (ng, ng + n, ng, ng + n)[side]
    if side < 2:
        alpha = np.full((n, 1), g.x2f[plane])
        beta = g.x3v[ng : ng + n].reshape(n, 1)
        cosine = g.cos_f2[ng : ng + n, plane, :].reshape(n, 1)
    else:
        alpha = g.x2v[ng : ng + n].reshape(n, 1)
        beta = np.full((n, 1), g.x3f[plane])
        cosine = g.cos_f3[plane, ng : ng + n, :].reshape(n, 1)
    return metric_at(alpha, beta), cosine


def _slab(t, g, side):
    ng, n = g.nghost, g.n
    plane = (ng, ng + n, ng, ng + n)[side]
    if side < 2:
        return t[:, ng : ng + n, plane, :]
    return t[:, plane, ng : ng + n, :]


def reconcile_seams(g, f2, f3):
    """reconcile_panel_seams: on each of the twelve seams the lower panel
    owns the face flux; the other panel takes it through the global
    Cartesian frame (reversed along the edge and negated where the
    connectivity says so).  f2/f3: lists of six flux arrays, in place."""
    payload = {}
    for p in range(6):
        for side in range(4):
            peer, peer_side, rev = EDGES[p][side]
            if p < peer:
                t = f2[p] if side < 2 else f3[p]
                rows = _slab(t, g, side).copy()
                m, cosine = _seam_operands(g, side)
                vel = rows[IVX : IVX + 3]
                vec_raise(vel, cosine)
                contra_to_cart(vel, m, p)
                rows[IVX : IVX + 3] = vel
                payload[(p, side)] = rows
    new = {}
    for p in range(6):
        for side in range(4):
            peer, peer_side, rev = EDGES[p][side]
            if p < peer:
                continue
            gg = payload[(peer, peer_side)].copy()
            if rev:
                gg = gg[:, ::-1].copy()
            if (side % 2) == (peer_side % 2):
                gg = -gg
            m, cosine = _seam_operands(g, side)
            vel = gg[IVX : IVX + 3]
            cart_to_contra(vel, m, p)
            vec_lower(vel, cosine)
            gg[IVX : IVX + 3] = vel
            new[(p, side)] = gg
    for (p, side), gg in new.items():
        t = f2[p] if side < 2 else f3[p]
        ng, n = g.nghost, g.n
        plane = (ng, ng + n, ng, ng + n)[side]
        if side < 2:
            t[:, ng : ng + n, plane, :] = gg
        else:
            t[:, plane, ng : ng + n, :] = gg


# ---------------------------------------------------------------------------
# EOS, divergence, CFL
# ---------------------------------------------------------------------------
def cons2prim(u, g):
    w = np.empty_like(u)
    w[IDN] = u[IDN]
    w[IVX : IVZ + 1] = u[IVX : IVZ + 1] / u[IDN]
    vec_raise(w[IVX : IVZ + 1], g.cos_c)
    return w


def prim2cons(w, g):
    u = np.empty_like(w)
    u[IDN] = w[IDN]
    u[IVX : IVZ + 1] = w[IVX : IVZ + 1] * w[IDN]
    vec_lower(u[IVX : IVZ + 1], g.cos_c)
    return u


def divergence(g, w, f2, f3):
    """CoordinateImpl::divergence with the gnomonic areas and volume, then
    GnomonicEquiangleImpl::forward's shallow-water metric sources."""
    ng, n = g.nghost, g.n
    sj, ej = ng, ng + n
    d = np.zeros_like(f2)
    a2 = g.area2[None]
    a3 = g.area3[None]
    d[:, :, sj:ej] += a2[:, :, sj + 1 : ej + 1] * f2[:, :, sj + 1 : ej + 1] - a2[:, :, sj:ej] * f2[:, :, sj:ej]
    d[:, sj:ej] += a3[:, sj + 1 : ej + 1] * f3[:, sj + 1 : ej + 1] - a3[:, sj:ej] * f3[:, sj:ej]
    d = d / g.vol[None]
    pr = 0.5 * (w[IDN] * w[IDN])
    rho = w[IDN]
    v2, v3 = w[IVY], w[IVZ]
    sine2 = g.sin_c * g.sin_c
    d[IVY] -= g.x_ov_rD * (pr + rho * v3 * v3 * sine2)
    d[IVZ] -= g.y_ov_rC * (pr + rho * v2 * v2 * sine2)
    return d


def x1_mass_source(g, w, dt):
    """GnomonicEquiangleImpl::x1_mass_source_ with m1 = rho v1 into a zero
    array: du[v_i] -= dt m1 v_i / r with the covariant v_2 = v2 + c v3,
    v_3 = v3 + c v2."""
    src = np.zeros_like(w)
    c = g.cos_c
    mdt = (-dt) * ((w[IDN] * w[IVX]) / g.x1v[None, None, :])
    src[IVY] = src[IVY] + mdt * (w[IVY] + w[IVZ] * c)
    src[IVZ] = src[IVZ] + mdt * (w[IVZ] + w[IVY] * c)
    return src


def max_time_step(g, w):
    """HydroImpl::max_time_step on one panel (x2 and x3 resolved)."""
    sub = g.interior()
    cs = np.sqrt(w[IDN])
    dt2 = (g.width2 / (np.abs(w[IVY]) + cs))[sub].min()
    dt3 = (g.width3 / (np.abs(w[IVZ]) + cs))[sub].min()
    return min(1.0e9, dt2, dt3)


# ---------------------------------------------------------------------------
# the six-panel mesh
# ---------------------------------------------------------------------------
@dataclass
class Mesh:
    """Mesh::forward over the six panels, one block each (the stage-entry
    ghost exchange is done at the end of the previous stage and at
    initialisation, which is the same order of operations)."""

    grid: Grid
    shock: bool = False
    cfl: float = 0.9
    u: list = None
    u0: list = None
    w: list = None
    flux: dict = field(default_factory=dict)

    def initialize(self, ws):
        g = self.grid
        ws = [np.array(w, dtype=np.float64) for w in ws]
        exchange_ghosts(g, ws, conserved=False)
        self.w = ws
        self.u = [prim2cons(w, g) for w in ws]
        exchange_ghosts(g, self.u, conserved=True)
        return self.u

    def fluxes(self):
        """The stage's x2/x3 face fluxes of all six panels (after the face
        state sync and the seam reconciliation)."""
        g = self.grid
        self.w = [cons2prim(u, g) for u in self.u]
        wlr2 = [fv.reconstruct(w, 2, self.shock) for w in self.w]
        wlr3 = [fv.reconstruct(w, 1, self.shock) for w in self.w]
        exchange_face_states(g, wlr2, wlr3)
        f2 = [solve(wlr2[p], 2, g) for p in range(6)]
        f3 = [solve(wlr3[p], 1, g) for p in range(6)]
        reconcile_seams(g, f2, f3)
        self.flux = {"x2": f2, "x3": f3, "wlr2": wlr2, "wlr3": wlr3}
        return f2, f3

    def forward(self, dt, stage):
        g = self.grid
        if stage == 0:
            self.u0 = [u.copy() for u in self.u]
        f2, f3 = self.fluxes()
        sub = (slice(None),) + g.interior()
        for p in range(6):
            div = divergence(g, self.w[p], f2[p], f3[p])
            du = np.zeros_like(div)
            du[sub] = -dt * div[sub]
            src = x1_mass_source(g, self.w[p], dt)
            du[sub] += src[sub]
            self.u[p] = fv.rk_combine(self.u0[p], self.u[p], du, stage)
        exchange_ghosts(g, self.u, conserved=True)
        self._w_stale = True

    def check_redo(self):
        """w = prim(u) of the accepted state on every panel
        (meshblock_redo.cpp:107-112)."""
        self.w = [cons2prim(u, self.grid) for u in self.u]
        self._w_stale = False
        return 0

    def max_time_step(self):
        if getattr(self, "_w_stale", False):
            self.check_redo()
        return self.cfl * min(max_time_step(self.grid, w) for w in self.w)

    def mass(self):
        """The total mass, sum over panels of phi times cell volume."""
        sub = self.grid.interior()
        return sum(float((u[IDN][sub] * self.grid.vol[sub]).sum()) for u in self.u)


def lonlat(face, alpha, beta):
    """cs_ab_to_lonlat (src/coord/cubed_sphere_utils.cpp:115-146)."""
    x = _map(math.tan, alpha)
    y = _map(math.tan, beta)
    r = np.sqrt(x * x + y * y + 1)
    if face in (0, 1, 2, 4):
        lon = np.asarray(alpha, dtype=np.float64) + (0.0, 0.5 * math.pi, math.pi, None, 1.5 * math.pi)[face]
        lat = _map(math.atan, y / np.sqrt(1.0 + x * x))
    elif face == 3:
        lon = np.vectorize(math.atan2)(x, -y)
        lat = _map(math.asin, 1.0 / r)
    else:
        lon = np.vectorize(math.atan2)(x, y)
        lat = -_map(math.asin, 1.0 / r)
    lon = lon + np.where(lon < 0.0, 2.0 * math.pi, 0.0)
    return lon, lat
