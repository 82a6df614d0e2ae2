"""Passive scalars (tracers) of a Cartesian block: the oracle restatement of
ScalarImpl::forward (src/scalar/scalar.cpp:104-218), the implicit tracer
transfer with its donor cap (MeshBlockImpl::stage_advance,
src/mesh/meshblock_advance.cpp:212-236; donor_cap_cpu,
src/implicit/implicit_dispatch.cpp:54-75), the scalar RK combine and
primitive (meshblock_advance.cpp:338-341; set_scalar_primitive,
src/mesh/meshblock.cpp:67-79) and the scalar boundary fill (kScalar,
src/bc/bc_func.cpp:10-35).

Test infrastructure: only tests/, the smoke check and the bench's CPU leg use
it, as the checker.

Scope: dry ideal gas (the dry fraction is 1), no upper bound, no dry sources,
no radiating boundary for the tracers.  The positivity limiter is the
scalars' instance of PositivityLimiter (src/hydro/flux_positivity.cpp:
39-268) without the energy carry and the settling part.

Pins (tests/test_oracle_scalar.py): the scaled WENO5 against the unscaled
one at vscale and its exact homogeneity; a uniform mixing ratio stays
uniform, and r = 1 follows the gas density, through the face ratio, the
covariance, the positivity limiter and the donor cap; conservation of
sum s V with reflecting and periodic walls; the donor cap against a brute
force of its sequential rule; theta against its closed form on a one-cell
drain; the covariance row against the product formula on a linear field.
"""

from __future__ import annotations

import numpy as np

from . import fv
from .fv import ILT, IRT, IDN, IVY, IVZ

EPS_WENO = 1e-6
ROUNDOFF_ULP = 4096.0


# ---------------------------------------------------------------------------
# reconstruction: WENO5 with scale = true, faces clamped at 0
# ---------------------------------------------------------------------------
def weno5_point_scaled(phi_src, c):
    """weno5_point with scale = true (recon_point.h:69-83, 113-147): vscale =
    mean |src| over the 5 points, the weights of src / vscale, 0 where
    vscale = 0."""
    T = phi_src[0].dtype.type
    s = np.abs(phi_src[0])
    for k in range(1, 5):
        s = s + np.abs(phi_src[k])
    vscale = s / T(5)
    with np.errstate(divide="ignore", invalid="ignore"):
        safe = np.where(vscale != 0, vscale, T(1))
        phi = [p / safe for p in phi_src]
    c = c.astype(phi[0].dtype)
    rows = [c[5 * r : 5 * r + 5] for r in range(9)]
    d = c[45:48]
    dot = fv._vvdot
    p0, p1, p2 = dot(phi, rows[0]), dot(phi, rows[1]), dot(phi, rows[2])
    sq = lambda x: x * x  # noqa: E731
    b0 = T(13.0 / 12.0) * sq(dot(phi, rows[3])) + T(0.25) * sq(dot(phi, rows[4]))
    b1 = T(13.0 / 12.0) * sq(dot(phi, rows[5])) + T(0.25) * sq(dot(phi, rows[6]))
    b2 = T(13.0 / 12.0) * sq(dot(phi, rows[7])) + T(0.25) * sq(dot(phi, rows[8]))
    a0 = d[0] / sq(b0 + T(EPS_WENO))
    a1 = d[1] / sq(b1 + T(EPS_WENO))
    a2 = d[2] / sq(b2 + T(EPS_WENO))
    out = vscale * (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2)
    return np.where(vscale != 0, out, T(0))


def recon(q, dim):
    """ReconstructImpl::forward of a module without a hydro parent
    (reconstruct.cpp:262-266): every row with the scaled WENO5 on faces
    il..iu+1 (the ghost faces copy the end faces, _apply_inplace), then
    clamped at 0.  q: (nrow, nc3, nc2, nc1); dim the tensor axis (3 = x1).
    Returns (2, nrow, nc3, nc2, nc1): [ILT] the left state of face i (from
    cell i - 1), [IRT] the right state (from cell i)."""
    n = q.shape[dim]
    ng = 3
    il, iu = ng, n - ng
    win = fv._windows(q, dim, n - 4)
    outl = weno5_point_scaled(win, fv.WENO5_CM)
    outr = weno5_point_scaled(win, fv.WENO5_CP)
    out = np.zeros((2,) + q.shape, dtype=q.dtype)
    sl = lambda s: fv._sl(q.ndim, dim, s)  # noqa: E731
    out[IRT][sl(slice(il - 1, iu + 1))] = outl
    out[ILT][sl(slice(il, iu + 2))] = outr
    for s in (IRT, ILT):
        lo = out[s][sl(slice(il, il + 1))].copy()
        hi = out[s][sl(slice(iu, iu + 1))].copy()
        out[s][sl(slice(0, il))] = lo
        out[s][sl(slice(iu + 1, n))] = hi
    return np.maximum(out, 0.0)


def face_r(rho, s, dim):
    """face_r (scalar.cpp:114-119): r = q_face / rho_face where the
    reconstructed density is positive, else the donor cell's r (cell i - 1
    on the left, cell i on the right)."""
    rf = recon(rho, dim)
    sf = recon(s, dim)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = s / rho
        donor = np.stack([np.roll(r, 1, axis=dim), r])
        return np.where(rf > 0.0, sf / rf, donor)


def upwind(rl, rr, mass):
    """The upwind solver (riemann.cpp:164-171): mass (ui rl + (1 - ui) rr),
    ui = [mass > 0]."""
    ui = (mass > 0.0).astype(rl.dtype)
    return mass * (ui * rl + (1.0 - ui) * rr)


def covariance(rl, rr, wl, wr, dim, grid):
    """HydroImpl::scalar_flux_covariance, Cartesian dry (hydro_forward.cpp:
    233-297): sigma1^2 rho D1[u_n] D1[q] on the x1 cells 1..n1-2 of the x2
    (dim 2) or x3 (dim 1) faces, with q = (rl + rr)/2 and wbar = (wl + wr)/2
    of the hydro face states."""
    n1 = rl.shape[-1]
    q = 0.5 * (rl * 1.0 + rr * 1.0)
    wbar = 0.5 * (wl + wr)
    un = wbar[IVY if dim == 2 else IVZ]
    x1v = grid.centers(1)
    dx1f = np.full(n1, grid.dx(1))
    s2 = dx1f * dx1f / 12.0
    rho = wbar[IDN]
    dx1 = x1v[2:n1] - x1v[0 : n1 - 2]
    cell = lambda a: a[..., 1 : n1 - 1]  # noqa: E731
    d1 = lambda a: (a[..., 2:n1] - a[..., 0 : n1 - 2]) / dx1  # noqa: E731
    dflx = np.zeros_like(q)
    dflx[..., 1 : n1 - 1] = cell(s2) * cell(rho) * d1(un) * d1(q)
    return dflx


def prescribe_walls(flux, axis, bcs, grid):
    """boundary.prescribed on the scalar register (scalar.cpp:165-174): every
    tracer row is 0 on the reflecting planes of direction `axis`."""
    nx = (grid.nx1, grid.nx2, grid.nx3)[axis - 1]
    dim = 4 - axis
    for side, plane in ((0, grid.nghost), (1, grid.nghost + nx)):
        if bcs[2 * (axis - 1) + side] == "reflecting":
            flux[fv._sl(4, dim, slice(plane, plane + 1))] = 0.0


# ---------------------------------------------------------------------------
# positivity limiter (flux_positivity.cpp), the scalars' instance
# ---------------------------------------------------------------------------
def positivity_theta(u, fluxes, grid, dt):
    """flux_positivity_theta (flux_positivity.cpp:39-78): theta_i =
    min(1, avail_i / (dt out_i)), out_i the outgoing A F over the faces of
    interior cell i, avail = relu(u) V (1 - 4096 eps); 1 where nothing
    drains (ghosts included).  fluxes: {axis: F}."""
    out = np.zeros_like(u)
    k, j, i = grid.interior()
    area = {1: grid.dx(3) * grid.dx(2), 2: grid.dx(3) * grid.dx(1), 3: grid.dx(2) * grid.dx(1)}
    for axis in (1, 2, 3):
        if axis not in fluxes:
            continue
        dim = 4 - axis
        lo, hi = ((i.start, i.stop - 1), (j.start, j.stop - 1), (k.start, k.stop - 1))[axis - 1]
        nf = hi - lo + 2
        af = area[axis] * fluxes[axis][fv._sl(4, dim, slice(lo, lo + nf))]
        upper = np.maximum(af[fv._sl(4, dim, slice(1, nf))], 0.0)
        lower = np.maximum(-af[fv._sl(4, dim, slice(0, nf - 1))], 0.0)
        out[fv._sl(4, dim, slice(lo, hi + 1))] += upper + lower
    margin = ROUNDOFF_ULP * np.finfo(u.dtype).eps
    vol = grid.dx(3) * (grid.dx(2) * grid.dx(1))
    avail = np.maximum(u, 0.0) * vol * (1.0 - margin)
    drain = out * dt
    tiny = np.finfo(drain.dtype).tiny
    with np.errstate(over="ignore"):
        return np.where(drain > 0.0, np.minimum(avail / np.maximum(drain, tiny), 1.0), 1.0)


def positivity_share(q):
    """flux_positivity_share (flux_positivity.cpp:80-88): phi = clamp(min of
    the neighbours' q / q, 0, 1) along x1 and every resolved x2/x3, 0 where
    q <= 0 (np.roll: the tensor's roll, wrapping)."""
    low = np.minimum(np.roll(q, 1, 3), np.roll(q, -1, 3))
    for dim in (2, 1):
        if q.shape[dim] > 1:
            low = np.minimum(low, np.minimum(np.roll(q, 1, dim), np.roll(q, -1, dim)))
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(q > 0.0, np.clip(low / q, 0.0, 1.0), 0.0)


def positivity_relax(theta, td, phi):
    """flux_positivity_relax (flux_positivity.cpp:143-149)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where((td < 1.0) & (td > 0.0), theta + phi * (np.minimum(theta / td, 1.0) - theta), theta)


def positivity_factor(theta, flux, axis, grid):
    """flux_positivity_factor (flux_positivity.cpp:90-101): the donor's
    theta on faces lo..hi+1 (the lower cell where F > 0), 1 elsewhere."""
    dim = 4 - axis
    k, j, i = grid.interior()
    lo, hi = ((i.start, i.stop - 1), (j.start, j.stop - 1), (k.start, k.stop - 1))[axis - 1]
    f = flux[fv._sl(4, dim, slice(lo, hi + 2))]
    factor = np.ones_like(flux)
    factor[fv._sl(4, dim, slice(lo, hi + 2))] = np.where(
        f > 0.0, theta[fv._sl(4, dim, slice(lo - 1, hi + 1))], theta[fv._sl(4, dim, slice(lo, hi + 2))]
    )
    return factor


def boundaries(q, bcs):
    """The boundary functions on a kScalar field (bc_func.cpp:10-35): the
    hydro's rules without the normal-velocity sign."""
    for f in range(6):
        dim = 3 - f // 2
        kind = bcs[f]
        if q.shape[dim] == 1:
            continue
        if kind == "outflow":
            raise ValueError("scalar: a radiating boundary is not restated for the tracers")
        fv.bc_gather(q, dim, kind, f % 2 == 1, flip=False)


def limit(s, rho, fluxes, mass, grid, dt, bcs):
    """PositivityLimiter::apply for the scalars (flux_positivity.cpp:186-268,
    scalar.cpp:193-205): theta of s relaxed by the dry-air theta and the
    neighbours' share, its ghosts filled as a scalar, each flux scaled by its
    donor's theta.  Returns the scaled fluxes and theta."""
    theta = positivity_theta(s, fluxes, grid, dt)
    td = positivity_theta(rho, {a: m for a, m in mass.items()}, grid, dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = positivity_share(s / rho)
    theta = positivity_relax(theta, td, phi)
    boundaries(theta, bcs)
    out = {a: f * positivity_factor(theta, f, a, grid) for a, f in fluxes.items()}
    return out, theta


# ---------------------------------------------------------------------------
# the stage
# ---------------------------------------------------------------------------
def fluxes(s, rho, hydro_flux, grid, bcs, w=None, shock23=False, cov=False):
    """The scalar face fluxes of one stage (scalar.cpp:114-174): the upwind
    flux on the hydro's dry-mass row of each direction, plus the covariance
    row on x2/x3 when the hydro's covariance is on (w: the hydro primitives
    whose x2/x3 face states it uses); walls prescribed.  Returns ({axis: F},
    {axis: mass row}, cov_used)."""
    out, mass = {}, {}
    used = False
    for axis in (1, 2, 3):
        if axis not in hydro_flux:
            continue
        dim = 4 - axis
        m = hydro_flux[axis][IDN][None]
        rlr = face_r(rho, s, dim)
        f = upwind(rlr[ILT], rlr[IRT], m)
        if axis > 1 and cov and grid.nc(1) >= 3:
            wlr = fv.reconstruct(w, dim, shock23)
            f = f + covariance(rlr[ILT], rlr[IRT], wlr[ILT], wlr[IRT], dim, grid)
            used = True
        prescribe_walls(f, axis, bcs, grid)
        out[axis] = f
        mass[axis] = m
    return out, mass, used


def donor_cap(P, held, il, iu):
    """donor_cap_cpu (implicit_dispatch.cpp:54-75): per column, sequentially
    upward, the transfer through the face below cell i capped by what the
    donor below holds (its own plus its inflow) and by what cell i holds."""
    P = P.copy()
    avail = held[..., il].copy()
    for i in range(il + 1, iu + 1):
        P[..., i] = np.maximum(np.minimum(P[..., i], avail), -held[..., i])
        avail = held[..., i] + P[..., i]
    return P


def implicit_transfer(ds, s, r, moved, grid):
    """The implicit tracer transfer (meshblock_advance.cpp:215-233): the
    donor's r times the dry-gas transfer M through the face below each cell
    (upward: the cell below), capped, added as (P - P_above) / V.  moved:
    (nc3, nc2, nc1) with M at the interior cells, 0 elsewhere."""
    nc1 = r.shape[-1]
    M = moved[None]
    r_below = np.concatenate([np.zeros_like(r[..., :1]), r[..., : nc1 - 1]], axis=-1)
    P = np.where(M > 0.0, r_below, r) * M
    vol = grid.dx(3) * (grid.dx(2) * grid.dx(1))
    held = np.maximum(s + ds, 0.0) * vol
    ng = grid.nghost
    P = donor_cap(P, held, ng, ng + grid.nx1 - 1)
    P_above = np.concatenate([P[..., 1:], np.zeros_like(P[..., :1])], axis=-1)
    return ds + (P - P_above) / vol
