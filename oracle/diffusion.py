"""CPU FP64 restatement of isotropic Cartesian diffusion (test infrastructure
only: the product path never imports this module).

Follows DiffusionImpl (src/forcing/diffusion.cpp) for a Cartesian block on a
uniform grid, kinematic coefficients (dynamic: false), no x1 profiles, no
prescribed wall temperatures, stress-free walls:

* ``cartesian_fluxes``   -- DiffusionImpl::cartesian_fluxes (:557-706):
  viscous stress with the -(2/3) div v bulk term, the face-mean shear term,
  the face density (wall faces of a reflecting x1 wall extrapolated from the
  two nearest active cells, :138-185), the viscous work in the energy row and
  the Fourier flux -kappa * (rho cv)_face * dT/dn.
* ``post_fluxes``        -- the energy row of every face flux added to the
  hydro flux totals, then the prescribed wall faces zeroed again
  (diffusion_cubed_sphere.cpp:365-379; hydro_forward.cpp:930-935).
* ``add_momentum``       -- DiffusionImpl::add (:532-547): du[0:IPR] -= dt *
  div(F)[0:IPR] (the energy row enters only through the posts).
* ``max_time_step``      -- DiffusionImpl::max_time_step (:708-756), the
  kinematic branch: dx_min^2 / (2 ndim max(nu, kappa)).

Cell widths and centre distances are the uniform dx of each axis.
"""

from __future__ import annotations

import numpy as np

from .fv import IDN, IPR, IVX, divergence, prescribe_walls

RGAS = 8.31446261815324  # kintera::constants::Rgas (J/mol/K)

# idir 0, 1, 2 (x1, x2, x3) -> numpy axis of a (nc3, nc2, nc1) field
_AXIS = {0: 2, 1: 1, 2: 0}


def _shift(a, axis, by):
    """a[..., i + by, ...] along `axis` (periodic roll; only interior-safe
    index ranges are read from the result)."""
    return np.roll(a, -by, axis=axis)


def _cd(v, grid, idir):
    """centered_derivative (:73-95): (v[i+1] - v[i-1]) / (2 dx)."""
    ax = _AXIS[idir]
    return (_shift(v, ax, 1) - _shift(v, ax, -1)) / (2.0 * grid.dx(idir + 1))


def _face_average(v, idir):
    """face_average at face i (lower face of cell i): 0.5 (v[i] + v[i-1])
    (face_mean :109-124, uniform branch)."""
    return 0.5 * (v + _shift(v, _AXIS[idir], -1))


def _face_normal_derivative(v, grid, idir):
    """(v[i] - v[i-1]) / dx (:97-107)."""
    return (v - _shift(v, _AXIS[idir], -1)) / grid.dx(idir + 1)


def _face_coefficient(v, grid, idir, wall_lo, wall_hi):
    """face_coefficient (:164-185): the two-cell average, a reflecting x1
    wall face extrapolated from the two nearest active cells,
    (1 + t) v_near - t v_next with t = 1/2, falling back to v_near when the
    extrapolation is not positive (extrapolate_to_wall :138-162)."""
    out = _face_average(v, idir)
    if idir != 0 or grid.nx1 + 1 < 3:
        return out
    lo, hi = grid.nghost, grid.nghost + grid.nx1  # faces lo .. hi
    t = 0.5
    if wall_lo:
        near, nxt = v[..., lo], v[..., lo + 1]
        wall = (1.0 + t) * near - t * nxt
        out[..., lo] = np.where(wall > 0.0, wall, near)
    if wall_hi:
        near, nxt = v[..., hi - 1], v[..., hi - 2]
        wall = (1.0 + t) * near - t * nxt
        out[..., hi] = np.where(wall > 0.0, wall, near)
    return out


def active_dirs(grid):
    return [d for d in (0, 1, 2) if grid.nc(d + 1) > 1]


def cartesian_fluxes(w, grid, nu, kappa, gammad, weight, bcs):
    """Face fluxes {axis: F (5, nc3, nc2, nc1)} of the active axes; F at
    index i of an axis is the flux through the lower face of cell i.  Only
    the interior faces (all faces along the axis, interior across it) are
    meaningful; the rest is zero."""
    rd = RGAS / weight
    cv = rd / (gammad - 1.0)
    act = active_dirs(grid)
    temp = w[IPR] / (w[IDN] * rd)
    rho_cv = w[IDN] * np.full_like(temp, cv)
    div_vel = None
    if nu > 0.0:
        div_vel = np.zeros_like(w[IDN])
        for d in act:
            div_vel = div_vel + _cd(w[IVX + d], grid, d)
    wall_lo = bcs[0] == "reflecting"
    wall_hi = bcs[1] == "reflecting"
    out = {}
    for d in act:
        wl, wh = (wall_lo, wall_hi) if d == 0 else (False, False)
        flux = np.zeros_like(w)
        rho_face = _face_coefficient(w[IDN], grid, d, wl, wh)
        if nu > 0.0:
            div_face = _face_average(div_vel, d)
            for ivar in range(3):
                stress = _face_normal_derivative(w[IVX + ivar], grid, d)
                if ivar == d:
                    stress = 2.0 * stress - (2.0 / 3.0) * div_face
                elif ivar in act:
                    cd = _cd(w[IVX + d], grid, ivar)
                    stress = stress + 0.5 * (cd + _shift(cd, _AXIS[d], -1))
                mflux = -nu * rho_face * stress
                flux[IVX + ivar] = mflux
                flux[IPR] = flux[IPR] + _face_average(w[IVX + ivar], d) * mflux
        if kappa > 0.0:
            dtdn = _face_normal_derivative(temp, grid, d)
            flux[IPR] = flux[IPR] - kappa * _face_coefficient(rho_cv, grid, d, wl, wh) * dtdn
        out[d + 1] = _restrict_faces(flux, grid, d)
    return out


def _restrict_faces(flux, grid, d):
    """Zero everything outside the face region of axis d+1: faces lo..hi
    along the axis, interior cells across it."""
    k, j, i = grid.interior()
    sl = [k, j, i]
    ax = _AXIS[d]
    s = sl[ax]
    sl[ax] = slice(s.start, s.stop + 1)
    out = np.zeros_like(flux)
    idx = (slice(None),) + tuple(sl)
    out[idx] = flux[idx]
    return out


def post_fluxes(totals, fdiff, grid, bcs):
    """The energy row of the diffusion fluxes added to the hydro totals, the
    wall faces prescribed again."""
    tdim = {1: 3, 2: 2, 3: 1}
    for axis, f in fdiff.items():
        totals[axis][IPR] = totals[axis][IPR] + f[IPR]
        prescribe_walls(totals[axis], tdim[axis], bcs, axis, grid)


def add_momentum(du, fdiff, grid, dt):
    """du[0:IPR] -= dt * div(F)[0:IPR] on the interior cells."""
    div = divergence(grid, fdiff)
    sub = (slice(0, IPR),) + grid.interior()
    du[sub] = du[sub] - dt * div[sub]


def max_time_step(grid, nu, kappa):
    act = active_dirs(grid)
    if not act:
        return float("inf")
    coeff = max(nu, kappa)
    if coeff == 0.0:
        return float("inf")
    dx_min = min(grid.dx(d + 1) for d in act)
    return dx_min * dx_min / (2.0 * len(act) * coeff)
