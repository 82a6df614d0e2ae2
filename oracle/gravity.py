"""CPU FP64 oracle of the gravity path of the x1 sweep (numpy).

TEST INFRASTRUCTURE ONLY (see oracle/fv.py).  Restates, for a uniform
Cartesian block that owns its whole x1 column (nb1 = 1), conversion order 2,
no species, no solids, FP64 storage:

  * the hydrostatic x1 reference: hydro_ref_x1_impl
    (src/hydro/hydro_ref_x1_impl.h:24-272) and X1Reference::evaluate
    (src/recon/x1_reference.cpp:75-250);
  * the well-balanced x1 reconstruction WellBalancedX1::forward
    (src/recon/well_balanced_x1.cpp:16-154);
  * const gravity ConstGravityImpl::forward_split
    (src/forcing/const_gravity.cpp:30-55);
  * the face-based gravity work and the corrected potential-energy work
    (src/hydro/hydro_forward.cpp:1049-1124; src/hydro/gravity_work.hpp).

Pinning: tests/test_oracle_gravity.py checks the reference scan against the
discrete hydrostatic identity psf_lo - psf_hi = g rho dx, the six-face cell
reference against the closed form of an isothermal column, the rest state
(zero flux divergence plus source at hydrostatic equilibrium: the
well-balanced property the reference asserts, tests/test_fp32_policy.py:65),
the slope stencil against exact derivatives of quadratics, the face work
against its telescoping sum, and the face work plus the corrected
potential-energy term against the exact gravity work of a cubic mass flux
(Gauss quadrature).
"""

from __future__ import annotations

import numpy as np

from . import fv

IDN, IVX, IVY, IVZ, IPR = fv.IDN, fv.IVX, fv.IVY, fv.IVZ, fv.IPR
ILT, IRT = fv.ILT, fv.IRT
DBL_MIN = np.finfo(np.float64).tiny

W6 = (11.0 / 1440.0, -31.0 / 480.0, 401.0 / 720.0, 401.0 / 720.0, -31.0 / 480.0, 11.0 / 1440.0)
W6E = (
    (95.0 / 288.0, 1427.0 / 1440.0, -133.0 / 240.0, 241.0 / 720.0, -173.0 / 1440.0, 3.0 / 160.0),
    (-3.0 / 160.0, 637.0 / 1440.0, 511.0 / 720.0, -43.0 / 240.0, 77.0 / 1440.0, -11.0 / 1440.0),
)


def _wall_rop(r0, r1, k):
    """hydro_ref_x1_wall_rop (hydro_ref_x1_impl.h:82-90), integer k."""
    q = r0 / r1
    qk = q
    for _ in range(1, k):
        qk = qk * q
    e = np.where(r1 > r0, r0 * qk, r0 + k * (r0 - r1))
    return np.where(e > 0, e, r0)


def _rop_smooth(w, i, jlo, jhi, ext_lo, ext_hi):
    """hydro_ref_x1_rop_smooth (hydro_ref_x1_impl.h:104-134), uniform grid."""

    def rop(j):
        return w[IDN][..., j] / w[IPR][..., j]

    if not (ext_lo and i - 2 < jlo) and not (ext_hi and i + 2 > jhi):
        return rop(min(max(i, jlo), jhi))
    v = []
    for m in range(-2, 3):
        j = i + m
        if j < jlo:
            v.append(_wall_rop(rop(jlo), rop(jlo + 1), jlo - j) if ext_lo else rop(jlo))
        elif j > jhi:
            v.append(_wall_rop(rop(jhi), rop(jhi - 1), j - jhi) if ext_hi else rop(jhi))
        else:
            v.append(rop(j))
    return (v[0] + 4.0 * v[1] + 6.0 * v[2] + 4.0 * v[3] + v[4]) / 16.0


def hydro_ref_x1(w, dx1f, iu, grav, phys_in=True, phys_out=True, wall_clamp=True):
    """hydro_ref_x1_impl with a local anchor (nb1 = 1) on a uniform grid.
    w: (nvar, nc3, nc2, nc1); dx1f: (nc1,); grav > 0 (downward magnitude).
    Returns psf_lo, psf_hi, pref, dsf, dref, each (nc3, nc2, nc1)."""
    nc1 = w.shape[-1]
    rho = w[IDN]
    shp = rho.shape
    psf_lo = np.empty(shp)
    psf_hi = np.empty(shp)
    # scan (hydro_ref_x1_scan_impl, :57-78; pressure_scan_impl, :31-55)
    rdt_top = w[IPR][..., iu] / rho[..., iu]
    top = w[IPR][..., iu] * np.exp(-grav * 0.5 * dx1f[iu] / rdt_top)
    face = top
    for i in range(iu, -1, -1):
        dp = grav * rho[..., i] * dx1f[i]
        lo = face + dp
        hi = face
        psf_lo[..., i] = np.where(lo > DBL_MIN, lo, DBL_MIN)
        psf_hi[..., i] = np.where(hi > DBL_MIN, hi, DBL_MIN)
        face = lo
    face = top
    for i in range(iu + 1, nc1):
        dp = grav * rho[..., i] * dx1f[i]
        lo = face
        hi = lo - dp
        psf_lo[..., i] = np.where(lo > DBL_MIN, lo, DBL_MIN)
        psf_hi[..., i] = np.where(hi > DBL_MIN, hi, DBL_MIN)
        face = hi

    def facev(f):
        return psf_lo[..., f] if f < nc1 else psf_hi[..., nc1 - 1]

    pref = np.empty(shp)
    dsf = np.empty(shp)
    dref = np.empty(shp)
    il = nc1 - 1 - iu
    thin = (iu - il + 1) < 5
    fits = iu >= 4
    jlo = il if (wall_clamp and phys_in) else 0
    jhi = iu if (wall_clamp and phys_out) else nc1 - 1
    ext_lo = wall_clamp and phys_in and il + 1 <= iu
    ext_hi = wall_clamp and phys_out and iu - 1 >= il
    for i in range(nc1):
        # hydro_ref_x1_cell_impl (:155-240), uniform branch
        wall_in = wall_clamp and phys_in and il <= i < il + 2
        wall_out = wall_clamp and phys_out and iu - 2 < i <= iu
        clamp_in = wall_in and fits and not (thin and phys_out)
        clamp_out = wall_out and fits and not (thin and phys_in)
        lo = psf_lo[..., i]
        hi = psf_hi[..., i]
        cell = 0.5 * (lo + hi)
        lower = np.minimum(lo, hi)
        upper = np.maximum(lo, hi)

        def use_six(start, weights, reverse=False):
            nonlocal cell
            value = np.zeros(shp[:-1])
            for m in range(6):
                value = value + weights[5 - m if reverse else m] * facev(start + m)
            cell = np.where((value >= lower) & (value <= upper), value, cell)

        if 2 <= i < nc1 - 2 and not wall_in and not wall_out:
            use_six(i - 2, W6)
        if clamp_in:
            use_six(il, W6E[i - il])
        if clamp_out:
            s = iu + 1 - 5
            use_six(s, W6E[4 - (i - s)], True)
        if not phys_in and i < 2:
            use_six(0, W6E[i])
        if not phys_out and i >= nc1 - 2:
            use_six(nc1 - 5, W6E[4 - (i - (nc1 - 5))], True)
        pref[..., i] = cell
        rs = _rop_smooth(w, i, jlo, jhi, ext_lo, ext_hi)
        rf = rs
        if i > 0:
            rf = 0.5 * (_rop_smooth(w, i - 1, jlo, jhi, ext_lo, ext_hi) + rs)
        dref[..., i] = cell * rs
        dsf[..., i] = lo * rf
    return psf_lo, psf_hi, pref, dsf, dref


def x1_reference(w, grid, grav1, refl=(True, True), wall_clamp=True):
    """X1Reference::evaluate, nb1 = 1, conversion order 2, uniform weno5/cp5
    (x1_reference.cpp:75-250).  Returns psf_lo, pref, dsf_lr (2, ...), dref."""
    nc1 = w.shape[-1]
    ng = grid.nghost
    iu = nc1 - ng - 1
    dx1f = np.full(nc1, grid.dx(1))
    psf_lo, psf_hi, pref, dsf, dref = hydro_ref_x1(w, dx1f, iu, -grav1, True, True, wall_clamp)
    dsf_lr = np.stack([dsf, dsf])
    il, iu5 = 3, nc1 - 3
    lin = np.stack([dsf, dsf])
    outl, outr = fv.interp_apply(dref, dref.ndim - 1, "cp5")
    lin[IRT][..., il - 1 : iu5 + 1] = outl
    lin[ILT][..., il : iu5 + 2] = outr
    dl = dref[..., 0 : nc1 - 1]
    dr = dref[..., 1:nc1]
    n = iu5 + 3 - il
    for s in (ILT, IRT):
        face = lin[s][..., il - 1 : il - 1 + n]
        lo = np.minimum(dl, dr)[..., il - 2 : il - 2 + n]
        hi = np.maximum(dl, dr)[..., il - 2 : il - 2 + n]
        d = dsf_lr[s][..., il - 1 : il - 1 + n]
        dsf_lr[s][..., il - 1 : il - 1 + n] = np.where((face >= lo) & (face <= hi), face, d)
    for f in (0, 1):
        if not refl[f]:
            continue
        i = ng if f == 0 else iu + 1
        mean = 0.5 * (dsf_lr[ILT][..., i] + dsf_lr[IRT][..., i])
        dsf_lr[ILT][..., i] = mean
        dsf_lr[IRT][..., i] = mean
    return psf_lo, pref, dsf_lr, dref


def wb_reconstruct_x1(w, grid, grav1, shock, mirror=(True, True), refl=(True, True)):
    """WellBalancedX1::forward, non-split FP64 (well_balanced_x1.cpp:16-154),
    with its reference.  Returns the x1 face states wlr (2, nvar, ...)."""
    psf_lo, pref, dsf_lr, dref = x1_reference(w, grid, grav1, refl)
    wx1 = w.copy()
    density = w[IDN].copy()
    wx1[IPR] = wx1[IPR] - pref
    wx1[IDN] = wx1[IDN] - dref
    ng = grid.nghost
    nc1 = w.shape[-1]
    iu = nc1 - ng - 1
    il = ng
    for c in (IPR, IDN):
        if mirror[0]:
            wx1[c][..., il - ng : il] = wx1[c][..., il : il + ng][..., ::-1]
        if mirror[1]:
            wx1[c][..., iu + 1 : iu + 1 + ng] = wx1[c][..., iu + 1 - ng : iu + 1][..., ::-1]
    wtmp = fv.reconstruct(wx1, 3, shock)
    pl = wtmp[ILT][IPR] + psf_lo
    pr = wtmp[IRT][IPR] + psf_lo
    wtmp[ILT][IPR] = np.where(pl > 0.0, pl, psf_lo)
    wtmp[IRT][IPR] = np.where(pr > 0.0, pr, psf_lo)
    dl = wtmp[ILT][IDN] + dsf_lr[ILT]
    dr = wtmp[IRT][IDN] + dsf_lr[IRT]
    rho_below = np.concatenate([density[..., 0:1], density[..., 0 : nc1 - 1]], axis=-1)
    wtmp[ILT][IDN] = np.where(dl > 0.0, dl, rho_below)
    wtmp[IRT][IDN] = np.where(dr > 0.0, dr, density)
    return wtmp


# ---------------------------------------------------------------------------
# gravity work (src/hydro/gravity_work.hpp)
# ---------------------------------------------------------------------------
def centroid_slope_weights(x):
    """centroid_slope_stencil (gravity_work.hpp:88-117): (idx, wt) of the
    three-point slope at each of the n points x."""
    n = x.shape[0]
    hm = x[1 : n - 1] - x[0 : n - 2]
    hp = x[2:n] - x[1 : n - 1]
    a = x[1:2] - x[0:1]
    b = x[2:3] - x[1:2]
    c = x[n - 2 : n - 1] - x[n - 3 : n - 2]
    d = x[n - 1 : n] - x[n - 2 : n - 1]
    wt0 = np.concatenate([-(2.0 * a + b) / (a * (a + b)), -hp / (hm * (hm + hp)), d / (c * (c + d))])
    wt1 = np.concatenate([(a + b) / (a * b), (hp - hm) / (hm * hp), -((c + d) / (c * d))])
    wt2 = np.concatenate([-(a / ((a + b) * b)), hm / (hp * (hm + hp)), (c + 2.0 * d) / ((c + d) * d)])
    j = np.clip(np.arange(n), 1, n - 2) - 1
    return (j, j + 1, j + 2), (wt0, wt1, wt2)


def centroid_slope(q, x):
    idx, wt = centroid_slope_weights(x)
    s = wt[0] * q[..., idx[0]]
    s = s + wt[1] * q[..., idx[1]]
    s = s + wt[2] * q[..., idx[2]]
    return s


def x1_variance_plain(x1f):
    n = x1f.shape[0] - 1
    h = x1f[1 : n + 1] - x1f[0:n]
    return h * h / 12.0


def gravity_energy_correction(grid, w, flux1, dt, grav1, non_hydrostatic=1.0):
    """gravity_energy_correction of HydroImpl::forward, Cartesian, nb1 = 1
    (hydro_forward.cpp:1049-1124): face_gravity_work + corrected PE work -
    the cell-centred gravity work, on the interior x1 cells (all x2/x3)."""
    ng = grid.nghost
    nc1 = w.shape[-1]
    is_, ie = ng, nc1 - ng
    x1f = grid.faces(1)
    x1v = grid.centers(1)
    a1 = grid.dx(3) * grid.dx(2)
    vol = grid.dx(3) * (grid.dx(2) * grid.dx(1))
    vmf1 = flux1[IDN].copy()
    vmd = (a1 * vmf1[..., is_ + 1 : ie + 1] - a1 * vmf1[..., is_:ie]) / vol
    face_mass = a1 * vmf1
    centres = x1v[is_:ie]
    hi = grav1 * (x1f[is_ + 1 : ie + 1] - centres)
    lo = grav1 * (x1f[is_:ie] - centres)
    gw = (hi * face_mass[..., is_ + 1 : ie + 1] - lo * face_mass[..., is_:ie]) / vol
    fgw = dt * gw
    drho = -dt * vmd
    gvar = grav1 * x1_variance_plain(x1f[is_ : ie + 1])
    fgw = fgw + gvar * centroid_slope(drho, x1v[is_:ie])
    ogw = dt * w[IDN][..., is_:ie] * w[IVX][..., is_:ie] * grav1 * non_hydrostatic
    return fgw - ogw


def const_gravity(du, w, dt, grav1, non_hydrostatic=1.0):
    """ConstGravityImpl::forward(_split) with rho1 = w[IDN]
    (const_gravity.cpp:30-55), on the whole array."""
    du[IVX] = du[IVX] + dt * w[IDN] * grav1 * non_hydrostatic
    du[IPR] = du[IPR] + dt * w[IDN] * w[IVX] * grav1 * non_hydrostatic
    return du
