"""CPU FP64 oracle of the Cartesian finite-volume stage (numpy).

TEST INFRASTRUCTURE ONLY.  This module restates the reference snapy CPU path
for the dry ideal-gas Cartesian decks: it is the checker the parity tests,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` compare
the HIP path against.  Nothing in the product imports it.

Each function follows the reference file:line it cites and keeps the order of
the reference's tensor operations, so an FP64 run agrees with the reference CPU
path to round-off.  Arrays are laid out like the reference tensors:
``(nvar, nc3, nc2, nc1)`` with x1 the last (contiguous) axis.

Pinning (tests/test_oracle_*.py):
  * weno5 / cp5 rows: the reference's independent closed forms
    (tests/interp_simple.hpp:110-150), polynomial exactness, ENO selection;
  * hllc: the reference's face-pressure closed form
    (tests/test_riemann.cpp:169-229), consistency F(w, w) = F(w), and the
    Sod exact solution for the whole scheme;
  * lmars: consistency, mass-flux upwinding, the formula of
    tests/test_fp32_typed_constants.cpp:264-270;
  * divergence: telescoping sums and exact linear fluxes;
  * SSP-RK3: the stability polynomial 1 + z + z^2/2 + z^3/6;
  * boundary functions and the CFL: brute-force loops;
  * the outflow (radiating) ghost: linear acoustics on moving states (a
    supersonic outflow copies the interior, a supersonic inflow keeps the
    background, a subsonic face keeps exactly the leaving acoustic, entropy
    and vortical waves), the mirror symmetry of its two faces and the
    0.99 floor scaling;
  * LayerHeat (integer depth): flux * area * dt of energy into the wall
    cells, evenly, nothing elsewhere.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

IDN, IVX, IVY, IVZ, IPR = 0, 1, 2, 3, 4
ILT, IRT = 0, 1

# ---------------------------------------------------------------------------
# reconstruction tables (src/recon/reconstruct.cpp:64-89)
# ---------------------------------------------------------------------------
_WENO5_ROWS = np.array(
    [
        [-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0, 0.0, 0.0],
        [0.0, 1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0, 0.0],
        [0.0, 0.0, 11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0],
        [1.0, -2.0, 1.0, 0.0, 0.0],
        [1.0, -4.0, 3.0, 0.0, 0.0],
        [0.0, 1.0, -2.0, 1.0, 0.0],
        [0.0, -1.0, 0.0, 1.0, 0.0],
        [0.0, 0.0, 1.0, -2.0, 1.0],
        [0.0, 0.0, 3.0, -4.0, 1.0],
    ]
)
_WENO5_D = np.array([0.3, 0.6, 0.1])
WENO5_CM = np.concatenate([_WENO5_ROWS.ravel(), _WENO5_D])
WENO5_CP = np.concatenate([_WENO5_ROWS[:, ::-1].ravel(), _WENO5_D])
CP5_CM = np.array([-1.0 / 20.0, 9.0 / 20.0, 47.0 / 60.0, -13.0 / 60.0, 1.0 / 30.0])
CP5_CP = CP5_CM[::-1].copy()


def _vvdot(phi, c):
    """sum_k phi[k] * c[k], accumulated from 0 in k order (recon_point.h:44-51)."""
    out = np.zeros_like(phi[0])
    for k in range(len(c)):
        out = out + phi[k] * c.dtype.type(c[k])
    return out


def weno5_point(phi, c):
    """weno5_point with scale=false (recon_point.h:113-147).

    phi: list of 5 arrays (the stencil window); c: 48-entry table.
    """
    T = phi[0].dtype.type
    c = c.astype(phi[0].dtype)
    rows = [c[5 * r : 5 * r + 5] for r in range(9)]
    d = c[45:48]
    p0 = _vvdot(phi, rows[0])
    p1 = _vvdot(phi, rows[1])
    p2 = _vvdot(phi, rows[2])
    sq = lambda x: x * x  # noqa: E731  (SNAP_SQR)
    beta0 = T(13.0 / 12.0) * sq(_vvdot(phi, rows[3])) + T(0.25) * sq(_vvdot(phi, rows[4]))
    beta1 = T(13.0 / 12.0) * sq(_vvdot(phi, rows[5])) + T(0.25) * sq(_vvdot(phi, rows[6]))
    beta2 = T(13.0 / 12.0) * sq(_vvdot(phi, rows[7])) + T(0.25) * sq(_vvdot(phi, rows[8]))
    eps = T(1e-6)
    a0 = d[0] / sq(beta0 + eps)
    a1 = d[1] / sq(beta1 + eps)
    a2 = d[2] / sq(beta2 + eps)
    vscale = T(1.0)
    return vscale * (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2)


def poly_point(phi, c):
    """poly_point<T,5> (recon_point.h:53-61)."""
    return _vvdot(phi, c.astype(phi[0].dtype))


def _windows(w, axis, n_out):
    """The 5 shifted views w[k : k + n_out] along axis (the iterator of
    x1_uniform_iterator, src/recon/x1_rows.hpp:50-61)."""
    return [np.take(w, np.arange(k, k + n_out), axis=axis) for k in range(5)]


def interp_apply(w, axis, scheme):
    """InterpImpl::forward for a stencil scheme: left (cm) and right (cp)
    outputs of length n-4 along axis (reconstruct.cpp:119-167)."""
    n = w.shape[axis]
    phi = _windows(w, axis, n - 4)
    if scheme == "weno5":
        return weno5_point(phi, WENO5_CM), weno5_point(phi, WENO5_CP)
    if scheme == "cp5":
        return poly_point(phi, CP5_CM), poly_point(phi, CP5_CP)
    raise ValueError(scheme)


def _apply_inplace(axis, w, scheme, wlr, lr_axis_offset):
    """_apply_inplace (reconstruct.cpp:215-229): faces il..iu+1 of cells
    il..iu; the ghost faces copy the end values.

    w has the variable axis first; ``axis`` is the spatial axis in w;
    wlr[s] arrays have the same layout as w."""
    n = w.shape[axis]
    nghost = 3
    il, iu = nghost, n - nghost
    outl, outr = interp_apply(w, axis, scheme)
    sl = [slice(None)] * w.ndim

    def put(arr, lo, hi, val):
        s = list(sl)
        s[axis] = slice(lo, hi)
        arr[tuple(s)] = val

    def get(arr, i):
        s = list(sl)
        s[axis] = slice(i, i + 1)
        return arr[tuple(s)]

    put(wlr[IRT], il - 1, iu + 1, outl)
    put(wlr[ILT], il, iu + 2, outr)
    for s in (IRT, ILT):
        lo = get(wlr[s], il).copy()
        hi = get(wlr[s], iu).copy()
        put(wlr[s], 0, il, lo)
        put(wlr[s], iu + 1, n, hi)


def reconstruct(w, dim, shock):
    """ReconstructImpl::forward without floors (reconstruct.cpp:252-333;
    decks with limiter: false).  dim: 3 = x1, 2 = x2, 1 = x3 (tensor axis).
    Returns wlr of shape (2, nvar, nc3, nc2, nc1)."""
    nvar = w.shape[0]
    wlr = np.zeros((2,) + w.shape, dtype=w.dtype)
    if shock:
        _apply_inplace(dim, w, "weno5", wlr, 0)
        return wlr
    # density (weno5), then velocity/pressure (cp5), then species (weno5)
    sub = [wlr[0][IDN : IDN + 1], wlr[1][IDN : IDN + 1]]
    _apply_inplace(dim, w[IDN : IDN + 1], "weno5", sub, 0)
    ln = min(IPR, nvar - 1)
    sub = [wlr[0][IVX : IVX + ln], wlr[1][IVX : IVX + ln]]
    _apply_inplace(dim, w[IVX : IVX + ln], "cp5", sub, 0)
    return wlr


# ---------------------------------------------------------------------------
# Riemann solvers (src/riemann/riemann_kernels.h)
# ---------------------------------------------------------------------------
def face_axes(dim):
    """face_axes (riemann_kernels.h:36-40): normal row IPR - dim, then the
    tangential rows in cyclic order."""
    ivx = IPR - dim
    return ivx, IVX + ((ivx - IVX) + 1) % 3, IVX + ((ivx - IVX) + 2) % 3


def hllc(wl, wr, gamma, dim):
    """hllc_impl, FP64 branch, dry, no background (riemann_kernels.h:56-194)
    with its EOS face pass (riemann.cpp:232-260).  Returns (flux, cp)."""
    T = wl.dtype.type
    g = T(gamma)
    el = wl[IPR] / (g - T(1))
    er = wr[IPR] / (g - T(1))
    cl = np.sqrt(g * wl[IPR] / wl[IDN])
    cr = np.sqrt(g * wr[IPR] / wr[IDN])
    ivx, ivy, ivz = face_axes(dim)
    sq = lambda x: x * x  # noqa: E731
    el = el + T(0.5) * wl[IDN] * (sq(wl[IVX]) + sq(wl[IVY]) + sq(wl[IVZ]))
    er = er + T(0.5) * wr[IDN] * (sq(wr[IVX]) + sq(wr[IVY]) + sq(wr[IVZ]))
    rhoa = T(0.5) * (wl[IDN] + wr[IDN])
    ca = T(0.5) * (cl + cr)
    pmid = T(0.5) * (wl[IPR] + wr[IPR] + (wl[ivx] - wr[ivx]) * rhoa * ca)
    shock_l = pmid / wl[IPR] - T(1.0)
    shock_r = pmid / wr[IPR] - T(1.0)
    with np.errstate(invalid="ignore"):
        ql = np.where(pmid <= wl[IPR], T(1.0), np.sqrt(T(1.0) + (g + 1) / (2 * g) * shock_l))
        qr = np.where(pmid <= wr[IPR], T(1.0), np.sqrt(T(1.0) + (g + 1) / (2 * g) * shock_r))
    al = wl[ivx] - cl * ql
    ar = wr[ivx] + cr * qr
    bp = np.where(ar > 0, ar, T(1.0e-10))
    bm = np.where(al < 0, al, -T(1.0e-10))
    vxl = wl[ivx] - al
    vxr = wr[ivx] - ar
    tl = wl[IPR] + vxl * wl[IDN] * wl[ivx]
    tr = wr[IPR] + vxr * wr[IDN] * wr[ivx]
    ml = wl[IDN] * vxl
    mr = -(wr[IDN] * vxr)
    am = (tl - tr) / (ml + mr)
    cp = (ml * tr + mr * tl) / (ml + mr)
    cp = np.where(cp > 0, cp, T(0.0))
    vxl = wl[ivx] - bm
    vxr = wr[ivx] - bp
    fl = np.empty_like(wl)
    fr = np.empty_like(wr)
    one = T(1.0)  # dry_fraction with no species
    fl[IDN] = wl[IDN] * vxl * one
    fr[IDN] = wr[IDN] * vxr * one
    fl[ivx] = wl[IDN] * wl[ivx] * vxl + wl[IPR]
    fr[ivx] = wr[IDN] * wr[ivx] * vxr + wr[IPR]
    fl[ivy] = wl[IDN] * wl[ivy] * vxl
    fr[ivy] = wr[IDN] * wr[ivy] * vxr
    fl[ivz] = wl[IDN] * wl[ivz] * vxl
    fr[ivz] = wr[IDN] * wr[ivz] * vxr
    fl[IPR] = el * vxl + wl[IPR] * wl[ivx]
    fr[IPR] = er * vxr + wr[IPR] * wr[ivx]
    pos = am >= 0
    with np.errstate(divide="ignore", invalid="ignore"):
        sl = np.where(pos, am / (am - bm), T(0.0))
        sr = np.where(pos, T(0.0), -am / (bp - am))
        sm = np.where(pos, -bm / (am - bm), bp / (bp - am))
    flx = np.empty_like(wl)
    flx[IDN] = sl * fl[IDN] + sr * fr[IDN]
    flx[ivx] = sl * fl[ivx] + sr * fr[ivx] + sm * cp
    flx[ivy] = sl * fl[ivy] + sr * fr[ivy]
    flx[ivz] = sl * fl[ivz] + sr * fr[ivz]
    flx[IPR] = sl * fl[IPR] + sr * fr[IPR] + sm * cp * am
    return flx, cp


def lmars(wl, wr, gamma, dim):
    """lmars_impl, dry (riemann_kernels.h:197-255); e/rho and gamma from the
    EOS face pass (riemann.cpp:262-268).  Returns (flux, pbar)."""
    T = wl.dtype.type
    g = T(gamma)
    hl = (wl[IPR] / (g - T(1))) / wl[IDN]
    hr = (wr[IPR] / (g - T(1))) / wr[IDN]
    ivx, ivy, ivz = face_axes(dim)
    wli = [wl[IDN], wl[ivx], wl[ivy], wl[ivz], wl[IPR]]
    wri = [wr[IDN], wr[ivx], wr[ivy], wr[ivz], wr[IPR]]
    sq = lambda x: x * x  # noqa: E731
    hl = hl + (T(0.5) * (sq(wli[1]) + sq(wli[2]) + sq(wli[3])) + wli[4] / wli[0])
    hr = hr + (T(0.5) * (sq(wri[1]) + sq(wri[2]) + sq(wri[3])) + wri[4] / wri[0])
    rhobar = T(0.5) * (wli[0] + wri[0])
    gamma_bar = T(0.5) * (g + g)
    cbar = np.sqrt(T(0.5) * gamma_bar * (wli[4] + wri[4]) / rhobar)
    pbar = T(0.5) * (wli[4] + wri[4]) + T(0.5) * (rhobar * cbar) * (wli[1] - wri[1])
    ubar = T(0.5) * (wli[1] + wri[1]) + T(0.5) / (rhobar * cbar) * (wli[4] - wri[4])
    up = ubar > 0
    flx = np.empty_like(wl)
    one = T(1.0)
    flx[IDN] = np.where(up, ubar * wli[0] * one, ubar * wri[0] * one)
    flx[ivx] = np.where(up, ubar * wli[0] * wli[1] + pbar, ubar * wri[0] * wri[1] + pbar)
    flx[ivy] = np.where(up, ubar * wli[0] * wli[2], ubar * wri[0] * wri[2])
    flx[ivz] = np.where(up, ubar * wli[0] * wli[3], ubar * wri[0] * wri[3])
    flx[IPR] = np.where(up, ubar * wli[0] * hl, ubar * wri[0] * hr)
    return flx, pbar


# ---------------------------------------------------------------------------
# EOS (src/eos/ideal_gas_impl.h:16-35, src/eos/ideal_gas.cpp:49-94)
# ---------------------------------------------------------------------------
def cons2prim(u, gammad):
    w = np.empty_like(u)
    w[IDN] = u[IDN]
    w[IVX] = u[IVX] / w[IDN]
    w[IVY] = u[IVY] / w[IDN]
    w[IVZ] = u[IVZ] / w[IDN]
    ke = 0.5 * (w[IVX] * u[IVX] + w[IVY] * u[IVY] + w[IVZ] * u[IVZ])
    ie = u[IPR] - ke
    w[IPR] = (gammad - 1.0) * ie
    return w


def prim2cons(w, gammad):
    u = np.empty_like(w)
    u[IDN] = w[IDN]
    u[IVX] = w[IVX] * w[IDN]
    u[IVY] = w[IVY] * w[IDN]
    u[IVZ] = w[IVZ] * w[IDN]
    ke = 0.5 * ((w[IVX] * u[IVX] + w[IVY] * u[IVY]) + w[IVZ] * u[IVZ])
    u[IPR] = w[IPR] / (gammad - 1.0)
    u[IPR] = u[IPR] + ke
    return u


# ---------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    nx1: int
    nx2: int
    nx3: int
    x1min: float = 0.0
    x1max: float = 1.0
    x2min: float = 0.0
    x2max: float = 1.0
    x3min: float = 0.0
    x3max: float = 1.0
    nghost: int = 3

    def nc(self, d):
        nx = (self.nx1, self.nx2, self.nx3)[d - 1]
        return nx + 2 * self.nghost if nx > 1 else 1

    @property
    def shape(self):
        return (self.nc(3), self.nc(2), self.nc(1))

    def dx(self, d):
        lo, hi, n = [
            (self.x1min, self.x1max, self.nx1),
            (self.x2min, self.x2max, self.nx2),
            (self.x3min, self.x3max, self.nx3),
        ][d - 1]
        return (hi - lo) / n

    def faces(self, d):
        """block_faces_ / global_x1f for a uniform axis
        (src/coord/coordinate.cpp:369-380): linspace over the ghosts."""
        lo, hi, n = [
            (self.x1min, self.x1max, self.nx1),
            (self.x2min, self.x2max, self.nx2),
            (self.x3min, self.x3max, self.nx3),
        ][d - 1]
        dx = (hi - lo) / n
        ng = self.nghost if n > 1 else 0
        return np.linspace(lo - ng * dx, hi + ng * dx, n + 2 * ng + 1)

    def centers(self, d):
        f = self.faces(d)
        return 0.5 * (f[:-1] + f[1:])

    def interior(self):
        """part({0,0,0}, exterior=false): the interior slices of (nc3,nc2,nc1)."""
        out = []
        for d in (3, 2, 1):
            nx = (self.nx1, self.nx2, self.nx3)[d - 1]
            lo = self.nghost if nx > 1 else 0
            out.append(slice(lo, lo + nx))
        return tuple(out)


def flux_covariance(wl, wr, dim, grid, gammad):
    """HydroImpl::_flux_covariance, Cartesian dry rows (hydro_forward.cpp:
    58-231, :196-213): sigma1^2 rho d1[h] d1[u_n] on the x1 cells 1..n1-2 of
    the x2/x3 faces, wbar = (wl + wr)/2, h = (W->I + p)/rho."""
    n1 = wl.shape[-1]
    wbar = 0.5 * (wl + wr)
    p = wbar[IPR]
    enth = p / (gammad - 1.0) + p
    un = wbar[IVY if dim == 2 else IVZ]
    x1v = grid.centers(1)
    dx1f = np.full(n1, grid.dx(1))
    s2 = dx1f * dx1f / 12.0
    rho = wbar[IDN]
    dx1 = x1v[2:n1] - x1v[0 : n1 - 2]
    cell = lambda q: q[..., 1 : n1 - 1]  # noqa: E731
    d1 = lambda q: (q[..., 2:n1] - q[..., 0 : n1 - 2]) / dx1  # noqa: E731
    s2rho = cell(s2) * cell(rho)
    dun = d1(un)
    h = enth / rho
    return s2rho * d1(h) * dun


def prescribe_walls(flux, tdim, bcs, axis, grid):
    """boundary.prescribed on the wall planes of one direction
    (FluxRegister::apply_prescribed, src/hydro/flux_register.cpp:707-748;
    supply_prescribed, src/layout/seam_register.cpp:184-220): at a reflecting
    face (a FaceKind::wall, src/mesh/meshblock_options.cpp:188-209) the rows
    other than the momentum rows {IVX, IVY, IVZ} are set to zero."""
    nx = (grid.nx1, grid.nx2, grid.nx3)[axis - 1]
    planes = []
    if bcs[2 * (axis - 1)] == "reflecting":
        planes.append(grid.nghost)
    if bcs[2 * (axis - 1) + 1] == "reflecting":
        planes.append(grid.nghost + nx)
    for p in planes:
        for row in (IDN, IPR):
            flux[row][_sl(3, tdim - 1, slice(p, p + 1))] = 0.0


def divergence(grid, flux):
    """CoordinateImpl::divergence, Cartesian uniform (coordinate.cpp:618-662,
    areas/volume :534-556).  flux = {1: F1, 2: F2, 3: F3} (missing = inactive)."""
    dx1, dx2, dx3 = grid.dx(1), grid.dx(2), grid.dx(3)
    like = next(iter(flux.values()))
    dflx = np.zeros_like(like)
    k, j, i = grid.interior()
    if 1 in flux:
        a = dx3 * dx2
        f = flux[1]
        si, ei = i.start, i.stop
        dflx[..., si:ei] += a * f[..., si + 1 : ei + 1] - a * f[..., si:ei]
    if 2 in flux:
        a = dx3 * dx1
        f = flux[2]
        sj, ej = j.start, j.stop
        dflx[:, :, sj:ej, :] += a * f[:, :, sj + 1 : ej + 1, :] - a * f[:, :, sj:ej, :]
    if 3 in flux:
        a = dx2 * dx1
        f = flux[3]
        sk, ek = k.start, k.stop
        dflx[:, sk:ek] += a * f[:, sk + 1 : ek + 1] - a * f[:, sk:ek]
    vol = dx3 * (dx2 * dx1)
    return dflx / vol


# ---------------------------------------------------------------------------
# boundary functions (src/bc/bc_func.cpp) and apply_boundaries
# (src/mesh/meshblock_exchange.cpp:214-287)
# ---------------------------------------------------------------------------
def _sl(ndim, axis, s):
    idx = [slice(None)] * ndim
    idx[axis] = s
    return tuple(idx)


def bc_gather(var, dim, kind, outer, ng=3, flip=True):
    """reflecting / periodic / extrapolation on a (nvar, nc3, nc2, nc1) field;
    dim is the tensor axis (3 = x1, 2 = x2, 1 = x3) (bc_func.cpp:10-65).
    flip=False: a kScalar field (no normal-velocity row to negate)."""
    nc = var.shape[dim]
    if nc == 1:
        return
    nd = var.ndim
    if kind == "reflecting":
        if not outer:
            var[_sl(nd, dim, slice(0, ng))] = np.flip(var[_sl(nd, dim, slice(ng, 2 * ng))], axis=dim)
            if flip:
                var[4 - dim][_sl(nd - 1, dim - 1, slice(0, ng))] *= -1
        else:
            var[_sl(nd, dim, slice(nc - ng, nc))] = np.flip(
                var[_sl(nd, dim, slice(nc - 2 * ng, nc - ng))], axis=dim
            )
            if flip:
                var[4 - dim][_sl(nd - 1, dim - 1, slice(nc - ng, nc))] *= -1
    elif kind == "periodic":
        if not outer:
            var[_sl(nd, dim, slice(0, ng))] = var[_sl(nd, dim, slice(nc - 2 * ng, nc - ng))]
        else:
            var[_sl(nd, dim, slice(nc - ng, nc))] = var[_sl(nd, dim, slice(ng, 2 * ng))]
    elif kind == "extrapolation":
        if not outer:
            var[_sl(nd, dim, slice(0, ng))] = var[_sl(nd, dim, slice(ng, ng + 1))]
        else:
            var[_sl(nd, dim, slice(nc - ng, nc))] = var[_sl(nd, dim, slice(nc - ng - 1, nc - ng))]
    elif kind in ("none", "custom"):
        return
    else:
        raise ValueError(kind)


def radiating(var, dim, bref, gammad, dfloor, pfloor, outer, ng=3):
    """outflow (radiating about the saved background, bc_func.cpp:110-220),
    ideal gas, Cartesian, no tracers; on a primitive field."""
    nc = var.shape[dim]
    if nc == 1:
        return
    axis = 4 - dim
    vn = IVX + axis - 1
    src = nc - ng - 1 if outer else ng
    dst = nc - ng if outer else 0
    sign = 1.0 if outer else -1.0
    nd = var.ndim
    interior = lambda t: t[_sl(t.ndim, dim, slice(src, src + 1))]  # noqa: E731
    ghost = lambda t: t[_sl(t.ndim, dim, slice(dst, dst + ng))]  # noqa: E731
    w = var.copy()
    b = bref.copy()
    sound = lambda t: np.sqrt(gammad * t[IPR] / t[IDN])  # noqa: E731
    with np.errstate(invalid="ignore", divide="ignore"):
        c0 = interior(sound(b)[None])[0]
        c = interior(sound(w)[None])[0]
    ok = np.isfinite(c) & (c > 0)
    wi_ = interior(w)
    wi_[...] = np.where(ok[None], wi_, interior(b))
    c = np.where(ok, c, c0)
    wi = interior(w)
    bi = interior(b)
    d = wi - bi
    un = sign * wi[vn]
    advect = un >= 0
    impedance = bi[IDN] * c0
    plus = np.where(un + c >= 0, sign * d[vn] + d[IPR] / impedance, 0.0)
    minus = np.where(un - c >= 0, sign * d[vn] - d[IPR] / impedance, 0.0)
    entropy = np.where(advect, d[IDN] - d[IPR] / (c0 * c0), 0.0)
    d = d * advect[None]
    d[vn] = sign * 0.5 * (plus + minus)
    d[IPR] = 0.5 * impedance * (plus - minus)
    d[IDN] = entropy + d[IPR] / (c0 * c0)
    bg = ghost(b)
    alpha = np.ones_like(bg[IDN])

    def lower(base, delta, floor):
        nonlocal alpha
        with np.errstate(divide="ignore", invalid="ignore"):
            bound = np.where(delta < 0, (base - floor) / np.where(delta < 0, -delta, 1.0), 1.0)
        alpha = np.minimum(alpha, np.where(bound < 1, 0.99 * bound, bound))

    lower(bg[IDN], d[IDN], dfloor)
    lower(bg[IPR], d[IPR], pfloor)
    alpha = np.clip(alpha, 0.0, 1.0)
    ghost(var)[...] = bg + alpha[None] * d
    del nd


def apply_boundaries(q, bcs, gammad, bref=None, primitive=False, dfloor=1e-10, pfloor=1e-10):
    """MeshBlockImpl::apply_boundaries (meshblock_exchange.cpp:214-287);
    bcs = 6 names in the order x1-inner, x1-outer, x2-inner, ..., x3-outer.
    Works in place on q (conserved unless primitive)."""
    w = None
    pending = []

    def flush():
        nonlocal pending
        if not pending:
            return
        u = prim2cons(w, gammad)
        for f in pending:
            dim = 3 - f // 2
            nc = q.shape[dim]
            start = nc - 3 if f % 2 else 0
            q[_sl(4, dim, slice(start, start + 3))] = u[_sl(4, dim, slice(start, start + 3))]
        pending = []

    for f in range(6):
        dim = 3 - f // 2
        kind = bcs[f]
        if q.shape[dim] == 1:
            continue
        if kind == "outflow":
            if primitive:
                w = q
            elif not pending:
                w = cons2prim(q.copy(), gammad)
            radiating(w, dim, bref, gammad, dfloor, pfloor, f % 2 == 1)
            if not primitive:
                pending.append(f)
        else:
            flush()
            bc_gather(q, dim, kind, f % 2 == 1)
    flush()


# ---------------------------------------------------------------------------
# CFL (HydroImpl::max_time_step explicit branch, src/hydro/hydro.cpp:372-396;
# MeshBlockImpl::global_time_step, src/mesh/meshblock_advance.cpp:65-72)
# ---------------------------------------------------------------------------
def max_time_step(grid, w, gammad, scheme=0, cfl=0.9, advection_cfl=1.0):
    """Explicit branch, or with an implicit scheme the icorr branch
    (hydro.cpp:326-370): an implicit direction is bounded by advection,
    |v| * cfl / advection_cfl."""
    cs = np.sqrt(gammad * w[IPR] / w[IDN])
    sub = grid.interior()
    dt = [1.0e9, 1.0e9, 1.0e9]
    adv = cfl / advection_cfl
    if cs.shape[2] > 1:
        if scheme and (scheme & 1) and not (cs.shape[0] == 1 and cs.shape[1] == 1):
            with np.errstate(divide="ignore"):
                dt[0] = (grid.dx(1) / (np.abs(w[IVX]) * adv))[sub].min()
        else:
            dt[0] = (grid.dx(1) / (np.abs(w[IVX]) + cs))[sub].min()
    if cs.shape[1] > 1:
        dt[1] = (grid.dx(2) / (np.abs(w[IVY]) + cs))[sub].min()
    if cs.shape[0] > 1:
        dt[2] = (grid.dx(3) / (np.abs(w[IVZ]) + cs))[sub].min()
    return min(dt)


def layer_heat(du, grid, dt, flux, depth, top=False):
    """LayerHeat with an integer depth on a uniform x1 (layer_heat.cpp:78-85):
    the energy rows of the `depth` wall cells (bottom: the first interior
    cells; top: the last) gain dt * flux / (dx1 * depth), in place."""
    il = grid.nghost
    iu = grid.nghost + grid.nx1 - 1
    i0 = iu + 1 - depth if top else il
    du[IPR][..., i0 : i0 + depth] += dt * flux / (grid.dx(1) * depth)
    return du


# SSP-RK3 (harp "rk3"): u <- w0*u0 + w1*u + w2*du per stage; the weights
# wght2 = {1, 1/4, 2/3} are pinned by tests/test_forcing.cpp:1317-1321
RK3_WEIGHTS = ((1.0, 0.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0))


def step_field_fold(S, T, w1, w2, dt):
    """FluxRegister::end_stage (flux_register.cpp:761-777): the FP64 step
    field S <- w1 S + w2 dt T, as S.mul_(w1) then the axpy with
    w2dt = dt.scaled(w2), the stage total T widened to FP64."""
    w2dt = dt * w2
    return S * w1 + np.asarray(T, dtype=np.float64) * w2dt


def rk_combine(u0, u, du, stage):
    w0, w1, w2 = RK3_WEIGHTS[stage]
    return w0 * u0 + w1 * u + w2 * du


# ---------------------------------------------------------------------------
# one block
# ---------------------------------------------------------------------------
@dataclass
class Deck:
    grid: Grid
    gammad: float = 1.4
    riemann: str = "hllc"
    shock_x1: bool = True
    shock_x23: bool = True
    bcs: tuple = ("outflow",) * 6
    cfl: float = 0.9
    dfloor: float = 1e-10
    pfloor: float = 1e-10
    grav1: float = 0.0
    implicit_scheme: int = 0
    flux_covariance: bool = False
    bot_heat_flux: float = 0.0
    bot_heat_depth: int = 1
    # forcing/diffusion (kinematic, isotropic); weight is the EOS molar mass
    nu_iso: float = 0.0
    kappa_iso: float = 0.0
    weight: float = 0.02897025302492418
    # passive scalars (tracers), transported as oracle.scalar restates
    nscalar: int = 0


@dataclass
class Block:
    """The explicit dry Cartesian stage of MeshBlockImpl::forward
    (src/mesh/meshblock_advance.cpp:88-105, 145-404; HydroImpl::forward
    src/hydro/hydro_forward.cpp:299-1235 without gravity and implicit)."""

    deck: Deck
    u: np.ndarray = None
    u0: np.ndarray = None
    w: np.ndarray = None
    bref: np.ndarray = None
    flux: dict = field(default_factory=dict)
    face_pressure1: np.ndarray = None
    # passive scalars: s = rho r (conserved), its stage backup, r = s / rho
    s: np.ndarray = None
    s0: np.ndarray = None
    r: np.ndarray = None

    def initialize(self, w, r=None):
        """initialize_local (src/mesh/meshblock.cpp:306-343) + prim->cons;
        with scalars, s = u[IDN] r on every cell (finalize_initialization,
        meshblock.cpp:380-383)."""
        d = self.deck
        w = w.copy()
        if "outflow" in d.bcs:
            self.bref = w.copy()
        apply_boundaries(w, d.bcs, d.gammad, self.bref, primitive=True, dfloor=d.dfloor, pfloor=d.pfloor)
        self.w = w
        self._w_stale = False
        self.u = prim2cons(w, d.gammad)
        if d.nscalar:
            self.r = np.array(r, dtype=np.float64, copy=True)
            self.s = self.u[IDN][None] * self.r
        return self.u

    def _scalar_tendency(self, dt, cov):
        """ds of the stage before the implicit transfer (scalar.cpp:104-218)."""
        from . import scalar

        d = self.deck
        g = d.grid
        rho = self.u[IDN][None]
        fl, mass, used = scalar.fluxes(self.s, rho, self.flux, g, d.bcs, w=self.w, shock23=d.shock_x23, cov=cov)
        if used:
            fl, _ = scalar.limit(self.s, rho, fl, mass, g, dt, d.bcs)
        self.scalar_flux = fl
        div = divergence(g, fl)
        ds = np.zeros_like(div)
        sub = (slice(None),) + g.interior()
        ds[sub] = -dt * div[sub]
        return ds

    def tendency(self, dt):
        """du of one stage (hydro_forward.cpp steps 1, 2, 4, 5, 6)."""
        d = self.deck
        g = d.grid
        self.w = cons2prim(self.u, d.gammad)
        flux = {}
        for axis, tdim in ((1, 3), (2, 2), (3, 1)):
            if g.nc(axis) == 1:
                continue
            shock = d.shock_x1 if axis == 1 else d.shock_x23
            if axis == 1 and d.grav1 != 0.0:
                # the well-balanced x1 path whenever gravity is on
                # (hydro_forward.cpp:479-546)
                from . import gravity

                mirror = tuple(b != "outflow" for b in d.bcs[0:2])
                refl = tuple(b == "reflecting" for b in d.bcs[0:2])
                wlr = gravity.wb_reconstruct_x1(self.w, g, d.grav1, shock, mirror, refl)
            else:
                wlr = reconstruct(self.w, tdim, shock)
            solver = hllc if d.riemann == "hllc" else lmars
            f, p = solver(wlr[ILT], wlr[IRT], d.gammad, tdim)
            if axis > 1 and d.flux_covariance and g.nc(1) >= 3:
                f[IPR][..., 1 : g.nc(1) - 1] += flux_covariance(wlr[ILT], wlr[IRT], tdim, g, d.gammad)
            prescribe_walls(f, tdim, d.bcs, axis, g)
            flux[axis] = f
            if axis == 1:
                self.face_pressure1 = p
        fdiff = None
        if d.nu_iso > 0.0 or d.kappa_iso > 0.0:
            # the diffusion posts: the energy row joins the totals before the
            # divergence (hydro_forward.cpp:930-935)
            from . import diffusion

            fdiff = diffusion.cartesian_fluxes(self.w, g, d.nu_iso, d.kappa_iso, d.gammad, d.weight, d.bcs)
            diffusion.post_fluxes(flux, fdiff, g, d.bcs)
        self.flux = flux
        if d.nscalar:
            self._ds = self._scalar_tendency(dt, d.flux_covariance)
        div = divergence(g, flux)
        du = np.zeros_like(div)
        sub = (slice(None),) + g.interior()
        du[sub] = -dt * div[sub]
        if d.grav1 != 0.0:
            from . import gravity

            gravity.const_gravity(du, self.w, dt, d.grav1)
        if fdiff is not None:
            # DiffusionImpl::add: the momentum rows (diffusion.cpp:532-547)
            from . import diffusion

            diffusion.add_momentum(du, fdiff, g, dt)
        if d.grav1 != 0.0:
            if d.bot_heat_flux != 0.0:
                layer_heat(du, g, dt, d.bot_heat_flux, d.bot_heat_depth)
            if 1 in flux:
                gec = gravity.gravity_energy_correction(g, self.w, flux[1], dt, d.grav1)
                du[IPR][..., g.nghost : g.nghost + g.nx1] += gec
        if d.implicit_scheme in (1, 9):
            # the VIC with the stage weight inside the solve
            # (hydro_forward.cpp:1137-1180)
            du = self._implicit(du, dt * RK3_WEIGHTS[self._stage][2])
        return du

    def _implicit(self, du, dt_corr):
        from . import implicit

        g = self.deck.grid
        ng = g.nghost
        k, j, i = g.interior()
        n = g.nx1
        wc = self.w[:, k, j, ng - 1 : ng + n + 1]  # (5, nk, nj, n + 2)
        ncol = wc.shape[1] * wc.shape[2]
        wcol = np.moveaxis(wc, 0, -1).reshape(ncol, n + 2, 5)
        dcol = np.moveaxis(du[:, k, j, ng : ng + n], 0, -1).reshape(ncol, n, 5)
        x1f = g.faces(1)[ng : ng + n + 1]
        x1v = g.centers(1)[ng : ng + n]
        solve = implicit.vic_partial if self.deck.implicit_scheme == 1 else implicit.vic_full
        out, moved = solve(
            dcol, wcol, self.deck.gammad, self.deck.grav1, dt_corr, g.dx(1), x1f, x1v,
            g.dx(3) * g.dx(2), g.dx(3) * (g.dx(2) * g.dx(1)), return_moved=True,
        )
        du = du.copy()
        du[:, k, j, ng : ng + n] = np.moveaxis(out.reshape(wc.shape[1], wc.shape[2], n, 5), -1, 0)
        if self.deck.nscalar:
            # the dry-gas transfer through the face below each interior cell
            from . import scalar

            mv = np.zeros(g.shape)
            mv[k, j, ng : ng + n] = moved[:, :n].reshape(wc.shape[1], wc.shape[2], n)
            self._ds = scalar.implicit_transfer(self._ds, self.s, self.r, mv, g)
        return du

    def forward(self, dt, stage):
        d = self.deck
        if stage == 0:
            self.u0 = self.u.copy()
            if d.nscalar:
                self.s0 = self.s.copy()
        self._stage = stage
        du = self.tendency(dt)
        self.u = rk_combine(self.u0, self.u, du, stage)
        apply_boundaries(self.u, d.bcs, d.gammad, self.bref, dfloor=d.dfloor, pfloor=d.pfloor)
        if d.nscalar:
            from . import scalar

            # the scalar combine, ghosts, then r = s / rho of the new state
            # (meshblock_advance.cpp:338-341, meshblock_exchange.cpp:195-200)
            s = rk_combine(self.s0, self.s, self._ds, stage)
            scalar.boundaries(s, d.bcs)
            self.s = s
            self.r = s / self.u[IDN][None]
        self._w_stale = True

    def check_redo(self):
        """The primitives of the accepted state, w = prim(u) on every cell
        (MeshBlockImpl::local_redo_flag_tensor, meshblock_redo.cpp:107-112);
        the oracle models no retry causes."""
        self.w = cons2prim(self.u, self.deck.gammad)
        self._w_stale = False
        return 0

    def max_time_step(self):
        """From prim(u): the driver's check_redo refreshes the primitives
        after every step (examples/shock.cpp:60-72); a step not followed by
        one refreshes them here."""
        d = self.deck
        if getattr(self, "_w_stale", False):
            self.check_redo()
        dt = max_time_step(d.grid, self.w, d.gammad, d.implicit_scheme, d.cfl)
        if d.nu_iso > 0.0 or d.kappa_iso > 0.0:
            # the sources' bound joins before the CFL factor (hydro.cpp:396)
            from . import diffusion

            dt = min(dt, diffusion.max_time_step(d.grid, d.nu_iso, d.kappa_iso))
        return d.cfl * dt


def sod_deck(nx=256):
    """The C1 Sod deck: x in [0,1], nx cells, WENO5 shock + HLLC + RK3,
    outflow, CFL 0.9 (examples/shock.yaml with a 1D grid)."""
    g = Grid(nx, 1, 1, 0.0, 1.0, -0.5, 0.5, -0.5, 0.5)
    return Deck(grid=g)


def sod_initial(grid, x0=0.5, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1)):
    w = np.zeros((5,) + grid.shape)
    x1v = grid.centers(1)[None, None, :]
    w[IDN] = np.where(x1v < x0, left[0], right[0])
    w[IVX] = np.where(x1v < x0, left[1], right[1])
    w[IPR] = np.where(x1v < x0, left[2], right[2])
    return w


def run(block, tlim, nlim=10**9):
    """The example driver loop (examples/shock.cpp:55-75) with the last step
    clipped to tlim."""
    t = 0.0
    n = 0
    while t < tlim and n < nlim:
        dt = block.max_time_step()
        dt = min(dt, tlim - t)
        for s in range(3):
            block.forward(dt, s)
        t += dt
        n += 1
    return t, n


def sod_exact(x, t, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), gamma=1.4, x0=0.5):
    """Exact Riemann solution of the Sod problem (Toro, Riemann Solvers and
    Numerical Methods for Fluid Dynamics, ch. 4): Newton on the star pressure."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl = math.sqrt(gamma * pl / rl)
    cr = math.sqrt(gamma * pr / rr)

    def fk(p, rk, pk, ck):
        if p > pk:
            a = 2.0 / ((gamma + 1) * rk)
            b = (gamma - 1) / (gamma + 1) * pk
            f = (p - pk) * math.sqrt(a / (p + b))
            df = math.sqrt(a / (b + p)) * (1 - (p - pk) / (2 * (b + p)))
        else:
            f = 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)
            df = 1.0 / (rk * ck) * (p / pk) ** (-(gamma + 1) / (2 * gamma))
        return f, df

    p = 0.5 * (pl + pr)
    for _ in range(100):
        f1, d1 = fk(p, rl, pl, cl)
        f2, d2 = fk(p, rr, pr, cr)
        dp = (f1 + f2 + ur - ul) / (d1 + d2)
        p = max(p - dp, 1e-12)
        if abs(dp) < 1e-15 * p:
            break
    f1, _ = fk(p, rl, pl, cl)
    f2, _ = fk(p, rr, pr, cr)
    us = 0.5 * (ul + ur) + 0.5 * (f2 - f1)
    out = np.zeros((3, len(x)))
    g = gamma
    for n, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if p > pl:  # left shock
                rs = rl * ((p / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pl + 1))
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * p / pl + (g - 1) / (2 * g))
                out[:, n] = (rl, ul, pl) if s < sl else (rs, us, p)
            else:
                rs = rl * (p / pl) ** (1 / g)
                cs = cl * (p / pl) ** ((g - 1) / (2 * g))
                sh, st = ul - cl, us - cs
                if s < sh:
                    out[:, n] = (rl, ul, pl)
                elif s > st:
                    out[:, n] = (rs, us, p)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    out[:, n] = (rl * (c / cl) ** (2 / (g - 1)), u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if p > pr:  # right shock
                rs = rr * ((p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1))
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
                out[:, n] = (rr, ur, pr) if s > sr else (rs, us, p)
            else:
                rs = rr * (p / pr) ** (1 / g)
                cs = cr * (p / pr) ** ((g - 1) / (2 * g))
                sh, st = ur + cr, us + cs
                if s > sh:
                    out[:, n] = (rr, ur, pr)
                elif s < st:
                    out[:, n] = (rs, us, p)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    out[:, n] = (rr * (c / cr) ** (2 / (g - 1)), u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
