"""CPU FP64 restatement of the shallow-water stage on a Cartesian grid (test
infrastructure only: the product path never imports this module).

The state rows are (phi, phi v1, phi v2, phi v3), phi = g h the geopotential
height; the primitive rows (phi, v1, v2, v3).

* ``cons2prim`` / ``prim2cons`` -- ShallowWaterImpl::_cons2prim / _prim2cons
  (src/eos/shallow_water.cpp:31-55; _mom_from_vel_,
  src/eos/equation_of_state.cpp:290-295), Cartesian (no metric raise).
* ``shallow_roe``   -- shallow_roe_impl (src/riemann/riemann_kernels.h:
  447-495) with shallow_axes (:44-47) and the Roe sound speed
  sqrt((phi_l + phi_r) / 2) (src/eos/shallow_water_impl.h:12-15).
* ``max_time_step`` -- HydroImpl::max_time_step (src/hydro/hydro.cpp:372-397)
  with the shallow-water sound speed sqrt(phi) (shallow_water.cpp:20-22).
* ``Block``         -- the explicit stage (reconstruct, Riemann, divergence,
  SSP-RK3 combine, boundaries) of src/mesh/meshblock_advance.cpp with the
  shallow-water EOS; no sources on a Cartesian grid.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import fv

IDN, IVX, IVY, IVZ = 0, 1, 2, 3


def cons2prim(u):
    w = np.empty_like(u)
    w[IDN] = u[IDN]
    w[IVX : IVZ + 1] = u[IVX : IVZ + 1] / u[IDN]
    return w


def prim2cons(w):
    u = np.empty_like(w)
    u[IDN] = w[IDN]
    u[IVX : IVZ + 1] = w[IVX : IVZ + 1] * w[IDN]
    return u


def shallow_axes(dim, dir_yz=False):
    """(normal, tangential, third) velocity rows of a sweep along tensor
    axis dim (3 = x1, 2 = x2, 1 = x3)."""
    if dir_yz:
        return (IVY if dim == 2 else IVZ, IVZ if dim == 2 else IVY, IVX)
    return (IVX if dim == 3 else IVY, IVY if dim == 3 else IVX, IVZ)


def shallow_roe(wl, wr, dim, dir_yz=False):
    ivx, ivy, ivz = shallow_axes(dim, dir_yz)
    sqrtdl = np.sqrt(wl[IDN])
    sqrtdr = np.sqrt(wr[IDN])
    isdlpdr = 1.0 / (sqrtdl + sqrtdr)
    ubar = (wl[ivx] * sqrtdl + wr[ivx] * sqrtdr) * isdlpdr
    vbar = (wl[ivy] * sqrtdl + wr[ivy] * sqrtdr) * isdlpdr
    cbar = np.sqrt(0.5 * (wl[IDN] + wr[IDN]))
    del0 = wr[IDN] - wl[IDN]
    delx = wr[ivx] - wl[ivx]
    dely = wr[ivy] - wl[ivy]
    hbar = np.sqrt(wl[IDN] * wr[IDN])
    a1 = 0.5 * (cbar * del0 - hbar * delx) / cbar
    a2 = hbar * dely
    a3 = 0.5 * (cbar * del0 + hbar * delx) / cbar
    zero = np.zeros_like(a1)
    wave0 = [zero] * 4
    wave1 = [zero] * 4
    wave2 = [zero] * 4
    wave0[IDN] = a1
    wave0[ivx] = a1 * (ubar - cbar)
    wave0[ivy] = a1 * vbar
    wave1[ivy] = a2
    wave2[IDN] = a3
    wave2[ivx] = a3 * (ubar + cbar)
    wave2[ivy] = a3 * vbar
    speed0 = np.abs(ubar - cbar)
    speed1 = np.abs(ubar)
    speed2 = np.abs(ubar + cbar)
    flx = np.zeros_like(wl)
    flx[IDN] = 0.5 * (wl[IDN] * wl[ivx] + wr[IDN] * wr[ivx])
    flx[ivx] = 0.5 * (
        wl[IDN] * wl[ivx] * wl[ivx] + 0.5 * wl[IDN] * wl[IDN] + wr[IDN] * wr[ivx] * wr[ivx] + 0.5 * wr[IDN] * wr[IDN]
    )
    flx[ivy] = 0.5 * (wl[IDN] * wl[ivx] * wl[ivy] + wr[IDN] * wr[ivx] * wr[ivy])
    flx[ivz] = 0.0
    for v in range(4):
        flx[v] = flx[v] - 0.5 * (speed0 * wave0[v] + speed1 * wave1[v] + speed2 * wave2[v])
    return flx


def physical_flux(w, dim, dir_yz=False):
    """The exact shallow-water flux of a state along a sweep (the Riemann
    flux of two equal states)."""
    ivx, ivy, ivz = shallow_axes(dim, dir_yz)
    f = np.zeros_like(w)
    f[IDN] = w[IDN] * w[ivx]
    f[ivx] = w[IDN] * w[ivx] * w[ivx] + 0.5 * w[IDN] * w[IDN]
    f[ivy] = w[IDN] * w[ivx] * w[ivy]
    return f


def max_time_step(grid, w):
    cs = np.sqrt(w[IDN])
    sub = grid.interior()
    dt = [1.0e9, 1.0e9, 1.0e9]
    for axis, row in ((1, IVX), (2, IVY), (3, IVZ)):
        if grid.nc(axis) > 1:
            dt[axis - 1] = (grid.dx(axis) / (np.abs(w[row]) + cs))[sub].min()
    return min(dt)


def _prescribe_walls(flux, tdim, bcs, axis, grid):
    """A reflecting face keeps only the momentum rows (the mass row zero)."""
    nx = (grid.nx1, grid.nx2, grid.nx3)[axis - 1]
    planes = []
    if bcs[2 * (axis - 1)] == "reflecting":
        planes.append(grid.nghost)
    if bcs[2 * (axis - 1) + 1] == "reflecting":
        planes.append(grid.nghost + nx)
    for p in planes:
        flux[IDN][fv._sl(3, tdim - 1, slice(p, p + 1))] = 0.0


@dataclass
class Deck:
    grid: fv.Grid
    shock: bool = False
    bcs: tuple = ("periodic",) * 6
    cfl: float = 0.9


@dataclass
class Block:
    deck: Deck
    u: np.ndarray = None
    u0: np.ndarray = None
    w: np.ndarray = None
    flux: dict = field(default_factory=dict)

    def initialize(self, w):
        # the copy-type ghost fills are row-generic (fv.bc_gather negates the
        # normal velocity row of each axis); no outflow for shallow water
        w = w.copy()
        fv.apply_boundaries(w, self.deck.bcs, 1.4, primitive=True)
        self.w = w
        self.u = prim2cons(w)
        return self.u

    def tendency(self, dt):
        d = self.deck
        g = d.grid
        self.w = cons2prim(self.u)
        flux = {}
        for axis, tdim in ((1, 3), (2, 2), (3, 1)):
            if g.nc(axis) == 1:
                continue
            wlr = fv.reconstruct(self.w, tdim, d.shock)
            f = shallow_roe(wlr[fv.ILT], wlr[fv.IRT], tdim)
            _prescribe_walls(f, tdim, d.bcs, axis, g)
            flux[axis] = f
        self.flux = flux
        div = fv.divergence(g, flux)
        du = np.zeros_like(div)
        sub = (slice(None),) + g.interior()
        du[sub] = -dt * div[sub]
        return du

    def forward(self, dt, stage):
        """MeshBlockImpl::forward (meshblock_advance.cpp:88-105), as
        fv.Block.forward with the shallow-water stage."""
        if stage == 0:
            self.u0 = self.u.copy()
        du = self.tendency(dt)
        self.u = fv.rk_combine(self.u0, self.u, du, stage)
        fv.apply_boundaries(self.u, self.deck.bcs, 1.4)
        self._w_stale = True

    def check_redo(self):
        """w = prim(u) of the accepted state (meshblock_redo.cpp:107-112)."""
        self.w = cons2prim(self.u)
        self._w_stale = False
        return 0

    def max_time_step(self):
        """From the primitives of the accepted state, as the block's
        vars['hydro_w'] after check_redo (meshblock_advance.cpp:74-85)."""
        if getattr(self, "_w_stale", False):
            self.check_redo()
        return self.deck.cfl * max_time_step(self.deck.grid, self.w)
