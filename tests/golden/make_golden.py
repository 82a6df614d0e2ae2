"""Write the golden fixtures of the parity tests from the CPU FP64 oracle.

Run from the repository root:  python tests/golden/make_golden.py

Each fixture is a compressed .npz of plain float64/int arrays (no pickles)
holding the inputs of one case and the oracle's outputs, so that the GPU
tests (tests/test_gpu_golden.py) compare the HIP path against frozen
numbers and tests/test_golden.py checks on the CPU that the oracle still
produces them:

* sod_c1.npz: the Sod run of the shock deck (256 cells, WENO5 shock +
  HLLC + SSP-RK3, CFL 0.9, outflow, t_end = 0.25): the final conserved
  state and every step's dt;
* sweeps.npz: one x1 and one x2 sweep (face flux totals, and the x1 face
  pressure) of a seeded random state on a 16 x 12 block, for HLLC with
  the WENO5 shock reconstruction and for LMARS with CP5 velocity/pressure;
* vic_columns.npz: one call of the vertical implicit correction, scheme 9
  (full 5x5 blocks) and scheme 1 (partial 3x3 blocks), on the columns of a
  16 x 6 stratified block with random increments;
* swe_seams.npz: the cubed-sphere shallow-water panel-edge ghost fill
  (interpolation orders 2 and 4, conserved) and one stage's x2/x3 face
  fluxes with the seams reconciled, on 8 x 8 panels;
* tracer_block.npz: two passive tracers (a sharp slab and a height ramp)
  on a 16 x 12 stratified block with gravity, diffusion, the covariance
  with its positivity limiter and the scheme-9 implicit correction with
  the donor-capped tracer transfer: the initial state, every step's dt and
  the final hydro and tracer states after 2 RK3 steps.
"""

from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import cubed as oc  # noqa: E402
from oracle import fv, implicit  # noqa: E402

RD, CP = 287.0, 1004.0
GAMMA_DRY = CP / (CP - RD)


# ---------------------------------------------------------------------------
# inputs
# ---------------------------------------------------------------------------
def random_state(grid, seed, amp=0.1):
    rng = np.random.default_rng(seed)
    shp = (5,) + grid.shape
    w = np.zeros(shp)
    w[fv.IDN] = 1.0 + 0.3 * rng.random(shp[1:])
    w[fv.IPR] = 1.0 + 0.3 * rng.random(shp[1:])
    for v in (fv.IVX, fv.IVY, fv.IVZ):
        w[v] = amp * (rng.random(shp[1:]) - 0.5)
    return w


def stratified(grid, seed, noise):
    """A dry isentropic column (theta = 300 K) with the density current's
    cold bubble and seeded velocity noise."""
    rng = np.random.default_rng(seed)
    z = grid.centers(1)[None, None, :]
    x = grid.centers(2)[None, :, None]
    t = 300.0 - 9.81 * z / CP + 0.0 * x
    p = 1.0e5 * (t / 300.0) ** (CP / RD)
    L = np.sqrt((x / 4000.0) ** 2 + ((z - 3000.0) / 2000.0) ** 2)
    t = t - np.where(L <= 1.0, 15.0 * (1 + np.cos(np.pi * L)) / 2, 0.0)
    w = np.zeros((5,) + grid.shape)
    w[fv.IPR] = p
    w[fv.IDN] = p / (RD * t)
    for v in (fv.IVX, fv.IVY):
        w[v] = noise * (rng.random(grid.shape) - 0.5)
    return w


def hill_state(g, seed):
    """A seeded Gaussian hill of phi = g h on the sphere plus a small
    solid-body wind, per panel (4, nc, nc, 1), velocities contravariant."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(50.0, 200.0)
    rad = rng.uniform(3.0e5, 1.0e6) / 6.371e6
    lat = math.asin(rng.uniform(-1.0, 1.0))
    lon = rng.uniform(0.0, 2.0 * math.pi)
    c = np.array([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)])
    omega = rng.normal(size=3) * 2e-6
    ws = []
    for f in range(6):
        xyz = np.empty((3, g.nc, g.nc))
        for k in range(g.nc):
            for j in range(g.nc):
                xyz[:, k, j] = oc.ab_to_xyz(f, g.x2v[j], g.x3v[k])
        ang = np.arccos(np.clip(np.tensordot(c, xyz, axes=1), -1, 1))
        w = np.zeros((4, g.nc, g.nc, 1))
        w[0, :, :, 0] = 9.81 * (1000.0 + amp * np.exp(-((ang / rad) ** 2)))
        v = np.cross(omega[:, None, None], xyz, axis=0) * 6.371e6
        vel = np.stack([v[2], v[0], v[1]])
        al = np.broadcast_to(g.x2v[None, :], (g.nc, g.nc))
        be = np.broadcast_to(g.x3v[:, None], (g.nc, g.nc))
        oc.cart_to_contra(vel, oc.metric_at(al, be), f)
        w[1:, :, :, 0] = vel
        w[1] = 0.0
        ws.append(w)
    return np.stack(ws)


# ---------------------------------------------------------------------------
# the cases (each returns a dict of arrays)
# ---------------------------------------------------------------------------
def sod_c1():
    d = fv.sod_deck(256)
    b = fv.Block(d)
    b.initialize(fv.sod_initial(d.grid))
    t, dts = 0.0, []
    while t < 0.25:
        dt = min(b.max_time_step(), 0.25 - t)
        dts.append(dt)
        for s in range(3):
            b.forward(dt, s)
        t += dt
    return {"u": b.u, "dts": np.array(dts)}


SWEEP_GRID = dict(nx1=16, nx2=12, nx3=1, x1min=0.0, x1max=2.0, x2min=-1.0, x2max=1.0, x3min=0.0, x3max=3.0)
SWEEP_CASES = (("hllc", True), ("lmars", False))


def sweep_grid():
    a = SWEEP_GRID
    return fv.Grid(a["nx1"], a["nx2"], a["nx3"], a["x1min"], a["x1max"], a["x2min"], a["x2max"], a["x3min"],
                   a["x3max"])


def sweeps():
    g = sweep_grid()
    w = random_state(g, seed=17)
    out = {"w": w}
    for riemann, shock in SWEEP_CASES:
        solver = fv.hllc if riemann == "hllc" else fv.lmars
        for axis, tdim in ((1, 3), (2, 2)):
            wlr = fv.reconstruct(w, tdim, shock)
            f, p = solver(wlr[fv.ILT], wlr[fv.IRT], 1.4, tdim)
            out[f"{riemann}_flux_x{axis}"] = f
            if axis == 1:
                out[f"{riemann}_pface_x1"] = p
    return out


VIC_GRID = dict(nx1=16, nx2=6, nx3=1, x1min=0.0, x1max=800.0, x2min=-400.0, x2max=400.0, x3min=0.0, x3max=300.0)
VIC_DT = 0.7
VIC_GRAV = -9.81


def vic_grid():
    a = VIC_GRID
    return fv.Grid(a["nx1"], a["nx2"], a["nx3"], a["x1min"], a["x1max"], a["x2min"], a["x2max"], a["x3min"],
                   a["x3max"])


def vic_columns():
    g = vic_grid()
    w = stratified(g, seed=5, noise=0.5)
    rng = np.random.default_rng(9)
    du = np.zeros_like(w)
    sub = (slice(None),) + g.interior()
    du[sub] = 1e-3 * (rng.random(du[sub].shape) - 0.5) * np.array([1, 1, 1, 1, 1e3])[:, None, None, None]
    k, j, _ = g.interior()
    n, ng = g.nx1, g.nghost
    wc = w[:, k, j, ng - 1 : ng + n + 1]
    ncol = wc.shape[1] * wc.shape[2]
    wcol = np.moveaxis(wc, 0, -1).reshape(ncol, n + 2, 5)
    dcol = np.moveaxis(du[:, k, j, ng : ng + n], 0, -1).reshape(ncol, n, 5)
    a1, vol = g.dx(3) * g.dx(2), g.dx(3) * (g.dx(2) * g.dx(1))
    x1f, x1v = g.faces(1)[ng : ng + n + 1], g.centers(1)[ng : ng + n]
    out = {"w": w, "du": du}
    for scheme, solve in ((9, implicit.vic_full), (1, implicit.vic_partial)):
        out[f"delta_scheme{scheme}"] = solve(dcol, wcol, GAMMA_DRY, VIC_GRAV, VIC_DT, g.dx(1), x1f, x1v, a1, vol)
    return out


SWE_N = 8


def swe_seams():
    out = {}
    for order in (2, 4):
        g = oc.Grid(SWE_N, interp_order=order)
        rng = np.random.default_rng(40 + order)
        us = []
        for _ in range(6):
            u = rng.normal(size=(4, g.nc, g.nc, 1))
            u[0] += 10.0
            us.append(u)
        out[f"ghost_in_order{order}"] = np.stack(us)
        oc.exchange_ghosts(g, us, True)
        out[f"ghost_out_order{order}"] = np.stack(us)
    g = oc.Grid(SWE_N)
    ws = hill_state(g, 11)
    ref = oc.Mesh(g)
    ref.initialize([w.copy() for w in ws])
    f2, f3 = ref.fluxes()
    out["w"] = ws
    out["flux2"] = np.stack(f2)
    out["flux3"] = np.stack(f3)
    return out


TRACER_GRID = dict(nx1=16, nx2=12, nx3=1, x1min=0.0, x1max=6400.0, x2min=-6400.0, x2max=6400.0, x3min=0.0,
                   x3max=100.0)
TRACER_NU = 25.0


def tracer_grid():
    a = TRACER_GRID
    return fv.Grid(a["nx1"], a["nx2"], a["nx3"], a["x1min"], a["x1max"], a["x2min"], a["x2max"], a["x3min"],
                   a["x3max"])


def tracer_deck():
    return fv.Deck(grid=tracer_grid(), gammad=GAMMA_DRY, riemann="lmars", shock_x1=False, shock_x23=False,
                   bcs=("reflecting",) * 6, grav1=VIC_GRAV, cfl=0.9, implicit_scheme=9, flux_covariance=True,
                   nu_iso=TRACER_NU, kappa_iso=TRACER_NU, nscalar=2)


def tracer_block():
    d = tracer_deck()
    g = d.grid
    w = stratified(g, seed=6, noise=0.5)
    x2 = g.centers(2)[None, :, None] + np.zeros(g.shape)
    z = g.centers(1)[None, None, :] + np.zeros(g.shape)
    r = np.stack([np.where(np.abs(x2) < 2000.0, 1.0, 0.0), 0.2 + z / (g.x1max - g.x1min)])
    b = fv.Block(d)
    b.initialize(w, r=r)
    dts = []
    for _ in range(2):
        dt = b.max_time_step()
        dts.append(dt)
        for s in range(3):
            b.forward(dt, s)
    return {"w": w, "r": r, "dts": np.array(dts), "u": b.u, "s": b.s}


CASES = {"sod_c1": sod_c1, "sweeps": sweeps, "vic_columns": vic_columns, "swe_seams": swe_seams,
         "tracer_block": tracer_block}


def path(name):
    return os.path.join(HERE, name + ".npz")


def load(name):
    with np.load(path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def main(names=None):
    for name in names or CASES:
        arrays = CASES[name]()
        np.savez_compressed(path(name), **arrays)
        print(f"{name}: {', '.join(f'{k}{tuple(v.shape)}' for k, v in arrays.items())}")


if __name__ == "__main__":
    main(sys.argv[1:] or None)
