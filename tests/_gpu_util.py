"""Helpers of the GPU parity tests: build descriptors and call the C-ABI."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from oracle import fv
from snapy_amd import _capi


def grid_desc(g: fv.Grid, nvar=5):
    return _capi.Grid(
        nvar,
        g.nghost,
        g.nx1,
        g.nx2,
        g.nx3,
        g.nc(1),
        g.nc(2),
        g.nc(3),
        g.dx(1),
        g.dx(2),
        g.dx(3),
        g.x1min,
        g.x2min,
        g.x3min,
    )


def phys_desc(gammad=1.4, riemann="hllc", shock_x1=True, shock_x23=True, bcs=("outflow",) * 6, grav1=0.0):
    b = (ctypes.c_int32 * 6)(*[_capi.BC[n] for n in bcs])
    return _capi.Physics(gammad, grav1, 1e-10, 1e-10, _capi.RIEMANN[riemann], int(shock_x1), int(shock_x23), b)


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return num / den if den > 0 else num


def random_state(grid: fv.Grid, seed=0, amp=0.1):
    """A positive smooth-plus-noise primitive state on the whole block."""
    rng = np.random.default_rng(seed)
    shp = (5,) + grid.shape
    w = np.zeros(shp)
    w[fv.IDN] = 1.0 + 0.3 * rng.random(shp[1:])
    w[fv.IPR] = 1.0 + 0.3 * rng.random(shp[1:])
    for v in (fv.IVX, fv.IVY, fv.IVZ):
        w[v] = amp * (rng.random(shp[1:]) - 0.5)
    return w
