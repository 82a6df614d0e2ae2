"""Initial states of the benchmark decks (the reference's example drivers
build them in C++: examples/shock.cpp, examples/straka.cpp; the 3D
convection state follows the same isentropic-base recipe)."""

from __future__ import annotations

import numpy as np
import torch

RD, CP = 287.0, 1004.0


def sod(block, x0=0.5, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1)):
    """Sod tube on x1 (examples/shock.cpp:37-47 with a 1D grid)."""
    x1v = block.module("coord").buffer("x1v").numpy()[None, None, :]
    w = np.zeros(block.shape())
    w[0] = np.where(x1v < x0, left[0], right[0])
    w[1] = np.where(x1v < x0, left[1], right[1])
    w[4] = np.where(x1v < x0, left[2], right[2])
    return torch.from_numpy(w)


def _grav(block, grav):
    """The gravity a state is balanced for: the given magnitude, else the
    block's own x1 gravity (the reference's example drivers read it from the
    deck, examples/straka.cpp:62), else 9.81."""
    if grav is not None:
        return float(grav)
    g1 = float(getattr(block.options.hydro, "grav1", 0.0) or 0.0)
    return -g1 if g1 != 0.0 else 9.81


def isentropic(block, theta=300.0, psfc=1.0e5, grav=None):
    """Dry isentropic column: T = theta - g z / cp, p = psfc (T/theta)^(cp/Rd)."""
    grav = _grav(block, grav)
    z = block.module("coord").buffer("x1v").numpy()[None, None, :]
    t = theta - grav * z / CP + 0.0 * np.zeros(block.shape()[1:])
    p = psfc * (t / theta) ** (CP / RD)
    return z, t, p


def convection3d(block, seed=2024, amp=0.1, layers=10, theta=300.0, psfc=1.0e5, grav=None):
    """3D dry convection: isentropic base, theta noise U(-amp, amp) K in the
    lowest `layers` interior layers from a seeded generator, at rest."""
    z, t, p = isentropic(block, theta, psfc, grav)
    c = block.options.coord
    ng = c.nghost
    rng = np.random.default_rng(seed)
    dth = np.zeros(block.shape()[1:])
    noise = rng.uniform(-amp, amp, size=(c.nx3, c.nx2, layers))
    dth[ng : ng + c.nx3, ng : ng + c.nx2, ng : ng + layers] = noise
    # theta' at fixed p: T' = theta' (p / psfc)^(Rd/cp)
    t = t + dth * (p / psfc) ** (RD / CP)
    w = np.zeros(block.shape())
    w[4] = p
    w[0] = p / (RD * t)
    return torch.from_numpy(w)


def density_current(block, dT=-15.0, theta=300.0, psfc=1.0e5, grav=None, xc=0.0, xr=4000.0, zc=3000.0, zr=2000.0):
    """The dry density current (examples/straka.cpp): a cold bubble
    dT (1 + cos(pi L)) / 2 for L <= 1 applied to T at fixed p."""
    z, t, p = isentropic(block, theta, psfc, grav)
    x = block.module("coord").buffer("x2v").numpy()[None, :, None]
    L = np.sqrt(((x - xc) / xr) ** 2 + ((z - zc) / zr) ** 2)
    t = t + np.where(L <= 1.0, dT * (1.0 + np.cos(np.pi * L)) / 2.0, 0.0)
    w = np.zeros(block.shape())
    w[4] = p
    w[0] = p / (RD * t)
    return torch.from_numpy(w)


def swe_hill(mesh, seed, radius=6.371e6, grav=9.81, h0=1000.0):
    """The batched shallow-water initial state: phi = g (h0 + A exp(-(d/R)^2))
    at rest, A ~ U(50, 200) m, R ~ U(300, 1000) km, d the great-circle
    distance to a centre with latitude asin(U(-1, 1)) and longitude
    U(0, 2 pi), all drawn from `seed`.  Returns one hydro_w per panel."""
    import math

    from .cubed import sphere_points

    rng = np.random.default_rng(seed)
    amp = rng.uniform(50.0, 200.0)
    rad = rng.uniform(3.0e5, 1.0e6)
    lat = math.asin(rng.uniform(-1.0, 1.0))
    lon = rng.uniform(0.0, 2.0 * math.pi)
    c = np.array([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)])
    out = []
    for panel in range(6):
        xyz = sphere_points(mesh.grid, panel)
        d = radius * np.arccos(np.clip(np.tensordot(c, xyz, axes=1), -1.0, 1.0))
        w = np.zeros(mesh.block_shape())
        w[0, :, :, 0] = grav * (h0 + amp * np.exp(-((d / rad) ** 2)))
        out.append(torch.from_numpy(w))
    return out

