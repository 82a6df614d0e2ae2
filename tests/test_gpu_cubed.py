"""The cubed-sphere shallow-water stage on the GPU against the oracle
(oracle/cubed.py), bit for bit: the panel-edge ghost fill, one stage's seam
face fluxes, and whole RK3 steps of several members in one launch; the
hipGraph replay against eager; and, at the benchmark's 192 x 192 panels,
the size-independent properties (mass conserved, a lake at rest, the
polar symmetry)."""

import math
import os

import numpy as np
import pytest
import torch

from oracle import cubed as oc

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IDN, IVX, IVY, IVZ = 0, 1, 2, 3


def _mesh(n, order=2, members=1, execution="eager"):
    from snapy_amd.cubed import CubedSphereMesh
    from snapy_amd.options import MeshOptions

    mo = MeshOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "swe_cubed.yaml"))
    mo.block.coord.nx2 = mo.block.coord.nx3 = n
    mo.block.coord.interp_order = order
    mo.block.intg.execution = execution
    m = CubedSphereMesh(mo, members=members)
    m.to("cuda")
    return m


def _xyz(g, face):
    out = np.empty((3, g.nc, g.nc))
    for k in range(g.nc):
        for j in range(g.nc):
            out[:, k, j] = oc.ab_to_xyz(face, g.x2v[j], g.x3v[k])
    return out


def _hill_state(g, seed):
    """A seeded Gaussian hill of h (C3's initial state, at rest) plus a
    small seeded wind so the velocity rows are exercised."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(50.0, 200.0)
    rad = rng.uniform(3.0e5, 1.0e6) / 6.371e6
    lat = math.asin(rng.uniform(-1.0, 1.0))
    lon = rng.uniform(0.0, 2.0 * math.pi)
    c = np.array([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)])
    omega = rng.normal(size=3) * 2e-6
    ws = []
    for f in range(6):
        xyz = _xyz(g, f)
        ang = np.arccos(np.clip(np.tensordot(c, xyz, axes=1), -1, 1))
        w = np.zeros((4, g.nc, g.nc, 1))
        w[IDN, :, :, 0] = 9.81 * (1000.0 + amp * np.exp(-((ang / rad) ** 2)))
        v = np.cross(omega[:, None, None], xyz, axis=0) * 6.371e6
        vel = np.stack([v[2], v[0], v[1]])
        al = np.broadcast_to(g.x2v[None, :], (g.nc, g.nc))
        be = np.broadcast_to(g.x3v[:, None], (g.nc, g.nc))
        oc.cart_to_contra(vel, oc.metric_at(al, be), f)
        w[IVX:, :, :, 0] = vel
        w[IVX] = 0.0
        ws.append(w)
    return ws


def _not_corner(g):
    m = np.ones((g.nc, g.nc), bool)
    for a in (slice(0, g.nghost), slice(g.nghost + g.n, g.nc)):
        for b in (slice(0, g.nghost), slice(g.nghost + g.n, g.nc)):
            m[a, b] = False
    return m


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("conserved", [0, 1])
def test_ghost_fill_is_the_oracles(order, conserved):
    import ctypes

    from snapy_amd import _capi

    n = 12
    print(f"cs ghosts n={n} order={order} conserved={conserved}", flush=True)
    m = _mesh(n, order)
    g = oc.Grid(n, interp_order=order)
    rng = np.random.default_rng(3 + order + conserved)
    us = []
    for f in range(6):
        u = rng.normal(size=(4, g.nc, g.nc, 1))
        u[IDN] += 10.0
        us.append(u)
    q = torch.from_numpy(np.stack(us)).cuda().contiguous()
    lib = _capi.load()
    _capi.check(lib.snapy_cs_ghosts(ctypes.byref(m._desc), conserved, _capi.ptr(q), _capi.stream_of(q.device)),
                "cs_ghosts")
    out = q.cpu().numpy()
    oc.exchange_ghosts(g, us, bool(conserved))
    mask = _not_corner(g)
    for f in range(6):
        np.testing.assert_array_equal(out[f][:, mask], us[f][:, mask])


def test_stage_seam_fluxes_are_the_oracles():
    n = 12
    print(f"cs fluxes n={n}", flush=True)
    g = oc.Grid(n)
    ws = _hill_state(g, 11)
    ref = oc.Mesh(g)
    ref.initialize([w.copy() for w in ws])
    f2, f3 = ref.fluxes()
    m = _mesh(n)
    vars = [{"hydro_w": torch.from_numpy(w)} for w in ws]
    m.initialize(vars)
    for name, fn in m._stage_calls(0)[:4]:
        from snapy_amd import _capi

        _capi.check(fn(), name)
    got2, got3 = m._f2.cpu().numpy(), m._f3.cpu().numpy()
    ng = g.nghost
    i, fa = slice(ng, ng + n), slice(ng, ng + n + 1)
    for f in range(6):
        np.testing.assert_array_equal(got2[f][:, i, fa], f2[f][:, i, fa])
        np.testing.assert_array_equal(got3[f][:, fa, i], f3[f][:, fa, i])


@pytest.mark.parametrize("order", [2, 4])
def test_steps_of_two_members_are_the_oracles(order):
    n, steps = 12, 3
    print(f"cs steps n={n} order={order} members=2", flush=True)
    g = oc.Grid(n, interp_order=order)
    states = [_hill_state(g, 21), _hill_state(g, 22)]
    refs = []
    for ws in states:
        r = oc.Mesh(g)
        r.initialize([w.copy() for w in ws])
        refs.append(r)
    m = _mesh(n, order, members=2)
    vars = [{"hydro_w": torch.from_numpy(w)} for ws in states for w in ws]
    m.initialize(vars)
    for _ in range(steps):
        dts = [r.max_time_step() for r in refs]
        got = m.max_time_step(vars)
        np.testing.assert_array_equal(got, dts)
        for s in range(3):
            m.forward(vars, torch.tensor(dts, dtype=torch.float64, device="cuda"), s)
            for r, dt in zip(refs, dts):
                r.forward(dt, s)
        assert m.check_redo(vars) == 0
    u = m._u.cpu().numpy()
    mask = _not_corner(g)
    for k, r in enumerate(refs):
        for f in range(6):
            np.testing.assert_array_equal(u[6 * k + f][:, mask], r.u[f][:, mask])
    # the views the caller holds are the resident state
    assert vars[7]["hydro_u"].data_ptr() == m._u[7].data_ptr()


def test_graph_replay_is_eager():
    n = 16
    print(f"cs graph vs eager n={n}", flush=True)
    g = oc.Grid(n)
    ws = _hill_state(g, 5)
    out = []
    for execution in ("eager", "hip-graph"):
        m = _mesh(n, execution=execution)
        vars = [{"hydro_w": torch.from_numpy(w)} for w in ws]
        m.initialize(vars)
        for _ in range(4):
            dt = m.device_time_step(vars)
            for s in range(3):
                m.forward(vars, dt, s)
        torch.cuda.synchronize()
        out.append(m._u.cpu().numpy())
        if execution == "hip-graph":
            rep = m.execution_report()
            assert rep["captures"] == 3 and rep["replays"] >= 6
        m.release()
    np.testing.assert_array_equal(out[0], out[1])


def test_full_size_properties():
    """192 x 192 panels, four members: mass conserved, a lake at rest stays
    at rest, and a polar state keeps the cube's quarter-turn symmetry."""
    n = 192
    print(f"cs full size n={n} members=4", flush=True)
    m = _mesh(n, members=4, execution="hip-graph")
    g = m.grid
    nc = g.nc
    X = np.broadcast_to(g.x2v[None, :], (nc, nc))
    Y = np.broadcast_to(g.x3v[:, None], (nc, nc))
    vars = []
    for member in range(4):
        for f in range(6):
            w = np.zeros((4, nc, nc, 1))
            if member == 0:
                w[IDN] = 9.81 * 1000.0  # lake at rest
            else:
                # hills on both poles, member-scaled: symmetric under quarter
                # turns about the polar axis
                xx = np.tan(X)
                yy = np.tan(Y)
                r = np.sqrt(1 + xx * xx + yy * yy)
                z = [yy / r, yy / r, yy / r, 1 / r, yy / r, -1 / r][f]
                w[IDN, :, :, 0] = 9.81 * (1000.0 + 50.0 * member * np.exp(-((1 - z) / 0.05)) +
                                          20.0 * np.exp(-((1 + z) / 0.05)))
            vars.append({"hydro_w": torch.from_numpy(w)})
    m.initialize(vars)
    m0 = m.mass().cpu().numpy()
    for _ in range(10):
        dt = m.device_time_step(vars)
        for s in range(3):
            m.forward(vars, dt, s)
    assert m.check_redo(vars) == 0
    m1 = m.mass().cpu().numpy()
    np.testing.assert_allclose(m1, m0, rtol=1e-13, atol=0)
    i = slice(g.ng, g.ng + n)
    u = m._u[:, :, i, i, 0].cpu().numpy()
    c = math.sqrt(9.81 * 1000.0)
    assert np.abs(u[:6, IDN] / (9.81 * 1000.0) - 1.0).max() < 1e-12
    assert np.abs(u[:6, IVY:] / u[:6, IDN:IDN + 1]).max() < 1e-10 * c
    for member in (1, 2, 3):
        b = 6 * member
        scale = np.array([np.abs(u[b][IDN]).max()] + [np.abs(u[b][IVY:]).max()] * 3)[:, None, None]
        assert scale[IVY] > 0
        # the panels' frame tables differ in the last bits, so the copies
        # agree to rounding accumulated over the steps
        for f in (1, 2, 4):
            np.testing.assert_allclose(u[b + f] / scale, u[b] / scale, rtol=0, atol=5e-11)
        for f in (3, 5):
            phi = u[b + f][IDN]
            np.testing.assert_allclose(phi.T[:, ::-1], phi, rtol=1e-12)
    m.release()
