"""Pins of the cubed-sphere oracle (oracle/cubed.py) by what the geometry and
the physics fix, each chosen so that a plausible slip (a wrong edge
orientation, a sign in a velocity table, a dropped metric term) fails:

* torch's linspace restated (to rounding); the six panels' solid angles sum
  to 4 pi; the radial basis vector lands on the unit position of the face
  (two independent tables, cs_face_xyz_from_tan and the local-to-Cartesian
  velocity table); the frame rotations invert each other;
* the connectivity is an involution and its edges meet on the sphere;
* the ghost fill of a field linear in the Cartesian position, and of a
  solid-body wind, converges at second (linear) or fourth (cubic) order to
  the exact values at the ghost centres;
* a lake at rest stays at rest (the metric sources balance the pressure
  flux); mass is conserved; seam face fluxes are single valued;
* a state symmetric under the cube's rotations about the polar axis stays
  symmetric: the four equatorial panels agree, and the polar panels keep
  their four-fold symmetry.
"""

import math

import numpy as np
import pytest
import torch

from oracle import cubed

IDN, IVX, IVY, IVZ = 0, 1, 2, 3


@pytest.mark.parametrize("steps", [2, 3, 7, 8, 15, 39, 199, 205])
def test_linspace_restatement_is_torch(steps):
    """torch's scalar linspace rule; its vectorised loop can round a middle
    element differently by an ulp, so the match is to rounding."""
    lo, hi = -0.25 * math.pi - 0.1, 0.25 * math.pi + 0.1
    ref = torch.linspace(lo, hi, steps, dtype=torch.float64).numpy()
    out = cubed.torch_linspace(lo, hi, steps)
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-16)
    assert out[0] == lo and out[-1] == hi


def test_solid_angles_sum_to_four_pi():
    g = cubed.Grid(24)
    sub = g.interior()
    total = 6.0 * g.solid_angle[sub].sum()
    assert abs(total - 4.0 * math.pi) < 1e-12


def test_metric_identities():
    g = cubed.Grid(12)
    for c, s in ((g.cos_c, g.sin_c), (g.cos_f2, g.sin_f2), (g.cos_f3, g.sin_f3)):
        np.testing.assert_allclose(c * c + s * s, 1.0, rtol=0, atol=1e-15)
    # the centre of the panel is orthogonal
    assert abs(g.cos_c[g.nghost + 6, g.nghost + 6, 0]) < 1e-15 or g.n % 2 == 0


@pytest.mark.parametrize("face", range(6))
def test_radial_basis_is_the_unit_position(face):
    rng = np.random.default_rng(face)
    alpha = rng.uniform(-0.7, 0.7, 20)
    beta = rng.uniform(-0.7, 0.7, 20)
    up = np.zeros((3, 20))
    up[0] = 1.0
    cubed.contra_to_cart(up, cubed.metric_at(alpha, beta), face)
    for n in range(20):
        x, y, z = cubed.ab_to_xyz(face, alpha[n], beta[n])
        np.testing.assert_allclose(up[:, n], (z, x, y), rtol=0, atol=1e-15)


@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("conserved", [False, True])
def test_rotations_invert_each_other(face, conserved):
    rng = np.random.default_rng(10 + face)
    alpha = rng.uniform(-0.9, 0.9, 15)
    beta = rng.uniform(-0.9, 0.9, 15)
    x, y = np.tan(alpha), np.tan(beta)
    cos = -x * y / (np.sqrt(1 + x * x) * np.sqrt(1 + y * y)) if conserved else None
    m_to = cubed.rotation_matrix(face, alpha, beta, True, cos)
    m_from = cubed.rotation_matrix(face, alpha, beta, False, cos)
    prod = np.einsum("ijn,jkn->ikn", m_from, m_to)
    np.testing.assert_allclose(prod, np.broadcast_to(np.eye(3)[..., None], prod.shape), atol=1e-13)
    # cart_to_contra undoes contra_to_cart
    v = rng.normal(size=(3, 15))
    w = v.copy()
    m = cubed.metric_at(alpha, beta)
    cubed.contra_to_cart(w, m, face)
    cubed.cart_to_contra(w, m, face)
    np.testing.assert_allclose(w, v, atol=1e-13)


def test_connectivity_is_an_involution_and_edges_meet():
    n = 10
    d = math.pi / (2 * n)
    for f in range(6):
        for s in range(4):
            nf, ns, rev = cubed.EDGES[f][s]
            assert cubed.EDGES[nf][ns] == (f, s, rev)

            # the along-edge point j of side s of face f is the point j (or
            # n - 1 - j when reversed) of side ns of face nf
            def edge_point(face, side, j):
                along = -math.pi / 4 + (j + 0.5) * d
                perp = math.pi / 4 * (1 if side in (1, 3) else -1)
                ab = (perp, along) if side < 2 else (along, perp)
                return np.array(cubed.ab_to_xyz(face, *ab))

            for j in (0, 3, n - 1):
                jj = n - 1 - j if rev else j
                np.testing.assert_allclose(edge_point(f, s, j), edge_point(nf, ns, jj), atol=1e-14)


def test_ghost_sources_stay_in_the_source_interior():
    for n in (8, 16, 48, 192):
        u = cubed.ghost_usrc(n, 3)
        u0 = np.floor(u).astype(int)
        assert u0.min() >= 0 and (u0 + 1).max() <= n - 1
        # the slide along the edge is towards the edge midpoint, symmetric
        np.testing.assert_allclose(u, (n - 1) - u[:, ::-1], atol=1e-9)


def _cell_xyz(g, face):
    """(3, nc3, nc2) unit positions of a panel's cell centres (ghosts too)."""
    out = np.empty((3, g.nc, g.nc))
    for k in range(g.nc):
        for j in range(g.nc):
            out[:, k, j] = cubed.ab_to_xyz(face, g.x2v[j], g.x3v[k])
    return out


def _edge_ghost_mask(g):
    m = np.zeros((g.nc, g.nc), bool)
    ng, n = g.nghost, g.n
    m[ng : ng + n, :ng] = m[ng : ng + n, ng + n :] = True
    m[:ng, ng : ng + n] = m[ng + n :, ng : ng + n] = True
    return m


def _linear_field_error(n, order):
    g = cubed.Grid(n, interp_order=order)
    coef = np.array([0.3, -0.7, 0.5])
    ws, exact = [], []
    for f in range(6):
        xyz = _cell_xyz(g, f)
        val = 2.0 + np.tensordot(coef, xyz, axes=1)
        w = np.zeros((4, g.nc, g.nc, 1))
        w[IDN, :, :, 0] = val
        exact.append(val)
        interior = np.zeros_like(val, dtype=bool)
        interior[g.interior()[:2]] = True
        w[IDN, :, :, 0] = np.where(interior, val, np.nan)
        ws.append(w)
    cubed.exchange_ghosts(g, ws, conserved=False)
    mask = _edge_ghost_mask(g)
    return max(np.abs(ws[f][IDN, :, :, 0] - exact[f])[mask].max() for f in range(6))


@pytest.mark.parametrize("order,rate", [(2, 3.3), (4, 9.0)])
def test_ghost_fill_of_a_linear_field_converges(order, rate):
    e1 = _linear_field_error(8, order)
    e2 = _linear_field_error(16, order)
    assert np.isfinite(e1) and np.isfinite(e2)
    assert e2 < 2e-3
    assert e1 / e2 > rate


def _wind_error(n, conserved):
    """A solid-body rotation about an oblique axis, in each panel's
    contravariant components; returns the ghost-fill error."""
    g = cubed.Grid(n)
    omega = np.array([0.4, -0.2, 0.9])  # (x, y, z)
    ws, exact = [], []
    for f in range(6):
        xyz = _cell_xyz(g, f)
        v = np.cross(omega[:, None, None], xyz, axis=0)  # (vx, vy, vz)
        vel = np.stack([v[2], v[0], v[1]])  # (z, x, y) storage order
        al = np.broadcast_to(g.x2v[None, :], (g.nc, g.nc))
        be = np.broadcast_to(g.x3v[:, None], (g.nc, g.nc))
        cubed.cart_to_contra(vel, cubed.metric_at(al, be), f)
        w = np.zeros((4, g.nc, g.nc, 1))
        w[IDN] = 5.0
        w[IVX : IVZ + 1, :, :, 0] = vel
        if conserved:
            w = cubed.prim2cons(w, g)
        exact.append(w.copy())
        mask = np.zeros((g.nc, g.nc), bool)
        mask[g.interior()[:2]] = True
        w[:, ~mask] = np.nan
        ws.append(w)
    cubed.exchange_ghosts(g, ws, conserved=conserved)
    mask = _edge_ghost_mask(g)
    return max(np.abs(ws[f][:, mask] - exact[f][:, mask]).max() for f in range(6))


@pytest.mark.parametrize("conserved", [False, True])
def test_ghost_fill_of_a_solid_body_wind_converges(conserved):
    e1 = _wind_error(8, conserved)
    e2 = _wind_error(16, conserved)
    assert e2 < (5.0 if conserved else 1.0) * 5e-3  # momenta carry phi = 5
    assert e1 / e2 > 3.3


def _hill(g, face, lat0, amp, rad=0.5):
    out = np.empty((g.nc, g.nc))
    c = np.array([math.cos(lat0), 0.0, math.sin(lat0)])
    xyz = _cell_xyz(g, face)
    ang = np.arccos(np.clip(np.tensordot(c, xyz, axes=1), -1, 1))
    out[:] = 1000.0 + amp * np.exp(-((ang / rad) ** 2))
    return out


def _mesh(n, phis, order=2):
    g = cubed.Grid(n, radius=6.371e6, interp_order=order)
    m = cubed.Mesh(g)
    ws = []
    for f in range(6):
        w = np.zeros((4, g.nc, g.nc, 1))
        w[IDN, :, :, 0] = 9.81 * phis(g, f)
        ws.append(w)
    m.initialize(ws)
    return g, m


def _step(m, nsteps):
    for _ in range(nsteps):
        dt = m.max_time_step()
        for s in range(3):
            m.forward(dt, s)


def test_lake_at_rest_on_the_sphere():
    g, m = _mesh(8, lambda g, f: np.full((g.nc, g.nc), 1000.0))
    _step(m, 3)
    sub = (slice(None),) + g.interior()
    c = math.sqrt(9.81 * 1000.0)
    for f in range(6):
        u = m.u[f][sub]
        assert np.abs(u[IDN] / (9.81 * 1000.0) - 1.0).max() < 1e-13
        assert np.abs(u[IVY:] / u[IDN]).max() < 1e-11 * c


def test_mass_is_conserved_and_seam_fluxes_are_single_valued():
    g, m = _mesh(8, lambda g, f: _hill(g, f, 0.3, 120.0))
    m0 = m.mass()
    _step(m, 3)
    assert abs(m.mass() - m0) < 1e-13 * abs(m0)
    f2, f3 = m.flux["x2"], m.flux["x3"]
    ng, n = g.nghost, g.n
    plane = (ng, ng + n, ng, ng + n)
    for p in range(6):
        for s in range(4):
            q, qs, rev = cubed.EDGES[p][s]
            a = cubed._slab(f2[p] if s < 2 else f3[p], g, s)[IDN]
            b = cubed._slab(f2[q] if qs < 2 else f3[q], g, qs)[IDN]
            if rev:
                b = b[::-1]
            sign = -1.0 if (s % 2) == (qs % 2) else 1.0
            np.testing.assert_array_equal(a, sign * b)
            assert plane[s] in (ng, ng + n)


def test_polar_symmetry_is_kept():
    """Hills on both poles: every quarter turn about the polar axis maps the
    state onto itself, so panels 0, 1, 2, 4 hold the same arrays and panels
    3 and 5 are four-fold symmetric."""

    def phis(g, f):
        return _hill(g, f, math.pi / 2, 150.0) + _hill(g, f, -math.pi / 2, 60.0) - 1000.0

    g, m = _mesh(8, phis)
    _step(m, 4)
    sub = (slice(None),) + g.interior()
    u = [m.u[f][sub][..., 0] for f in range(6)]
    # phi by its own scale, the momentum rows (the radial one at rounding
    # level) by the largest momentum
    scale = np.array([np.abs(u[0][IDN]).max()] + [np.abs(u[0][IVY:]).max()] * 3)[:, None, None]
    for f in (1, 2, 4):
        np.testing.assert_allclose(u[f] / scale, u[0] / scale, rtol=0, atol=1e-12)
    # +Z: (alpha, beta) -> (-beta, alpha) is a quarter turn; -Z likewise
    for f in (3, 5):
        phi = u[f][IDN]
        np.testing.assert_allclose(phi.T[:, ::-1], phi, rtol=1e-13)
    # the hills have moved: the test is not of a resting state
    assert np.abs(u[0][IVY:]).max() > 1e-3 * scale[IDN, 0, 0]
