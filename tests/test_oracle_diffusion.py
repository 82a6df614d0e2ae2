"""Pins of the diffusion oracle (oracle/diffusion.py) against closed forms:
a plane shear carries a uniform stress, a sine profile decays at nu k^2
(second order in dx), the bulk viscosity gives the 4/3 factor of a
compressional profile, the conductive flux is -kappa rho cv dT/dx, the
viscous work is v_face * flux, the x1 wall face extrapolates the density,
and the time-step bound is dx^2 / (2 ndim nu)."""

import numpy as np
import pytest

from oracle import diffusion, fv

WEIGHT = 0.02897025302492418
GAMMA = 1.4
RD = diffusion.RGAS / WEIGHT
CV = RD / (GAMMA - 1.0)
PERIODIC = ("periodic",) * 6
WALLS = ("reflecting", "reflecting") + ("periodic",) * 4


def base_state(g, rho=1.2, p=1.0e5):
    w = np.zeros((5,) + g.shape)
    w[fv.IDN] = rho
    w[fv.IPR] = p
    return w


def fill(w, g, bcs):
    fv.apply_boundaries(w, bcs, GAMMA, None, primitive=True)
    return w


def test_uniform_state_has_no_flux():
    g = fv.Grid(8, 8, 1, 0.0, 800.0, 0.0, 800.0, 0.0, 100.0)
    w = fill(base_state(g), g, PERIODIC)
    w[fv.IVY] = 3.0
    f = diffusion.cartesian_fluxes(w, g, 75.0, 75.0, GAMMA, WEIGHT, PERIODIC)
    # a uniform velocity carries no stress; a uniform temperature no heat
    assert np.all(f[1][fv.IVX : fv.IPR + 1] == 0.0)
    assert np.all(f[2][fv.IVX : fv.IPR + 1] == 0.0)


def test_plane_shear_uniform_stress_and_work():
    """v2 = a x1 with rho const: tau_12 = nu rho a on every x1 face, its work
    v2_face tau, and zero momentum divergence away from the boundaries."""
    g = fv.Grid(12, 6, 1, 0.0, 1200.0, 0.0, 600.0, 0.0, 100.0)
    a, nu, rho = 0.01, 75.0, 1.2
    w = base_state(g, rho=rho)
    x1v = g.centers(1)[None, None, :]
    w[fv.IVY] = a * x1v + 0.0 * w[fv.IDN]
    f = diffusion.cartesian_fluxes(w, g, nu, 0.0, GAMMA, WEIGHT, PERIODIC)
    ng = g.nghost
    faces = slice(ng, ng + g.nx1 + 1)
    js = slice(ng, ng + g.nx2)
    f1 = f[1][fv.IVY][0, js, faces]
    np.testing.assert_allclose(f1, -nu * rho * a, rtol=1e-12)
    x1f = g.faces(1)[faces]
    np.testing.assert_allclose(
        f[1][fv.IPR][0, js, faces], np.broadcast_to((a * x1f) * (-nu * rho * a), (g.nx2, g.nx1 + 1)), rtol=1e-10
    )
    # the x2 faces carry the symmetric part d v1/d x2 = 0 plus the mean of
    # d v2/d x1 = a along the face: tau_21 = nu rho a in the IVX row
    f2 = f[2][fv.IVX][0, slice(ng, ng + g.nx2 + 1), slice(ng + 1, ng + g.nx1 - 1)]
    np.testing.assert_allclose(f2, -nu * rho * a, rtol=1e-12)
    du = np.zeros_like(w)
    diffusion.add_momentum(du, f, g, 1.0)
    inner = du[fv.IVY][0, js, ng + 1 : ng + g.nx1 - 1]
    assert np.max(np.abs(inner)) < 1e-13 * nu * rho * a


def _sine_decay_error(nx, bulk):
    """Return the max error of the momentum tendency against the closed
    form nu rho c v''(x) for a sine profile of v1 (bulk: c = 4/3) or v2
    (c = 1), rho const, periodic x1."""
    L, nu, rho, k = 1000.0, 75.0, 1.2, 2.0 * np.pi / 1000.0
    g = fv.Grid(nx, 1, 1, 0.0, L, 0.0, 100.0, 0.0, 100.0)
    w = base_state(g, rho=rho)
    x1v = g.centers(1)[None, None, :]
    row = fv.IVX if bulk else fv.IVY
    w[row] = 0.5 * np.sin(k * x1v) + 0.0 * w[fv.IDN]
    f = diffusion.cartesian_fluxes(w, g, nu, 0.0, GAMMA, WEIGHT, PERIODIC)
    du = np.zeros_like(w)
    diffusion.add_momentum(du, f, g, 1.0)
    c = 4.0 / 3.0 if bulk else 1.0
    exact = -nu * rho * c * 0.5 * k * k * np.sin(k * x1v)
    sub = (0, 0, slice(g.nghost, g.nghost + nx))
    return np.max(np.abs(du[row][sub] - exact[0, 0, g.nghost : g.nghost + nx]))


@pytest.mark.parametrize("bulk", [False, True])
def test_sine_profile_decays_at_nu_k2_second_order(bulk):
    e1 = _sine_decay_error(32, bulk)
    e2 = _sine_decay_error(64, bulk)
    # a sign error or a dropped 4/3 leaves an O(1) error; second order
    # halves the spacing and quarters the error
    assert e1 < 0.02 * 75.0 * 1.2 * (4 / 3) * 0.5 * (2 * np.pi / 1000.0) ** 2
    assert 3.6 < e1 / e2 < 4.4


def test_fourier_flux_linear_temperature():
    """T = T0 + b x1 with rho const: the energy flux is -kappa rho cv b on
    every x1 face, so the conductive tendency of the interior vanishes."""
    g = fv.Grid(10, 1, 1, 0.0, 500.0, 0.0, 100.0, 0.0, 100.0)
    kappa, rho, b = 75.0, 1.1, -0.0065
    w = base_state(g, rho=rho)
    x1v = g.centers(1)[None, None, :]
    temp = 300.0 + b * x1v
    w[fv.IPR] = rho * RD * temp + 0.0 * w[fv.IDN]
    f = diffusion.cartesian_fluxes(w, g, 0.0, kappa, GAMMA, WEIGHT, PERIODIC)
    ng = g.nghost
    faces = slice(ng + 1, ng + g.nx1)
    np.testing.assert_allclose(f[1][fv.IPR][0, 0, faces], -kappa * rho * CV * b, rtol=1e-9)


def test_x1_wall_face_extrapolates_density():
    g = fv.Grid(8, 1, 1, 0.0, 800.0, 0.0, 100.0, 0.0, 100.0)
    w = base_state(g)
    ng = g.nghost
    rho = np.array([1.2, 1.1, 1.0, 0.9, 0.8, 0.7, 0.6, 0.5])
    w[fv.IDN][0, 0, ng : ng + 8] = rho
    fill(w, g, WALLS)
    x1v = g.centers(1)[None, None, :]
    w[fv.IVY] = 0.01 * x1v + 0.0 * w[fv.IDN]
    nu = 75.0
    f = diffusion.cartesian_fluxes(w, g, nu, 0.0, GAMMA, WEIGHT, WALLS)
    # stress at the wall faces: v2 ghosts mirror the interior, so the
    # face-normal derivative there is (v2(lo) - v2(lo-1)) / dx
    lo, hi = ng, ng + 8
    stress_lo = (w[fv.IVY][0, 0, lo] - w[fv.IVY][0, 0, lo - 1]) / g.dx(1)
    rho_lo = 1.5 * rho[0] - 0.5 * rho[1]
    assert f[1][fv.IVY][0, 0, lo] == pytest.approx(-nu * rho_lo * stress_lo, rel=1e-14)
    stress_hi = (w[fv.IVY][0, 0, hi] - w[fv.IVY][0, 0, hi - 1]) / g.dx(1)
    rho_hi = 1.5 * rho[-1] - 0.5 * rho[-2]
    assert f[1][fv.IVY][0, 0, hi] == pytest.approx(-nu * rho_hi * stress_hi, rel=1e-14)
    # a face with a non-positive extrapolation takes the nearest cell
    w[fv.IDN][0, 0, lo + 1] = 5.0 * rho[0]
    f = diffusion.cartesian_fluxes(w, g, nu, 0.0, GAMMA, WEIGHT, WALLS)
    assert f[1][fv.IVY][0, 0, lo] == pytest.approx(-nu * rho[0] * stress_lo, rel=1e-14)


def test_posts_skip_wall_faces():
    g = fv.Grid(8, 4, 1, 0.0, 800.0, 0.0, 400.0, 0.0, 100.0)
    w = base_state(g)
    x1v = g.centers(1)[None, None, :]
    w[fv.IPR] = 1.2 * RD * (300.0 - 0.01 * x1v) + 0.0 * w[fv.IDN]
    fill(w, g, WALLS)
    f = diffusion.cartesian_fluxes(w, g, 0.0, 75.0, GAMMA, WEIGHT, WALLS)
    totals = {1: np.zeros_like(w), 2: np.zeros_like(w)}
    diffusion.post_fluxes(totals, f, g, WALLS)
    ng = g.nghost
    assert np.all(totals[1][fv.IPR][0, :, ng] == 0.0)
    assert np.all(totals[1][fv.IPR][0, :, ng + 8] == 0.0)
    inner = totals[1][fv.IPR][0, ng : ng + 4, ng + 1 : ng + 8]
    np.testing.assert_array_equal(inner, f[1][fv.IPR][0, ng : ng + 4, ng + 1 : ng + 8])
    assert np.all(inner != 0.0)


def test_time_step_bound():
    g = fv.Grid(64, 512, 1, 0.0, 6400.0, -25600.0, 25600.0, 0.0, 100.0)
    assert diffusion.max_time_step(g, 75.0, 75.0) == pytest.approx(100.0**2 / (2 * 2 * 75.0), rel=1e-15)
    assert diffusion.max_time_step(g, 0.0, 0.0) == float("inf")
