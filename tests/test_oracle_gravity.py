"""Pins of the gravity oracle (CPU): the discrete hydrostatic identity of the
scan, the isothermal closed form of the cell reference, the rest state of a
hydrostatic column (well-balanced, tests/test_fp32_policy.py:65 asserts
|m1| < 2e-8 for the reference), the slope stencil on quadratics, and the
telescoping of the face gravity work."""

import numpy as np

from oracle import fv, gravity


def isothermal(g, T0=300.0, Rd=287.0, g1=-9.8):
    x1v = g.centers(1)
    p = 1e5 * np.exp(g1 * x1v / (Rd * T0))
    w = np.zeros((5,) + g.shape)
    w[4] = p
    w[0] = w[4] / (Rd * T0)
    return w


def test_scan_is_discretely_hydrostatic():
    g = fv.Grid(40, 3, 1, 0.0, 8000.0, 0.0, 300.0)
    w = isothermal(g)
    dx = np.full(g.nc(1), g.dx(1))
    lo, hi, pref, dsf, dref = gravity.hydro_ref_x1(w, dx, g.nc(1) - 4, 9.8)
    np.testing.assert_allclose(lo - hi, 9.8 * w[0] * g.dx(1), rtol=1e-12)
    # faces are continuous: the upper face of cell i is the lower of i + 1
    np.testing.assert_array_equal(hi[..., :-1], lo[..., 1:])
    # an isothermal column: the cell reference is the cell's log-mean
    # pressure to the six-face order, and dref = pref * rho/p
    ratio = lo / hi
    logmean = 9.8 * w[0] * g.dx(1) / np.log(ratio)
    sl = (slice(None), slice(None), slice(5, -5))
    np.testing.assert_allclose(pref[sl], logmean[sl], rtol=1e-9)
    np.testing.assert_allclose(dref[sl], pref[sl] / (287.0 * 300.0), rtol=1e-9)


def test_hydrostatic_rest_stays_at_rest():
    g = fv.Grid(32, 16, 1, 0.0, 6400.0, -800.0, 800.0, 0.0, 100.0)
    w = isothermal(g)
    d = fv.Deck(grid=g, gammad=1004 / 717, riemann="lmars", shock_x1=False, shock_x23=False,
                bcs=("reflecting",) * 6, grav1=-9.8)
    b = fv.Block(d)
    b.initialize(w)
    for _ in range(5):
        dt = b.max_time_step()
        for s in range(3):
            b.forward(dt, s)
    wn = fv.cons2prim(b.u, d.gammad)
    sub = g.interior()
    assert np.abs(wn[1][sub]).max() < 2e-8
    assert np.abs(wn[2][sub]).max() < 2e-8


def test_centroid_slope_exact_on_quadratics():
    x = np.cumsum(np.array([0.0, 1.0, 1.3, 0.7, 2.0, 1.1, 0.9]))
    q = 3.0 - 2.0 * x + 0.5 * x * x
    s = gravity.centroid_slope(q[None], x)[0]
    np.testing.assert_allclose(s, -2.0 + x, rtol=1e-12, atol=1e-12)


def test_face_gravity_work_of_a_uniform_flux():
    # hi_i F - lo_i F = g (x1f_{i+1} - x1f_i) F: a uniform upward mass flux
    # through a resting column does g F H of work on it (H the column
    # height); no density change, so no corrected-PE part, and no cell work
    g = fv.Grid(12, 1, 1, 0.0, 1200.0)
    w = isothermal(g)
    w[1] = 0.0
    flux = np.zeros((5,) + g.shape)
    flux[0] = 0.7
    gec = gravity.gravity_energy_correction(g, w, flux, 1.0, -9.8)
    np.testing.assert_allclose(gec.sum() * g.dx(1), -9.8 * 0.7 * 1200.0, rtol=1e-12)


def test_corrected_pe_work_is_the_exact_work_of_a_cubic_flux():
    """With a mass flux F(x) cubic along x1, the density change inside a
    cell is the quadratic q = -dt F'(x).  The face work plus the corrected
    potential-energy term (g dx^2/12 times the slope of the cell changes)
    is then the exact gravity work g dt <F> over the cell:
    g int F dx = g [(x - xc) F] - g int (x - xc) F' dx, and for a quadratic
    q the last term is q'(xc) dx^2/12, which the three-point slope of the
    cell averages gives exactly.  <F> is taken by 2-point Gauss-Legendre
    quadrature (exact for cubics).  The cell-centred work dt rho v g is
    subtracted on top (zero here: the cells are at rest)."""
    g = fv.Grid(10, 2, 1, 0.0, 1000.0, 0.0, 200.0)
    w = isothermal(g)
    w[1] = 0.0
    x1f = g.faces(1)
    F = lambda x: 0.3 + 2e-3 * x - 3e-6 * x * x + 2e-9 * x ** 3  # noqa: E731
    flux = np.zeros((5,) + g.shape)
    flux[0] = F(x1f[: g.nc(1)])
    dt, grav = 0.8, -9.8
    gec = gravity.gravity_energy_correction(g, w, flux, dt, grav)
    ng, n = g.nghost, g.nx1
    xc = g.centers(1)[ng : ng + n]
    xi = g.dx(1) / (2.0 * np.sqrt(3.0))
    mean_f = 0.5 * (F(xc - xi) + F(xc + xi))
    np.testing.assert_allclose(gec[0, 0], grav * dt * mean_f, rtol=1e-12)
    # the cell-centred work rho v g dt comes off
    w[1] = 0.25
    gec2 = gravity.gravity_energy_correction(g, w, flux, dt, grav)
    np.testing.assert_allclose(gec2 - gec, -dt * w[0][..., ng : ng + n] * 0.25 * grav, rtol=1e-12)
