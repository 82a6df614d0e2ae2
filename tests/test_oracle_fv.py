"""Pins of the explicit-stage oracle (CPU): each check is an independent
statement of what the reference computes (its own closed forms and test
formulas) or of what the mathematics fixes, so a dropped term, a wrong sign
or a transposed stencil in oracle/fv.py fails one of them."""

import numpy as np
import pytest

from oracle import fv


# ---- reconstruction -------------------------------------------------------
def interp_weno5(m2, m1, c, p1, p2):
    """The reference's independent closed form (tests/interp_simple.hpp:
    131-150): the left-edge value of the centre cell."""
    sqr = lambda x: x * x  # noqa: E731
    p0 = (1.0 / 3.0) * c + (5.0 / 6.0) * m1 - (1.0 / 6.0) * m2
    q1 = (-1.0 / 6.0) * p1 + (5.0 / 6.0) * c + (1.0 / 3.0) * m1
    q2 = (1.0 / 3.0) * p2 - (7.0 / 6.0) * p1 + (11.0 / 6.0) * c
    b0 = 13.0 / 12.0 * sqr(c - 2.0 * m1 + m2) + 0.25 * sqr(3.0 * c - 4.0 * m1 + m2)
    b1 = 13.0 / 12.0 * sqr(p1 - 2.0 * c + m1) + 0.25 * sqr(p1 - m1)
    b2 = 13.0 / 12.0 * sqr(p2 - 2.0 * p1 + c) + 0.25 * sqr(p2 - 4.0 * p1 + 3.0 * c)
    a0 = 0.3 / sqr(b0 + 1e-6)
    a1 = 0.6 / sqr(b1 + 1e-6)
    a2 = 0.1 / sqr(b2 + 1e-6)
    return (a0 * p0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2)


def interp_cp5(m2, m1, c, p1, p2):
    """tests/interp_simple.hpp:110-115."""
    return -1.0 / 20.0 * m2 + 9.0 / 20.0 * m1 + 47.0 / 60.0 * c - 13.0 / 60.0 * p1 + 1.0 / 30.0 * p2


def test_weno5_and_cp5_match_reference_closed_forms():
    rng = np.random.default_rng(0)
    v = rng.random((5, 1000)) * 2 - 1
    phi = list(v)
    got_m = fv.weno5_point(phi, fv.WENO5_CM)
    got_p = fv.weno5_point(phi, fv.WENO5_CP)
    np.testing.assert_allclose(got_m, interp_weno5(*v), rtol=1e-12, atol=1e-14)
    # the right edge is the left-edge form on the mirrored stencil
    np.testing.assert_allclose(got_p, interp_weno5(*v[::-1]), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(fv.poly_point(phi, fv.CP5_CM), interp_cp5(*v), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(fv.poly_point(phi, fv.CP5_CP), interp_cp5(*v[::-1]), rtol=1e-13, atol=1e-15)


def test_cp5_is_exact_on_quartic_cell_averages():
    # averages of x^4 over unit cells centred at -2..2; edge value at -1/2
    def avg(c):
        a, b = c - 0.5, c + 0.5
        return (b**5 - a**5) / 5.0

    v = [np.array([avg(c)]) for c in (-2, -1, 0, 1, 2)]
    assert abs(fv.poly_point(v, fv.CP5_CM)[0] - 0.5**4) < 1e-13
    assert abs(fv.poly_point(v, fv.CP5_CP)[0] - 0.5**4) < 1e-13


def test_weno5_picks_the_smooth_stencil_at_a_jump():
    # a step between cells 3 and 4 of the window: the left edge of the centre
    # cell (cell 2) must come from the smooth left stencil, not the jump
    v = [np.array([x]) for x in (1.0, 1.0, 1.0, 5.0, 5.0)]
    assert abs(fv.weno5_point(v, fv.WENO5_CM)[0] - 1.0) < 1e-6


def test_reconstruct_faces_and_dummy_copy():
    w = np.random.default_rng(1).random((5, 1, 1, 20)) + 1.0
    wlr = fv.reconstruct(w, 3, True)
    # face 5: left state is the right edge of cell 4, right state the left
    # edge of cell 5
    L = fv.weno5_point([w[:, :, :, k] for k in range(2, 7)], fv.WENO5_CP)
    R = fv.weno5_point([w[:, :, :, k] for k in range(3, 8)], fv.WENO5_CM)
    np.testing.assert_array_equal(wlr[fv.ILT][..., 5], L)
    np.testing.assert_array_equal(wlr[fv.IRT][..., 5], R)
    # ghost faces copy the end values (reconstruct.cpp:225-228)
    np.testing.assert_array_equal(wlr[fv.IRT][..., 0], wlr[fv.IRT][..., 3])
    np.testing.assert_array_equal(wlr[fv.ILT][..., 19], wlr[fv.ILT][..., 17])


# ---- Riemann ---------------------------------------------------------------
def states(n=200, seed=77):
    rng = np.random.default_rng(seed)
    wl = rng.random((5, n))
    wr = rng.random((5, n))
    wl[0] += 1.0
    wr[0] += 1.0
    wl[4] += 1.0
    wr[4] += 1.0
    wl[1:4] = (wl[1:4] - 0.5) * 0.1
    wr[1:4] = (wr[1:4] - 0.5) * 0.1
    return wl, wr


def test_hllc_face_pressure_closed_form():
    """tests/test_riemann.cpp:169-229 restated."""
    g = 1.4
    wl, wr = states()
    _, cp = fv.hllc(wl, wr, g, 3)
    cl = np.sqrt(g * wl[4] / wl[0])
    cr = np.sqrt(g * wr[4] / wr[0])
    rhoa = 0.5 * (wl[0] + wr[0])
    ca = 0.5 * (cl + cr)
    pmid = 0.5 * (wl[4] + wr[4] + (wl[1] - wr[1]) * rhoa * ca)
    ql = np.where(pmid <= wl[4], 1.0, np.sqrt(1.0 + (g + 1) / (2 * g) * (pmid / wl[4] - 1.0)))
    qr = np.where(pmid <= wr[4], 1.0, np.sqrt(1.0 + (g + 1) / (2 * g) * (pmid / wr[4] - 1.0)))
    al = wl[1] - cl * ql
    ar = wr[1] + cr * qr
    vxl = wl[1] - al
    vxr = wr[1] - ar
    tl = wl[4] + vxl * wl[0] * wl[1]
    tr = wr[4] + vxr * wr[0] * wr[1]
    ml = wl[0] * vxl
    mr = -(wr[0] * vxr)
    expected = np.maximum((ml * tr + mr * tl) / (ml + mr), 0.0)
    np.testing.assert_allclose(cp, expected, rtol=1e-10, atol=1e-10)


def physical_flux(w, g, dim):
    ivx, ivy, ivz = fv.face_axes(dim)
    rho, un, p = w[0], w[ivx], w[4]
    E = p / (g - 1) + 0.5 * rho * (w[1] ** 2 + w[2] ** 2 + w[3] ** 2)
    f = np.empty_like(w)
    f[0] = rho * un
    f[ivx] = rho * un * un + p
    f[ivy] = rho * un * w[ivy]
    f[ivz] = rho * un * w[ivz]
    f[4] = (E + p) * un
    return f


@pytest.mark.parametrize("solver", [fv.hllc, fv.lmars])
@pytest.mark.parametrize("dim", [3, 2, 1])
def test_solvers_are_consistent(solver, dim):
    wl, _ = states(seed=3)
    f, _ = solver(wl, wl.copy(), 1.4, dim)
    np.testing.assert_allclose(f, physical_flux(wl, 1.4, dim), rtol=1e-12, atol=1e-12)


def test_lmars_upwinds_mass():
    wl, wr = states(seed=5)
    f, pbar = fv.lmars(wl, wr, 1.4, 3)
    rhobar = 0.5 * (wl[0] + wr[0])
    cbar = np.sqrt(0.5 * 1.4 * (wl[4] + wr[4]) / rhobar)
    ubar = 0.5 * (wl[1] + wr[1]) + 0.5 / (rhobar * cbar) * (wl[4] - wr[4])
    np.testing.assert_allclose(f[0], np.where(ubar > 0, ubar * wl[0], ubar * wr[0]), rtol=1e-13)
    np.testing.assert_allclose(pbar, 0.5 * (wl[4] + wr[4]) + 0.5 * rhobar * cbar * (wl[1] - wr[1]), rtol=1e-13)


# ---- divergence, RK, boundaries, CFL ----------------------------------------
def test_divergence_telescopes_and_is_exact_for_linear_fluxes():
    g = fv.Grid(8, 6, 4, 0.0, 2.0, 0.0, 3.0, 0.0, 1.0)
    rng = np.random.default_rng(2)
    flux = {a: rng.random((5,) + g.shape) for a in (1, 2, 3)}
    div = fv.divergence(g, flux)
    sub = (slice(None),) + g.interior()
    vol = g.dx(3) * (g.dx(2) * g.dx(1))
    k, j, i = g.interior()
    total = (div[sub] * vol).sum(axis=(1, 2, 3))
    boundary = (
        (flux[1][:, k, j, i.stop] - flux[1][:, k, j, i.start]).sum(axis=(1, 2)) * g.dx(3) * g.dx(2)
        + (flux[2][:, k, j.stop, i] - flux[2][:, k, j.start, i]).sum(axis=(1, 2)) * g.dx(3) * g.dx(1)
        + (flux[3][:, k.stop, j, i] - flux[3][:, k.start, j, i]).sum(axis=(1, 2)) * g.dx(2) * g.dx(1)
    )
    np.testing.assert_allclose(total, boundary, rtol=1e-12)
    # F1 = x1 at the faces: div = 1
    f1 = np.zeros((5,) + g.shape)
    f1[:] = g.faces(1)[:-1][None, None, None, :]
    d = fv.divergence(g, {1: f1})
    np.testing.assert_allclose(d[sub], 1.0, rtol=1e-12)


def test_ssp_rk3_stability_polynomial():
    # du = lambda u dt per stage: one step multiplies u by 1 + z + z^2/2 + z^3/6
    z = -0.37
    u0 = np.array([1.0])
    u = u0.copy()
    for s in range(3):
        u = fv.rk_combine(u0, u, z * u, s)
    assert abs(u[0] - (1 + z + z * z / 2 + z**3 / 6)) < 1e-15


def test_boundary_index_maps():
    g = fv.Grid(6, 5, 4)
    q = np.random.default_rng(4).random((5,) + g.shape)
    r = q.copy()
    fv.bc_gather(r, 3, "reflecting", False)
    for gi in range(3):
        np.testing.assert_array_equal(r[[0, 2, 3, 4], ..., gi], q[[0, 2, 3, 4], ..., 5 - gi])
        np.testing.assert_array_equal(r[1, ..., gi], -q[1, ..., 5 - gi])
    r = q.copy()
    fv.bc_gather(r, 2, "periodic", True)
    nc = g.nc(2)
    for gi in range(3):
        np.testing.assert_array_equal(r[:, :, nc - 3 + gi], q[:, :, 3 + gi])
    r = q.copy()
    fv.bc_gather(r, 1, "extrapolation", False)
    for gi in range(3):
        np.testing.assert_array_equal(r[:, gi], q[:, 3])


def test_outflow_keeps_a_resting_background():
    g = fv.Grid(10, 1, 1)
    w = np.zeros((5,) + g.shape)
    w[0] = 1.2
    w[4] = 0.7
    b = w.copy()
    fv.radiating(w, 3, b, 1.4, 1e-10, 1e-10, True)
    fv.radiating(w, 3, b, 1.4, 1e-10, 1e-10, False)
    np.testing.assert_array_equal(w, b)


def test_cfl_brute_force():
    g = fv.Grid(7, 5, 1, 0.0, 1.4, 0.0, 2.0)
    rng = np.random.default_rng(6)
    w = rng.random((5,) + g.shape) + 0.5
    dt = fv.max_time_step(g, w, 1.4)
    best = 1e9
    k, j, i = g.interior()
    for jj in range(j.start, j.stop):
        for ii in range(i.start, i.stop):
            c = np.sqrt(1.4 * w[4, 0, jj, ii] / w[0, 0, jj, ii])
            best = min(best, g.dx(1) / (abs(w[1, 0, jj, ii]) + c), g.dx(2) / (abs(w[2, 0, jj, ii]) + c))
    assert dt == best


def test_sod_converges_to_the_exact_solution():
    errs = []
    for nx in (64, 128):
        d = fv.sod_deck(nx)
        b = fv.Block(d)
        b.initialize(fv.sod_initial(d.grid))
        fv.run(b, 0.1)
        w = fv.cons2prim(b.u, d.gammad)[:, 0, 0, 3:-3]
        ex = fv.sod_exact(d.grid.centers(1)[3:-3], 0.1)
        errs.append(np.abs(w[0] - ex[0]).mean())
    assert errs[1] < 0.7 * errs[0]


@pytest.mark.parametrize("dim", [2, 3])
def test_flux_covariance_is_the_face_average_of_linear_profiles(dim):
    """The covariance term makes the energy flux of an x2/x3 face the average
    over the face's x1 extent: with uniform density and h, u_n linear in x1,
    rho <u_n h> - rho u_n(x) h(x) = (dx1^2 / 12) rho h' u_n' exactly.  The
    average is taken by 2-point Gauss-Legendre quadrature, exact for the
    quadratic product.  The face states enter as their mean (wl != wr), and
    the other tangential velocity does not enter."""
    gamma = 1.4
    g = fv.Grid(10, 6, 4, 0.0, 500.0, 0.0, 600.0, 0.0, 400.0)
    x1 = g.centers(1)
    dx1 = g.dx(1)
    rho0, p0, a, u0, b = 1.2, 9.0e4, -11.0, 3.0, 0.02
    w = np.zeros((5,) + g.shape)
    w[fv.IDN] = rho0
    w[fv.IPR] = p0 + a * x1
    un_row = fv.IVY if dim == 2 else fv.IVZ
    other = fv.IVZ if dim == 2 else fv.IVY
    w[un_row] = u0 + b * x1
    w[other] = np.sin(x1 / 37.0)  # must not enter
    rng = np.random.default_rng(3)
    d = 1e-3 * rng.standard_normal(w.shape)
    d[fv.IDN] = 0.0
    cov = fv.flux_covariance(w + d, w - d, dim, g, gamma)
    k = gamma / (gamma - 1.0)
    h = lambda x: k * (p0 + a * x) / rho0  # noqa: E731
    un = lambda x: u0 + b * x  # noqa: E731
    xi = dx1 / (2.0 * np.sqrt(3.0))
    xc = x1[1:-1]
    avg = 0.5 * (un(xc - xi) * h(xc - xi) + un(xc + xi) * h(xc + xi))
    expected = rho0 * (avg - un(xc) * h(xc))
    got = cov[0, 0]
    assert np.allclose(got, np.broadcast_to(expected, got.shape), rtol=1e-9, atol=0.0)
    # and the closed form
    assert np.allclose(got, dx1 * dx1 / 12.0 * rho0 * (k * a / rho0) * b, rtol=1e-9, atol=0.0)


# ---------------------------------------------------------------------------
# the outflow (radiating) ghost on moving states: linear acoustics fixes it.
# About the background (rho0, p0, c0) a perturbation d splits into the
# outgoing/incoming acoustic waves W+- = s*dv_n +- dp/(rho0 c0), the entropy
# wave S = drho - dp/c0^2 and the tangential (vortical) waves; the ghost
# keeps the waves that leave the domain and drops the ones that enter.
# ---------------------------------------------------------------------------
def _radiate(wi, outer, gamma=1.4, rho0=1.1, p0=0.9, floor=1e-10):
    g = fv.Grid(10, 1, 1)
    b = np.zeros((5,) + g.shape)
    b[fv.IDN] = rho0
    b[fv.IPR] = p0
    w = b.copy()
    w[:, 0, 0, 3:-3] = np.asarray(wi)[:, None]
    fv.radiating(w, 3, b, gamma, floor, floor, outer)
    ghost = w[:, 0, 0, -3:] if outer else w[:, 0, 0, :3]
    return ghost, b[:, 0, 0, 0], np.sqrt(gamma * p0 / rho0)


def _waves(d, sign, rho0, c0):
    wp = sign * d[fv.IVX] + d[fv.IPR] / (rho0 * c0)
    wm = sign * d[fv.IVX] - d[fv.IPR] / (rho0 * c0)
    s = d[fv.IDN] - d[fv.IPR] / (c0 * c0)
    return wp, wm, s


@pytest.mark.parametrize("outer", [True, False])
def test_outflow_supersonic_outflow_copies_the_interior(outer):
    sign = 1.0 if outer else -1.0
    c0 = np.sqrt(1.4 * 0.9 / 1.1)
    wi = [1.13, sign * 3.0 * c0, 0.21, -0.17, 0.95]
    ghost, _, _ = _radiate(wi, outer)
    np.testing.assert_allclose(ghost, np.broadcast_to(np.array(wi)[:, None], ghost.shape), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("outer", [True, False])
def test_outflow_supersonic_inflow_keeps_the_background(outer):
    sign = 1.0 if outer else -1.0
    c0 = np.sqrt(1.4 * 0.9 / 1.1)
    wi = [1.13, -sign * 3.0 * c0, 0.21, -0.17, 0.95]
    ghost, bg, _ = _radiate(wi, outer)
    np.testing.assert_array_equal(ghost, np.broadcast_to(bg[:, None], ghost.shape))


@pytest.mark.parametrize("outer", [True, False])
@pytest.mark.parametrize("mach", [0.4, -0.4])
def test_outflow_subsonic_keeps_only_the_leaving_waves(outer, mach):
    sign = 1.0 if outer else -1.0
    rho0, p0 = 1.1, 0.9
    c0 = np.sqrt(1.4 * p0 / rho0)
    wi = np.array([1.13, sign * mach * c0, 0.21, -0.17, 0.95])
    ghost, bg, _ = _radiate(wi, outer, rho0=rho0, p0=p0)
    d_int = wi - bg
    d_gh = ghost - bg[:, None]
    wp_i, wm_i, s_i = _waves(d_int, sign, rho0, c0)
    wp_g, wm_g, s_g = _waves(d_gh, sign, rho0, c0)
    # the wave moving at u_n + c leaves at either sign of a subsonic u_n
    np.testing.assert_allclose(wp_g, wp_i, rtol=1e-12, atol=1e-14)
    # the wave moving at u_n - c enters
    np.testing.assert_allclose(wm_g, 0.0, atol=1e-14)
    if mach > 0:
        # entropy and vorticity are carried out with the flow
        np.testing.assert_allclose(s_g, s_i, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(d_gh[fv.IVY], d_int[fv.IVY], rtol=1e-13)
        np.testing.assert_allclose(d_gh[fv.IVZ], d_int[fv.IVZ], rtol=1e-13)
    else:
        # an inflow brings the background's entropy and tangential motion
        np.testing.assert_allclose(s_g, 0.0, atol=1e-14)
        np.testing.assert_array_equal(d_gh[fv.IVY], 0.0)
        np.testing.assert_array_equal(d_gh[fv.IVZ], 0.0)


def test_outflow_inner_face_mirrors_the_outer_face():
    """Reversing x1 and the normal velocity swaps the two faces."""
    g = fv.Grid(10, 1, 1)
    rng = np.random.default_rng(12)
    b = np.zeros((5,) + g.shape)
    b[fv.IDN] = 1.0 + 0.2 * rng.random(g.shape)
    b[fv.IPR] = 0.8 + 0.2 * rng.random(g.shape)
    w = b + 0.05 * (rng.random((5,) + g.shape) - 0.5)
    flip = lambda q: np.concatenate([q[:1], -q[1:2], q[2:]])[..., ::-1].copy()  # noqa: E731
    a = w.copy()
    fv.radiating(a, 3, b, 1.4, 1e-10, 1e-10, True)
    m = flip(w)
    fv.radiating(m, 3, flip(b), 1.4, 1e-10, 1e-10, False)
    np.testing.assert_allclose(flip(m)[..., -3:], a[..., -3:], rtol=1e-14, atol=1e-15)


def test_outflow_ghost_respects_the_floors():
    """An outgoing expansion deep enough to drive the ghost's pressure
    below the floor is scaled back to 0.99 of the distance to it."""
    rho0, p0, floor = 1.0, 1.0, 1e-3
    c0 = np.sqrt(1.4 * p0 / rho0)
    wi = [1.0, 3.0 * c0, 0.0, 0.0, 0.5]
    g = fv.Grid(10, 1, 1)
    b = np.zeros((5,) + g.shape)
    b[fv.IDN] = rho0
    b[fv.IPR] = p0
    w = b.copy()
    w[:, 0, 0, 3:-3] = np.array(wi)[:, None]
    w[fv.IPR][..., 3:-3] = 0.5  # the interior itself stays a valid gas
    # the perturbation d = w - b of the source cell is kept whole (all
    # waves leave); a thinner background in the ghosts makes bg + d fall
    # below the floor there
    b2 = b.copy()
    b2[fv.IPR][..., -3:] = 0.4  # a thinner background in the ghosts
    fv.radiating(w, 3, b2, 1.4, floor, floor, True)
    ghost = w[:, 0, 0, -3:]
    dp = 0.5 - p0
    alpha = 0.99 * (0.4 - floor) / -dp
    np.testing.assert_allclose(ghost[fv.IPR], 0.4 + alpha * dp, rtol=1e-13)
    assert (ghost[fv.IPR] > floor).all()
    # one factor for every characteristic: the normal velocity too
    np.testing.assert_allclose(ghost[fv.IVX], alpha * 3.0 * c0, rtol=1e-13)


def test_layer_heat_adds_flux_times_area_of_energy_to_the_wall_cells():
    """bot-heat (top-cool) with an integer depth puts flux * A * dt of energy
    into the bottom (top) `depth` cells of every column, evenly, and nothing
    elsewhere (layer_heat.cpp:78-85)."""
    g = fv.Grid(12, 3, 2, 0.0, 600.0, 0.0, 300.0, 0.0, 200.0)
    area = g.dx(2) * g.dx(3)
    vol = area * g.dx(1)
    for depth in (1, 2, 5):
        for top in (False, True):
            du = np.zeros((5,) + g.shape)
            fv.layer_heat(du, g, 0.7, 100.0, depth, top)
            e = du[fv.IPR][g.interior()] * vol
            np.testing.assert_allclose(e.sum(-1), 0.7 * 100.0 * area, rtol=1e-14)
            cells = np.arange(g.nx1)
            hot = cells >= g.nx1 - depth if top else cells < depth
            assert (e[..., hot] > 0).all() and not e[..., ~hot].any()
            np.testing.assert_allclose(e[..., hot], 0.7 * 100.0 * area / depth, rtol=1e-14)
            # nothing in other rows, nor in the x1 ghosts (the x2/x3 ghost
            # columns take the same heating, as the reference's narrow along
            # x1 does; their du is never read)
            assert not du[[fv.IDN, fv.IVX, fv.IVY, fv.IVZ]].any()
            assert not du[fv.IPR][..., : g.nghost].any() and not du[fv.IPR][..., g.nghost + g.nx1 :].any()


def test_step_field_fold_is_the_weighted_stage_sum():
    """Three stages of S <- w1 S + w2 dt T with the SSP-RK3 weights give
    sum_s (w2_s prod_{t > s} w1_t) dt T_s (tests/test_flux_register.cpp:
    211-248), whatever S held before the step (w1 = 0 at stage 0)."""
    rng = np.random.default_rng(8)
    T = [rng.standard_normal((5, 3, 4)) for _ in range(3)]
    S = rng.standard_normal((5, 3, 4))
    dt = 0.73
    for s in range(3):
        _, w1, w2 = fv.RK3_WEIGHTS[s]
        S = fv.step_field_fold(S, T[s], w1, w2, dt)
    c = [1.0 * 0.25 * (2.0 / 3.0), 0.25 * (2.0 / 3.0), 2.0 / 3.0]
    ref = sum(c[s] * dt * T[s] for s in range(3))
    np.testing.assert_allclose(S, ref, rtol=1e-14, atol=1e-15)
