"""Pins of the oracle's passive scalars (oracle/scalar.py) on the CPU:
the scaled WENO5 against the unscaled one and its homogeneity; a uniform
mixing ratio stays uniform and r = 1 follows the gas density through a whole
stage (face ratio, covariance, positivity limiter, implicit transfer and
donor cap); conservation with walls; the donor cap against what its rule
guarantees; theta, the share and the covariance row against closed forms."""

import numpy as np
import pytest

from oracle import fv, scalar

CP, RD, GAMMA = 1004.0, 287.0, 1004.0 / (1004.0 - 287.0)


def _stencils(rng, n=200, scale=1.0):
    return [scale * (0.5 + rng.random(n)) for _ in range(5)]


def test_scaled_weno_is_homogeneous_and_reduces_to_the_unscaled_one():
    rng = np.random.default_rng(1)
    x = _stencils(rng, scale=1e-4)
    got = scalar.weno5_point_scaled(x, fv.WENO5_CM)
    # at vscale the scaled form is the unscaled WENO5 of x / vscale, times vscale
    vs = sum(np.abs(p) for p in x) / 5.0
    ref = vs * fv.weno5_point([p / vs for p in x], fv.WENO5_CM)
    assert np.allclose(got, ref, rtol=1e-14, atol=0)
    # exact homogeneity under a power of two; the unscaled form (eps = 1e-6
    # against data of 1e-4) is not homogeneous
    for lam in (2.0**-20, 2.0**7):
        assert np.array_equal(scalar.weno5_point_scaled([lam * p for p in x], fv.WENO5_CM), lam * got)
    unscaled = fv.weno5_point(x, fv.WENO5_CM)
    unscaled_big = fv.weno5_point([2.0**20 * p for p in x], fv.WENO5_CM)
    assert not np.allclose(unscaled_big, 2.0**20 * unscaled, rtol=1e-6)
    # a constant stencil gives the constant; a zero stencil gives 0
    c = [np.full(3, 0.37) for _ in range(5)]
    assert np.allclose(scalar.weno5_point_scaled(c, fv.WENO5_CP), 0.37, rtol=1e-15)
    z = [np.zeros(3) for _ in range(5)]
    assert np.array_equal(scalar.weno5_point_scaled(z, fv.WENO5_CP), np.zeros(3))


def test_recon_faces_are_clamped_and_ghost_faces_copy_the_ends():
    q = np.zeros((1, 1, 1, 20))
    q[..., :10] = 1.0  # a step down to 0: WENO undershoots below 0 there
    out = scalar.recon(q, 3)
    assert out.min() >= 0.0
    assert np.array_equal(out[fv.IRT][..., :3], np.repeat(out[fv.IRT][..., 3:4], 3, axis=-1))
    assert np.array_equal(out[fv.ILT][..., 18:], np.repeat(out[fv.ILT][..., 17:18], 2, axis=-1))
    # the left state of face i comes from cell i - 1: far left, exactly 1
    assert out[fv.ILT][..., 5] == 1.0 and out[fv.IRT][..., 5] == 1.0


def test_face_ratio_of_a_uniform_mixing_ratio_is_that_ratio():
    rng = np.random.default_rng(3)
    rho = 0.5 + rng.random((1, 1, 12, 16))
    for r0 in (0.3, 1.0, 7.5e-9):
        rlr = scalar.face_r(rho, r0 * rho, 3)
        assert np.allclose(rlr, r0, rtol=4e-15, atol=0)
        rlr = scalar.face_r(rho, r0 * rho, 2)
        assert np.allclose(rlr, r0, rtol=4e-15, atol=0)


def test_face_ratio_takes_the_donor_where_the_density_face_is_zero():
    rho = np.ones((1, 1, 1, 20))
    rho[..., 8:] = 0.0
    s = 0.25 * rho
    s[..., 10:] = 0.0
    s[..., 8:10] = 0.5  # tracer where the gas is gone: r = inf there
    with np.errstate(divide="ignore", invalid="ignore"):
        rlr = scalar.face_r(rho, s, 3)
    # face 12: both stencils (cells 9..13, 10..14) empty of gas, the donors'
    # r (0 / 0); face 10's left state comes from cell 9's stencil 7..11,
    # where rho has mass, so it is a ratio of faces, not the donor's inf
    assert np.isnan(rlr[fv.ILT][..., 12]) and np.isnan(rlr[fv.IRT][..., 12])
    assert np.isfinite(rlr[fv.ILT][..., 10])
    # face 5: in the full region, 0.25 exactly
    assert rlr[fv.ILT][..., 5] == 0.25 and rlr[fv.IRT][..., 5] == 0.25


def test_upwind_takes_the_upstream_state():
    m = np.array([2.0, -3.0, 0.0])
    rl = np.array([0.1, 0.2, 0.3])
    rr = np.array([0.7, 0.8, 0.9])
    assert np.array_equal(scalar.upwind(rl, rr, m), np.array([0.2, -2.4000000000000004, 0.0]))


def test_covariance_row_on_linear_fields():
    g = fv.Grid(10, 6, 1, 0.0, 1000.0, 0.0, 600.0, 0.0, 1.0)
    x1 = g.centers(1)[None, None, None, :]
    wl = np.zeros((5,) + g.shape)
    wl[fv.IDN] = 1.3
    wl[fv.IVY] = 0.02 * x1  # u_n linear in x1
    q = 1e-3 * x1 + np.zeros(g.shape)[None]
    got = scalar.covariance(q[0], q[0], wl, wl, 2, g)
    # sigma1^2 = dx1^2 / 12, the centred differences exact on linear fields
    want = g.dx(1) ** 2 / 12.0 * 1.3 * 0.02 * 1e-3
    assert np.allclose(got[..., 1:-1], want, rtol=1e-12)
    assert np.array_equal(got[..., [0, -1]], np.zeros_like(got[..., [0, -1]]))


def test_theta_closed_form_and_share():
    g = fv.Grid(4, 1, 1, 0.0, 4.0, 0.0, 1.0, 0.0, 1.0)
    u = np.full((1,) + g.shape, 2.0)
    f = np.zeros((1,) + g.shape)
    i0 = g.nghost
    f[..., i0 + 2] = 5.0  # out of cell i0 + 1 through its upper face
    f[..., i0 + 1] = -1.0  # and through its lower face
    th = scalar.positivity_theta(u, {1: f}, g, 0.5)
    vol = g.dx(3) * (g.dx(2) * g.dx(1))
    avail = 2.0 * vol * (1.0 - 4096.0 * np.finfo(float).eps)
    assert th[..., i0 + 1] == pytest.approx(avail / (0.5 * (5.0 + 1.0)), rel=1e-15)
    # cell i0 + 2 only receives (F > 0 at its lower face) and cell i0 only
    # receives (F < 0 at its upper face): nothing drains them
    assert th[..., i0 + 2] == 1.0 and th[..., i0] == 1.0
    assert np.array_equal(th[..., :i0], np.ones((1, 1, 1, i0)))
    q = np.array([1.0, 0.5, 2.0, 4.0, 3.0])[None, None, None, :]
    phi = scalar.positivity_share(q)
    # cell 2: neighbours 0.5 and 4 -> 0.5 / 2
    assert phi[..., 2] == 0.25 and phi[..., 1] == 1.0


def test_donor_cap_keeps_every_cell_non_negative_and_is_idle_when_feasible():
    rng = np.random.default_rng(11)
    n = 12
    held = rng.random((200, n)) * 2.0
    P = (rng.random((200, n)) - 0.5) * 6.0
    P[:, 0] = 0.0  # no transfer through the column's bottom face
    out = scalar.donor_cap(P, held, 0, n - 1)
    above = np.concatenate([out[:, 1:], np.zeros((200, 1))], axis=1)
    final = held + out - above
    assert final.min() >= -1e-15
    # conservation: the face transfers telescope
    assert np.allclose(final.sum(1), held.sum(1), rtol=1e-14)
    # a feasible transfer passes unchanged
    small = 1e-3 * (rng.random((200, n)) - 0.5)
    small[:, 0] = 0.0
    assert np.array_equal(scalar.donor_cap(small, held + 1.0, 0, n - 1), small)


def _deck(g, scheme, cov, bcs=("reflecting",) * 6):
    return fv.Deck(grid=g, gammad=GAMMA, riemann="lmars", shock_x1=False, shock_x23=False, bcs=bcs, grav1=-9.81,
                   cfl=0.9, implicit_scheme=scheme, flux_covariance=cov, nu_iso=25.0, kappa_iso=25.0, nscalar=1)


def _stratified(g, seed):
    from tests.test_gpu_gravity import stratified

    return stratified(g, seed=seed, noise=0.5)


def _run(blk, steps):
    for _ in range(steps):
        dt = blk.max_time_step()
        for s in range(3):
            blk.forward(dt, s)


@pytest.mark.parametrize("scheme", [0, 9, 1])
def test_unit_ratio_follows_the_gas_density(scheme):
    """r = 1: the tracer's fluxes are the mass fluxes, its implicit transfer
    the dry-gas transfer, so s stays the density and r stays 1."""
    g = fv.Grid(16, 12, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0)
    blk = fv.Block(_deck(g, scheme, True))
    blk.initialize(_stratified(g, 2), r=np.ones((1,) + g.shape))
    _run(blk, 3)
    sub = (slice(None),) + g.interior()
    assert np.abs(blk.r[sub] - 1.0).max() < 5e-14
    assert np.abs(blk.s[0] - blk.u[fv.IDN]).max() <= 5e-14 * blk.u[fv.IDN].max()


@pytest.mark.parametrize("scheme", [0, 9])
def test_uniform_ratio_stays_uniform_and_tracer_mass_is_conserved(scheme):
    g = fv.Grid(16, 12, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0)
    blk = fv.Block(_deck(g, scheme, True))
    blk.initialize(_stratified(g, 3), r=np.full((1,) + g.shape, 0.3))
    _run(blk, 3)
    sub = (slice(None),) + g.interior()
    assert np.abs(blk.r[sub] - 0.3).max() < 2e-14
    # a tracer step: conservation with reflecting walls
    blk = fv.Block(_deck(g, scheme, True))
    x2 = g.centers(2)[None, None, :, None] + np.zeros((1,) + g.shape)
    blk.initialize(_stratified(g, 3), r=np.where(np.abs(x2) < 2000.0, 1.0, 0.0))
    vol = g.dx(3) * (g.dx(2) * g.dx(1))
    m0 = blk.s[sub].sum() * vol
    _run(blk, 4)
    assert abs(blk.s[sub].sum() * vol - m0) <= 1e-14 * m0
    assert blk.s[sub].min() >= -1e-12 * blk.s[sub].max()


def test_the_limiter_is_what_keeps_a_tracer_step_non_negative(monkeypatch):
    """The same step without the limiter drains cells below zero (WENO faces
    above the donor's average), so the non-negativity above is the
    limiter's doing."""
    g = fv.Grid(16, 12, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0)
    x2 = g.centers(2)[None, None, :, None] + np.zeros((1,) + g.shape)
    r0 = np.where(np.abs(x2) < 2000.0, 1.0, 0.0)
    sub = (slice(None),) + g.interior()
    blk = fv.Block(_deck(g, 9, True))
    blk.initialize(_stratified(g, 3), r=r0)
    _run(blk, 4)
    assert blk.s[sub].min() >= 0.0
    monkeypatch.setattr(scalar, "limit", lambda s, rho, fl, *a, **k: (fl, None))
    blk = fv.Block(_deck(g, 9, True))
    blk.initialize(_stratified(g, 3), r=r0)
    _run(blk, 4)
    assert blk.s[sub].min() < -1e-12 * blk.s[sub].max()
