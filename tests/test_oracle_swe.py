"""Pins of the shallow-water oracle (oracle/swe.py): the Riemann flux of two
equal states is the physical flux, a flow supersonic across the face takes
the upwind flux (the Roe linearisation's defining property), a lake at rest
stays at rest, a periodic block conserves mass and momentum, and a mirror
image of the initial state gives the mirror image of the solution."""

import numpy as np
import pytest

from oracle import fv, swe


def test_equal_states_give_the_physical_flux():
    rng = np.random.default_rng(0)
    w = np.zeros((4, 1, 1, 50))
    w[0] = 5.0 + 10.0 * rng.random(w[0].shape)
    w[1:3] = rng.random(w[1:3].shape) - 0.5
    for dim in (3, 2):
        np.testing.assert_array_equal(swe.shallow_roe(w, w, dim), swe.physical_flux(w, dim))


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_supersonic_face_takes_the_upwind_flux(sign):
    """All three Roe speeds of one sign: the flux is the upwind side's
    physical flux (F_roe = F_L for flow to +x, F_R for flow to -x), which
    holds only with the exact Roe averages, the 1/2 factors and the signs."""
    rng = np.random.default_rng(1)
    n = 40
    wl = np.zeros((4, 1, 1, n))
    wr = np.zeros((4, 1, 1, n))
    wl[0] = 4.0 + rng.random(n)
    wr[0] = 4.0 + rng.random(n)
    wl[1] = sign * (6.0 + rng.random(n))  # |u| > sqrt(phi) ~ 2.2
    wr[1] = sign * (6.0 + rng.random(n))
    wl[2] = rng.random(n) - 0.5
    wr[2] = rng.random(n) - 0.5
    f = swe.shallow_roe(wl, wr, 3)
    up = swe.physical_flux(wl if sign > 0 else wr, 3)
    np.testing.assert_allclose(f, up, rtol=1e-13, atol=1e-12)


def _block(nx1, nx2, bcs, w):
    g = fv.Grid(nx1, nx2, 1, -10.0, 10.0, -5.0, 5.0, -0.5, 0.5)
    b = swe.Block(swe.Deck(grid=g, bcs=bcs))
    b.initialize(w(g))
    return g, b


def test_lake_at_rest():
    def w(g):
        out = np.zeros((4,) + g.shape)
        out[0] = 10.0
        return out

    g, b = _block(16, 8, ("reflecting",) * 2 + ("periodic",) * 4, w)
    for _ in range(5):
        dt = b.max_time_step()
        for s in range(3):
            b.forward(dt, s)
    sub = (slice(None),) + g.interior()
    assert np.all(b.u[sub][0] == 10.0)
    assert np.max(np.abs(b.u[sub][1:])) == 0.0


def _hump(g, cx=0.0, cy=0.0):
    x = g.centers(1)[None, None, :]
    y = g.centers(2)[None, :, None]
    out = np.zeros((4,) + g.shape)
    out[0] = 10.0 + 2.0 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2))
    out[1] = 0.5 + 0.0 * out[0]
    return out


def test_periodic_block_conserves():
    g, b = _block(32, 16, ("periodic",) * 6, _hump)
    sub = (slice(None),) + g.interior()
    m0 = b.u[sub].sum(axis=(1, 2, 3))
    for _ in range(10):
        dt = b.max_time_step()
        for s in range(3):
            b.forward(dt, s)
    m1 = b.u[sub].sum(axis=(1, 2, 3))
    assert abs(m1[0] - m0[0]) < 1e-12 * abs(m0[0])
    # no sources: momentum is conserved as well
    assert abs(m1[1] - m0[1]) < 1e-12 * abs(m0[0])
    assert abs(m1[2] - m0[2]) < 1e-12 * abs(m0[0])


def test_mirror_symmetry_in_x2():
    """A state symmetric about x2 = 0 stays symmetric, with v2 antisymmetric."""
    g, b = _block(24, 16, ("periodic",) * 2 + ("reflecting",) * 2 + ("periodic",) * 2, _hump)
    for _ in range(6):
        dt = b.max_time_step()
        for s in range(3):
            b.forward(dt, s)
    sub = g.interior()
    h = b.u[0][sub]
    m2 = b.u[2][sub]
    np.testing.assert_allclose(h, h[:, ::-1, :], rtol=1e-12)
    np.testing.assert_allclose(m2, -m2[:, ::-1, :], atol=1e-10)
