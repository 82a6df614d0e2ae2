"""Pins of the VIC oracle (CPU): the block sweep against a dense solve of the
same assembled system (Woodbury points included), mass conservation of the
redistribution, rest in -> rest out, and the explicit limit dt -> 0; for the
partial matrix (scheme 1) also its blocks as the full system with the
tangential substitution, the equality with the full solve without
tangential motion, the tangential momentum conservation and the donor
convention of the moved momentum."""

import numpy as np
import pytest

from oracle import fv, implicit

RD, CP = 287.0, 1004.0
GAMMA = CP / (CP - RD)


def column(n=12, ncol=3, seed=0, rest=False):
    g = fv.Grid(n, 1, 1, 0.0, 50.0 * n)
    x1f = g.faces(1)[3 : 3 + n + 1]
    x1v = g.centers(1)[3 : 3 + n]
    zc = g.centers(1)[2 : 3 + n + 1]
    t = 300.0 - 9.81 * zc / CP
    p = 1e5 * (t / 300.0) ** (CP / RD)
    rng = np.random.default_rng(seed)
    w = np.zeros((ncol, n + 2, 5))
    w[..., 0] = p / (RD * t)
    w[..., 4] = p
    if not rest:
        w[..., 1:4] = 0.5 * (rng.random((ncol, n + 2, 3)) - 0.5)
        w[..., 0] *= 1 + 0.01 * rng.random((ncol, n + 2))
    du = np.zeros((ncol, n, 5)) if rest else 1e-3 * (rng.random((ncol, n, 5)) - 0.5) * np.array([1, 1, 1, 1, 1e3])
    return g, w, du, x1f, x1v


def dense_delta(du, w, dt, g, x1f, x1v):
    n = du.shape[1]
    A = np.full(n + 1, g.dx(3) * g.dx(2))
    V = g.dx(3) * (g.dx(2) * g.dx(1))
    wl = 0.5 * A[0:n] * (x1v - x1f[0:n]) / V
    wh = 0.5 * A[1:] * (x1f[1:] - x1v) / V
    a, b, c = implicit.assemble(w, GAMMA, A, V, wl, wh, dt, -9.81)
    gvar = -9.81 * ((x1f[1:] - x1f[:-1]) ** 2 / 12.0)
    lo, mid, hi, t0, t1 = implicit.pe_rows(x1v, gvar)
    a[..., 4, 0] -= mid / dt
    b[..., 4, 0] -= lo / dt
    c[..., 4, 0] -= hi / dt
    rhs = du.copy()
    rhs[..., 4] -= implicit.apply_rows((lo, mid, hi, t0, t1), du[..., 0])
    rhs = rhs / dt
    out = []
    for col in range(du.shape[0]):
        M = np.zeros((5 * n, 5 * n))
        for i in range(n):
            M[5 * i : 5 * i + 5, 5 * i : 5 * i + 5] = a[col, i]
            if i > 0:
                M[5 * i : 5 * i + 5, 5 * (i - 1) : 5 * i] = b[col, i]
            if i < n - 1:
                M[5 * i : 5 * i + 5, 5 * (i + 1) : 5 * i + 10] = c[col, i]
        M[4, 5 * 2 + 0] += -t0 / dt
        M[5 * (n - 1) + 4, 5 * (n - 3) + 0] += -t1 / dt
        out.append(np.linalg.solve(M, rhs[col].ravel()).reshape(n, 5))
    return np.array(out)


def test_block_sweep_matches_dense_solve():
    g, w, du, x1f, x1v = column()
    dt = 2.0
    out = implicit.vic_full(du, w, GAMMA, -9.81, dt, g.dx(1), x1f, x1v, g.dx(3) * g.dx(2),
                            g.dx(3) * (g.dx(2) * g.dx(1)))
    ref = dense_delta(du, w, dt, g, x1f, x1v)
    # momentum rows are the solved increments themselves
    np.testing.assert_allclose(out[..., 1:4], ref[..., 1:4], rtol=1e-9, atol=1e-14)
    # the redistributed mass conserves the column total of the solve
    V = g.dx(3) * (g.dx(2) * g.dx(1))
    np.testing.assert_allclose((out[..., 0] * V).sum(-1), (du[..., 0] * V).sum(-1), rtol=0, atol=1e-13)


def test_rest_in_rest_out():
    g, w, du, x1f, x1v = column(rest=True)
    out = implicit.vic_full(du, w, GAMMA, -9.81, 5.0, g.dx(1), x1f, x1v, 1.0, g.dx(1))
    assert np.abs(out).max() < 1e-12


def test_small_dt_limit_is_explicit():
    g, w, du, x1f, x1v = column(seed=3)
    out = implicit.vic_full(du, w, GAMMA, -9.81, 1e-9, g.dx(1), x1f, x1v, 1.0, g.dx(1))
    np.testing.assert_allclose(out[..., 1:4], du[..., 1:4], rtol=1e-5, atol=1e-12)


# ---------------------------------------------------------------------------
# implicit-scheme 1 (the partial 3x3 matrix with the tangential fold)
# ---------------------------------------------------------------------------
def _geometry(g, x1f, x1v, n):
    A = np.full(n + 1, g.dx(3) * g.dx(2))
    V = g.dx(3) * (g.dx(2) * g.dx(1))
    wl = 0.5 * A[0:n] * (x1v - x1f[0:n]) / V
    wh = 0.5 * A[1:] * (x1f[1:] - x1v) / V
    return A, V, wl, wh


def test_partial_blocks_are_the_full_system_with_the_tangential_substitution():
    """Substituting dm_t,j = v_t,j drho_j + q_t,j into the rho, m1, E rows of
    the full 5x5 system gives the partial blocks and the fold's right-hand
    side: the rho column gains the m_t columns times the velocity of the
    cell each block multiplies, q_t moves to the right."""
    g, w, du, x1f, x1v = column(seed=5)
    n, dt = du.shape[1], 3.0
    A, V, wl, wh = _geometry(g, x1f, x1v, n)
    a5, b5, c5 = implicit.assemble(w, GAMMA, A, V, wl, wh, dt, -9.81)
    a3, b3, c3, corr = implicit.assemble_partial(w, du, GAMMA, A, V, wl, wh, dt, -9.81)
    R, T = [0, 1, 4], [2, 3]
    vt = w[:, 1 : n + 1][..., T]
    qt = du[..., T] - vt * du[..., [0]]
    ea, eb, ec = a5[..., R, :][..., :, R].copy(), b5[..., R, :][..., :, R].copy(), c5[..., R, :][..., :, R].copy()
    ecorr = np.zeros_like(corr)
    for i in range(n):
        ea[:, i, :, 0] += np.einsum("crk,ck->cr", a5[:, i][:, R][:, :, T], vt[:, i])
        ecorr[:, i] += np.einsum("crk,ck->cr", a5[:, i][:, R][:, :, T], qt[:, i])
        if i > 0:
            eb[:, i, :, 0] += np.einsum("crk,ck->cr", b5[:, i][:, R][:, :, T], vt[:, i - 1])
            ecorr[:, i] += np.einsum("crk,ck->cr", b5[:, i][:, R][:, :, T], qt[:, i - 1])
        if i < n - 1:
            ec[:, i, :, 0] += np.einsum("crk,ck->cr", c5[:, i][:, R][:, :, T], vt[:, i + 1])
            ecorr[:, i] += np.einsum("crk,ck->cr", c5[:, i][:, R][:, :, T], qt[:, i + 1])
    scale = np.abs(a5).max()
    for got, want in ((a3, ea), (b3, eb), (c3, ec)):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(corr, ecorr, rtol=1e-11, atol=1e-16)


def test_partial_sweep_matches_dense_solve():
    g, w, du, x1f, x1v = column(seed=6)
    n, dt = du.shape[1], 2.5
    A, V, wl, wh = _geometry(g, x1f, x1v, n)
    a, b, c, corr = implicit.assemble_partial(w, du, GAMMA, A, V, wl, wh, dt, -9.81)
    rng = np.random.default_rng(1)
    rhs = rng.normal(size=(du.shape[0], n, 3))
    y, z0, z1 = implicit.sweep_partial(a, b, c, rhs, corr)
    for col in range(du.shape[0]):
        M = np.zeros((3 * n, 3 * n))
        for i in range(n):
            M[3 * i : 3 * i + 3, 3 * i : 3 * i + 3] = a[col, i]
            if i > 0:
                M[3 * i : 3 * i + 3, 3 * (i - 1) : 3 * i] = b[col, i]
            if i < n - 1:
                M[3 * i : 3 * i + 3, 3 * (i + 1) : 3 * i + 6] = c[col, i]
        r = np.where(corr[col] != 0, rhs[col] - corr[col], rhs[col]).ravel()
        np.testing.assert_allclose(y[col].ravel(), np.linalg.solve(M, r), rtol=1e-9, atol=1e-12)
        e0 = np.zeros(3 * n)
        e0[2] = 1.0
        e1 = np.zeros(3 * n)
        e1[3 * n - 1] = 1.0
        np.testing.assert_allclose(z0[col].ravel(), np.linalg.solve(M, e0), rtol=1e-9, atol=1e-14)
        np.testing.assert_allclose(z1[col].ravel(), np.linalg.solve(M, e1), rtol=1e-9, atol=1e-14)


def test_inverse3_is_the_inverse():
    rng = np.random.default_rng(2)
    m = rng.normal(size=(50, 3, 3)) + 3 * np.eye(3)
    np.testing.assert_allclose(implicit.inverse3(m), np.linalg.inv(m), rtol=1e-12, atol=1e-14)


def test_partial_without_tangential_motion_is_the_full_solve():
    """With v_t = 0 and du_t = 0 the full system's m_t rows decouple with a
    zero solution, so both schemes solve the same rho, m1, E system."""
    g, w, du, x1f, x1v = column(seed=7)
    w[..., 2:4] = 0.0
    du[..., 2:4] = 0.0
    args = (GAMMA, -9.81, 4.0, g.dx(1), x1f, x1v, g.dx(3) * g.dx(2), g.dx(3) * (g.dx(2) * g.dx(1)))
    full = implicit.vic_full(du, w, *args)
    part = implicit.vic_partial(du, w, *args)
    np.testing.assert_allclose(part, full, rtol=1e-9, atol=1e-12 * np.abs(full).max())
    assert np.abs(part[..., 2:4]).max() == 0.0


def test_partial_conserves_mass_and_tangential_momentum():
    g, w, du, x1f, x1v = column(seed=8)
    V = g.dx(3) * (g.dx(2) * g.dx(1))
    out = implicit.vic_partial(du, w, GAMMA, -9.81, 3.0, g.dx(1), x1f, x1v, g.dx(3) * g.dx(2), V)
    for row in (0, 2, 3):
        np.testing.assert_allclose((out[..., row] * V).sum(-1), (du[..., row] * V).sum(-1), rtol=0,
                                   atol=1e-13 * np.abs(du[..., row] * V).sum())
    # and the moved mass did carry tangential momentum
    assert np.abs(out[..., 2:4] - du[..., 2:4]).max() > 0


def test_partial_rest_and_small_dt():
    g, w, du, x1f, x1v = column(rest=True)
    out = implicit.vic_partial(du, w, GAMMA, -9.81, 5.0, g.dx(1), x1f, x1v, 1.0, g.dx(1))
    assert np.abs(out).max() < 1e-12
    g, w, du, x1f, x1v = column(seed=3)
    out = implicit.vic_partial(du, w, GAMMA, -9.81, 1e-9, g.dx(1), x1f, x1v, 1.0, g.dx(1))
    np.testing.assert_allclose(out[..., [1]], du[..., [1]], rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_moved_mass_carries_the_velocity_of_the_cell_it_leaves(sign):
    """One face moves mass q: the mass update takes q from cell i and gives
    it to cell i + 1 (q > 0), or the reverse (q < 0); the momentum moved is q
    times the tangential velocity of the cell the mass leaves."""
    n, V = 5, 2.0
    w = np.zeros((1, n + 2, 5))
    w[0, 1 : n + 1, 2] = [1.0, 2.0, 3.0, 4.0, 5.0]
    w[0, 1 : n + 1, 3] = [-1.0, -2.0, -3.0, -4.0, -5.0]
    du = np.zeros((1, n, 5))
    moved = np.zeros((1, n + 1))
    q = 0.6 * sign
    moved[0, 3] = q  # the face between cells 2 and 3
    implicit.tangential_transfer(du, w, moved, V)
    src = 2 if q > 0 else 3
    for m in (2, 3):
        v = w[0, 1 + src, m]
        assert du[0, 2, m] == -q * v / V and du[0, 3, m] == q * v / V
        assert np.count_nonzero(du[0, :, m]) == 2
