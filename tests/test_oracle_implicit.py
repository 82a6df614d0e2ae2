"""Pins of the VIC oracle (CPU): the block sweep against a dense solve of the
same assembled system (Woodbury points included), mass conservation of the
redistribution, rest in -> rest out, and the explicit limit dt -> 0."""

import numpy as np

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
