"""Pins of the VIC oracle's assembly pieces (CPU), each by what the
mathematics of the Euler equations fixes, not by restating the formulas:

* ``flux_jacobian`` (flux_decomposition_impl.h:105-126) is dF/dU: central
  finite differences of the physical flux at random states, in all three
  sweep directions (the matrix is in the direction's rotated frame);
* ``eigen`` (flux_decomposition_impl.h:62-103): R and its stated inverse
  multiply to the identity, and R diag(u - c, u, u + c, u, u) R^-1 is the
  flux Jacobian of the same state;
* ``roe_average`` (flux_decomposition_impl.h:31-60): Roe's property
  A(U~)(U_R - U_L) = F(U_R) - F(U_L), symmetry in the two states, and
  A(U, U) = A(U);
* ``roe_abs_matrix``: |A| squared is A squared, |A| commutes with A, and
  |A| applied to A's right eigenvectors scales them by |lambda|.

Each check fails for a one-term mistake in the restated matrices (a
coefficient off in one entry, a wrong sign, an eigenvalue off by a factor),
which ``tests/test_oracle_implicit.py`` alone does not see: its dense solve
solves the same assembled matrix.
"""

import numpy as np
import pytest

from oracle import implicit

IDN, IVX, IVY, IVZ, IPR = 0, 1, 2, 3, 4


def random_prims(n, seed):
    rng = np.random.default_rng(seed)
    w = np.empty((n, 5))
    w[:, IDN] = 0.3 + 1.5 * rng.random(n)
    w[:, IVX:IVZ + 1] = 2.0 * (rng.random((n, 3)) - 0.5)
    w[:, IPR] = 0.2 + 2.0 * rng.random(n)
    return w


def rotate(w, dir):
    """The direction's frame: the normal velocity first, then the two
    tangential ones in the cyclic order the Jacobian uses."""
    out = w.copy()
    for k in range(3):
        out[..., IVX + k] = w[..., IVX + (k + dir) % 3]
    return out


def cons(w, gamma):
    u = np.empty_like(w)
    u[..., IDN] = w[..., IDN]
    for v in (IVX, IVY, IVZ):
        u[..., v] = w[..., IDN] * w[..., v]
    ke = 0.5 * w[..., IDN] * (w[..., IVX] ** 2 + w[..., IVY] ** 2 + w[..., IVZ] ** 2)
    u[..., IPR] = w[..., IPR] / (gamma - 1.0) + ke
    return u


def prim(u, gamma):
    w = np.empty_like(u)
    w[..., IDN] = u[..., IDN]
    for v in (IVX, IVY, IVZ):
        w[..., v] = u[..., v] / u[..., IDN]
    ke = 0.5 * (u[..., IVX] ** 2 + u[..., IVY] ** 2 + u[..., IVZ] ** 2) / u[..., IDN]
    w[..., IPR] = (gamma - 1.0) * (u[..., IPR] - ke)
    return w


def flux_x(u, gamma):
    """The Euler flux along the first velocity component."""
    w = prim(u, gamma)
    f = np.empty_like(u)
    vn = w[..., IVX]
    f[..., IDN] = u[..., IVX]
    f[..., IVX] = u[..., IVX] * vn + w[..., IPR]
    f[..., IVY] = u[..., IVY] * vn
    f[..., IVZ] = u[..., IVZ] * vn
    f[..., IPR] = (u[..., IPR] + w[..., IPR]) * vn
    return f


GAMMAS = (1.4, 1004.0 / 717.0, 5.0 / 3.0)


@pytest.mark.parametrize("gamma", GAMMAS)
@pytest.mark.parametrize("dir", [0, 1, 2])
def test_flux_jacobian_is_the_derivative_of_the_flux(gamma, dir):
    w = random_prims(16, seed=11 + dir)
    jac = implicit.flux_jacobian(gamma - 1.0, w, dir)
    wr = rotate(w, dir)
    u = cons(wr, gamma)
    for k in range(5):
        h = 1e-6 * np.maximum(np.abs(u[:, k]), 1.0)
        up, um = u.copy(), u.copy()
        up[:, k] += h
        um[:, k] -= h
        col = (flux_x(up, gamma) - flux_x(um, gamma)) / (2.0 * h[:, None])
        np.testing.assert_allclose(jac[:, :, k], col, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("gamma", GAMMAS)
def test_eigenvectors_invert_and_diagonalise_the_jacobian(gamma):
    gm1 = gamma - 1.0
    w = random_prims(24, seed=5)
    cs = implicit.sound_speed(w, gm1)
    R, Ri, lam = implicit.eigen(w, cs, gm1)
    eye = np.broadcast_to(np.eye(5), R.shape)
    np.testing.assert_allclose(R @ Ri, eye, atol=1e-12)
    np.testing.assert_allclose(Ri @ R, eye, atol=1e-12)
    u = w[:, IVX]
    signed = np.stack([u - cs, u, u + cs, u, u], -1)
    np.testing.assert_array_equal(lam, np.abs(signed))
    A = (R * signed[:, None, :]) @ Ri
    np.testing.assert_allclose(A, implicit.flux_jacobian(gm1, w), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("gamma", GAMMAS)
def test_roe_average_has_roes_property(gamma):
    gm1 = gamma - 1.0
    wl = random_prims(32, seed=7)
    wr = random_prims(32, seed=8)
    roe = implicit.roe_average(gm1, wl, wr)
    A = implicit.flux_jacobian(gm1, roe)
    du = cons(wr, gamma) - cons(wl, gamma)
    df = flux_x(cons(wr, gamma), gamma) - flux_x(cons(wl, gamma), gamma)
    np.testing.assert_allclose(np.einsum("nij,nj->ni", A, du), df, rtol=1e-10, atol=1e-10)
    # symmetric in the two states, and the state itself when they agree
    np.testing.assert_allclose(implicit.roe_average(gm1, wr, wl), roe, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(implicit.roe_average(gm1, wl, wl), wl, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("gamma", GAMMAS)
def test_roe_abs_matrix_is_the_absolute_value_of_the_jacobian(gamma):
    gm1 = gamma - 1.0
    wl = random_prims(20, seed=9)
    wr = random_prims(20, seed=10)
    absA = implicit.roe_abs_matrix(wl, wr, gamma)
    roe = implicit.roe_average(gm1, wl, wr)
    A = implicit.flux_jacobian(gm1, roe)
    scale = np.abs(A).max(axis=(1, 2))[:, None, None] ** 2
    np.testing.assert_allclose((absA @ absA) / scale, (A @ A) / scale, atol=1e-11)
    np.testing.assert_allclose((absA @ A) / scale, (A @ absA) / scale, atol=1e-11)
    # |A| r_k = |lambda_k| r_k on the Roe state's right eigenvectors
    cs = implicit.sound_speed(roe, gm1)
    R, _, lam = implicit.eigen(roe, cs, gm1)
    np.testing.assert_allclose(absA @ R, R * lam[:, None, :], rtol=1e-10, atol=1e-10)
    # and for equal states |A| is |A(U)|: positive semi-definite spectrum
    ev = np.linalg.eigvals(implicit.roe_abs_matrix(wl, wl, gamma))
    assert (ev.real > -1e-10).all()


def test_partial_blocks_take_the_pinned_rows():
    """The 3x3 blocks of scheme 1 use the same Jacobian and |A| as the full
    5x5 blocks (rows and columns rho, m1, E), so the pins above carry over:
    check on a random column that the partial blocks without tangential
    motion are exactly the full blocks' sub-blocks."""
    from tests.test_oracle_implicit import GAMMA, _geometry, column

    g, w, du, x1f, x1v = column(seed=21)
    w[..., IVY] = 0.0
    w[..., IVZ] = 0.0
    du[..., IVY] = 0.0
    du[..., IVZ] = 0.0
    n = du.shape[1]
    A, V, wl, wh = _geometry(g, x1f, x1v, n)
    a5, b5, c5 = implicit.assemble(w, GAMMA, A, V, wl, wh, 2.0, -9.81)
    a3, b3, c3, corr = implicit.assemble_partial(w, du, GAMMA, A, V, wl, wh, 2.0, -9.81)
    R = [IDN, IVX, IPR]
    np.testing.assert_array_equal(a3, a5[..., R, :][..., :, R])
    np.testing.assert_array_equal(b3, b5[..., R, :][..., :, R])
    np.testing.assert_array_equal(c3, c5[..., R, :][..., :, R])
    assert not corr.any()


def _column_residual(U, w, gamma, area, vol, work_lo, work_hi, grav):
    """The semi-discrete x1 residual dU/dt = -R(U) that the VIC linearises,
    written from its definition: Roe-type face fluxes
    F = (F(U_L) + F(U_R))/2 - |A|(U_R - U_L)/2 with |A| frozen at the base
    primitives w, the flux divergence (A+ F+ - A- F-)/V, the momentum
    source rho g and the energy source of the face gravity work
    g (2 work_hi F+_mass + 2 work_lo F-_mass); reflecting walls at both ends
    (the ghost is the mirror of the wall cell, normal momentum negated).
    U: (ncol, n, 5) conserved, interior cells only."""
    ncol, n, _ = U.shape
    bnd = np.array([1.0, -1.0, 1.0, 1.0, 1.0])
    Ug = np.concatenate([U[:, :1] * bnd, U, U[:, -1:] * bnd], axis=1)
    absA = implicit.roe_abs_matrix(w[:, :-1], w[:, 1:], gamma)  # (ncol, n + 1, 5, 5)
    fl = flux_x(Ug[:, :-1], gamma)
    fr = flux_x(Ug[:, 1:], gamma)
    F = 0.5 * (fl + fr) - 0.5 * np.einsum("cfij,cfj->cfi", absA, Ug[:, 1:] - Ug[:, :-1])
    div = (area[None, 1:, None] * F[:, 1:] - area[None, :-1, None] * F[:, :-1]) / vol
    S = np.zeros_like(U)
    S[..., IVX] = U[..., IDN] * grav
    S[..., IPR] = grav * (2.0 * work_hi * F[:, 1:, IDN] + 2.0 * work_lo * F[:, :-1, IDN])
    return div - S


def test_assembled_blocks_linearise_the_roe_flux_residual():
    """b_i, a_i - I/dt, c_i are the derivatives of R_i with respect to
    U_{i-1}, U_i, U_{i+1} (central differences of the residual above), in
    every row including the wall-merged end rows and the gravity-work energy
    rows.  Pins the halves, the areas and signs of the divergence, the
    gravity source and the wall merge of vic_assemble_impl<5>."""
    from tests.test_oracle_implicit import GAMMA, _geometry, column

    g, w, du, x1f, x1v = column(n=8, ncol=2, seed=31)
    # mirror ghosts, as the walls take them
    w[:, 0] = w[:, 1] * np.array([1.0, -1.0, 1.0, 1.0, 1.0])
    w[:, -1] = w[:, -2] * np.array([1.0, -1.0, 1.0, 1.0, 1.0])
    n, dt, grav = du.shape[1], 3.0, -9.81
    A, V, wl, wh = _geometry(g, x1f, x1v, n)
    a, b, c = implicit.assemble(w, GAMMA, A, V, wl, wh, dt, grav)
    U = cons(w[:, 1:-1], GAMMA)
    for col in range(U.shape[0]):
        for j in range(n):
            for k in range(5):
                h = 1e-6 * max(abs(U[col, j, k]), 1.0)
                up, um = U.copy(), U.copy()
                up[col, j, k] += h
                um[col, j, k] -= h
                d = (_column_residual(up, w, GAMMA, A, V, wl, wh, grav)
                     - _column_residual(um, w, GAMMA, A, V, wl, wh, grav))[col] / (2.0 * h)
                for i in range(n):
                    if i == j:
                        blk = a[col, i] - np.eye(5) / dt
                    elif i == j + 1:
                        blk = b[col, i]
                    elif i == j - 1:
                        blk = c[col, i]
                    else:
                        np.testing.assert_allclose(d[i], 0.0, atol=1e-9)
                        continue
                    scale = np.abs(blk).max() + 1e-3
                    np.testing.assert_allclose(d[i] / scale, blk[:, k] / scale, atol=2e-7,
                                               err_msg=f"col {col} cell {i} wrt cell {j} var {k}")


def test_pe_rows_are_the_slope_operator():
    """pe_rows/apply_rows hold gvar times the three-point centroid slope
    (pinned on quadratics in tests/test_oracle_gravity.py) as a banded
    matrix with its two wall third points."""
    from oracle import gravity

    x = np.cumsum(np.array([0.0, 1.0, 1.3, 0.7, 2.0, 1.1, 0.9, 1.2]))
    gvar = 0.1 + np.arange(x.size) * 0.01
    q = np.random.default_rng(2).random((3, x.size))
    rows = implicit.pe_rows(x, gvar)
    np.testing.assert_allclose(implicit.apply_rows(rows, q), gvar * gravity.centroid_slope(q, x), rtol=1e-13,
                               atol=1e-15)
