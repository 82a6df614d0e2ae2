"""CPU FP64 oracle of the vertical implicit correction, full 5x5 matrix
(implicit-scheme 9) on a Cartesian single-block column with x1 gravity.

TEST INFRASTRUCTURE ONLY (see oracle/fv.py).  Restates
ImplicitHydroImpl::forward_masked (src/implicit/implicit_hydro.cpp:301-623)
for no solids, dry ideal gas, FP64, no register posts:

  * the block assembly vic_assemble_impl<5> (src/implicit/
    vic_assemble_impl.h:45-204) with the flux decomposition
    (flux_decomposition_impl.h) and the gravity-work rows
    (vic_gravity_work_impl.h:24-36);
  * the corrected-PE coupling couple_radial_exact (implicit_hydro.cpp:
    408-446) and its Woodbury wall third points (vic_kernels.h:129-150);
  * the block-tridiagonal forward sweep and back substitution
    (forward_sweep_impl.h:31-132, vic_redistribute_impl.h:31-51);
  * the mass redistribution of the column (vic_constituent_column,
    vic_redistribute_impl.h:56-175; vic_redistribute_cell :208-229);
  * the post-solve remainder book_remainder_ (implicit_hydro.cpp:713-794).

The column arrays carry one ghost row at each end: w has shape
(ncol, nlayer + 2, 5); du has (ncol, nlayer, 5).  Pinning:
tests/test_oracle_implicit.py solves the same assembled system densely
(numpy.linalg.solve on the (5n)x(5n) matrix, Woodbury points included),
checks mass conservation of the redistribution, the rest state (zero du in,
zero du out), and the small-dt limit delta -> du.
"""

from __future__ import annotations

import numpy as np

IDN, IVX, IVY, IVZ, IPR = 0, 1, 2, 3, 4


def sound_speed(prim, gm1):
    return np.sqrt(prim[..., IPR] * (gm1 + 1.0) / prim[..., IDN])


def roe_average(gm1, wl, wr):
    """RoeAverage (flux_decomposition_impl.h:31-60)."""
    sqrtdl = np.sqrt(wl[..., IDN])
    sqrtdr = np.sqrt(wr[..., IDN])
    isdlpdr = 1.0 / (sqrtdl + sqrtdr)
    prim = np.empty_like(wl)
    prim[..., IDN] = sqrtdl * sqrtdr
    for v in (IVX, IVY, IVZ):
        prim[..., v] = (sqrtdl * wl[..., v] + sqrtdr * wr[..., v]) * isdlpdr
    sq = lambda x: x * x  # noqa: E731
    el = wl[..., IPR] / gm1 + 0.5 * wl[..., IDN] * (sq(wl[..., IVX]) + sq(wl[..., IVY]) + sq(wl[..., IVZ]))
    er = wr[..., IPR] / gm1 + 0.5 * wr[..., IDN] * (sq(wr[..., IVX]) + sq(wr[..., IVY]) + sq(wr[..., IVZ]))
    hbar = ((el + wl[..., IPR]) / sqrtdl + (er + wr[..., IPR]) / sqrtdr) * isdlpdr
    prim[..., IPR] = (
        (hbar - 0.5 * (sq(prim[..., IVX]) + sq(prim[..., IVY]) + sq(prim[..., IVZ])))
        * gm1
        / (gm1 + 1.0)
        * prim[..., IDN]
    )
    return prim


def eigen(prim, cs, gm1, dir=0):
    """Eigenvector / Eigenvalue (flux_decomposition_impl.h:62-103)."""
    r = prim[..., IDN]
    u = prim[..., IVX + dir]
    v = prim[..., IVX + (IVY - IVX + dir) % 3]
    w = prim[..., IVX + (IVZ - IVX + dir) % 3]
    p = prim[..., IPR]
    ke = 0.5 * (u * u + v * v + w * w)
    hp = (gm1 + 1.0) / gm1 * p / r
    h = hp + ke
    z, o = np.zeros_like(u), np.ones_like(u)
    R = np.stack(
        [
            np.stack([o, o, o, z, z], -1),
            np.stack([u - cs, u, u + cs, z, z], -1),
            np.stack([v, v, v, o, z], -1),
            np.stack([w, w, w, z, o], -1),
            np.stack([h - u * cs, ke, h + u * cs, v, w], -1),
        ],
        -2,
    )
    Ri = np.stack(
        [
            np.stack([(cs * ke + u * hp) / (2.0 * cs * hp), (-hp - cs * u) / (2.0 * cs * hp),
                      -v / (2.0 * hp), -w / (2.0 * hp), 1.0 / (2.0 * hp)], -1),
            np.stack([(hp - ke) / hp, u / hp, v / hp, w / hp, -1.0 / hp], -1),
            np.stack([(cs * ke - u * hp) / (2.0 * cs * hp), (hp - cs * u) / (2.0 * cs * hp),
                      -v / (2.0 * hp), -w / (2.0 * hp), 1.0 / (2.0 * hp)], -1),
            np.stack([-v, z, o, z, z], -1),
            np.stack([-w, z, z, o, z], -1),
        ],
        -2,
    )
    lam = np.stack([np.abs(u - cs), np.abs(u), np.abs(u + cs), np.abs(u), np.abs(u)], -1)
    return R, Ri, lam


def flux_jacobian(gm1, w, dir=0):
    """FluxJacobian (flux_decomposition_impl.h:105-126)."""
    v1 = w[..., IVX + dir]
    v2 = w[..., IVX + (IVY - IVX + dir) % 3]
    v3 = w[..., IVX + (IVZ - IVX + dir) % 3]
    rho = w[..., IDN]
    pres = w[..., IPR]
    s2 = v1 * v1 + v2 * v2 + v3 * v3
    c1 = ((gm1 - 1) * s2 / 2 - (gm1 + 1) / gm1 * pres / rho) * v1
    c2 = (gm1 + 1) / gm1 * pres / rho + s2 / 2 - gm1 * v1 * v1
    z, o = np.zeros_like(v1), np.ones_like(v1)
    return np.stack(
        [
            np.stack([z, o, z, z, z], -1),
            np.stack([gm1 * s2 / 2 - v1 * v1, (2.0 - gm1) * v1, -gm1 * v2, -gm1 * v3, gm1 * o], -1),
            np.stack([-v1 * v2, v2, v1, z, z], -1),
            np.stack([-v1 * v3, v3, z, v1, z], -1),
            np.stack([c1, c2, -gm1 * v2 * v1, -gm1 * v3 * v1, (gm1 + 1) * v1], -1),
        ],
        -2,
    )


def roe_abs_matrix(wl, wr, gamma):
    gm1 = 0.5 * (gamma + gamma) - 1.0
    prim = roe_average(gm1, wl, wr)
    cs = sound_speed(prim, gm1)
    R, Ri, lam = eigen(prim, cs, gm1)
    return (R * lam[..., None, :]) @ Ri


def assemble(w, gamma, area, vol, work_lo, work_hi, dt, grav):
    """a, b, c of every layer of every column (vic_assemble_impl<5>), no
    solids, both column ends physical walls (merged)."""
    ncol, nl2, _ = w.shape
    n = nl2 - 2
    gm1 = gamma - 1.0
    wm, wc, wp = w[:, 0:n], w[:, 1 : n + 1], w[:, 2 : n + 2]
    dprev = flux_jacobian(gm1, wm)
    dcurr = flux_jacobian(gm1, wc)
    dnext = flux_jacobian(gm1, wp)
    Am = roe_abs_matrix(wm, wc, gamma)
    Ap = roe_abs_matrix(wc, wp, gamma)
    Phi = np.zeros((5, 5))
    Phi[IVX, IDN] = grav
    Phi[IPR, IVX] = grav
    ai, aip = area[0:n], area[1 : n + 1]
    hiv = 0.5 / vol
    a = (Am * ai[:, None, None] + Ap * aip[:, None, None] + (aip - ai)[:, None, None] * dcurr) * hiv - Phi
    a = a + np.eye(5) * (1.0 / dt)
    b = -(Am + dprev) * ai[:, None, None] * hiv
    c = -(Ap - dnext) * aip[:, None, None] * hiv
    # gravity-work rows (vic_gravity_work_impl.h:24-36), e = IPR, m = IVX
    am, ap = Am[..., 0, :], Ap[..., 0, :]
    lo, up = work_lo[None, :, None], work_hi[None, :, None]
    a[..., IPR, :] = a[..., IPR, :] - grav * (up * ap - lo * am)
    b[..., IPR, :] = b[..., IPR, :] - grav * lo * am
    c[..., IPR, :] = c[..., IPR, :] + grav * up * ap
    a[..., IPR, IVX] = a[..., IPR, IVX] + grav * (1.0 - work_lo - work_hi)
    b[..., IPR, IVX] = b[..., IPR, IVX] - grav * work_lo
    c[..., IPR, IVX] = c[..., IPR, IVX] - grav * work_hi
    bnd = np.ones(5)
    bnd[IVX] = -1.0
    a[:, 0] = a[:, 0] + b[:, 0] * bnd
    a[:, n - 1] = a[:, n - 1] + c[:, n - 1] * bnd
    return a, b, c


def pe_rows(x1v, gvar):
    """GravityWork::update_rows (gravity_work.hpp:321-346): the corrected-PE
    operator's three bands and two wall third points on a whole column."""
    from .gravity import centroid_slope_weights

    n = x1v.shape[0]
    idx, wt = centroid_slope_weights(x1v)
    ar = np.arange(n)

    def band(off):
        at = ar + off
        s = wt[0] * (idx[0] == at)
        s = s + wt[1] * (idx[1] == at)
        s = s + wt[2] * (idx[2] == at)
        return s

    mid = gvar * band(0)
    lo = np.zeros(n)
    hi = np.zeros(n)
    lo[1:n] = band(-1)[1:n]
    hi[0 : n - 1] = band(1)[0 : n - 1]
    lo = lo * gvar
    hi = hi * gvar
    t0 = gvar[0] * band(2)[0]
    t1 = gvar[n - 1] * band(-2)[n - 1]
    return lo, mid, hi, t0, t1


def apply_rows(rows, q):
    lo, mid, hi, t0, t1 = rows
    n = q.shape[-1]
    s = mid * q
    s[..., 1:n] = s[..., 1:n] + lo[1:n] * q[..., 0 : n - 1]
    s[..., 0 : n - 1] = s[..., 0 : n - 1] + hi[0 : n - 1] * q[..., 1:n]
    s[..., 0] = s[..., 0] + t0 * q[..., 2]
    s[..., n - 1] = s[..., n - 1] + t1 * q[..., n - 3]
    return s


def block_solve(a, b, c, rhs):
    """Forward sweep + back substitution of the block-tridiagonal system
    b_i x_{i-1} + a_i x_i + c_i x_{i+1} = rhs_i (ForwardSweep,
    forward_sweep_impl.h:31-132; vic_backward_substitute)."""
    n = a.shape[1]
    ap = np.empty_like(a)
    d = np.empty_like(rhs)
    for i in range(n):
        ai = a[:, i]
        r = rhs[:, i]
        if i > 0:
            ai = ai - b[:, i] @ ap[:, i - 1]
            r = r - np.einsum("...ij,...j->...i", b[:, i], d[:, i - 1])
        ap[:, i] = np.linalg.solve(ai, c[:, i])
        d[:, i] = np.linalg.solve(ai, r[..., None])[..., 0]
    for i in range(n - 2, -1, -1):
        d[:, i] = d[:, i] - np.einsum("...ij,...j->...i", ap[:, i], d[:, i + 1])
    return d


def vic_full(du, w, gamma, grav1, dt, dx1, x1f, x1v, area1, vol):
    """ImplicitHydroImpl::forward_masked, scheme 9, dir 0, no solids, for
    columns of one block.  du: (ncol, n, 5) explicit increments (interior);
    w: (ncol, n + 2, 5); x1f: the n + 1 interior faces, x1v the n centres.
    Returns the corrected du (ncol, n, 5)."""
    ncol, n, _ = du.shape
    du0 = du.copy()
    du = du.copy()
    A = np.full(n + 1, area1)
    V = vol
    # GravityWork face weights (gravity_work.hpp:245-262)
    work_lo = 0.5 * A[0:n] * (x1v - x1f[0:n]) / V
    work_hi = 0.5 * A[1 : n + 1] * (x1f[1 : n + 1] - x1v) / V
    a, b, c = assemble(w, gamma, A, V, work_lo, work_hi, dt, grav1)
    # couple_radial_exact: rows on the mass column of the energy row
    from .gravity import x1_variance_plain

    gvar = grav1 * x1_variance_plain(x1f)
    rows = pe_rows(x1v, gvar)
    lo, mid, hi, t0, t1 = rows
    a[..., IPR, IDN] = a[..., IPR, IDN] - mid / dt
    b[..., IPR, IDN] = b[..., IPR, IDN] - lo / dt
    c[..., IPR, IDN] = c[..., IPR, IDN] - hi / dt
    mass0 = du0[..., IDN].copy()
    du[..., IPR] = du[..., IPR] - apply_rows(rows, mass0)
    rhs = du / dt
    # the Woodbury wall third points, solved here on the full system
    y = block_solve(a, b, c, rhs)
    e0 = np.zeros_like(rhs)
    e0[:, 0, IPR] = 1.0
    e1 = np.zeros_like(rhs)
    e1[:, n - 1, IPR] = 1.0
    z0 = block_solve(a, b, c, e0)
    z1 = block_solve(a, b, c, e1)
    k0 = -t0 / dt
    k1 = -t1 / dt
    i0, i1 = 2, n - 3
    m00 = 1.0 + k0 * z0[:, i0, 0]
    m01 = k0 * z1[:, i0, 0]
    m10 = k1 * z0[:, i1, 0]
    m11 = 1.0 + k1 * z1[:, i1, 0]
    r0 = k0 * y[:, i0, 0]
    r1 = k1 * y[:, i1, 0]
    det = m00 * m11 - m01 * m10
    al0 = (m11 * r0 - m01 * r1) / det
    al1 = (m00 * r1 - m10 * r0) / det
    delta = y - z0 * al0[:, None, None]
    delta = delta - z1 * al1[:, None, None]
    # redistribution (vic_constituent_column, dry)
    phi = (delta[..., 0] - du[..., IDN]) * V
    R = phi.sum(-1)
    S = (w[:, 1 : n + 1, IDN] * V).sum(-1)
    M = np.zeros((ncol, n + 1))
    acc = np.zeros(ncol)
    for i in range(n):
        M[:, i] = acc
        acc = acc - (phi[:, i] - R * (w[:, 1 + i, IDN] * V / S))
    mass_idn = np.zeros((ncol, n))
    moved = np.zeros((ncol, n + 1))
    avail = (w[:, 1, IDN] * 1.0 + du[:, 0, IDN]) * V
    avail = np.where(avail < 0, 0.0, avail)
    for i in range(n - 1):
        avail_up = (w[:, 2 + i, IDN] * 1.0 + du[:, i + 1, IDN]) * V
        avail_up = np.where(avail_up < 0, 0.0, avail_up)
        mf = M[:, i + 1]
        q = np.where(mf > 0, mf * 1.0, mf * 1.0)
        q = np.where(q > avail, avail, q)
        q = np.where(q < -avail_up, -avail_up, q)
        mass_idn[:, i] = mass_idn[:, i] - q / V
        mass_idn[:, i + 1] = mass_idn[:, i + 1] + q / V
        moved[:, i + 1] = q
        avail = avail_up + q
    out = du.copy()
    out[..., IDN] = du[..., IDN] + mass_idn
    out[..., IVX] = delta[..., 1]
    out[..., IVY] = delta[..., 2]
    out[..., IVZ] = delta[..., 3]
    out[..., IPR] = delta[..., 4]
    # book_remainder_ (implicit_hydro.cpp:713-794)
    requested = M
    mv = moved - requested
    projected = (requested[:, 0:n] - requested[:, 1 : n + 1]) / V
    phic = -grav1 * x1v
    projection_work = phic * (delta[..., 0] - du0[..., IDN] - projected)
    dp_lo = -grav1 * (x1f[0:n] - x1v)
    dp_hi = -grav1 * (x1f[1 : n + 1] - x1v)
    clamp_work = -(dp_hi * mv[:, 1 : n + 1] - dp_lo * mv[:, 0:n]) / V
    out[..., IPR] = out[..., IPR] + (projection_work + clamp_work)
    moved_rho = out[..., IDN] - du0[..., IDN]
    from .gravity import centroid_slope

    work = gvar * centroid_slope(moved_rho, x1v)
    work = work - apply_rows(rows, delta[..., 0] - mass0)
    out[..., IPR] = out[..., IPR] + work
    return out
