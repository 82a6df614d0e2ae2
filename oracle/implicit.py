"""CPU FP64 oracle of the vertical implicit correction on a Cartesian
single-block column with x1 gravity: the full 5x5 matrix (implicit-scheme
9, vic_full) and the partial 3x3 matrix with the tangential-momentum fold
(implicit-scheme 1, vic_partial).

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
zero du out), and the small-dt limit delta -> du.  The assembly pieces are
pinned by tests/test_oracle_vic_assembly.py from the mathematics of the
Euler equations: the flux Jacobian against finite differences of the flux,
the Roe eigen-system (R R^-1 = I, R diag(lambda) R^-1 = A, Roe's property),
|A| (|A|^2 = A^2, |A| r_k = |lambda_k| r_k), the assembled a, b, c as the
exact linearisation of the Roe-type residual with the gravity source, the
face gravity work and the wall mirrors, and the corrected-PE rows as the
slope operator.
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


def vic_full(du, w, gamma, grav1, dt, dx1, x1f, x1v, area1, vol, return_moved=False):
    """ImplicitHydroImpl::forward_masked, scheme 9, dir 0, no solids, for
    columns of one block.  du: (ncol, n, 5) explicit increments (interior);
    w: (ncol, n + 2, 5); x1f: the n + 1 interior faces, x1v the n centres.
    Returns the corrected du (ncol, n, 5); with return_moved also the
    clamped dry-gas transfer through the face below each cell (ncol, n + 1),
    the mass_fix row the tracer transfer reads (vic_redistribute_impl.h:
    80-90, 139-141)."""
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
    if return_moved:
        return out, moved
    return out


# ---------------------------------------------------------------------------
# implicit-scheme 1: the partial (3x3) matrix
# ---------------------------------------------------------------------------
ROWS3 = (IDN, IVX, IPR)  # rho, m1, E
TAN = (IVY, IVZ)  # the tangential momenta, dir 0


def _take(m, rows, cols):
    return m[..., rows, :][..., :, cols]


def assemble_partial(w, du, gamma, area, vol, work_lo, work_hi, dt, grav):
    """vic_assemble_impl<3> with the m_t fold (vic_assemble_impl.h:45-218):
    the 3x3 blocks (rows and columns rho, m1, E of the same 5x5 terms), the
    gravity-work rows (e = 2, m = 1), the wall merges, then the tangential
    momenta written dm_t,j = v_t,j drho_j + q_t,j: each block's rho column
    gains its m_t columns times the cell's velocity, and
    rhs_corr_i = aT q_i + bT q_{i-1} + cT q_{i+1} is returned for the sweep
    to subtract.  Columns of one block, no solids, both ends walls."""
    ncol, nl2, _ = w.shape
    n = nl2 - 2
    gm1 = gamma - 1.0
    wm, wc, wp = w[:, 0:n], w[:, 1 : n + 1], w[:, 2 : n + 2]
    dprev = flux_jacobian(gm1, wm)
    dcurr = flux_jacobian(gm1, wc)
    dnext = flux_jacobian(gm1, wp)
    Am = roe_abs_matrix(wm, wc, gamma)
    Ap = roe_abs_matrix(wc, wp, gamma)
    R3, T2 = list(ROWS3), list(TAN)
    AmN, ApN = _take(Am, R3, R3), _take(Ap, R3, R3)
    dpN, dcN, dnN = _take(dprev, R3, R3), _take(dcurr, R3, R3), _take(dnext, R3, R3)
    Phi = np.zeros((3, 3))
    Phi[1, 0] = grav
    Phi[2, 1] = grav
    ai, aip = area[0:n][:, None, None], area[1 : n + 1][:, None, None]
    hiv = 0.5 / vol
    a = (AmN * ai + ApN * aip + (aip - ai) * dcN) * hiv - Phi
    a = a + np.eye(3) * (1.0 / dt)
    b = -(AmN + dpN) * ai * hiv
    c = -(ApN - dnN) * aip * hiv
    am, ap = AmN[..., 0, :], ApN[..., 0, :]
    lo, up = work_lo[None, :, None], work_hi[None, :, None]
    a[..., 2, :] = a[..., 2, :] - grav * (up * ap - lo * am)
    b[..., 2, :] = b[..., 2, :] - grav * lo * am
    c[..., 2, :] = c[..., 2, :] + grav * up * ap
    a[..., 2, 1] = a[..., 2, 1] + grav * (1.0 - work_lo - work_hi)
    b[..., 2, 1] = b[..., 2, 1] - grav * work_lo
    c[..., 2, 1] = c[..., 2, 1] - grav * work_hi
    bnd = np.array([1.0, -1.0, 1.0])
    a[:, 0] = a[:, 0] + b[:, 0] * bnd
    a[:, n - 1] = a[:, n - 1] + c[:, n - 1] * bnd
    # the tangential columns of the same rows
    AmT, ApT = _take(Am, R3, T2), _take(Ap, R3, T2)
    dpT, dcT, dnT = _take(dprev, R3, T2), _take(dcurr, R3, T2), _take(dnext, R3, T2)
    aT = (AmT * ai + ApT * aip + (aip - ai) * dcT) * hiv
    bT = -(AmT + dpT) * ai * hiv
    cT = -(ApT - dnT) * aip * hiv
    aT[..., 2, :] = aT[..., 2, :] - grav * (up * ApT[..., 0, :] - lo * AmT[..., 0, :])
    bT[..., 2, :] = bT[..., 2, :] - grav * lo * AmT[..., 0, :]
    cT[..., 2, :] = cT[..., 2, :] + grav * up * ApT[..., 0, :]
    aT[:, 0] = aT[:, 0] + bT[:, 0]
    aT[:, n - 1] = aT[:, n - 1] + cT[:, n - 1]
    corr = np.zeros((ncol, n, 3))
    drho = du[..., IDN]
    vt = wc[..., T2]  # (ncol, n, 2)
    qt = du[..., T2] - vt * drho[..., None]

    def fold(blk, tb, j, i):
        for k in range(2):
            v = vt[:, j, k]
            q = qt[:, j, k]
            blk[:, i, :, 0] = np.where((v != 0)[:, None], blk[:, i, :, 0] + tb[:, i, :, k] * v[:, None], blk[:, i, :, 0])
            corr[:, i] = np.where((q != 0)[:, None], corr[:, i] + tb[:, i, :, k] * q[:, None], corr[:, i])

    for i in range(n):
        fold(a, aT, i, i)
        if i > 0:
            fold(b, bT, i - 1, i)
        if i < n - 1:
            fold(c, cT, i + 1, i)
    return a, b, c, corr


def _mm3(x, y):
    """3x3 products, each entry ((x0 y0 + x1 y1) + x2 y2)."""
    out = np.empty(np.broadcast_shapes(x.shape, y.shape))
    for r in range(3):
        for c in range(3):
            out[..., r, c] = x[..., r, 0] * y[..., 0, c] + x[..., r, 1] * y[..., 1, c] + x[..., r, 2] * y[..., 2, c]
    return out


def _mv3(x, v):
    return np.stack([x[..., r, 0] * v[..., 0] + x[..., r, 1] * v[..., 1] + x[..., r, 2] * v[..., 2] for r in range(3)],
                    -1)


def inverse3(m):
    """The 3x3 inverse by cofactors of Eigen 3.4 (Eigen/src/LU/
    InverseImpl.h, compute_inverse<..., 3>; the reference's FetchContent
    pin), which ForwardSweep calls for the partial matrix
    (forward_sweep_impl.h:108-121)."""
    g = lambda r, c: m[..., r, c]  # noqa: E731
    cof00 = g(1, 1) * g(2, 2) - g(1, 2) * g(2, 1)
    cof10 = g(2, 1) * g(0, 2) - g(2, 2) * g(0, 1)
    cof20 = g(0, 1) * g(1, 2) - g(0, 2) * g(1, 1)
    det = cof00 * g(0, 0) + cof10 * g(1, 0) + cof20 * g(2, 0)
    inv = 1.0 / det
    out = np.empty_like(m)
    out[..., 1, 0] = (g(1, 2) * g(2, 0) - g(1, 0) * g(2, 2)) * inv
    out[..., 1, 1] = (g(2, 2) * g(0, 0) - g(2, 0) * g(0, 2)) * inv
    out[..., 2, 0] = (g(1, 0) * g(2, 1) - g(1, 1) * g(2, 0)) * inv
    out[..., 1, 2] = (g(0, 2) * g(1, 0) - g(0, 0) * g(1, 2)) * inv
    out[..., 2, 1] = (g(2, 0) * g(0, 1) - g(2, 1) * g(0, 0)) * inv
    out[..., 2, 2] = (g(0, 0) * g(1, 1) - g(0, 1) * g(1, 0)) * inv
    out[..., 0, 0] = cof00 * inv
    out[..., 0, 1] = cof10 * inv
    out[..., 0, 2] = cof20 * inv
    return out


def sweep_partial(a, b, c, rhs, corr):
    """ForwardSweep<2> for N = 3 (forward_sweep_impl.h:31-132: the small
    matrix path, a_i^-1 by cofactors) with the two Woodbury unit right-hand
    sides (the E row at the first and last cell), then
    vic_backward_substitute.  Returns (delta, z0, z1)."""
    ncol, n = a.shape[:2]
    ap = np.empty_like(a)
    d = np.empty((ncol, n, 3))
    z0 = np.empty((ncol, n, 3))
    z1 = np.empty((ncol, n, 3))
    for i in range(n):
        # rhs(row) -= delta(row) only where the fold left a nonzero term
        r = np.where(corr[:, i] != 0, rhs[:, i] - corr[:, i], rhs[:, i])
        e0 = np.zeros((ncol, 3))
        e1 = np.zeros((ncol, 3))
        if i == 0:
            e0[:, 2] = 1.0
        if i == n - 1:
            e1[:, 2] = 1.0
        ai = a[:, i]
        if i > 0:
            ai = ai - _mm3(b[:, i], ap[:, i - 1])
        inv = inverse3(ai)
        if i > 0:
            d[:, i] = _mv3(inv, r - _mv3(b[:, i], d[:, i - 1]))
            z0[:, i] = _mv3(inv, e0 - _mv3(b[:, i], z0[:, i - 1]))
            z1[:, i] = _mv3(inv, e1 - _mv3(b[:, i], z1[:, i - 1]))
        else:
            d[:, i] = _mv3(inv, r)
            z0[:, i] = _mv3(inv, e0)
            z1[:, i] = _mv3(inv, e1)
        ap[:, i] = _mm3(inv, c[:, i])
    for i in range(n - 2, -1, -1):
        d[:, i] = d[:, i] - _mv3(ap[:, i], d[:, i + 1])
        z0[:, i] = z0[:, i] - _mv3(ap[:, i], z0[:, i + 1])
        z1[:, i] = z1[:, i] - _mv3(ap[:, i], z1[:, i + 1])
    return d, z0, z1


def tangential_transfer(du, w, moved, V):
    """vic_tangential_momentum_column (vic_redistribute_impl.h:177-205), in
    place on du: the total mass moved through the face above cell i (moved[:,
    i + 1] > 0: from cell i up to i + 1) carries the tangential velocity of
    the cell it leaves."""
    n = du.shape[1]
    for i in range(n - 1):
        q = moved[:, i + 1]
        for m in TAN:
            vd = np.where(q > 0, w[:, 1 + i, m], w[:, 2 + i, m])
            p = q * vd
            nz = p != 0
            du[:, i, m] = np.where(nz, du[:, i, m] - p / V, du[:, i, m])
            du[:, i + 1, m] = np.where(nz, du[:, i + 1, m] + p / V, du[:, i + 1, m])


def vic_partial(du, w, gamma, grav1, dt, dx1, x1f, x1v, area1, vol, return_moved=False):
    """ImplicitHydroImpl::forward_masked, scheme 1 (the partial 3x3 matrix
    with the m_t fold), dir 0, no solids, for the columns of one block.
    Arguments and result as vic_full."""
    ncol, n, _ = du.shape
    du0 = du.copy()
    du = du.copy()
    A = np.full(n + 1, area1)
    V = vol
    work_lo = 0.5 * A[0:n] * (x1v - x1f[0:n]) / V
    work_hi = 0.5 * A[1 : n + 1] * (x1f[1 : n + 1] - x1v) / V
    a, b, c, corr = assemble_partial(w, du, gamma, A, V, work_lo, work_hi, dt, grav1)
    from .gravity import x1_variance_plain

    gvar = grav1 * x1_variance_plain(x1f)
    rows = pe_rows(x1v, gvar)
    lo, mid, hi, t0, t1 = rows
    a[..., 2, 0] = a[..., 2, 0] - mid / dt
    b[..., 2, 0] = b[..., 2, 0] - lo / dt
    c[..., 2, 0] = c[..., 2, 0] - hi / dt
    mass0 = du0[..., IDN].copy()
    du[..., IPR] = du[..., IPR] - apply_rows(rows, mass0)
    rhs = np.stack([du[..., IDN] / dt, du[..., IVX] / dt, du[..., IPR] / dt], -1)
    y, z0, z1 = sweep_partial(a, b, c, rhs, corr)
    # vic_woodbury_column, N = 3
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
    # vic_constituent_column (dry), as vic_full
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
    tangential_transfer(du, w, moved, V)
    out = du.copy()
    out[..., IDN] = du[..., IDN] + mass_idn
    out[..., IVX] = delta[..., 1]
    out[..., IPR] = delta[..., 2]
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
    if return_moved:
        return out, moved
    return out
