"""Grid-only x1 factors of the gravity path, built once per block on the
host in FP64 with the reference's arithmetic and uploaded as one table.

Rows (each nc1 long, interior cells filled, ghosts zero):
  0 x1v
  1 grav1 * (x1f[i+1] - x1v[i])   face-work factor of the upper face
  2 grav1 * (x1f[i]   - x1v[i])   face-work factor of the lower face
    (gravity_work_from_face_mass, src/hydro/gravity_work.hpp:17-29)
  3 grav1 * h^2 / 12              <(x1 - r_c)^2> of a Cartesian cell
    (x1_variance, gravity_work.hpp:54-63; corrected_pe_factors :210-219)
  4-6 the three-point centroid-slope weights of each interior cell
    (centroid_slope_stencil, gravity_work.hpp:88-117)
  7 unused
"""

from __future__ import annotations

import torch


def centroid_slope_weights(x: torch.Tensor):
    n = x.shape[0]
    hm = x[1 : n - 1] - x[0 : n - 2]
    hp = x[2:n] - x[1 : n - 1]
    a = x[1:2] - x[0:1]
    b = x[2:3] - x[1:2]
    c = x[n - 2 : n - 1] - x[n - 3 : n - 2]
    d = x[n - 1 : n] - x[n - 2 : n - 1]
    wt0 = torch.cat([-(2.0 * a + b) / (a * (a + b)), -hp / (hm * (hm + hp)), d / (c * (c + d))])
    wt1 = torch.cat([(a + b) / (a * b), (hp - hm) / (hm * hp), -((c + d) / (c * d))])
    wt2 = torch.cat([-(a / ((a + b) * b)), hm / (hp * (hm + hp)), (c + 2.0 * d) / ((c + d) * d)])
    return wt0, wt1, wt2


def x1_gravity_table(x1f: torch.Tensor, x1v: torch.Tensor, nghost: int, nx1: int, grav1: float) -> torch.Tensor:
    nc1 = x1v.shape[0]
    tab = torch.zeros((8, nc1), dtype=torch.float64)
    is_, ie = nghost, nghost + nx1
    tab[0] = x1v
    centres = x1v[is_:ie]
    tab[1, is_:ie] = grav1 * (x1f[is_ + 1 : ie + 1] - centres)
    tab[2, is_:ie] = grav1 * (x1f[is_:ie] - centres)
    h = x1f[is_ + 1 : ie + 1] - x1f[is_:ie]
    tab[3, is_:ie] = grav1 * (h * h / 12.0)
    if nx1 >= 3:
        wt0, wt1, wt2 = centroid_slope_weights(x1v[is_:ie])
        tab[4, is_:ie] = wt0
        tab[5, is_:ie] = wt1
        tab[6, is_:ie] = wt2
    return tab


def band_rows(x1v_int: torch.Tensor, gvar: torch.Tensor):
    """GravityWork::update_rows (gravity_work.hpp:321-346): the corrected-PE
    operator's bands (lo, mid, hi) and wall third points (t0, t1)."""
    n = x1v_int.shape[0]
    wt0, wt1, wt2 = centroid_slope_weights(x1v_int)
    j = torch.arange(n).clamp(1, n - 2) - 1
    idx = (j, j + 1, j + 2)
    ar = torch.arange(n)

    def band(off):
        at = ar + off
        s = wt0 * (idx[0] == at).to(wt0.dtype)
        s = s + wt1 * (idx[1] == at).to(wt1.dtype)
        s = s + wt2 * (idx[2] == at).to(wt2.dtype)
        return s

    mid = gvar * band(0)
    lo = torch.zeros(n, dtype=torch.float64)
    hi = torch.zeros(n, dtype=torch.float64)
    lo[1:n] = band(-1)[1:n]
    hi[0 : n - 1] = band(1)[0 : n - 1]
    lo = lo * gvar
    hi = hi * gvar
    t0 = float(gvar[0] * band(2)[0])
    t1 = float(gvar[n - 1] * band(-2)[n - 1])
    return lo, mid, hi, t0, t1


def vic_table(x1f, x1v, nghost, nx1, grav1, area1, vol):
    """FP64 rows (nc1 each, interior filled) of the vertical implicit solve:
    0/1 the face-work weights 0.5 A (x1v - x1f) / V and 0.5 A (x1f - x1v) / V
    (GravityWork::update, gravity_work.hpp:245-262); 2/3/4 the corrected-PE
    bands; 5 the cell potential -grav1 x1v; 6/7 the face-minus-cell
    potentials.  Returns (table, t0, t1)."""
    nc1 = x1v.shape[0]
    tab = torch.zeros((8, nc1), dtype=torch.float64)
    is_, ie = nghost, nghost + nx1
    zc = x1v[is_:ie]
    tab[0, is_:ie] = 0.5 * area1 * (zc - x1f[is_:ie]) / vol
    tab[1, is_:ie] = 0.5 * area1 * (x1f[is_ + 1 : ie + 1] - zc) / vol
    h = x1f[is_ + 1 : ie + 1] - x1f[is_:ie]
    gvar = grav1 * (h * h / 12.0)
    lo, mid, hi, t0, t1 = band_rows(zc, gvar)
    tab[2, is_:ie] = lo
    tab[3, is_:ie] = mid
    tab[4, is_:ie] = hi
    tab[5, is_:ie] = -grav1 * zc
    tab[6, is_:ie] = -grav1 * (x1f[is_:ie] - zc)
    tab[7, is_:ie] = -grav1 * (x1f[is_ + 1 : ie + 1] - zc)
    return tab, t0, t1
