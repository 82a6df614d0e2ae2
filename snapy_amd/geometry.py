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
