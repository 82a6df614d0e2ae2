// The vertical implicit correction (VIC), full 5x5 matrix (implicit-scheme
// 9), in FP64 on gfx950: one thread owns one x1 column and runs the whole
// column program -- block assembly, the corrected-PE coupling, the
// block-tridiagonal forward sweep with two Woodbury right-hand sides, back
// substitution, the 2x2 Woodbury correction, the mass redistribution, the
// post-solve remainder -- and writes the RK-combined state.  The factored
// blocks and partial solutions live in a column-interleaved FP64 workspace
// (layer-major, columns fastest, so a wavefront's accesses are coalesced).
//
// Restates ImplicitHydroImpl::forward_masked (src/implicit/
// implicit_hydro.cpp:301-623) for a Cartesian block owning its whole x1
// column, no solids, dry ideal gas, no FluxRegister posts.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

constexpr int NW = 40;  // workspace doubles per layer: A' 25, d 5, z0 5, z1 5

// v-table rows (nc1 each)
enum { V_WLO = 0, V_WHI, V_KLO, V_KMID, V_KHI, V_PHI, V_DPLO, V_DPHI };
// x1 table rows used here (see gravity.hip)
enum { T_GVAR = 3, T_WT0 = 4, T_WT1 = 5, T_WT2 = 6 };

struct VicArgs {
  Range3 r;
  int n;  // layers (nx1)
  double gamma, grav, area, vol, t0, t1;
  double const* dt_dev;
  double dt_host;  // used when dt_dev is NULL
  double wstage;   // the stage weight inside the solve: dt_c = wstage * dt
  double w0, w1, w2;
  // NULL, or (nc3, nc2, nc1) FP64: the clamped dry-gas transfer through the
  // face below each interior cell (0 in a failed column)
  double* moved;
  // the workspace records hold the solved y, z0, z1 of every layer (the
  // cyclic-reduction solve) rather than the forward sweep's rows
  int solved;
};

using M5 = double[5][5];

// FluxJacobian (flux_decomposition_impl.h:105-126), dir 0
__device__ __forceinline__ void flux_jacobian(M5 f, double gm1, double const* w) {
  double v1 = w[IVX], v2 = w[IVY], v3 = w[IVZ], rho = w[IDN], pres = w[IPR];
  double s2 = v1 * v1 + v2 * v2 + v3 * v3;
  double c1 = ((gm1 - 1) * s2 / 2 - (gm1 + 1) / gm1 * pres / rho) * v1;
  double c2 = (gm1 + 1) / gm1 * pres / rho + s2 / 2 - gm1 * v1 * v1;
  f[0][0] = 0.; f[0][1] = 1.; f[0][2] = 0.; f[0][3] = 0.; f[0][4] = 0.;
  f[1][0] = gm1 * s2 / 2 - v1 * v1; f[1][1] = (2. - gm1) * v1;
  f[1][2] = -gm1 * v2; f[1][3] = -gm1 * v3; f[1][4] = gm1;
  f[2][0] = -v1 * v2; f[2][1] = v2; f[2][2] = v1; f[2][3] = 0.; f[2][4] = 0.;
  f[3][0] = -v1 * v3; f[3][1] = v3; f[3][2] = 0.; f[3][3] = v1; f[3][4] = 0.;
  f[4][0] = c1; f[4][1] = c2; f[4][2] = -gm1 * v2 * v1;
  f[4][3] = -gm1 * v3 * v1; f[4][4] = (gm1 + 1) * v1;
}

// |A| = R Lambda R^-1 at the Roe average of wl, wr (RoeAverage, SoundSpeed,
// Eigenvalue, Eigenvector: flux_decomposition_impl.h:13-103)
__device__ __forceinline__ void roe_abs(M5 out, double gm1, double const* wl,
                                        double const* wr) {
  double sdl = sqrt(wl[IDN]), sdr = sqrt(wr[IDN]);
  double is = 1.0 / (sdl + sdr);
  double pr[5];
  pr[IDN] = sdl * sdr;
  for (int v = IVX; v <= IVZ; ++v) pr[v] = (sdl * wl[v] + sdr * wr[v]) * is;
  double el = wl[IPR] / gm1 +
              0.5 * wl[IDN] * (wl[IVX] * wl[IVX] + wl[IVY] * wl[IVY] + wl[IVZ] * wl[IVZ]);
  double er = wr[IPR] / gm1 +
              0.5 * wr[IDN] * (wr[IVX] * wr[IVX] + wr[IVY] * wr[IVY] + wr[IVZ] * wr[IVZ]);
  double hbar = ((el + wl[IPR]) / sdl + (er + wr[IPR]) / sdr) * is;
  pr[IPR] = (hbar - 0.5 * (pr[IVX] * pr[IVX] + pr[IVY] * pr[IVY] + pr[IVZ] * pr[IVZ])) *
            gm1 / (gm1 + 1.) * pr[IDN];
  double cs = sqrt(pr[IPR] * (gm1 + 1.) / pr[IDN]);
  double r = pr[IDN], u = pr[IVX], v = pr[IVY], w = pr[IVZ], p = pr[IPR];
  double ke = 0.5 * (u * u + v * v + w * w);
  double hp = (gm1 + 1.) / gm1 * p / r;
  double h = hp + ke;
  double lam[5] = {fabs(u - cs), fabs(u), fabs(u + cs), fabs(u), fabs(u)};
  double R[5][5] = {{1., 1., 1., 0., 0.},
                    {u - cs, u, u + cs, 0., 0.},
                    {v, v, v, 1., 0.},
                    {w, w, w, 0., 1.},
                    {h - u * cs, ke, h + u * cs, v, w}};
  double Ri[5][5] = {
      {(cs * ke + u * hp) / (2. * cs * hp), (-hp - cs * u) / (2. * cs * hp),
       -v / (2. * hp), -w / (2. * hp), 1. / (2. * hp)},
      {(hp - ke) / hp, u / hp, v / hp, w / hp, -1. / hp},
      {(cs * ke - u * hp) / (2. * cs * hp), (hp - cs * u) / (2. * cs * hp),
       -v / (2. * hp), -w / (2. * hp), 1. / (2. * hp)},
      {-v, 0., 1., 0., 0.},
      {-w, 0., 0., 1., 0.}};
  for (int a = 0; a < 5; ++a)
    for (int b = 0; b < 5; ++b) {
      double s = 0.;
      for (int k = 0; k < 5; ++k) s += (R[a][k] * lam[k]) * Ri[k][b];
      out[a][b] = s;
    }
}

// Gaussian elimination of the augmented block [A | B] (NRHS columns of B)
// with scaled partial pivoting: each row is measured against its largest
// original entry, the pivot of step k is the row i >= k of largest
// |a_ik| / s_i, the rows below are eliminated, then x is back-substituted
// (the textbook algorithm, e.g. Golub & Van Loan, Matrix Computations,
// sec. 3.4).  A pivot whose scaled size is within 8 N eps of zero, or any
// non-finite value, fails the solve (the reference's singularity threshold).
// This is the serial cross-check of the lane-parallel elimination in
// vic_forward_kernel; the solution is returned in B.
template <int NRHS>
__device__ __forceinline__ bool lu_solve(M5 a, double (*b)[NRHS]) {
  constexpr int N = 5;
  double const eps_n = 8. * N * 2.220446049250313e-16;
  double rs[N];  // the rows' largest original magnitudes, swapped with them
  for (int r = 0; r < N; ++r) {
    double m = 0.;
    for (int c = 0; c < N; ++c) m = fmax(m, fabs(a[r][c]));
    if (!(m > 0.) || !isfinite(m)) return false;
    rs[r] = m;
  }
  for (int k = 0; k < N; ++k) {
    int p = k;
    double best = fabs(a[k][k]) / rs[k];
    for (int r = k + 1; r < N; ++r) {
      double q = fabs(a[r][k]) / rs[r];
      if (q > best) {
        best = q;
        p = r;
      }
    }
    if (!(best > eps_n) || !isfinite(best)) return false;
    if (p != k) {
      for (int c = 0; c < N; ++c) {
        double t = a[k][c];
        a[k][c] = a[p][c];
        a[p][c] = t;
      }
      for (int c = 0; c < NRHS; ++c) {
        double t = b[k][c];
        b[k][c] = b[p][c];
        b[p][c] = t;
      }
      double t = rs[k];
      rs[k] = rs[p];
      rs[p] = t;
    }
    double inv = 1. / a[k][k];
    for (int r = k + 1; r < N; ++r) {
      double m = a[r][k] * inv;
      for (int c = k + 1; c < N; ++c) a[r][c] -= m * a[k][c];
      for (int c = 0; c < NRHS; ++c) b[r][c] -= m * b[k][c];
    }
  }
  for (int k = N - 1; k >= 0; --k)
    for (int c = 0; c < NRHS; ++c) {
      double x = b[k][c];
      for (int j2 = k + 1; j2 < N; ++j2) x -= a[k][j2] * b[j2][c];
      b[k][c] = x / a[k][k];
      if (!isfinite(b[k][c])) return false;
    }
  return true;
}

__global__ void __launch_bounds__(64)
    vic_full_kernel(double const* __restrict__ w, double const* __restrict__ du,
                    double const* __restrict__ u0, double* __restrict__ u,
                    double const* __restrict__ vtab,
                    double const* __restrict__ xtab, double* __restrict__ ws,
                    int* __restrict__ status, VicArgs a) {
  Range3 const& r = a.r;
  int64_t col = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t ncol = int64_t(r.n2) * r.n3;
  if (col >= ncol) return;
  int j = r.lo2 + int(col % r.n2), k = r.lo3 + int(col / r.n2);
  int64_t row = k * r.s3 + j * r.s2 + r.lo1;  // interior cell 0 of the column
  int n = a.n, nc1 = r.nc1, lo1 = r.lo1;
  double const gm1 = a.gamma - 1.;
  double const grav = a.grav, A = a.area, V = a.vol;
  double const dt = dt_at(a.dt_dev, a.dt_host, r, k) * a.wstage;
  double const hiv = 0.5 / V;
  auto W = [&](int v, int i) { return w[v * r.sv + row + i]; };      // i in [-1, n]
  auto DU = [&](int v, int i) { return du[v * r.sv + row + i]; };    // i in [0, n)
  auto WS = [&](int l, int e) -> double& { return ws[(int64_t(l) * NW + e) * ncol + col]; };
  auto tab = [&](double const* t, int rowi, int i) { return t[rowi * nc1 + lo1 + i]; };
  // the corrected-PE rows applied to the explicit mass column at cell i
  auto rows_at = [&](int i, auto q) {
    double s = tab(vtab, V_KMID, i) * q(i);
    if (i > 0) s = s + tab(vtab, V_KLO, i) * q(i - 1);
    if (i < n - 1) s = s + tab(vtab, V_KHI, i) * q(i + 1);
    if (i == 0) s = s + a.t0 * q(2);
    if (i == n - 1) s = s + a.t1 * q(n - 3);
    return s;
  };
  auto mass0 = [&](int i) { return DU(IDN, i); };

  // ---- forward sweep ------------------------------------------------------
  double ap[5][5], dp[5], z0p[5], z1p[5];  // the previous layer's solution
  bool ok = true;
  for (int i = 0; i < n && ok; ++i) {
    double wl[5], wc[5], wr[5];
    for (int v = 0; v < 5; ++v) {
      wl[v] = W(v, i - 1);
      wc[v] = W(v, i);
      wr[v] = W(v, i + 1);
    }
    double am[5][5], apm[5][5], dc[5][5];
    roe_abs(am, 0.5 * (a.gamma + a.gamma) - 1., wl, wc);
    roe_abs(apm, 0.5 * (a.gamma + a.gamma) - 1., wc, wr);
    double bb[5][5], cc[5][5], aa[5][5];
    {
      double dprev[5][5];
      flux_jacobian(dprev, gm1, wl);
      for (int p = 0; p < 5; ++p)
        for (int q = 0; q < 5; ++q) bb[p][q] = -(am[p][q] + dprev[p][q]) * A * hiv;
    }
    {
      double dnext[5][5];
      flux_jacobian(dnext, gm1, wr);
      for (int p = 0; p < 5; ++p)
        for (int q = 0; q < 5; ++q) cc[p][q] = -(apm[p][q] - dnext[p][q]) * A * hiv;
    }
    flux_jacobian(dc, gm1, wc);
    for (int p = 0; p < 5; ++p)
      for (int q = 0; q < 5; ++q)
        aa[p][q] = (am[p][q] * A + apm[p][q] * A + (A - A) * dc[p][q]) * hiv;
    aa[IVX][IDN] -= grav;
    aa[IPR][IVX] -= grav;
    for (int p = 0; p < 5; ++p) aa[p][p] += 1. / dt;
    // gravity-work rows (vic_gravity_work_impl.h:24-36)
    double lo = tab(vtab, V_WLO, i), up = tab(vtab, V_WHI, i);
    for (int q = 0; q < 5; ++q) {
      aa[IPR][q] -= grav * (up * apm[0][q] - lo * am[0][q]);
      bb[IPR][q] -= grav * lo * am[0][q];
      cc[IPR][q] += grav * up * apm[0][q];
    }
    aa[IPR][IVX] += grav * (1. - lo - up);
    bb[IPR][IVX] -= grav * lo;
    cc[IPR][IVX] -= grav * up;
    // the wall mirror merges the boundary face: Bnd = (1, -1, 1, 1, 1)
    if (i == 0)
      for (int p = 0; p < 5; ++p)
        for (int q = 0; q < 5; ++q) aa[p][q] += bb[p][q] * (q == IVX ? -1. : 1.);
    if (i == n - 1)
      for (int p = 0; p < 5; ++p)
        for (int q = 0; q < 5; ++q) aa[p][q] += cc[p][q] * (q == IVX ? -1. : 1.);
    // couple_radial_exact (implicit_hydro.cpp:408-446)
    aa[IPR][IDN] -= tab(vtab, V_KMID, i) / dt;
    bb[IPR][IDN] -= tab(vtab, V_KLO, i) / dt;
    cc[IPR][IDN] -= tab(vtab, V_KHI, i) / dt;
    // right-hand sides: du/dt (energy without the explicit corrected PE),
    // and the unit E rows at the two wall cells
    double rhs[5];
    rhs[0] = DU(IDN, i) / dt;
    rhs[1] = DU(IVX, i) / dt;
    rhs[2] = DU(IVY, i) / dt;
    rhs[3] = DU(IVZ, i) / dt;
    rhs[4] = (DU(IPR, i) - rows_at(i, mass0)) / dt;
    double sol[5][8];
    for (int p = 0; p < 5; ++p) {
      for (int q = 0; q < 5; ++q) sol[p][q] = cc[p][q];
      sol[p][5] = rhs[p];
      sol[p][6] = (i == 0 && p == 4) ? 1. : 0.;
      sol[p][7] = (i == n - 1 && p == 4) ? 1. : 0.;
    }
    if (i > 0) {
      // a_i -= b_i A'_{i-1}; rhs -= b_i d_{i-1} (and the Woodbury columns)
      for (int p = 0; p < 5; ++p) {
        for (int q = 0; q < 5; ++q) {
          double s = 0.;
          for (int m = 0; m < 5; ++m) s += bb[p][m] * ap[m][q];
          aa[p][q] -= s;
        }
        double s5 = 0., s6 = 0., s7 = 0.;
        for (int m = 0; m < 5; ++m) {
          s5 += bb[p][m] * dp[m];
          s6 += bb[p][m] * z0p[m];
          s7 += bb[p][m] * z1p[m];
        }
        sol[p][5] -= s5;
        sol[p][6] -= s6;
        sol[p][7] -= s7;
      }
    }
    if (!lu_solve<8>(aa, sol)) {
      ok = false;
      break;
    }
    for (int p = 0; p < 5; ++p) {
      for (int q = 0; q < 5; ++q) {
        ap[p][q] = sol[p][q];
        WS(i, p * 5 + q) = sol[p][q];
      }
      dp[p] = sol[p][5];
      z0p[p] = sol[p][6];
      z1p[p] = sol[p][7];
      WS(i, 25 + p) = dp[p];
      WS(i, 30 + p) = z0p[p];
      WS(i, 35 + p) = z1p[p];
      if (!isfinite(dp[p])) ok = false;
    }
  }
  if (!ok) {
    // a failed column keeps its explicit increments (the reference undoes
    // the column with its status bits and flags a retry)
    vic_mark(status, col, false);
    for (int i = 0; i < n; ++i) {
      for (int v = 0; v < 5; ++v) {
        int64_t o = v * r.sv + row + i;
        u[o] = a.w0 * u0[o] + a.w1 * u[o] + a.w2 * du[o];
      }
      if (a.moved) a.moved[row + i] = 0.;
    }
    return;
  }
  vic_mark(status, col, true);
  // ---- back substitution (vic_backward_substitute) ------------------------
  for (int i = n - 2; i >= 0; --i)
    for (int p = 0; p < 5; ++p) {
      double sd = 0., s0 = 0., s1 = 0.;
      for (int q = 0; q < 5; ++q) {
        double aq = WS(i, p * 5 + q);
        sd += aq * WS(i + 1, 25 + q);
        s0 += aq * WS(i + 1, 30 + q);
        s1 += aq * WS(i + 1, 35 + q);
      }
      WS(i, 25 + p) -= sd;
      WS(i, 30 + p) -= s0;
      WS(i, 35 + p) -= s1;
    }
  // ---- Woodbury 2x2 (vic_woodbury_column) ---------------------------------
  {
    double k0 = -a.t0 / dt, k1 = -a.t1 / dt;
    int i0 = 2, i1 = n - 3;
    double m00 = 1. + k0 * WS(i0, 30), m01 = k0 * WS(i0, 35);
    double m10 = k1 * WS(i1, 30), m11 = 1. + k1 * WS(i1, 35);
    double r0 = k0 * WS(i0, 25), r1 = k1 * WS(i1, 25);
    double det = m00 * m11 - m01 * m10;
    double al0 = (m11 * r0 - m01 * r1) / det, al1 = (m00 * r1 - m10 * r0) / det;
    for (int i = 0; i < n; ++i)
      for (int p = 0; p < 5; ++p) {
        double y = WS(i, 25 + p);
        y -= WS(i, 30 + p) * al0;
        y -= WS(i, 35 + p) * al1;
        WS(i, 25 + p) = y;
      }
  }
  // ---- mass redistribution (vic_constituent_column, dry) ------------------
  // workspace reuse per layer: [0] phi, [1] M (face below), [2] mass change,
  // [3] moved q (face below)
  double R = 0., S = 0.;
  for (int i = 0; i < n; ++i) {
    double phi = (WS(i, 25) - DU(IDN, i)) * V;
    WS(i, 0) = phi;
    R += phi;
    S += W(IDN, i) * V;
  }
  {
    double M = 0.;
    for (int i = 0; i < n; ++i) {
      WS(i, 1) = M;
      M -= WS(i, 0) - R * (W(IDN, i) * V / S);
      WS(i, 2) = 0.;
      WS(i, 3) = 0.;
    }
  }
  {
    double avail = (W(IDN, 0) * 1. + DU(IDN, 0)) * V;
    if (avail < 0) avail = 0;
    for (int i = 0; i + 1 < n; ++i) {
      double avail_up = (W(IDN, i + 1) * 1. + DU(IDN, i + 1)) * V;
      if (avail_up < 0) avail_up = 0;
      double mf = WS(i + 1, 1);
      double q = mf > 0 ? mf * 1. : mf * 1.;
      if (q > avail) q = avail;
      if (q < -avail_up) q = -avail_up;
      WS(i, 2) -= q / V;
      WS(i + 1, 2) += q / V;
      WS(i + 1, 3) = q;
      avail = avail_up + q;
    }
  }
  // ---- outputs, remainder (book_remainder_) and the RK combine ------------
  auto moved_rho = [&](int i) { return (DU(IDN, i) + WS(i, 2)) - DU(IDN, i); };
  auto raw_minus_mass0 = [&](int i) { return WS(i, 25) - DU(IDN, i); };
  for (int i = 0; i < n; ++i) {
    if (a.moved) a.moved[row + i] = WS(i, 3);
    double out[5];
    out[IDN] = DU(IDN, i) + WS(i, 2);
    out[IVX] = WS(i, 26);
    out[IVY] = WS(i, 27);
    out[IVZ] = WS(i, 28);
    out[IPR] = WS(i, 29);
    double req_lo = WS(i, 1), req_hi = i + 1 < n ? WS(i + 1, 1) : 0.;
    double mv_lo = WS(i, 3) - req_lo;
    double mv_hi = (i + 1 < n ? WS(i + 1, 3) : 0.) - req_hi;
    double projected = (req_lo - req_hi) / V;
    double pw = tab(vtab, V_PHI, i) * (WS(i, 25) - DU(IDN, i) - projected);
    double cw = -(tab(vtab, V_DPHI, i) * mv_hi - tab(vtab, V_DPLO, i) * mv_lo) / V;
    out[IPR] = out[IPR] + (pw + cw);
    // corrected-PE work of the moved mass minus the operator's part
    int jj = (i < 1 ? 1 : (i > n - 2 ? n - 2 : i)) - 1;
    double s = tab(xtab, T_WT0, i) * moved_rho(jj);
    s = s + tab(xtab, T_WT1, i) * moved_rho(jj + 1);
    s = s + tab(xtab, T_WT2, i) * moved_rho(jj + 2);
    double work = tab(xtab, T_GVAR, i) * s;
    work = work - rows_at(i, raw_minus_mass0);
    out[IPR] = out[IPR] + work;
    for (int v = 0; v < 5; ++v) {
      int64_t o = v * r.sv + row + i;
      u[o] = a.w0 * u0[o] + a.w1 * u[o] + a.w2 * out[v];
    }
  }
}


// ---------------------------------------------------------------------------
// One column per 5 lanes, 12 columns per wavefront (lanes 60..63 idle).
// Lane p owns row p of every 5x5 block and of every right-hand side.  The
// face matrices |A| and the cell flux Jacobians are carried from one layer
// to the next, so each layer builds one Roe average and one Jacobian row.
// The block solve is Gauss-Jordan with the scaled partial pivoting of
// ludcmp; rows move between lanes with ds_bpermute shuffles.
// ---------------------------------------------------------------------------
constexpr int kSlots = 12;  // columns per wavefront

__device__ __forceinline__ double from_lane(double v, int src) { return __shfl(v, src); }

// one x1 cell's Roe-average inputs: w, sqrt(rho), (E + p) / sqrt(rho)
struct RoeCell {
  double w[5];
  double sd, hs;
};

__device__ __forceinline__ RoeCell roe_cell(double const* w, double gm1) {
  RoeCell c;
#pragma unroll
  for (int v = 0; v < 5; ++v) c.w[v] = w[v];
  c.sd = sqrt(w[IDN]);
  double e = w[IPR] / gm1 +
             0.5 * w[IDN] * (w[IVX] * w[IVX] + w[IVY] * w[IVY] + w[IVZ] * w[IVZ]);
  c.hs = (e + w[IPR]) / c.sd;
  return c;
}

// rows p and 0 of |A| = R Lambda R^-1 at the Roe average of cells l, r
// (RoeAverage, SoundSpeed, Eigenvalue, Eigenvector:
// flux_decomposition_impl.h:13-103); the R^-1 entries use the reciprocals
// of 2 cs hp and hp
__device__ __forceinline__ void roe_rows(double* rowp, double* row0, int p,
                                         double gm1, double gp1_gm1,
                                         RoeCell const& l, RoeCell const& r) {
  double is = 1.0 / (l.sd + r.sd);
  double rho = l.sd * r.sd;
  double u = (l.sd * l.w[IVX] + r.sd * r.w[IVX]) * is;
  double v = (l.sd * l.w[IVY] + r.sd * r.w[IVY]) * is;
  double w = (l.sd * l.w[IVZ] + r.sd * r.w[IVZ]) * is;
  double hbar = (l.hs + r.hs) * is;
  double pp = (hbar - 0.5 * (u * u + v * v + w * w)) * gm1 / (gm1 + 1.) * rho;
  double cs = sqrt(pp * (gm1 + 1.) / rho);
  double ke = 0.5 * (u * u + v * v + w * w);
  double hp = gp1_gm1 * pp / rho;
  double h = hp + ke;
  double lam0 = fabs(u - cs), lam1 = fabs(u), lam2 = fabs(u + cs);
  double ihp = 1. / hp, i2hp = 0.5 * ihp, i2c = 1. / (2. * cs * hp);
  double Ri[5][5] = {{(cs * ke + u * hp) * i2c, (-hp - cs * u) * i2c, -v * i2hp, -w * i2hp, i2hp},
                     {(hp - ke) * ihp, u * ihp, v * ihp, w * ihp, -ihp},
                     {(cs * ke - u * hp) * i2c, (hp - cs * u) * i2c, -v * i2hp, -w * i2hp, i2hp},
                     {-v, 0., 1., 0., 0.},
                     {-w, 0., 0., 1., 0.}};
  // row p of R (R = [1 1 1 0 0; u-c u u+c 0 0; v v v 1 0; w w w 0 1;
  // h-uc ke h+uc v w])
  double vt = p == 1 ? u : (p == 2 ? v : w);
  double R0 = p == 0 ? 1. : (p == 4 ? h - u * cs : (p == 1 ? u - cs : vt));
  double R1 = p == 0 ? 1. : (p == 4 ? ke : vt);
  double R2 = p == 0 ? 1. : (p == 4 ? h + u * cs : (p == 1 ? u + cs : vt));
  double R3 = p == 2 ? 1. : (p == 4 ? v : 0.);
  double R4 = p == 3 ? 1. : (p == 4 ? w : 0.);
  double a0 = R0 * lam0, a1 = R1 * lam1, a2 = R2 * lam2, a3 = R3 * lam1, a4 = R4 * lam1;
#pragma unroll
  for (int b = 0; b < 5; ++b) {
    rowp[b] = (((a0 * Ri[0][b] + a1 * Ri[1][b]) + a2 * Ri[2][b]) + a3 * Ri[3][b]) + a4 * Ri[4][b];
    row0[b] = (lam0 * Ri[0][b] + lam1 * Ri[1][b]) + lam2 * Ri[2][b];
  }
}

// row p of FluxJacobian (flux_decomposition_impl.h:105-126), dir 0
__device__ __forceinline__ void jacobian_row(double* f, int p, double gm1,
                                             double gp1_gm1, double const* w) {
  double v1 = w[IVX], v2 = w[IVY], v3 = w[IVZ];
  double s2 = v1 * v1 + v2 * v2 + v3 * v3;
  double hr = gp1_gm1 * w[IPR] / w[IDN];
  double c1 = ((gm1 - 1) * s2 / 2 - hr) * v1;
  double c2 = hr + s2 / 2 - gm1 * v1 * v1;
  bool r1 = p == 1, r2 = p == 2, r3 = p == 3, r4 = p == 4;
  f[0] = r1 ? gm1 * s2 / 2 - v1 * v1 : r2 ? -v1 * v2 : r3 ? -v1 * v3 : r4 ? c1 : 0.;
  f[1] = r1 ? (2. - gm1) * v1 : r2 ? v2 : r3 ? v3 : r4 ? c2 : 1.;
  f[2] = r1 ? -gm1 * v2 : r2 ? v1 : r4 ? -gm1 * v2 * v1 : 0.;
  f[3] = r1 ? -gm1 * v3 : r3 ? v1 : r4 ? -gm1 * v3 * v1 : 0.;
  f[4] = r1 ? gm1 : r4 ? (gm1 + 1) * v1 : 0.;
}

// Gauss-Jordan on the five lanes of a column group (lane base + p holds row
// p of the 5x5 matrix in ar and of the NX right-hand sides in x), scaled
// partial pivoting with ludcmp's pivot choice (the last maximum wins) and
// its singularity test (ok becomes false on a pivot not above tol or not
// finite); X is row p of the solution.  Columns left of a pivot are not
// needed again, and ar and x are consumed.
template <int NX>
__device__ __forceinline__ void gj5(double (&ar)[5], double (&x)[NX], int p, int base, double tol,
                                    bool& ok, double (&X)[NX]) {
  double big = 0.;
#pragma unroll
  for (int q = 0; q < 5; ++q) big = fmax(big, fabs(ar[q]));
  double scale = big > 0. ? 1. / big : 0.;
  unsigned used = 0;
  int piv[5];
#pragma unroll
  for (int jc = 0; jc < 5; ++jc) {
    double bv = (used >> p) & 1 ? -1. : scale * fabs(ar[jc]);
    double best = -1.;
    int bl = 0;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      double c = from_lane(bv, base + m);
      if (c >= best) {
        best = c;
        bl = m;
      }
    }
    if (!(best > tol) || !isfinite(best)) ok = false;
    used |= 1u << bl;
    piv[jc] = bl;
    int src = base + bl;
    double pj = from_lane(ar[jc], src);
    double fct = p == bl ? 0. : ar[jc] * (1. / pj);
#pragma unroll
    for (int q = jc + 1; q < 5; ++q) ar[q] -= fct * from_lane(ar[q], src);
#pragma unroll
    for (int t = 0; t < NX; ++t) x[t] -= fct * from_lane(x[t], src);
  }
  // this lane pivoted the unknown whose column it still holds; lane p takes
  // the solution row of unknown p from its pivot lane
  int myj = 0;
#pragma unroll
  for (int jc = 0; jc < 5; ++jc) myj = piv[jc] == p ? jc : myj;
  double dj = ar[0];
#pragma unroll
  for (int jc = 1; jc < 5; ++jc) dj = myj == jc ? ar[jc] : dj;
  double idj = 1. / dj;
  int srcp = piv[0];
#pragma unroll
  for (int jc = 1; jc < 5; ++jc) srcp = p == jc ? piv[jc] : srcp;
#pragma unroll
  for (int t = 0; t < NX; ++t) X[t] = from_lane(x[t] * idj, base + srcp);
}

// Row p of layer i's block equation of the full-matrix solve (scheme 9):
// the diagonal ar, the lower block br, the upper block crw and the
// right-hand side, from the Roe rows of the faces below (am, am0) and above
// (apm, ap0) and the flux-Jacobian rows of the cells below (fl) and above
// (fr); sd(v, l) is row v of the explicit increment at layer l, tab(t, row,
// l) the column table.  The (A - A) * FluxJacobian(cell) term of the
// diagonal is zero on the uniform Cartesian x1 grid.
template <typename TAB, typename SDF>
__device__ __forceinline__ void vic9_rows(VicArgs const& a, double const* vtab, int i, int p, double A,
                                          double hiv,
                                          double idt, double dt, double const (&am)[5],
                                          double const (&am0)[5], double const (&apm)[5],
                                          double const (&ap0)[5], double const (&fl)[5],
                                          double const (&fr)[5], TAB tab, SDF sd,
                                          double (&ar)[5], double (&br)[5], double (&crw)[5],
                                          double& rhs) {
  int const n = a.n;
  double const grav = a.grav;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    br[q] = -(am[q] + fl[q]) * A * hiv;
    crw[q] = -(apm[q] - fr[q]) * A * hiv;
    ar[q] = (am[q] * A + apm[q] * A) * hiv;
  }
  ar[IDN] -= p == IVX ? grav : 0.;
  ar[IVX] -= p == IPR ? grav : 0.;
#pragma unroll
  for (int q = 0; q < 5; ++q) ar[q] += q == p ? idt : 0.;
  rhs = sd(p, i) * idt;
  if (p == IPR) {
    // gravity-work rows (vic_gravity_work_impl.h:24-36)
    double lo = tab(vtab, V_WLO, i), up = tab(vtab, V_WHI, i);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      ar[q] -= grav * (up * ap0[q] - lo * am0[q]);
      br[q] -= grav * lo * am0[q];
      crw[q] += grav * up * ap0[q];
    }
    ar[IVX] += grav * (1. - lo - up);
    br[IVX] -= grav * lo;
    crw[IVX] -= grav * up;
  }
  // the wall mirror merges the boundary face: Bnd = (1, -1, 1, 1, 1)
  if (i == 0)
#pragma unroll
    for (int q = 0; q < 5; ++q) ar[q] += br[q] * (q == IVX ? -1. : 1.);
  if (i == n - 1)
#pragma unroll
    for (int q = 0; q < 5; ++q) ar[q] += crw[q] * (q == IVX ? -1. : 1.);
  if (p == IPR) {
    // couple_radial_exact (implicit_hydro.cpp:408-446)
    ar[IDN] -= tab(vtab, V_KMID, i) / dt;
    br[IDN] -= tab(vtab, V_KLO, i) / dt;
    crw[IDN] -= tab(vtab, V_KHI, i) / dt;
    double s = tab(vtab, V_KMID, i) * sd(IDN, i);
    if (i > 0) s = s + tab(vtab, V_KLO, i) * sd(IDN, i - 1);
    if (i < n - 1) s = s + tab(vtab, V_KHI, i) * sd(IDN, i + 1);
    if (i == 0) s = s + a.t0 * sd(IDN, 2);
    if (i == n - 1) s = s + a.t1 * sd(IDN, n - 3);
    rhs = (sd(IPR, i) - s) / dt;
  }
}

constexpr int kQ = 8;          // workspace doubles per lane per layer: A' row, d, z0, z1
constexpr int kLB = 8;         // layers per staged block of the forward sweep
constexpr int kWin = kLB + 4;  // staged layers [l0 - 2, l0 + kLB + 2)
constexpr int kOB = 16;        // layers per output block of the back kernel
constexpr int kOWin = kOB + 4;

// The lanes of a wavefront: 12 columns of 5 lanes.  Lanes 60..63 shadow
// lanes 0..3 of the last column and never write; columns past the grid's
// last are computed on the last column and never written.
struct WaveLanes {
  int slot, p, base;
  bool shadow;  // lanes 60..63
  bool live;    // owns a column of the grid
  int64_t col;  // the column computed on
};

__device__ __forceinline__ WaveLanes wave_lanes(int64_t ncol, unsigned bx) {
  WaveLanes L;
  int lane = threadIdx.x;
  L.shadow = lane >= 5 * kSlots;
  L.slot = L.shadow ? kSlots - 1 : lane / 5;
  L.p = L.shadow ? lane - 5 * kSlots : lane - 5 * L.slot;
  L.base = 5 * L.slot;
  int64_t c = int64_t(bx) * kSlots + L.slot;
  L.live = !L.shadow && c < ncol;
  L.col = c < ncol ? c : ncol - 1;
  return L;
}

__device__ __forceinline__ int64_t column_row(Range3 const& r, int64_t col) {
  int j = r.lo2 + int(col % r.n2), k = r.lo3 + int(col / r.n2);
  return k * r.s3 + j * r.s2 + r.lo1;
}

// Coalesced cooperative copy of layers [i0, i0 + CNT) of NV rows (from row
// v0) of a state array into LDS dst[v][slot][CNT] for the wavefront's
// columns; layer indices are clamped to the x1 row's ghost range [-3, n+2].
// All loads are issued before the first LDS store.
template <int NV, int CNT>
__device__ __forceinline__ void stage_layers(double* dst, double const* __restrict__ src,
                                             int64_t sv, int v0, int64_t const* srow,
                                             int i0, int n) {
  constexpr int TOT = NV * kSlots * CNT;
  constexpr int IT = (TOT + 63) / 64;
  double val[IT];
#pragma unroll
  for (int t = 0; t < IT; ++t) {
    int f = threadIdx.x + 64 * t;
    if (f < TOT) {
      int v = f / (kSlots * CNT);
      int h = f - v * (kSlots * CNT);
      int sl = h / CNT;
      int i = i0 + (h - sl * CNT);
      i = i < -3 ? -3 : (i > n + 2 ? n + 2 : i);
      val[t] = src[(v0 + v) * sv + srow[sl] + i];
    }
  }
#pragma unroll
  for (int t = 0; t < IT; ++t) {
    int f = threadIdx.x + 64 * t;
    if (f < TOT) dst[f] = val[t];
  }
}

// The forward sweep of the VIC: block assembly, the corrected-PE coupling,
// Gauss-Jordan on [a | c, rhs, e0, e1] and the forward elimination of the
// block-tridiagonal system.  The column state is staged through LDS in
// blocks of kLB layers; each lane writes its row of [A' | d | z0 | z1] to a
// lane-contiguous workspace ws[blk][layer][q][64].
struct VicForwardBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ du,
                                             double const* __restrict__ vtab,
                                             double* __restrict__ ws, int* __restrict__ status,
                                             VicArgs a) {
    Range3 const& r = a.r;
    int64_t const ncol = int64_t(r.n2) * r.n3;
    WaveLanes const L = wave_lanes(ncol, bid.x);
    int const p = L.p, base = L.base, slot = L.slot;
    __shared__ int64_t s_row[kSlots];
    __shared__ double s_w[5 * kSlots * kWin], s_du[5 * kSlots * kWin];
    if (threadIdx.x < kSlots) {
      int64_t c = int64_t(bid.x) * kSlots + threadIdx.x;
      s_row[threadIdx.x] = column_row(r, c < ncol ? c : ncol - 1);
    }
    int const n = a.n, nc1 = r.nc1, lo1 = r.lo1;
    double const gm1 = a.gamma - 1.;
    double const gp1_gm1 = (gm1 + 1) / gm1;
    double const A = a.area, V = a.vol;
    double const dt = dt_at(a.dt_dev, a.dt_host, r, r.lo3 + int(L.col / r.n2)) * a.wstage;
    double const idt = 1. / dt;
    double const hiv = 0.5 / V;
    double* wsl = ws + int64_t(bid.x) * n * kQ * 64 + threadIdx.x;
    auto WS = [&](int l, int q) -> double& { return wsl[(l * kQ + q) * 64]; };
    auto tab = [&](double const* t, int rowi, int i) { return t[rowi * nc1 + lo1 + i]; };
    int l0 = 0;  // first layer of the staged block
    auto SW = [&](int v, int i) { return s_w[(v * kSlots + slot) * kWin + (i - l0 + 2)]; };
    auto SD = [&](int v, int i) { return s_du[(v * kSlots + slot) * kWin + (i - l0 + 2)]; };
    auto stage = [&](int start) {
      __syncthreads();
      l0 = start;
      stage_layers<5, kWin>(s_w, w, r.sv, 0, s_row, start - 2, n);
      stage_layers<5, kWin>(s_du, du, r.sv, 0, s_row, start - 2, n);
      __syncthreads();
    };

    stage(0);
    // carried state: cell i, the face below it, the Jacobians of cells i-1, i
    double wt[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) wt[v] = SW(v, -1);
    RoeCell cl = roe_cell(wt, gm1);
    double fl[5], fc[5];
    jacobian_row(fl, p, gm1, gp1_gm1, wt);
#pragma unroll
    for (int v = 0; v < 5; ++v) wt[v] = SW(v, 0);
    RoeCell cc = roe_cell(wt, gm1);
    jacobian_row(fc, p, gm1, gp1_gm1, wt);
    double am[5], am0[5];
    roe_rows(am, am0, p, gm1, gp1_gm1, cl, cc);
    double Ap[5] = {0, 0, 0, 0, 0}, dpv = 0., z0v = 0., z1v = 0.;
    bool ok = true;
    double const tol = 8. * 5. * 2.220446049250313e-16;  // ludcmp's singularity test

    for (int i = 0; i < n; ++i) {
      if (i > 0 && i % kLB == 0) stage(i);
#pragma unroll
      for (int v = 0; v < 5; ++v) wt[v] = SW(v, i + 1);
      RoeCell cr = roe_cell(wt, gm1);
      double fr[5], apm[5], ap0[5];
      jacobian_row(fr, p, gm1, gp1_gm1, wt);
      roe_rows(apm, ap0, p, gm1, gp1_gm1, cc, cr);
      double ar[5], br[5], crw[5], rhs;
      vic9_rows(a, vtab, i, p, A, hiv, idt, dt, am, am0, apm, ap0, fl, fr, tab,
                [&](int v, int l) { return SD(v, l); }, ar, br, crw, rhs);
      // right-hand sides: c | du/dt | unit E at the bottom | unit E at the top
      double x[8];
#pragma unroll
      for (int q = 0; q < 5; ++q) x[q] = crw[q];
      x[5] = rhs;
      x[6] = (i == 0 && p == IPR) ? 1. : 0.;
      x[7] = (i == n - 1 && p == IPR) ? 1. : 0.;
      if (i > 0) {
        // a_i -= b_i A'_{i-1}, rhs -= b_i d_{i-1}, e0 -= b_i z0_{i-1}; the
        // top-wall column z1 is zero below the top layer
        double s[5] = {0, 0, 0, 0, 0}, s5 = 0., s6 = 0.;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          double bm = br[m];
#pragma unroll
          for (int q = 0; q < 5; ++q) s[q] += bm * from_lane(Ap[q], base + m);
          s5 += bm * from_lane(dpv, base + m);
          s6 += bm * from_lane(z0v, base + m);
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) ar[q] -= s[q];
        x[5] -= s5;
        x[6] -= s6;
      }
      double X[8];
      gj5(ar, x, p, base, tol, ok, X);
#pragma unroll
      for (int q = 0; q < 5; ++q) Ap[q] = X[q];
      dpv = X[5];
      z0v = X[6];
      z1v = X[7];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (!isfinite(X[t])) ok = false;
#pragma unroll
      for (int q = 0; q < 5; ++q) WS(i, q) = X[q];
      WS(i, 5) = dpv;
      WS(i, 6) = z0v;
      // slide the window up one layer
      cl = cc;
      cc = cr;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        fl[q] = fc[q];
        fc[q] = fr[q];
        am[q] = apm[q];
        am0[q] = ap0[q];
      }
    }
    // a column fails as a whole: its five lanes agree
    int okv = ok ? 1 : 0;
#pragma unroll
    for (int m = 0; m < 5; ++m) okv &= __shfl(okv, base + m);
    if (p == 0 && L.live) vic_mark(status, L.col, okv != 0);
    WS(n - 1, 7) = z1v;  // the top-wall column, zero below the top layer
  }
};

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
    vic_forward_kernel(double const* __restrict__ w, double const* __restrict__ du,
                       double const* __restrict__ vtab, double* __restrict__ ws,
                       int* __restrict__ status, VicArgs a) {
  VicForwardBody::run(block_index(), grid_size(), w, du, vtab, ws, status, a);
}

// The full-matrix solve (scheme 9) for grids of few columns, where the
// forward sweep's chain of layers cannot fill the chip: one workgroup per
// column, a five-lane group per layer (kSlots groups per wavefront; lanes
// 60..63 shadow lanes 0..3 of the last group and never write), the block
// equations of all layers assembled at once (vic9_rows on the same Roe and
// Jacobian rows as the forward sweep) and solved by cyclic reduction with
// the three right-hand sides of the Woodbury form (du / dt, the unit E rows
// at the bottom and the top).  At stride s, an equation l with (l + 1) a
// multiple of s is eliminated when (l + 1) / s is odd: its group inverts
// its diagonal block by gj5 (ludcmp's pivot choice and singularity test)
// into [P | Q | q] = D^-1 [B | C | R], kept in registers and in LDS; the
// equations kept at that stride fold in their neighbours' rows (D -= B Q -
// C P, B <- -B P, C <- -C Q, R -= B q + C q).  Every layer is eliminated at
// one stride; the solution then comes back down the strides, x = q - P
// x_below - Q x_above.  The records of the workspace become the solved y,
// z0, z1 of every layer (A' rows zero) in the forward sweep's layout, and
// vic_back_kernel reads them with VicArgs.solved.  The solution rounds
// differently from the sweep's (the same system, another elimination
// order).  Dynamic LDS: 65 n doubles.
struct VicCrBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ du,
                                             double const* __restrict__ vtab,
                                             double* __restrict__ ws, int* __restrict__ status,
                                             VicArgs a) {
    extern __shared__ double lds[];  // [layer][row][13]: P, Q, q; then x (3 per row)
    __shared__ int s_fail;
    Range3 const& r = a.r;
    int64_t const col = bid.x;
    int const n = a.n, nc1 = r.nc1, lo1 = r.lo1;
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool const shadow = lane >= 5 * kSlots;
    int const slot = shadow ? kSlots - 1 : lane / 5;
    int const p = shadow ? lane - 5 * kSlots : lane - 5 * slot;
    int const base = 5 * slot;
    int const li = wave * kSlots + slot;
    bool const lay = li < n;
    bool const own = lay && !shadow;
    int const i = lay ? li : n - 1;
    if (threadIdx.x == 0) s_fail = 0;
    int64_t const row = column_row(r, col);
    double const gm1 = a.gamma - 1.;
    double const gp1_gm1 = (gm1 + 1) / gm1;
    double const A = a.area, V = a.vol;
    double const dt = dt_at(a.dt_dev, a.dt_host, r, r.lo3 + int(col / r.n2)) * a.wstage;
    double const idt = 1. / dt;
    double const hiv = 0.5 / V;
    auto tab = [&](double const* t, int rowi, int l) { return t[rowi * nc1 + lo1 + l]; };
    auto sd = [&](int v, int l) { return du[v * r.sv + row + l]; };
    // ---- the layer's block equation (row p) --------------------------------
    double wm[5], wc[5], wp[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      wm[v] = w[v * r.sv + row + i - 1];
      wc[v] = w[v * r.sv + row + i];
      wp[v] = w[v * r.sv + row + i + 1];
    }
    RoeCell const cl = roe_cell(wm, gm1), cc = roe_cell(wc, gm1), cr = roe_cell(wp, gm1);
    double fl[5], fr[5], am[5], am0[5], apm[5], ap0[5];
    jacobian_row(fl, p, gm1, gp1_gm1, wm);
    jacobian_row(fr, p, gm1, gp1_gm1, wp);
    roe_rows(am, am0, p, gm1, gp1_gm1, cl, cc);
    roe_rows(apm, ap0, p, gm1, gp1_gm1, cc, cr);
    double D[5], Bm[5], Cm[5], R[3];
    vic9_rows(a, vtab, i, p, A, hiv, idt, dt, am, am0, apm, ap0, fl, fr, tab, sd, D, Bm, Cm, R[0]);
    R[1] = (i == 0 && p == IPR) ? 1. : 0.;
    R[2] = (i == n - 1 && p == IPR) ? 1. : 0.;
    // ---- reduction ------------------------------------------------------------
    bool ok = true;
    double const tol = 8. * 5. * 2.220446049250313e-16;  // ludcmp's singularity test
    double X[13];
#pragma unroll
    for (int t = 0; t < 13; ++t) X[t] = 0.;
    int se = 1;  // the stride this layer is eliminated at
    __syncthreads();
    int smax = 1;
    while (smax < n) smax *= 2;
    for (int s = 1; s <= smax; s *= 2) {
      bool const act = lay && (li + 1) % s == 0;
      bool const elim = act && ((li + 1) / s) % 2 == 1;
      bool const keep = act && !elim;
      if (elim) {
        double x[13];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          x[q] = Bm[q];
          x[5 + q] = Cm[q];
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) x[10 + t] = R[t];
        gj5<13>(D, x, p, base, tol, ok, X);
#pragma unroll
        for (int t = 0; t < 13; ++t)
          if (!isfinite(X[t])) ok = false;
        se = s;
        if (own)
#pragma unroll
          for (int t = 0; t < 13; ++t) lds[(li * 5 + p) * 13 + t] = X[t];
      }
      __syncthreads();
      if (keep) {
        // the neighbour below (li - s >= 0 here) and above (li + s, when a layer)
        double const* Lw = lds + (li - s) * 65;
        double nB[5] = {0., 0., 0., 0., 0.}, nC[5] = {0., 0., 0., 0., 0.};
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          double bm = Bm[m];
#pragma unroll
          for (int q = 0; q < 5; ++q) {
            nB[q] -= bm * Lw[m * 13 + q];
            D[q] -= bm * Lw[m * 13 + 5 + q];
          }
#pragma unroll
          for (int t = 0; t < 3; ++t) R[t] -= bm * Lw[m * 13 + 10 + t];
        }
        if (li + s < n) {
          double const* Uw = lds + (li + s) * 65;
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            double cm = Cm[m];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
              D[q] -= cm * Uw[m * 13 + q];
              nC[q] -= cm * Uw[m * 13 + 5 + q];
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) R[t] -= cm * Uw[m * 13 + 10 + t];
          }
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          Bm[q] = nB[q];
          Cm[q] = nC[q];
        }
      }
    }
    if (!ok) s_fail = 1;
    __syncthreads();
    // ---- the solution, down the strides ----------------------------------------
    // x of layer l, row m, right-hand side t at lds[(l * 5 + m) * 13 + t]
    for (int s = smax; s >= 1; s /= 2) {
      double xr[3] = {0., 0., 0.};
      bool const mine = lay && se == s;
      if (mine) {
#pragma unroll
        for (int t = 0; t < 3; ++t) xr[t] = X[10 + t];
        if (li - s >= 0) {
          double const* Lw = lds + (li - s) * 65;
#pragma unroll
          for (int m = 0; m < 5; ++m)
#pragma unroll
            for (int t = 0; t < 3; ++t) xr[t] -= X[m] * Lw[m * 13 + t];
        }
        if (li + s < n) {
          double const* Uw = lds + (li + s) * 65;
#pragma unroll
          for (int m = 0; m < 5; ++m)
#pragma unroll
            for (int t = 0; t < 3; ++t) xr[t] -= X[5 + m] * Uw[m * 13 + t];
        }
      }
      if (mine && own)
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          if (!isfinite(xr[t])) s_fail = 1;
          lds[(li * 5 + p) * 13 + t] = xr[t];
        }
      __syncthreads();
    }
    // ---- the records: A' rows zero, then y, z0, z1 ----------------------------
    if (own) {
      int64_t const bx = col / kSlots;
      int const ln = 5 * int(col - bx * kSlots) + p;
      double* rec = ws + (bx * n + li) * kQ * 64 + ln;
#pragma unroll
      for (int q = 0; q < 5; ++q) rec[q * 64] = 0.;
#pragma unroll
      for (int t = 0; t < 3; ++t) rec[(5 + t) * 64] = lds[(li * 5 + p) * 13 + t];
    }
    if (threadIdx.x == 0) vic_mark(status, col, s_fail == 0);
  }
};

__global__ void __launch_bounds__(768)
    vic_cr_kernel(double const* __restrict__ w, double const* __restrict__ du,
                  double const* __restrict__ vtab, double* __restrict__ ws,
                  int* __restrict__ status, VicArgs a) {
  VicCrBody::run(block_index(), grid_size(), w, du, vtab, ws, status, a);
}

// The back half of the VIC, with the lanes and workspace of
// vic_forward_kernel: back substitution, the Woodbury correction, the dry
// mass redistribution, the post-solve remainder and the RK combine.  Every
// global access is coalesced: the workspace is lane-contiguous, the column
// state is staged through LDS and the outputs are transposed through LDS
// before the combine.
// PAIR: the state is the FP32 pair (u, c) of the FP32 state form, combined
// in FP64; otherwise FP64 u0, u.
template <bool PAIR>
struct VicBackBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ du,
                                             void const* __restrict__ u0v,
                                             float const* __restrict__ c0, void* __restrict__ uv,
                                             float* __restrict__ cc,
                                             double const* __restrict__ vtab,
                                             double const* __restrict__ xtab,
                                             double* __restrict__ ws,
                                             int const* __restrict__ status, VicArgs a) {
    Range3 const& r = a.r;
    int64_t const ncol = int64_t(r.n2) * r.n3;
    WaveLanes const L = wave_lanes(ncol, bid.x);
    int const p = L.p, base = L.base, slot = L.slot;
    __shared__ int64_t s_row[kSlots];
    __shared__ int s_live[kSlots];
    __shared__ double s_buf[5 * kSlots * kOB];  // staged mass/density, then outputs
    __shared__ double s_mw[kSlots * kOWin];     // staged du row IDN
    if (threadIdx.x < kSlots) {
      int64_t c = int64_t(bid.x) * kSlots + threadIdx.x;
      s_row[threadIdx.x] = column_row(r, c < ncol ? c : ncol - 1);
      s_live[threadIdx.x] = c < ncol;
    }
    __syncthreads();
    int const n = a.n, nc1 = r.nc1, lo1 = r.lo1;
    int64_t const row = s_row[slot];
    double const V = a.vol;
    double const dt = dt_at(a.dt_dev, a.dt_host, r, r.lo3 + int(L.col / r.n2)) * a.wstage;
    double* wsw = ws + int64_t(bid.x) * n * kQ * 64;
    double* wsl = wsw + threadIdx.x;
    double* wsb = wsw + base;  // lane 0 of the column
    auto WS = [&](int l, int q) -> double& { return wsl[(l * kQ + q) * 64]; };
    auto WSB = [&](int l, int q) -> double& { return wsb[(l * kQ + q) * 64]; };
    auto tab = [&](double const* t, int rowi, int i) { return t[rowi * nc1 + lo1 + i]; };
    bool const failed = (status[L.col] & kVicStageFailed) != 0;
    constexpr int B = 4;  // layers per batch of workspace loads

    // ---- back substitution (vic_backward_substitute) ------------------------
    int const i0 = 2, i1 = n - 3;
    double dpv = WS(n - 1, 5), z0v = WS(n - 1, 6), z1v = WS(n - 1, 7);
    double d_i0 = 0., z0_i0 = 0., z1_i0 = 0., d_i1 = 0., z0_i1 = 0., z1_i1 = 0.;
    if (a.solved) {
      // the records are the solution: the Woodbury rows read directly
      d_i0 = WSB(i0, 5);
      z0_i0 = WSB(i0, 6);
      z1_i0 = WSB(i0, 7);
      d_i1 = WSB(i1, 5);
      z0_i1 = WSB(i1, 6);
      z1_i1 = WSB(i1, 7);
    }
    for (int ib = a.solved ? -1 : n - 2; ib >= 0; ib -= B) {
      double Ab[B][5], db[B], zb[B];
#pragma unroll
      for (int t = 0; t < B; ++t) {
        int i = ib - t > 0 ? ib - t : 0;
#pragma unroll
        for (int q = 0; q < 5; ++q) Ab[t][q] = WS(i, q);
        db[t] = WS(i, 5);
        zb[t] = WS(i, 6);
      }
#pragma unroll
      for (int t = 0; t < B; ++t) {
        int i = ib - t;
        if (i >= 0) {
          double sd = 0., s0 = 0., s1 = 0.;
#pragma unroll
          for (int q = 0; q < 5; ++q) {
            double aq = Ab[t][q];
            sd += aq * from_lane(dpv, base + q);
            s0 += aq * from_lane(z0v, base + q);
            s1 += aq * from_lane(z1v, base + q);
          }
          dpv = db[t] - sd;
          z0v = zb[t] - s0;
          z1v = 0. - s1;  // the forward z1 is zero below the top layer
          WS(i, 5) = dpv;
          WS(i, 6) = z0v;
          WS(i, 7) = z1v;
          if (i == i0) {
            d_i0 = from_lane(dpv, base);
            z0_i0 = from_lane(z0v, base);
            z1_i0 = from_lane(z1v, base);
          }
          if (i == i1) {
            d_i1 = from_lane(dpv, base);
            z0_i1 = from_lane(z0v, base);
            z1_i1 = from_lane(z1v, base);
          }
        }
      }
    }
    // ---- Woodbury 2x2 (vic_woodbury_column); every lane solves it -----------
    double k0 = -a.t0 / dt, k1 = -a.t1 / dt;
    double m00 = 1. + k0 * z0_i0, m01 = k0 * z1_i0;
    double m10 = k1 * z0_i1, m11 = 1. + k1 * z1_i1;
    double r0 = k0 * d_i0, r1 = k1 * d_i1;
    double det = m00 * m11 - m01 * m10;
    double al0 = (m11 * r0 - m01 * r1) / det, al1 = (m00 * r1 - m10 * r0) / det;
    // the corrected increment of this lane's row: y = d - z0 al0 - z1 al1
    auto ywb = [&](int l) {
      double y = WS(l, 5);
      y -= WS(l, 6) * al0;
      y -= WS(l, 7) * al1;
      return y;
    };
    // ---- mass redistribution (vic_constituent_column, dry) ------------------
    // lane 0 of each column scans; lane 0's dead A' row keeps per layer
    // [0] M at the face below, [1] the mass change, [2] the moved q below,
    // and its [5] becomes the corrected mass increment
    constexpr int SB = kOB + 1;
    double* s_m0 = s_buf;
    double* s_rho = s_buf + kSlots * SB;
    double R = 0., S = 0.;
    for (int b0 = 0; b0 < n; b0 += kOB) {
      __syncthreads();
      stage_layers<1, SB>(s_m0, du, r.sv, IDN, s_row, b0, n);
      stage_layers<1, SB>(s_rho, w, r.sv, IDN, s_row, b0, n);
      __syncthreads();
      if (p == 0) {
#pragma unroll
        for (int t = 0; t < kOB; ++t)
          if (b0 + t < n) {
            R += (ywb(b0 + t) - s_m0[slot * SB + t]) * V;
            S += s_rho[slot * SB + t] * V;
          }
      }
    }
    double M = 0., avail = 0., carry = 0.;
    for (int b0 = 0; b0 < n; b0 += kOB) {
      __syncthreads();
      stage_layers<1, SB>(s_m0, du, r.sv, IDN, s_row, b0, n);
      stage_layers<1, SB>(s_rho, w, r.sv, IDN, s_row, b0, n);
      __syncthreads();
      if (p == 0) {
        if (b0 == 0) {
          avail = (s_rho[slot * SB] * 1. + s_m0[slot * SB]) * V;
          if (avail < 0) avail = 0;
          WS(0, 2) = 0.;
        }
#pragma unroll 8
        for (int t = 0; t < kOB; ++t) {
          int i = b0 + t;
          if (i < n) {
            double y0 = ywb(i);
            WS(i, 5) = y0;  // read back as the mass row by the other lanes
            WS(i, 0) = M;
            double phi = (y0 - s_m0[slot * SB + t]) * V;
            M -= phi - R * (s_rho[slot * SB + t] * V / S);
            if (i + 1 < n) {
              double avail_up = (s_rho[slot * SB + t + 1] * 1. + s_m0[slot * SB + t + 1]) * V;
              if (avail_up < 0) avail_up = 0;
              double mf = M;
              double q = mf > 0 ? mf * 1. : mf * 1.;
              if (q > avail) q = avail;
              if (q < -avail_up) q = -avail_up;
              WS(i, 1) = carry - q / V;
              carry = 0. + q / V;
              WS(i + 1, 2) = q;
              avail = avail_up + q;
            } else {
              WS(i, 1) = carry;
            }
          }
        }
      }
    }
    // lane 0's rows become visible to the column's other lanes
    __threadfence();
    // ---- outputs, the remainder (book_remainder_), the RK combine -----------
    double* s_out = s_buf;
    for (int b0 = 0; b0 < n; b0 += kOB) {
      __syncthreads();
      stage_layers<1, kOWin>(s_mw, du, r.sv, IDN, s_row, b0 - 2, n);
      __syncthreads();
      auto M0 = [&](int i) { return s_mw[slot * kOWin + (i - b0 + 2)]; };
      auto moved_rho = [&](int i) { return (M0(i) + WSB(i, 1)) - M0(i); };
      auto raw_minus_mass0 = [&](int i) { return WSB(i, 5) - M0(i); };
#pragma unroll 8
      for (int t = 0; t < kOB; ++t) {
        int i = b0 + t;
        double out = 0.;
        if (i < n) {
          if (p == IDN) {
            out = M0(i) + WSB(i, 1);
            if (a.moved && L.live) a.moved[row + i] = failed ? 0. : WSB(i, 2);
          } else if (p == IPR) {
            out = ywb(i);
            double req_lo = WSB(i, 0), req_hi = i + 1 < n ? WSB(i + 1, 0) : 0.;
            double mv_lo = WSB(i, 2) - req_lo;
            double mv_hi = (i + 1 < n ? WSB(i + 1, 2) : 0.) - req_hi;
            double projected = (req_lo - req_hi) / V;
            double pw = tab(vtab, V_PHI, i) * (WSB(i, 5) - M0(i) - projected);
            double cw = -(tab(vtab, V_DPHI, i) * mv_hi - tab(vtab, V_DPLO, i) * mv_lo) / V;
            out = out + (pw + cw);
            // corrected-PE work of the moved mass minus the operator's part
            int jj = (i < 1 ? 1 : (i > n - 2 ? n - 2 : i)) - 1;
            double s = tab(xtab, T_WT0, i) * moved_rho(jj);
            s = s + tab(xtab, T_WT1, i) * moved_rho(jj + 1);
            s = s + tab(xtab, T_WT2, i) * moved_rho(jj + 2);
            double work = tab(xtab, T_GVAR, i) * s;
            double rs = tab(vtab, V_KMID, i) * raw_minus_mass0(i);
            if (i > 0) rs = rs + tab(vtab, V_KLO, i) * raw_minus_mass0(i - 1);
            if (i < n - 1) rs = rs + tab(vtab, V_KHI, i) * raw_minus_mass0(i + 1);
            if (i == 0) rs = rs + a.t0 * raw_minus_mass0(2);
            if (i == n - 1) rs = rs + a.t1 * raw_minus_mass0(n - 3);
            work = work - rs;
            out = out + work;
          } else {
            out = ywb(i);
          }
          // a failed column keeps its explicit increments (the reference
          // undoes the column with its status bits and flags a retry)
          if (failed) out = du[p * r.sv + row + i];
        }
        if (!L.shadow) s_out[(p * kSlots + slot) * kOB + t] = out;
      }
      __syncthreads();
      // the RK combine u = w0 u0 + w1 u + w2 du' over the block, coalesced,
      // in chunks whose loads are all issued before their stores
      constexpr int TOT = 5 * kSlots * kOB;
      constexpr int CH = 5;  // elements per lane per chunk
      for (int f0 = 0; f0 < TOT; f0 += 64 * CH) {
        double x0[CH], x1[CH];
        int64_t oo[CH];
#pragma unroll
        for (int t = 0; t < CH; ++t) {
          int f = f0 + threadIdx.x + 64 * t;
          oo[t] = -1;
          x0[t] = 0.;
          x1[t] = 0.;
          if (f < TOT) {
            int v = f / (kSlots * kOB), h = f - v * (kSlots * kOB);
            int sl = h / kOB, i = b0 + (h - sl * kOB);
            if (i < n && s_live[sl]) {
              oo[t] = v * r.sv + s_row[sl] + i;
              if constexpr (PAIR) {
                x0[t] = double(static_cast<float const*>(u0v)[oo[t]]) + double(c0[oo[t]]);
                x1[t] = double(static_cast<float*>(uv)[oo[t]]) + double(cc[oo[t]]);
              } else {
                x0[t] = static_cast<double const*>(u0v)[oo[t]];
                x1[t] = static_cast<double*>(uv)[oo[t]];
              }
            }
          }
        }
#pragma unroll
        for (int t = 0; t < CH; ++t)
          if (oo[t] >= 0) {
            double x = a.w0 * x0[t] + a.w1 * x1[t] + a.w2 * s_out[f0 + threadIdx.x + 64 * t];
            if constexpr (PAIR) {
              float h = float(x);
              static_cast<float*>(uv)[oo[t]] = h;
              cc[oo[t]] = float(x - double(h));
            } else {
              static_cast<double*>(uv)[oo[t]] = x;
            }
          }
      }
    }
  }
};

template <bool PAIR>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
    vic_back_kernel(double const* __restrict__ w, double const* __restrict__ du,
                    void const* __restrict__ u0v, float const* __restrict__ c0,
                    void* __restrict__ uv, float* __restrict__ cc,
                    double const* __restrict__ vtab,
                    double const* __restrict__ xtab, double* __restrict__ ws,
                    int const* __restrict__ status, VicArgs a) {
  VicBackBody<PAIR>::run(block_index(), grid_size(), w, du, u0v, c0, uv, cc, vtab, xtab, ws, status,
                         a);
}

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

int64_t snapy_vic_workspace(const snapy_grid* g) {
  if (!g) return 0;
  // the group kernels: kQ doubles per lane per layer for whole wavefronts;
  // the serial kernel: NW per column per layer
  int64_t ncol = int64_t(g->nx2) * g->nx3;
  int64_t group = (ncol + kSlots - 1) / kSlots * 64 * kQ * g->nx1;
  int64_t serial = int64_t(NW) * g->nx1 * ncol;
  return group > serial ? group : serial;
}

// the cyclic-reduction solve's reach: columns (beyond them the forward
// sweep's wavefronts fill the chip) and layers (a workgroup of at most 12
// wavefronts, the kernel's launch bound, and 65 doubles of LDS per layer)
constexpr int64_t kCrMaxCols = 1536;
constexpr int kCrMaxLayers = 144;

static int vic_launch(const snapy_grid* g, const snapy_physics* p,
                      const double* w, const double* du, const void* u0,
                      const float* c0, void* u, float* c, const double* vtab,
                      const double* x1tab, double t0, double t1, double dt,
                      const double* dt_dev, double wstage, double w0,
                      double w1, double w2, double* ws, int* status,
                      double* moved, void* stream, bool serial) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !du || !u0 || !u || !vtab || !x1tab || !ws || !status)
    return fail(-1, "vic_full: NULL argument");
  bool pair = c0 != nullptr;
  if (pair && (!c || serial)) return fail(-1, "vic_full: the FP32 pair needs c0 and c");
  if (g->nx1 < 5) return fail(-9, "vic_full: needs >= 5 x1 cells");
  VicArgs a;
  a.r = ranges(*g);
  a.n = g->nx1;
  a.gamma = p->gammad;
  a.grav = p->grav1;
  a.area = g->dx3 * g->dx2;
  a.vol = g->dx3 * (g->dx2 * g->dx1);
  a.t0 = t0;
  a.t1 = t1;
  a.dt_dev = dt_dev;
  a.dt_host = dt;
  a.wstage = wstage;
  a.w0 = w0;
  a.w1 = w1;
  a.w2 = w2;
  a.moved = moved;
  a.solved = 0;
  int64_t ncol = int64_t(g->nx2) * g->nx3;
  // few columns: the cyclic-reduction solve (one workgroup per column, a
  // layer per lane group) instead of the forward sweep.  A recorded launch
  // takes the same solve as the block's own, so a block steps to the same
  // bits alone and grouped.
  bool const cr = !serial && ncol <= kCrMaxCols && g->nx1 <= kCrMaxLayers;
  if (cr) {
    unsigned nw = unsigned((g->nx1 + kSlots - 1) / kSlots);
    size_t lds = size_t(65) * g->nx1 * sizeof(double);
    launch_flex<VicCrBody, 768, 1>(vic_cr_kernel, dim3(unsigned(ncol)), dim3(64 * nw), lds,
                                    as_stream(stream), w, du, vtab, ws, status, a);
    if (int e = launch_status("vic_cr")) return e;
    a.solved = 1;
    unsigned nb = blocks_for(ncol, kSlots);
    if (pair)
      launch<VicBackBody<true>, 64, 3>(vic_back_kernel<true>, dim3(nb), dim3(64), 0,
                                       as_stream(stream), w, du, u0, c0, u, c, vtab, x1tab, ws,
                                       status, a);
    else
      launch<VicBackBody<false>, 64, 3>(vic_back_kernel<false>, dim3(nb), dim3(64), 0,
                                        as_stream(stream), w, du, u0, c0, u, c, vtab, x1tab, ws,
                                        status, a);
    return launch_status("vic_full");
  }
  if (serial)
    hipLaunchKernelGGL(vic_full_kernel, dim3(blocks_for(ncol, 64)), dim3(64),
                       0, as_stream(stream), w, du,
                       static_cast<double const*>(u0), static_cast<double*>(u),
                       vtab, x1tab, ws, status, a);
  else {
    unsigned nb = blocks_for(ncol, kSlots);
    launch<VicForwardBody, 64, 2>(vic_forward_kernel, dim3(nb), dim3(64), 0, as_stream(stream),
                                  w, du, vtab, ws, status, a);
    if (int e = launch_status("vic_forward")) return e;
    if (pair)
      launch<VicBackBody<true>, 64, 3>(vic_back_kernel<true>, dim3(nb), dim3(64), 0,
                                       as_stream(stream), w, du, u0, c0, u, c, vtab, x1tab, ws,
                                       status, a);
    else
      launch<VicBackBody<false>, 64, 3>(vic_back_kernel<false>, dim3(nb), dim3(64), 0,
                                        as_stream(stream), w, du, u0, c0, u, c, vtab, x1tab, ws,
                                        status, a);
  }
  return launch_status("vic_full");
}

int snapy_vic_full(const snapy_grid* g, const snapy_physics* p, const double* w,
                   const double* du, const double* u0, double* u,
                   const double* vtab, const double* x1tab, double t0,
                   double t1, double dt, const double* dt_dev, double wstage,
                   double w0, double w1, double w2, double* ws, int* status,
                   double* moved, void* stream) {
  return vic_launch(g, p, w, du, u0, nullptr, u, nullptr, vtab, x1tab, t0,
                    t1, dt, dt_dev, wstage, w0, w1, w2, ws, status, moved,
                    stream, false);
}

int snapy_vic_full_f32(const snapy_grid* g, const snapy_physics* p,
                       const double* w, const double* du, const float* u0,
                       const float* c0, float* u, float* c,
                       const double* vtab, const double* x1tab, double t0,
                       double t1, double dt, const double* dt_dev,
                       double wstage, double w0, double w1, double w2,
                       double* ws, int* status, double* moved, void* stream) {
  if (!c0 || !c) return fail(-1, "vic_full_f32: NULL compensation");
  return vic_launch(g, p, w, du, u0, c0, u, c, vtab, x1tab, t0, t1, dt,
                    dt_dev, wstage, w0, w1, w2, ws, status, moved, stream,
                    false);
}

int snapy_vic_full_serial(const snapy_grid* g, const snapy_physics* p,
                          const double* w, const double* du, const double* u0,
                          double* u, const double* vtab, const double* x1tab,
                          double t0, double t1, double dt,
                          const double* dt_dev, double wstage, double w0,
                          double w1, double w2, double* ws, int* status,
                          double* moved, void* stream) {
  return vic_launch(g, p, w, du, u0, nullptr, u, nullptr, vtab, x1tab, t0,
                    t1, dt, dt_dev, wstage, w0, w1, w2, ws, status, moved,
                    stream, true);
}

}  // extern "C"
