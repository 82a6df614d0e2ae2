// The vertical implicit correction with the partial matrix (implicit-scheme
// 1), FP64 on gfx950: 3x3 blocks in (rho, m1, E), the tangential momenta
// written dm_t = v_t drho + q_t and folded into the rho columns and the
// right-hand side, the corrected-PE coupling with its two Woodbury wall
// points, the mass redistribution with the tangential momentum each moved
// mass carries, the post-solve remainder and the RK combine.
//
// Two kernels:
//   * vicp_forward_kernel: one thread per column, 64 columns per block, the
//     state staged through LDS a few layers at a time with coalesced loads:
//     each layer's blocks assembled in registers (vic_assemble_cell,
//     src/implicit/vic_kernels.h:66-105, the face matrices and Jacobians
//     carried up from the layer below), the forward sweep, the Woodbury
//     2x2 and the back substitution;
//   * vicp_finish_kernel<PAIR>: a few columns per workgroup, cells across
//     the threads: the redistribution (its serial chains one thread per
//     column, in the column's order), the outputs and the combine.
//
// Restates ImplicitHydroImpl::forward_masked (src/implicit/
// implicit_hydro.cpp:301-623, scheme 1) for a Cartesian block owning its
// whole x1 column, no solids, dry ideal gas, no FluxRegister posts:
// vic_assemble_impl<3> (vic_assemble_impl.h:45-218), ForwardSweep's small
// matrix path (forward_sweep_impl.h:104-121), vic_backward_substitute,
// vic_woodbury_column, vic_constituent_column and
// vic_tangential_momentum_column (vic_redistribute_impl.h:31-205),
// book_remainder_ (implicit_hydro.cpp:713-794).
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

// doubles per layer record: A' 9, y 3, z0 3.  The second Woodbury column's
// right-hand side is nonzero only in the top layer, so its forward-sweep
// vector is zero below the top layer and only the top layer's (kept in
// registers) is needed.
constexpr int kRec = 15;
constexpr int kF = kRec;
// record slots: A' = a^-1 c 0..8, y 9..11, z0 12..14
enum { R_A = 0, R_Y = 9, R_Z0 = 12 };

enum { V_WLO = 0, V_WHI, V_KLO, V_KMID, V_KHI, V_PHI, V_DPLO, V_DPHI };
enum { T_GVAR = 3, T_WT0 = 4, T_WT1 = 5, T_WT2 = 6 };

struct PartArgs {
  Range3 r;
  int n;
  double gamma, grav, area, vol, t0, t1;
  double const* dt_dev;
  double dt_host;
  double wstage;
  double w0, w1, w2;
  // NULL, or (nc3, nc2, nc1) FP64: the clamped dry-gas transfer through the
  // face below each interior cell (0 in a failed column)
  double* moved;
  // NULL, or the primitives (the w the solve reads): the FP32 pair finish
  // then writes the primitives of the combined state into the interior
  double* wout;
};

// the stage dt of column col (the member's along a member x3 axis)
__device__ __forceinline__ double stage_dt(PartArgs const& a, int64_t col) {
  return dt_at(a.dt_dev, a.dt_host, a.r, a.r.lo3 + int(col / a.r.n2)) * a.wstage;
}

// rows {rho, m1, E} of FluxJacobian (flux_decomposition_impl.h:105-126), dir
// 0; gp_g = (gm1 + 1) / gm1, pr = p / rho (the caller's, formed once per
// cell)
__device__ __forceinline__ void jac_rows(double f[3][5], double gm1, double gp_g,
                                         double const* w, double pr) {
  double v1 = w[IVX], v2 = w[IVY], v3 = w[IVZ];
  double s2 = v1 * v1 + v2 * v2 + v3 * v3;
  double c1 = ((gm1 - 1) * s2 * 0.5 - gp_g * pr) * v1;
  double c2 = gp_g * pr + s2 * 0.5 - gm1 * v1 * v1;
  f[0][0] = 0.; f[0][1] = 1.; f[0][2] = 0.; f[0][3] = 0.; f[0][4] = 0.;
  f[1][0] = gm1 * s2 * 0.5 - v1 * v1; f[1][1] = (2. - gm1) * v1;
  f[1][2] = -gm1 * v2; f[1][3] = -gm1 * v3; f[1][4] = gm1;
  f[2][0] = c1; f[2][1] = c2; f[2][2] = -gm1 * v2 * v1;
  f[2][3] = -gm1 * v3 * v1; f[2][4] = (gm1 + 1) * v1;
}

// rows {rho, m1, E} of |A| = R Lambda R^-1 at the Roe average of wl, wr
// (flux_decomposition_impl.h:13-103).  The divisions by the constants of
// the gas and by 2 c hp and hp are taken as products with reciprocals
// formed once (seven divisions per face instead of twenty); the rows agree
// with the reference's to rounding.
struct GasK {
  double gm1, igm1, gp_g, g_gp;  // gm1, 1 / gm1, (gm1 + 1) / gm1, gm1 / (gm1 + 1)
};

__device__ __forceinline__ void roe_rows3(double out[3][5], GasK const& K, double const* wl,
                                          double const* wr) {
  double sdl = sqrt(wl[IDN]), sdr = sqrt(wr[IDN]);
  double is = 1.0 / (sdl + sdr);
  double pr[5];
  pr[IDN] = sdl * sdr;
  for (int v = IVX; v <= IVZ; ++v) pr[v] = (sdl * wl[v] + sdr * wr[v]) * is;
  double el = wl[IPR] * K.igm1 +
              0.5 * wl[IDN] * (wl[IVX] * wl[IVX] + wl[IVY] * wl[IVY] + wl[IVZ] * wl[IVZ]);
  double er = wr[IPR] * K.igm1 +
              0.5 * wr[IDN] * (wr[IVX] * wr[IVX] + wr[IVY] * wr[IVY] + wr[IVZ] * wr[IVZ]);
  double hbar = ((el + wl[IPR]) / sdl + (er + wr[IPR]) / sdr) * is;
  pr[IPR] = (hbar - 0.5 * (pr[IVX] * pr[IVX] + pr[IVY] * pr[IVY] + pr[IVZ] * pr[IVZ])) *
            K.g_gp * pr[IDN];
  double r = pr[IDN], u = pr[IVX], v = pr[IVY], w = pr[IVZ], p = pr[IPR];
  double hp = K.gp_g * p / r;
  double cs = sqrt(hp * K.gm1);  // p (gm1 + 1) / rho
  double ke = 0.5 * (u * u + v * v + w * w);
  double h = hp + ke;
  double ihp = 1. / hp;
  double i2hp = 0.5 * ihp;
  double i2cshp = i2hp / cs;
  double lam[5] = {fabs(u - cs), fabs(u), fabs(u + cs), fabs(u), fabs(u)};
  double R[3][5] = {{1., 1., 1., 0., 0.},
                    {u - cs, u, u + cs, 0., 0.},
                    {h - u * cs, ke, h + u * cs, v, w}};
  double Ri[5][5] = {
      {(cs * ke + u * hp) * i2cshp, (-hp - cs * u) * i2cshp, -v * i2hp, -w * i2hp, i2hp},
      {(hp - ke) * ihp, u * ihp, v * ihp, w * ihp, -ihp},
      {(cs * ke - u * hp) * i2cshp, (hp - cs * u) * i2cshp, -v * i2hp, -w * i2hp, i2hp},
      {-v, 0., 1., 0., 0.},
      {-w, 0., 0., 1., 0.}};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 5; ++b) {
      double s = 0.;
      for (int k = 0; k < 5; ++k) s += (R[a][k] * lam[k]) * Ri[k][b];
      out[a][b] = s;
    }
}

// the 5x5 row/column indices the partial blocks keep: rho, m1, E
__device__ __forceinline__ int full_col(int s) { return s == 2 ? IPR : s; }

// ---------------------------------------------------------------------------
// assembly of one cell's blocks (vic_assemble_impl<3>)
// ---------------------------------------------------------------------------
// the grid-only factors of cell i and the two explicit masses the wall
// cells' corrected-PE rows read
struct CellTab {
  double wlo, whi, klo, kmid, khi;
  double d2, dn3;  // du_rho at cells 2 and n - 3
};

// am, ap: rows {rho, m1, E} of |A| at the lower / upper face; dp, dc, dn:
// the flux Jacobian rows of cells i - 1, i, i + 1; wl, wc, wr and dl, dm, dr
// their primitives and explicit increments.  Writes the blocks a, b, c and
// the right-hand side with the fold's correction subtracted.
__device__ __forceinline__ void assemble_cell(
    double const (&am)[3][5], double const (&ap)[3][5], double const (&dp)[3][5],
    double const (&dc)[3][5], double const (&dn)[3][5], double const* wl, double const* wc,
    double const* wr, double const* dl, double const* dm, double const* dr, CellTab const& t,
    int i, int n, double A, double hiv, double grav, double dt, double t0, double t1,
    double (&aa)[3][3], double (&bb)[3][3], double (&cc)[3][3], double (&rhs)[3]) {
  for (int p = 0; p < 3; ++p)
    for (int s = 0; s < 3; ++s) {
      int q = full_col(s);
      aa[p][s] = (am[p][q] * A + ap[p][q] * A + (A - A) * dc[p][q]) * hiv;
      bb[p][s] = -(am[p][q] + dp[p][q]) * A * hiv;
      cc[p][s] = -(ap[p][q] - dn[p][q]) * A * hiv;
    }
  aa[1][0] -= grav;  // - Phi (Phi(m1, rho) = Phi(E, m1) = grav)
  aa[2][1] -= grav;
  double const idt = 1. / dt;
  for (int p = 0; p < 3; ++p) aa[p][p] += idt;
  double lo = t.wlo, up = t.whi;
  for (int s = 0; s < 3; ++s) {
    int q = full_col(s);
    aa[2][s] -= grav * (up * ap[0][q] - lo * am[0][q]);
    bb[2][s] -= grav * lo * am[0][q];
    cc[2][s] += grav * up * ap[0][q];
  }
  aa[2][1] += grav * (1. - lo - up);
  bb[2][1] -= grav * lo;
  cc[2][1] -= grav * up;
  bool merged_lo = i == 0, merged_hi = i == n - 1;
  if (merged_lo)
    for (int p = 0; p < 3; ++p)
      for (int s = 0; s < 3; ++s) aa[p][s] += bb[p][s] * (s == 1 ? -1. : 1.);
  if (merged_hi)
    for (int p = 0; p < 3; ++p)
      for (int s = 0; s < 3; ++s) aa[p][s] += cc[p][s] * (s == 1 ? -1. : 1.);
  // the tangential columns (m2, m3) of the same rows, and the fold
  double aT[3][2], bT[3][2], cT[3][2];
  for (int p = 0; p < 3; ++p)
    for (int kk = 0; kk < 2; ++kk) {
      int q = IVY + kk;
      aT[p][kk] = (am[p][q] * A + ap[p][q] * A + (A - A) * dc[p][q]) * hiv;
      bT[p][kk] = -(am[p][q] + dp[p][q]) * A * hiv;
      cT[p][kk] = -(ap[p][q] - dn[p][q]) * A * hiv;
    }
  for (int kk = 0; kk < 2; ++kk) {
    int q = IVY + kk;
    aT[2][kk] -= grav * (up * ap[0][q] - lo * am[0][q]);
    bT[2][kk] -= grav * lo * am[0][q];
    cT[2][kk] += grav * up * ap[0][q];
  }
  if (merged_lo)
    for (int p = 0; p < 3; ++p)
      for (int kk = 0; kk < 2; ++kk) aT[p][kk] += bT[p][kk];
  if (merged_hi)
    for (int p = 0; p < 3; ++p)
      for (int kk = 0; kk < 2; ++kk) aT[p][kk] += cT[p][kk];
  double corr[3] = {0., 0., 0.};
  // cell jc = 0, 1, 2: i - 1, i, i + 1
  auto fold = [&](double (*blk)[3], double (*tb)[2], int jc) {
    double const* dj = jc == 0 ? dl : (jc == 1 ? dm : dr);
    double const* wj = jc == 0 ? wl : (jc == 1 ? wc : wr);
    double drho = dj[IDN];
    for (int kk = 0; kk < 2; ++kk) {
      double v = wj[IVY + kk];
      double q = dj[IVY + kk] - v * drho;
      if (v != 0.)
        for (int p = 0; p < 3; ++p) blk[p][0] += tb[p][kk] * v;
      if (q != 0.)
        for (int p = 0; p < 3; ++p) corr[p] += tb[p][kk] * q;
    }
  };
  fold(aa, aT, 1);
  if (!merged_lo) fold(bb, bT, 0);
  if (!merged_hi) fold(cc, cT, 2);
  // couple_radial_exact (implicit_hydro.cpp:408-446): E row, rho column
  aa[2][0] -= t.kmid * idt;
  bb[2][0] -= t.klo * idt;
  cc[2][0] -= t.khi * idt;
  // right-hand side: du / dt, the energy without the explicit corrected PE
  double rows = t.kmid * dm[IDN];
  if (i > 0) rows = rows + t.klo * dl[IDN];
  if (i < n - 1) rows = rows + t.khi * dr[IDN];
  if (i == 0) rows = rows + t0 * t.d2;
  if (i == n - 1) rows = rows + t1 * t.dn3;
  rhs[0] = dm[IDN] * idt;
  rhs[1] = dm[IVX] * idt;
  rhs[2] = (dm[IPR] - rows) * idt;
  for (int p = 0; p < 3; ++p)
    if (corr[p] != 0.) rhs[p] -= corr[p];
}

// ---------------------------------------------------------------------------
// the column program: one thread per column
// ---------------------------------------------------------------------------
// ludcmp's singularity test for a 3x3 block with implicit row scaling
// (src/math/ludcmp.h:31-95): false if the matrix is singular or non-finite
__device__ __forceinline__ bool ludcmp3_ok(double const (&m)[3][3]) {
  double a[3][3], vv[3], scale[3];
  const double tol = 8. * 3. * 2.220446049250313e-16;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j] = m[i][j];
      if (!isfinite(a[i][j])) return false;
    }
  for (int i = 0; i < 3; ++i) {
    double big = 0.;
    for (int j = 0; j < 3; ++j) big = fabs(a[i][j]) > big ? fabs(a[i][j]) : big;
    if (big == 0.) return false;
    scale[i] = big;
    vv[i] = 1.0 / big;
  }
  for (int j = 0; j < 3; ++j) {
    for (int i = 0; i < j; ++i) {
      double sum = a[i][j];
      for (int k = 0; k < i; ++k) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
    }
    double big = 0.;
    int imax = j;
    for (int i = j; i < 3; ++i) {
      double sum = a[i][j];
      for (int k = 0; k < j; ++k) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
      double dum = vv[i] * fabs(sum);
      if (!isfinite(dum)) return false;
      if (dum >= big) {
        big = dum;
        imax = i;
      }
    }
    if (j != imax) {
      for (int k = 0; k < 3; ++k) {
        double d = a[imax][k];
        a[imax][k] = a[j][k];
        a[j][k] = d;
      }
      vv[imax] = vv[j];
      double t = scale[imax];
      scale[imax] = scale[j];
      scale[j] = t;
    }
    if (fabs(a[j][j]) / scale[j] <= tol) return false;
    if (j != 2) {
      double dum = 1.0 / a[j][j];
      for (int i = j + 1; i < 3; ++i) a[i][j] *= dum;
    }
  }
  return true;
}

// Eigen's 3x3 inverse by cofactors (Eigen 3.4, InverseImpl.h)
__device__ __forceinline__ void inverse3(double const (&m)[3][3], double (&o)[3][3]) {
  double cof00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  double cof10 = m[2][1] * m[0][2] - m[2][2] * m[0][1];
  double cof20 = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  double det = cof00 * m[0][0] + cof10 * m[1][0] + cof20 * m[2][0];
  double inv = 1.0 / det;
  o[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv;
  o[1][1] = (m[2][2] * m[0][0] - m[2][0] * m[0][2]) * inv;
  o[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv;
  o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
  o[2][1] = (m[2][0] * m[0][1] - m[2][1] * m[0][0]) * inv;
  o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
  o[0][0] = cof00 * inv;
  o[0][1] = cof10 * inv;
  o[0][2] = cof20 * inv;
}

__device__ __forceinline__ void mv3(double const (&m)[3][3], double const* v, double* o) {
  for (int r = 0; r < 3; ++r) o[r] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2];
}

// record value e of layer l of column col: ws[(l * kRec + e) * ncol + col], so
// a wavefront of consecutive columns reads one value as 64 consecutive doubles
struct Rec {
  double* p;
  int64_t s;
  __device__ double& operator[](int e) const { return p[e * s]; }
};

constexpr int kLB = 4;  // layers the producer stages per chunk (the layer
                        // above each of the chunk's layers)
constexpr int kH = 30;           // hand-off per column and layer: m 9, b 9, c 9, rhs 3
constexpr int kYL = 16;           // layers of the solution gathered before a write
constexpr int kYS = 3 * kYL + 1;  // LDS row of one column (padded)
// LDS: during the sweep, the producer's staging (w and du, 5 x kLB x 64
// each) and the hand-off (kH x 64); during the back substitution, the
// gather (64 x kYS)
constexpr int kFwdLds = 2 * 5 * kLB * 64 + kH * 64;
constexpr int kLds = 64 * kYS > kFwdLds ? 64 * kYS : kFwdLds;

// The column program's serial part for 64 columns, two wavefronts:
//   * the producer (wavefront 0), one lane per column, stages the state
//     through LDS (coalesced along x1) and assembles layer i + 1's blocks
//     (vic_assemble_cell, the upper face's |A| rows and the Jacobian rows
//     carried up from the layer below, so each face is assembled once);
//   * the consumer (wavefront 1), one lane per column, meanwhile runs the
//     forward sweep of layer i (ForwardSweep<2>, the small-matrix path),
//     writing A' = a^-1 c and the partial solutions y, z0 to the
//     lane-contiguous records, and accumulating row 0 of the back
//     substitution's layer 2 (the Woodbury row at the lower wall).
// Layer i's blocks pass from producer to consumer through LDS; two
// workgroup barriers per layer order the hand-off.  Both wavefronts then
// run the back substitution's loop: the consumer solves the Woodbury 2x2
// after the top three layers and then the one combined column, the gather
// of the solution (kYL layers in LDS) is written by both as runs along x1
// into yc, column-major (yc[(col * 3 + p) * n + l]) for the finish kernel.
// The role is a scalar value, so each wavefront takes one branch, and both
// branches pass the same barriers (1 + 2 n in the sweep, 2 per gather
// flush).
struct VicpForwardBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ du,
                                             double const* __restrict__ vtab,
                                             double* __restrict__ ws, double* __restrict__ yc,
                                             int* __restrict__ status, PartArgs a) {
    __shared__ double s_buf[kLds];
    __shared__ int64_t s_row[64];
    __shared__ int s_act[64];
    double* s_w = s_buf;  // [v][ll][cc]
    double* s_du = s_buf + 5 * kLB * 64;
    double* s_h = s_buf + 2 * 5 * kLB * 64;  // [e][cc]
    Range3 const& r = a.r;
    int ncol = r.n2 * r.n3;  // < 2^30 (checked by the launcher)
    int const lane = threadIdx.x & 63;
    bool const producer = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0;
    int col0 = bid.x * 64;
    int col = col0 + lane;
    bool live = col < ncol;
    int n = a.n, nc1 = r.nc1, lo1 = r.lo1;
    double const A = a.area, V = a.vol;
    double const dt = stage_dt(a, live ? col : 0);
    auto REC = [&](int l) { return Rec{ws + int64_t(l) * kRec * ncol + col, ncol}; };
    if (producer) {
      // the first interior cell of each of the block's columns
      // (column 0's for a lane past the grid, so its staging loads stay in
      // bounds)
      int cl = live ? col : 0;
      int jj = r.lo2 + cl % r.n2, kk = r.lo3 + cl / r.n2;
      s_row[lane] = kk * r.s3 + jj * r.s2 + r.lo1;
    }
    __syncthreads();

    // the consumer's Woodbury inputs for the back substitution: row 0 of
    // layer 2's solution for y and z0, and for z1 with its top-layer vector
    double acc[2] = {0., 0.}, acc_z1 = 0., z1top[3] = {0., 0., 0.};
    if (producer) {
      // ---- producer: staging and assembly ------------------------------------
      double const gm1 = a.gamma - 1.;
      double const gme = 0.5 * (a.gamma + a.gamma) - 1.;
      GasK const K{gme, 1. / gme, (gme + 1.) / gme, gme / (gme + 1.)};
      double const gp_g = (gm1 + 1.) / gm1;
      double const hiv = 0.5 / V;
      auto tab = [&](int rowi, int l) { return vtab[rowi * nc1 + lo1 + l]; };
      double d2 = 0., dn3 = 0.;
      if (live) {
        int64_t row = s_row[lane];
        d2 = du[IDN * r.sv + row + 2];
        dn3 = du[IDN * r.sv + row + n - 3];
      }
      // stage layers c0 + 1 .. c0 + kLB of w and du (the layer above each of
      // the chunk's layers; the layer itself and the one below are carried in
      // registers) for the 64 columns, lanes along x1 within a column
      // (coalesced runs); only this wavefront reads the staging area.  Every
      // load is unconditional and in bounds (the layer index clamped to the
      // ghost layers of w and the interior of du; a column past the grid reads
      // column 0), so a chunk's loads are in flight together; the out-of-range
      // values are then replaced.
      auto stage = [&](int c0) {
        double xw[kLB][5], xd[kLB][5];
#pragma unroll
        for (int t = 0; t < kLB; ++t) {
          int e = lane + 64 * t;
          int cc = e / kLB, layer = c0 + 1 + e % kLB;
          int64_t grow = s_row[cc];
          int lw = layer > n ? n : layer;
          int ld = layer > n - 1 ? n - 1 : layer;
          for (int v = 0; v < 5; ++v) {
            xw[t][v] = w[v * r.sv + grow + lw];
            xd[t][v] = du[v * r.sv + grow + ld];
          }
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < kLB; ++t) {
          int e = lane + 64 * t;
          int cc = e / kLB, ll = e % kLB, layer = c0 + 1 + ll;
          bool in = col0 + cc < ncol;
          for (int v = 0; v < 5; ++v) {
            s_w[(v * kLB + ll) * 64 + cc] = in && layer <= n ? xw[t][v] : 1.;
            s_du[(v * kLB + ll) * 64 + cc] = in && layer < n ? xd[t][v] : 0.;
          }
        }
        __builtin_amdgcn_wave_barrier();
      };
      // carried from the layer below: the state and increments of cells i - 1
      // and i, and the lower face's |A| rows (the Jacobian rows, cheap, are
      // formed again each layer)
      double wl[5], wc[5], dl[5], dmv[5], fa_lo[3][5], prl, prc;
      {
        int64_t row = s_row[lane];
        for (int v = 0; v < 5; ++v) {
          double x0 = w[v * r.sv + row - 1], x1 = w[v * r.sv + row], d0 = du[v * r.sv + row];
          wl[v] = live ? x0 : 1.;
          wc[v] = live ? x1 : 1.;
          dl[v] = 0.;
          dmv[v] = live ? d0 : 0.;
        }
        roe_rows3(fa_lo, K, wl, wc);
        prl = wl[IPR] / wl[IDN];
        prc = wc[IPR] / wc[IDN];
      }
      // assemble layer i from the carried cells and the staged layer above,
      // and hand it to the consumer
      auto assemble = [&](int i) {
        int ll = i % kLB;
        if (ll == 0) stage(i);
        double wr[5], dr[5];
        for (int v = 0; v < 5; ++v) {
          wr[v] = s_w[(v * kLB + ll) * 64 + lane];
          dr[v] = s_du[(v * kLB + ll) * 64 + lane];
        }
        double fa_hi[3][5], jp[3][5], jc[3][5], jn[3][5];
        roe_rows3(fa_hi, K, wc, wr);
        double prr = wr[IPR] / wr[IDN];
        jac_rows(jp, gm1, gp_g, wl, prl);
        jac_rows(jc, gm1, gp_g, wc, prc);
        jac_rows(jn, gm1, gp_g, wr, prr);
        CellTab ct;
        ct.wlo = tab(V_WLO, i);
        ct.whi = tab(V_WHI, i);
        ct.klo = tab(V_KLO, i);
        ct.kmid = tab(V_KMID, i);
        ct.khi = tab(V_KHI, i);
        ct.d2 = d2;
        ct.dn3 = dn3;
        double m[3][3], bb[3][3], c3[3][3], rhs[3];
        assemble_cell(fa_lo, fa_hi, jp, jc, jn, wl, wc, wr, dl, dmv, dr, ct, i, n, A, hiv, a.grav,
                      dt, a.t0, a.t1, m, bb, c3, rhs);
        for (int v = 0; v < 5; ++v) {
          wl[v] = wc[v];
          wc[v] = wr[v];
          dl[v] = dmv[v];
          dmv[v] = dr[v];
        }
        prl = prc;
        prc = prr;
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 5; ++q) fa_lo[p][q] = fa_hi[p][q];
        for (int p = 0; p < 3; ++p) {
          for (int q = 0; q < 3; ++q) {
            s_h[(3 * p + q) * 64 + lane] = m[p][q];
            s_h[(9 + 3 * p + q) * 64 + lane] = bb[p][q];
            s_h[(18 + 3 * p + q) * 64 + lane] = c3[p][q];
          }
          s_h[(27 + p) * 64 + lane] = rhs[p];
        }
      };
      assemble(0);
      __syncthreads();  // layer 0 handed off
      for (int i = 0; i < n; ++i) {
        __syncthreads();  // the consumer has taken layer i
        if (i + 1 < n) assemble(i + 1);
        __syncthreads();  // layer i + 1 handed off
      }
    } else {
      // ---- consumer: the forward sweep ---------------------------------------
      double ap[3][3], yp[3], z0p[3];
      bool ok = live;
      // row 0 of the back substitution's layer 2 for y, z0, z1, accumulated
      // over the sweep: x_2 = sum_{j >= 2} r_j zeta_j with r_2 = e0 and
      // r_{j+1} = -r_j A'_j (the recursion x_j = zeta_j - A'_j x_{j+1}
      // unrolled), so the Woodbury rows are known when the back substitution
      // starts
      double rv[3] = {1., 0., 0.};
      double acc_y = 0., acc_z0 = 0.;
      __syncthreads();  // layer 0 handed off
      for (int i = 0; i < n; ++i) {
        double m[3][3], bb[3][3], c3[3][3], rhs[3];
        for (int p = 0; p < 3; ++p) {
          for (int q = 0; q < 3; ++q) {
            m[p][q] = s_h[(3 * p + q) * 64 + lane];
            bb[p][q] = s_h[(9 + 3 * p + q) * 64 + lane];
            c3[p][q] = s_h[(18 + 3 * p + q) * 64 + lane];
          }
          rhs[p] = s_h[(27 + p) * 64 + lane];
        }
        __syncthreads();  // layer i taken
        // the factorisation and the singularity test of the pivot block are
        // independent work (the test only decides the column's status), so
        // both run every layer; a failed column's values are never read
        {
          double e0[3] = {0., 0., i == 0 ? 1. : 0.};
          if (i > 0)
            for (int p = 0; p < 3; ++p)
              for (int s = 0; s < 3; ++s)
                m[p][s] -= bb[p][0] * ap[0][s] + bb[p][1] * ap[1][s] + bb[p][2] * ap[2][s];
          bool piv = ludcmp3_ok(m);
          double inv[3][3];
          inverse3(m, inv);
          double y[3], z0[3];
          if (i > 0) {
            double t[3], s0[3], by[3], b0[3];
            mv3(bb, yp, by);
            mv3(bb, z0p, b0);
            for (int p = 0; p < 3; ++p) {
              t[p] = rhs[p] - by[p];
              s0[p] = e0[p] - b0[p];
            }
            mv3(inv, t, y);
            mv3(inv, s0, z0);
          } else {
            mv3(inv, rhs, y);
            mv3(inv, e0, z0);
          }
          for (int p = 0; p < 3; ++p)
            for (int s = 0; s < 3; ++s)
              ap[p][s] = inv[p][0] * c3[0][s] + inv[p][1] * c3[1][s] + inv[p][2] * c3[2][s];
          bool fin = true;
          for (int p = 0; p < 3; ++p) {
            if (!isfinite(y[p])) fin = false;
            for (int s = 0; s < 3; ++s)
              if (!isfinite(ap[p][s])) fin = false;
            yp[p] = y[p];
            z0p[p] = z0[p];
          }
          if (i >= 2) {
            acc_y += rv[0] * y[0] + rv[1] * y[1] + rv[2] * y[2];
            acc_z0 += rv[0] * z0[0] + rv[1] * z0[1] + rv[2] * z0[2];
            if (i == n - 1) {
              // the second Woodbury column: right-hand side e2 in the top
              // layer only, so its sweep vector is inv e2 there
              for (int p = 0; p < 3; ++p) z1top[p] = inv[p][2];
              acc_z1 = rv[0] * z1top[0] + rv[1] * z1top[1] + rv[2] * z1top[2];
            }
            double rn[3];
            for (int s = 0; s < 3; ++s)
              rn[s] = -(rv[0] * ap[0][s] + rv[1] * ap[1][s] + rv[2] * ap[2][s]);
            for (int s = 0; s < 3; ++s) rv[s] = rn[s];
          }
          ok = ok && piv && fin;
          if (ok) {
            Rec rec = REC(i);
            for (int p = 0; p < 3; ++p) {
              for (int s = 0; s < 3; ++s) rec[R_A + 3 * p + s] = ap[p][s];
              rec[R_Y + p] = y[p];
              rec[R_Z0 + p] = z0[p];
            }
          }
        }
        __syncthreads();  // layer i + 1 handed off
      }
      acc[0] = acc_y;
      acc[1] = acc_z0;
      // a failed column keeps its explicit increments (the finish kernel
      // reads the flag; the reference undoes the column with its status bits
      // and flags a retry)
      if (live) vic_mark(status, col, ok);
      s_act[lane] = ok;
    }

    // ---- back substitution (vic_backward_substitute) + Woodbury -------------
    // (vic_woodbury_column).  The top three layers are solved for the three
    // columns y, z0, z1; row 0 of layer n - 3 and the swept row 0 of layer 2
    // give the 2x2 and its coefficients al0, al1; every lower layer then
    // solves the one combined column, zeta = zeta_y - al0 zeta_z0 (the z1
    // vector is zero below the top layer), since the substitution is linear.
    // The staging and hand-off areas become the gather (the last hand-off was
    // read before the sweep's last barrier); both wavefronts run the loop, the
    // consumer solves, both write the gathered layers.
    bool act = !producer && live && s_act[lane];
    double* s_y = s_buf;  // [cc][p * kYL + layer % kYL], row kYS
    double ytop[3][3];    // the combined solution of layers n - 1, n - 2, n - 3
    double xc[3] = {0., 0., 0.};
    double al0 = 0., al1 = 0.;
    for (int t = 0; t < 3; ++t)
      for (int p = 0; p < 3; ++p) ytop[t][p] = 0.;
    double nx[kRec];
    for (int e = 0; e < kRec; ++e) nx[e] = 0.;
    if (act) {
      double xy[3][3], x0[3][3], x1[3][3];
      for (int t = 0; t < 3; ++t) {
        Rec rec = REC(n - 1 - t);
        double cy[3], cz[3];
        for (int p = 0; p < 3; ++p) {
          cy[p] = rec[R_Y + p];
          cz[p] = rec[R_Z0 + p];
        }
        if (t == 0) {
          for (int p = 0; p < 3; ++p) {
            xy[0][p] = cy[p];
            x0[0][p] = cz[p];
            x1[0][p] = z1top[p];
          }
        } else {
          double A3[3][3];
          for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) A3[p][q] = rec[R_A + 3 * p + q];
          double sy[3], s0[3], s1[3];
          mv3(A3, xy[t - 1], sy);
          mv3(A3, x0[t - 1], s0);
          mv3(A3, x1[t - 1], s1);
          for (int p = 0; p < 3; ++p) {
            xy[t][p] = cy[p] - sy[p];
            x0[t][p] = cz[p] - s0[p];
            x1[t][p] = 0. - s1[p];
          }
        }
      }
      // the record of layer n - 4, loaded while the 2x2 is solved
      {
        Rec rec = REC(n - 4);
        for (int e = 0; e < kRec; ++e) nx[e] = rec[e];
      }
      double k0 = -a.t0 / dt, k1 = -a.t1 / dt;
      double m00 = 1. + k0 * acc[1], m01 = k0 * acc_z1;
      double m10 = k1 * x0[2][0], m11 = 1. + k1 * x1[2][0];
      double r0 = k0 * acc[0], r1 = k1 * xy[2][0];
      double det = m00 * m11 - m01 * m10;
      al0 = (m11 * r0 - m01 * r1) / det;
      al1 = (m00 * r1 - m10 * r0) / det;
      for (int t = 0; t < 3; ++t)
        for (int p = 0; p < 3; ++p) {
          double y = xy[t][p];
          y -= x0[t][p] * al0;
          y -= x1[t][p] * al1;
          ytop[t][p] = y;
        }
      for (int p = 0; p < 3; ++p) xc[p] = ytop[2][p];
    }
    for (int i = n - 1; i >= 0; --i) {
      if (!producer) {
        double yv[3];
        if (i >= n - 3) {
          int t = n - 1 - i;
          for (int p = 0; p < 3; ++p) yv[p] = t == 0 ? ytop[0][p] : (t == 1 ? ytop[1][p] : ytop[2][p]);
        } else {
          double cur[kRec];
          for (int e = 0; e < kRec; ++e) cur[e] = nx[e];
          // the record of the layer below, loaded while this layer is solved
          if (act && i > 0) {
            Rec rec = REC(i - 1);
            for (int e = 0; e < kRec; ++e) nx[e] = rec[e];
          }
          double A3[3][3];
          for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) A3[p][q] = cur[R_A + 3 * p + q];
          double sx[3];
          mv3(A3, xc, sx);
          for (int p = 0; p < 3; ++p) {
            double z = cur[R_Y + p] - cur[R_Z0 + p] * al0;
            xc[p] = z - sx[p];
            yv[p] = xc[p];
          }
        }
        double* o = s_y + lane * kYS + (i % kYL);
        for (int p = 0; p < 3; ++p) o[p * kYL] = yv[p];
      }
      if (i % kYL == 0) {
        // layers i .. i + cnt - 1 of the block's columns, lanes along x1
        int cnt = n - i < kYL ? n - i : kYL;
        __syncthreads();
        for (int t = threadIdx.x; t < 64 * 3 * kYL; t += 128) {
          int ll = t % kYL, pe = t / kYL, cc = pe / 3, e = pe % 3;
          if (ll < cnt && col0 + cc < ncol && s_act[cc])
            yc[(int64_t(col0 + cc) * 3 + e) * n + i + ll] = s_y[cc * kYS + e * kYL + ll];
        }
        __syncthreads();
      }
    }
  }
};

__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
    vicp_forward_kernel(double const* __restrict__ w, double const* __restrict__ du,
                        double const* __restrict__ vtab, double* __restrict__ ws,
                        double* __restrict__ yc, int* __restrict__ status, PartArgs a) {
  VicpForwardBody::run(block_index(), grid_size(), w, du, vtab, ws, yc, status, a);
}

// the rest of the column program for kFinCols columns per workgroup
// (kFinThreads threads): every cell's inputs of the redistribution formed
// across the threads (the state reads and writes run along x1, coalesced),
// the serial chains of vic_constituent_column one thread per column in the
// column's order (the sums, the prefix of the requested face transfers,
// the availability clamps), then every cell's outputs, the tangential
// transfers, the remainder and the RK combine.  Dynamic LDS: 3 nc (n + 1)
// doubles for nc <= kFinCols columns.
constexpr int kFinCols = 8;
constexpr int kFinThreads = 256;

template <bool PAIR>
struct VicpFinishBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* w,
                                             double const* __restrict__ du,
                                             void const* __restrict__ u0v,
                                             float const* __restrict__ c0, void* __restrict__ uv,
                                             float* __restrict__ cc,
                                             double const* __restrict__ vtab,
                                             double const* __restrict__ xtab,
                                             double const* __restrict__ yc,
                                             int const* __restrict__ status, PartArgs a, int nc) {
    extern __shared__ double lds[];
    __shared__ int64_t s_row[kFinCols];
    __shared__ int s_in[kFinCols], s_ok[kFinCols];
    Range3 const& r = a.r;
    int64_t ncol = int64_t(r.n2) * r.n3;
    int64_t col0 = int64_t(bid.x) * nc;
    int const tid = threadIdx.x;
    int n = a.n, nc1 = r.nc1, lo1 = r.lo1, L = n + 1;
    double const V = a.vol;
    // per column c: phi_i, then M_i (the requested transfer through the face
    // below cell i); (w_rho V)_i; (w_rho + du_rho)_i V, then q_i (the moved
    // transfer through the face below cell i)
    double* s_A = lds;
    double* s_B = lds + nc * L;
    double* s_C = lds + 2 * nc * L;
    if (tid < nc) {
      int64_t col = col0 + tid;
      bool in = col < ncol;
      int64_t cl = in ? col : 0;
      int j = r.lo2 + int(cl % r.n2), k = r.lo3 + int(cl / r.n2);
      s_row[tid] = k * r.s3 + j * r.s2 + r.lo1;
      s_in[tid] = in;
      s_ok[tid] = in && (status[col] & kVicStageFailed) == 0;
    }
    __syncthreads();
    int const cells = nc * n;
    for (int e = tid; e < cells; e += kFinThreads) {
      int c = e / n, l = e - c * n;
      if (!s_ok[c]) continue;
      int64_t row = s_row[c];
      double y0 = yc[((col0 + c) * 3) * n + l];
      double d0 = du[IDN * r.sv + row + l], wr = w[IDN * r.sv + row + l];
      s_A[c * L + l] = (y0 - d0) * V;
      s_B[c * L + l] = wr * V;
      s_C[c * L + l] = (wr * 1. + d0) * V;
    }
    __syncthreads();
    // ---- the serial chains of vic_constituent_column, one thread per column,
    // in the column's order: the column sums R and S, then, face by face, the
    // requested transfer M (the prefix of phi_i - R (w_rho V)_i / S) and the
    // moved transfer q clamped by what the cells on either side hold
    if (tid < nc && s_ok[tid]) {
      double* A_ = s_A + tid * L;
      double* B_ = s_B + tid * L;
      double* C_ = s_C + tid * L;
      constexpr int kB = 8;
      double R = 0., S = 0.;
      int i0 = 0;
      for (; i0 + kB <= n; i0 += kB) {
        double ta[kB], tb[kB];
#pragma unroll
        for (int t = 0; t < kB; ++t) {
          ta[t] = A_[i0 + t];
          tb[t] = B_[i0 + t];
        }
#pragma unroll
        for (int t = 0; t < kB; ++t) {
          R += ta[t];
          S += tb[t];
        }
      }
      for (; i0 < n; ++i0) {
        R += A_[i0];
        S += B_[i0];
      }
      double M = 0.;
      double avail = C_[0];
      if (avail < 0) avail = 0;
      C_[0] = 0.;
      // one face: M_i stored, M_{i+1} formed, then the face above cell i
      // (i + 1 < n) clamped
      auto face = [&](int i, double phi, double wv, double av_up) {
        A_[i] = M;
        double wvp = phi - R * (wv / S);
        M -= wvp;
        if (i + 1 < n) {
          double avail_up = av_up;
          if (avail_up < 0) avail_up = 0;
          double mf = M;
          double q = mf > 0 ? mf * 1. : mf * 1.;
          if (q > avail) q = avail;
          if (q < -avail_up) q = -avail_up;
          C_[i + 1] = q;
          avail = avail_up + q;
        }
      };
      i0 = 0;
      for (; i0 + kB < n; i0 += kB) {
        double ta[kB], tb[kB], tc[kB];
#pragma unroll
        for (int t = 0; t < kB; ++t) {
          ta[t] = A_[i0 + t];
          tb[t] = B_[i0 + t];
          tc[t] = C_[i0 + t + 1];
        }
#pragma unroll
        for (int t = 0; t < kB; ++t) face(i0 + t, ta[t], tb[t], tc[t]);
      }
      for (; i0 < n; ++i0) face(i0, A_[i0], B_[i0], i0 + 1 < n ? C_[i0 + 1] : 0.);
      A_[n] = 0.;
      C_[n] = 0.;
    }
    __syncthreads();
    // ---- outputs, the tangential transfers, remainder, RK combine ------------
    for (int e = tid; e < cells; e += kFinThreads) {
      int c = e / n, i = e - c * n;
      if (!s_in[c]) continue;
      int64_t row = s_row[c];
      auto W = [&](int v, int l) { return w[v * r.sv + row + l]; };
      auto DU = [&](int v, int l) { return du[v * r.sv + row + l]; };
      auto tab = [&](double const* t, int rowi, int l) { return t[rowi * nc1 + lo1 + l]; };
      auto combine = [&](int v, double x) {
        int64_t o = v * r.sv + row + i;
        if constexpr (PAIR) {
          double y = a.w0 * (double(static_cast<float const*>(u0v)[o]) + double(c0[o])) +
                     a.w1 * (double(static_cast<float*>(uv)[o]) + double(cc[o])) + a.w2 * x;
          float h = float(y);
          static_cast<float*>(uv)[o] = h;
          cc[o] = float(y - double(h));
        } else {
          double* u = static_cast<double*>(uv);
          u[o] = a.w0 * static_cast<double const*>(u0v)[o] + a.w1 * u[o] + a.w2 * x;
        }
      };
      if (a.moved) a.moved[row + i] = s_ok[c] ? s_C[c * L + i] : 0.;
      if (!s_ok[c]) {
        for (int v = 0; v < 5; ++v) combine(v, DU(v, i));
        continue;
      }
      double const* ycol = yc + ((col0 + c) * 3) * n;
      double const* s_M = s_A + c * L;
      double const* s_q = s_C + c * L;
      // the mass change of cell l: the face below adds, the face above removes
      auto dm = [&](int l) { return (0. + s_q[l] / V) - s_q[l + 1] / V; };
      auto moved_rho = [&](int l) {
        double d0 = DU(IDN, l);
        return (d0 + dm(l)) - d0;
      };
      auto raw_minus_mass0 = [&](int l) { return ycol[l] - DU(IDN, l); };
      auto rows_at = [&](int l, auto q) {
        double s = tab(vtab, V_KMID, l) * q(l);
        if (l > 0) s = s + tab(vtab, V_KLO, l) * q(l - 1);
        if (l < n - 1) s = s + tab(vtab, V_KHI, l) * q(l + 1);
        if (l == 0) s = s + a.t0 * q(2);
        if (l == n - 1) s = s + a.t1 * q(n - 3);
        return s;
      };
      double out[5];
      double y0 = ycol[i], d0 = DU(IDN, i);
      out[IDN] = d0 + dm(i);
      out[IVX] = ycol[n + i];
      out[IPR] = ycol[2 * n + i];
      // vic_tangential_momentum_column: the face below adds (its upper cell),
      // then the face above removes (its lower cell), in the column's order
      double q_lo = s_q[i], q_hi = s_q[i + 1];
      for (int m = IVY; m <= IVZ; ++m) {
        double x = DU(m, i);
        if (i > 0) {
          double pv = q_lo * W(m, q_lo > 0 ? i - 1 : i);
          if (pv != 0.) x = x + pv / V;
        }
        if (i + 1 < n) {
          double pv = q_hi * W(m, q_hi > 0 ? i : i + 1);
          if (pv != 0.) x = x - pv / V;
        }
        out[m] = x;
      }
      double req_lo = s_M[i], req_hi = s_M[i + 1];
      double mv_lo = s_q[i] - req_lo;
      double mv_hi = s_q[i + 1] - req_hi;
      double projected = (req_lo - req_hi) / V;
      double pw = tab(vtab, V_PHI, i) * (y0 - d0 - projected);
      double cw = -(tab(vtab, V_DPHI, i) * mv_hi - tab(vtab, V_DPLO, i) * mv_lo) / V;
      out[IPR] = out[IPR] + (pw + cw);
      int jj = (i < 1 ? 1 : (i > n - 2 ? n - 2 : i)) - 1;
      double s = tab(xtab, T_WT0, i) * moved_rho(jj);
      s = s + tab(xtab, T_WT1, i) * moved_rho(jj + 1);
      s = s + tab(xtab, T_WT2, i) * moved_rho(jj + 2);
      double work = tab(xtab, T_GVAR, i) * s;
      work = work - rows_at(i, raw_minus_mass0);
      out[IPR] = out[IPR] + work;
      for (int v = 0; v < 5; ++v) combine(v, out[v]);
    }
    if constexpr (PAIR) {
      if (a.wout) {
        // the primitives of the combined pair, as cons2prim_f32 forms them
        // (pair_value, cons2prim5), once every thread of the workgroup has
        // read its columns' w and stored its pairs
        __syncthreads();
        float const* uf = static_cast<float const*>(uv);
        for (int e = tid; e < cells; e += kFinThreads) {
          int c = e / n, i = e - c * n;
          if (!s_in[c]) continue;
          int64_t o = s_row[c] + i;
          double uu[5], ww[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) uu[v] = double(uf[v * r.sv + o]) + double(cc[v * r.sv + o]);
          cons2prim5(uu, ww, a.gamma);
#pragma unroll
          for (int v = 0; v < 5; ++v) a.wout[v * r.sv + o] = ww[v];
        }
      }
    }
  }
};

template <bool PAIR>
__global__ void __launch_bounds__(kFinThreads)
    vicp_finish_kernel(double const* w, double const* __restrict__ du,
                       void const* __restrict__ u0v, float const* __restrict__ c0,
                       void* __restrict__ uv, float* __restrict__ cc,
                       double const* __restrict__ vtab, double const* __restrict__ xtab,
                       double const* __restrict__ yc, int const* __restrict__ status, PartArgs a,
                       int nc) {
  VicpFinishBody<PAIR>::run(block_index(), grid_size(), w, du, u0v, c0, uv, cc, vtab, xtab, yc,
                            status, a, nc);
}

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

static int vicp_launch(const snapy_grid* g, const snapy_physics* p, const double* w,
                       const double* du, const void* u0, const float* c0, void* u,
                       float* c, const double* vtab, const double* x1tab, double t0,
                       double t1, double dt, const double* dt_dev, double wstage,
                       double w0, double w1, double w2, double* ws, int* status,
                       double* moved, void* stream, double* wout = nullptr) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !du || !u0 || !u || !vtab || !x1tab || !ws || !status)
    return fail(-1, "vic_partial: NULL argument");
  bool pair = c0 != nullptr;
  if (pair && !c) return fail(-1, "vic_partial: the FP32 pair needs c0 and c");
  if (g->nx1 < 5) return fail(-9, "vic_partial: needs >= 5 x1 cells");
  PartArgs a;
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
  a.wout = wout;
  int64_t ncol = int64_t(g->nx2) * g->nx3;
  if (ncol > (int64_t(1) << 30)) return fail(-9, "vic_partial: too many columns");
  // the records (kF per layer, lane-contiguous) then yc (3 per layer)
  double* yc = ws + int64_t(kF) * g->nx1 * ncol;
  launch<VicpForwardBody, 128, 2>(vicp_forward_kernel, dim3(blocks_for(ncol, 64)), dim3(128), 0,
                                  as_stream(stream), w, du, vtab, ws, yc, status, a);
  if (int e = launch_status("vic_partial_forward")) return e;
  // columns per finish workgroup: the tuned value (default kFinCols), fewer
  // for a tall column
  int nc = int(tuned(SNAPY_TUNE_VIC_COLS, g, pair ? SNAPY_F32 : SNAPY_F64, kFinCols));
  if (nc > kFinCols) nc = kFinCols;
  while (nc > 1 && size_t(3) * nc * (g->nx1 + 1) * sizeof(double) > 64 * 1024) nc /= 2;
  size_t lds = size_t(3) * nc * (g->nx1 + 1) * sizeof(double);
  if (lds > 64 * 1024) return fail(-9, "vic_partial: the column does not fit the LDS");
  dim3 fgrid(unsigned((ncol + nc - 1) / nc));
  if (pair)
    launch<VicpFinishBody<true>, kFinThreads, 1>(vicp_finish_kernel<true>, fgrid, dim3(kFinThreads),
                                                 lds, as_stream(stream), w, du, u0, c0, u, c, vtab,
                                                 x1tab, yc, status, a, nc);
  else
    launch<VicpFinishBody<false>, kFinThreads, 1>(vicp_finish_kernel<false>, fgrid,
                                                  dim3(kFinThreads), lds, as_stream(stream), w, du,
                                                  u0, c0, u, c, vtab, x1tab, yc, status, a, nc);
  return launch_status("vic_partial");
}

int64_t snapy_vic_partial_workspace(const snapy_grid* g) {
  if (!g) return 0;
  // the lane-contiguous records and the column-major solution
  return int64_t(kRec + 3) * g->nx1 * g->nx2 * g->nx3;
}

int snapy_vic_partial(const snapy_grid* g, const snapy_physics* p, const double* w,
                      const double* du, const double* u0, double* u, const double* vtab,
                      const double* x1tab, double t0, double t1, double dt,
                      const double* dt_dev, double wstage, double w0, double w1,
                      double w2, double* ws, int* status, double* moved, void* stream) {
  return vicp_launch(g, p, w, du, u0, nullptr, u, nullptr, vtab, x1tab, t0, t1, dt, dt_dev,
                     wstage, w0, w1, w2, ws, status, moved, stream);
}

int snapy_vic_partial_f32(const snapy_grid* g, const snapy_physics* p, const double* w,
                          const double* du, const float* u0, const float* c0, float* u,
                          float* c, const double* vtab, const double* x1tab, double t0,
                          double t1, double dt, const double* dt_dev, double wstage,
                          double w0, double w1, double w2, double* ws, int* status,
                          double* moved, void* stream) {
  if (!c0 || !c) return fail(-1, "vic_partial_f32: NULL compensation");
  return vicp_launch(g, p, w, du, u0, c0, u, c, vtab, x1tab, t0, t1, dt, dt_dev, wstage, w0,
                     w1, w2, ws, status, moved, stream);
}

int snapy_vic_partial_prim_f32(const snapy_grid* g, const snapy_physics* p, double* w,
                               const double* du, const float* u0, const float* c0, float* u,
                               float* c, const double* vtab, const double* x1tab, double t0,
                               double t1, double dt, const double* dt_dev, double wstage,
                               double w0, double w1, double w2, double* ws, int* status,
                               double* moved, void* stream) {
  if (!c0 || !c) return fail(-1, "vic_partial_prim_f32: NULL compensation");
  if (g && g->nvar != 5) return fail(-6, "vic_partial_prim_f32: ideal gas needs nvar 5");
  return vicp_launch(g, p, w, du, u0, c0, u, c, vtab, x1tab, t0, t1, dt, dt_dev, wstage, w0,
                     w1, w2, ws, status, moved, stream, w);
}

}  // extern "C"
