// Passive scalars (tracers) of a Cartesian block, FP64 on gfx950: the face
// fluxes (ScalarImpl::forward, src/scalar/scalar.cpp:104-174), the scalars'
// positivity limiter (PositivityLimiter, src/hydro/flux_positivity.cpp:
// 39-268, as scalar.cpp:193-205 calls it), the divergence, the implicit
// tracer transfer with its donor cap (src/mesh/meshblock_advance.cpp:
// 212-233, src/implicit/implicit_dispatch.cu:69-87), the RK combine and the
// primitive r = s / rho (meshblock_advance.cpp:338-341,
// src/mesh/meshblock.cpp:67-79).
//
// The tracer density s = rho_dry r is stored FP64 in both state forms, as
// are its fluxes (the reference keeps the scalar register FP64,
// scalar.cpp:75-79).  The dry density of the stage is the primitive density
// w[IDN] (cons2prim copies u[IDN]; in the FP32 form its ghosts are the ones
// the primitive boundary fill leaves).
//
// Kernels:
//   * scalar_flux_kernel<DIR, TF>: one thread per face of direction DIR:
//     the scaled WENO5 faces of rho and s (clamped at 0), their ratio (the
//     donor cell's r where the density face is 0), the upwind flux on the
//     hydro's mass flux, on x2/x3 the covariance row (the face states of
//     the x1 neighbours formed again in the thread), the wall faces 0;
//   * scalar_theta_kernel<TF>: one thread per cell: theta of every scalar
//     row relaxed by the dry air's theta and the neighbours' share (1 on
//     ghosts; the boundary fill then sets the ghosts);
//   * scalar_update_kernel<COMBINE>: one thread per interior cell: the
//     divergence of the theta-scaled fluxes, ds = -dt div, and without an
//     implicit transfer the RK combine;
//   * scalar_transfer_kernel: one thread per (row, column): the donor-capped
//     transfer P of the implicit solve's dry-gas transfer, sequential up the
//     column, and the RK combine;
//   * scalar_prim_kernel<TU>: r = s / rho of the new state (interior).
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

// clamp_min(0) keeping -0 and NaN (torch's clamp)
__device__ __forceinline__ double clamp0(double x) { return x < 0. ? 0. : x; }
// torch.minimum / numpy minimum: NaN if either is NaN
__device__ __forceinline__ double nmin(double a, double b) {
  if (a != a || b != b) return a + b;
  return a < b ? a : b;
}
// numpy maximum(x, 0) of the drains
__device__ __forceinline__ double relu(double x) {
  if (x != x) return x;
  return x > 0. ? x : 0.;
}

// WENO5 with scale = true (recon_point.h:69-83, 113-147): vscale = mean |v|,
// the weights of v / vscale, 0 where vscale is 0; faces clamped at 0
// (reconstruct.cpp:262-266)
__device__ __forceinline__ double weno5s_plus(double v0, double v1, double v2, double v3,
                                              double v4) {
  double sum = fabs(v0);
  sum += fabs(v1);
  sum += fabs(v2);
  sum += fabs(v3);
  sum += fabs(v4);
  double vs = sum / 5.;
  if (vs == 0.) return 0.;
  double num, den;
  weno5_plus_parts(v0 / vs, v1 / vs, v2 / vs, v3 / vs, v4 / vs, num, den);
  return clamp0(vs * num / den);
}

__device__ __forceinline__ double weno5s_minus(double v0, double v1, double v2, double v3,
                                               double v4) {
  double sum = fabs(v0);
  sum += fabs(v1);
  sum += fabs(v2);
  sum += fabs(v3);
  sum += fabs(v4);
  double vs = sum / 5.;
  if (vs == 0.) return 0.;
  double num, den;
  weno5_minus_parts(v0 / vs, v1 / vs, v2 / vs, v3 / vs, v4 / vs, num, den);
  return clamp0(vs * num / den);
}

// face_r (scalar.cpp:114-119) at the face below cell c along stride st:
// the ratio of the scaled faces, the donor cell's r where the density face
// is not positive (cell c - 1 on the left, cell c on the right)
__device__ __forceinline__ void face_ratio(double const* rho, double const* s, int64_t st,
                                           double& rl, double& rr) {
  double p[6], q[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    p[t] = rho[(t - 3) * st];
    q[t] = s[(t - 3) * st];
  }
  double rfl = weno5s_plus(p[0], p[1], p[2], p[3], p[4]);
  double rfr = weno5s_minus(p[1], p[2], p[3], p[4], p[5]);
  double sfl = weno5s_plus(q[0], q[1], q[2], q[3], q[4]);
  double sfr = weno5s_minus(q[1], q[2], q[3], q[4], q[5]);
  rl = rfl > 0. ? sfl / rfl : q[2] / p[2];
  rr = rfr > 0. ? sfr / rfr : q[3] / p[3];
}

// the hydro's face state of row v at the face below cell c (the x2/x3
// reconstruction: weno5 for the density and with shock, else cp5)
__device__ __forceinline__ void hydro_face(double const* q, int64_t st, bool weno, double& l,
                                           double& r) {
  double c0 = q[-3 * st], c1 = q[-2 * st], c2 = q[-st], c3 = q[0], c4 = q[st], c5 = q[2 * st];
  if (weno) {
    l = weno5_plus(c0, c1, c2, c3, c4);
    r = weno5_minus(c1, c2, c3, c4, c5);
  } else {
    l = cp5_plus(c0, c1, c2, c3, c4);
    r = cp5_minus(c1, c2, c3, c4, c5);
  }
}

struct FluxArgs {
  Range3 r;
  int ns;
  int nf;          // faces along the sweep: n + 1
  int64_t st;      // stride along the sweep
  int cov;         // the covariance row on x2/x3
  int shock;       // the hydro's x2/x3 reconstruction is weno5 on every row
  int wall_lo, wall_hi;
  double s2;       // sigma1^2 = dx1^2 / 12
};

// one thread per face: faces lo..lo+n along DIR, the interior across it
template <int DIR, typename TF>
struct ScalarFluxBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ s,
                                             TF const* __restrict__ hflux,
                                             double const* __restrict__ x1v,
                                             double* __restrict__ sflux, FluxArgs a) {
    Range3 const& r = a.r;
    int n1 = DIR == 1 ? a.nf : r.n1, n2 = DIR == 2 ? a.nf : r.n2, n3 = DIR == 3 ? a.nf : r.n3;
    int64_t tot = int64_t(n1) * n2 * n3;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int i = r.lo1 + int(idx % n1);
    int64_t rest = idx / n1;
    int j = r.lo2 + int(rest % n2);
    int k = r.lo3 + int(rest / n2);
    int f = DIR == 1 ? i - r.lo1 : (DIR == 2 ? j - r.lo2 : k - r.lo3);
    int64_t c = k * r.s3 + j * r.s2 + i;
    bool wall = (f == 0 && a.wall_lo) || (f == a.nf - 1 && a.wall_hi);
    double m = double(hflux[IDN * r.sv + c]);
    double const* rho = w + IDN * r.sv;
    // the covariance's hydro parts, shared by every row: rho at the face and
    // the normal velocity of the x1 neighbours' faces
    double crho = 0., d1un = 0., dx1c = 1.;
    bool cov = DIR > 1 && a.cov;
    if (cov) {
      int vn = DIR == 2 ? IVY : IVZ;
      double rl, rr, ulm, urm, ulp, urp;
      hydro_face(rho + c, a.st, true, rl, rr);
      crho = 0.5 * (rl + rr);
      hydro_face(w + vn * r.sv + c - 1, a.st, a.shock != 0, ulm, urm);
      hydro_face(w + vn * r.sv + c + 1, a.st, a.shock != 0, ulp, urp);
      dx1c = x1v[i + 1] - x1v[i - 1];
      d1un = (0.5 * (ulp + urp) - 0.5 * (ulm + urm)) / dx1c;
    }
    for (int v = 0; v < a.ns; ++v) {
      double const* sv = s + v * r.sv;
      double rl, rr;
      face_ratio(rho + c, sv + c, a.st, rl, rr);
      // the upwind solver (riemann.cpp:164-171)
      double ui = m > 0. ? 1. : 0.;
      double F = m * (ui * rl + (1. - ui) * rr);
      if (cov) {
        // HydroImpl::scalar_flux_covariance (hydro_forward.cpp:233-297):
        // sigma1^2 rho D1[u_n] D1[q], q the mean of the face ratios
        double lm, rm, lp, rp;
        face_ratio(rho + c - 1, sv + c - 1, a.st, lm, rm);
        face_ratio(rho + c + 1, sv + c + 1, a.st, lp, rp);
        double qm = 0.5 * (lm * 1. + rm * 1.), qp = 0.5 * (lp * 1. + rp * 1.);
        double d1q = (qp - qm) / dx1c;
        F = F + a.s2 * crho * d1un * d1q;
      }
      // boundary.prescribed: every tracer row is 0 on a wall face
      sflux[v * r.sv + c] = wall ? 0. : F;
    }
  }
};

template <int DIR, typename TF>
__global__ void __launch_bounds__(256)
    scalar_flux_kernel(double const* __restrict__ w, double const* __restrict__ s,
                       TF const* __restrict__ hflux, double const* __restrict__ x1v,
                       double* __restrict__ sflux, FluxArgs a) {
  ScalarFluxBody<DIR, TF>::run(block_index(), grid_size(), w, s, hflux, x1v, sflux, a);
}

// ---------------------------------------------------------------------------
// positivity limiter
// ---------------------------------------------------------------------------
struct ThetaArgs {
  Range3 r;
  int ns;
  int act[3];
  double area[3], vol;
  double margin;  // 4096 eps
  double const* dt_dev;
  double dt;
};

// the outgoing A F of the cell at offset c over the active directions, in
// direction order (flux_positivity.cpp:46-59)
template <typename F>
__device__ __forceinline__ double drain_of(ThetaArgs const& a, int64_t c, F const& flux) {
  int64_t st[3] = {1, a.r.s2, a.r.s3};
  double out = 0.;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!a.act[d]) continue;
    double hi = a.area[d] * flux(d, c + st[d]);
    double lo = a.area[d] * flux(d, c);
    out += relu(hi) + relu(-lo);
  }
  return out;
}

__device__ __forceinline__ double theta_of(double x, double out, double vol, double margin,
                                           double dt) {
  double avail = (x > 0. ? x : 0.) * vol * (1. - margin);
  double drain = out * dt;
  double tiny = 2.2250738585072014e-308;
  if (!(drain > 0.)) return 1.;
  double t = avail / (drain > tiny ? drain : tiny);
  return nmin(t, 1.);
}

template <typename TF>
struct ScalarThetaBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ s,
                                             TF const* __restrict__ h1, TF const* __restrict__ h2,
                                             TF const* __restrict__ h3,
                                             double const* __restrict__ f1,
                                             double const* __restrict__ f2,
                                             double const* __restrict__ f3,
                                             double* __restrict__ theta, ThetaArgs a) {
    Range3 const& r = a.r;
    int64_t tot = int64_t(r.nc1) * r.nc2 * r.nc3;
    int64_t c = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (c >= tot) return;
    int i = int(c % r.nc1);
    int64_t rest = c / r.nc1;
    int j = int(rest % r.nc2);
    int k = int(rest / r.nc2);
    bool inner = i >= r.lo1 && i < r.lo1 + r.n1 && j >= r.lo2 && j < r.lo2 + r.n2 &&
                 k >= r.lo3 && k < r.lo3 + r.n3;
    if (!inner) {
      for (int v = 0; v < a.ns; ++v) theta[v * r.sv + c] = 1.;
      return;
    }
    double dt = dt_at(a.dt_dev, a.dt, r, k);
    TF const* hf[3] = {h1, h2, h3};
    double const* sf[3] = {f1, f2, f3};
    // the dry air's theta from the mass fluxes
    double rho = w[IDN * r.sv + c];
    double outr = drain_of(a, c, [&](int d, int64_t o) { return double(hf[d][IDN * r.sv + o]); });
    double td = theta_of(rho, outr, a.vol, a.margin, dt);
    for (int v = 0; v < a.ns; ++v) {
      int64_t vo = v * r.sv;
      double x = s[vo + c];
      double outs = drain_of(a, c, [&](int d, int64_t o) { return sf[d][vo + o]; });
      double th = theta_of(x, outs, a.vol, a.margin, dt);
      // flux_positivity_share (flux_positivity.cpp:80-88): q = s / rho, the
      // neighbours along x1, then x2 and x3 when resolved
      auto q = [&](int64_t o) { return s[vo + o] / w[IDN * r.sv + o]; };
      double qc = q(c);
      // a single-cell axis rolls onto itself: the cell's own q
      double low = r.nc1 > 1 ? nmin(q(c - 1), q(c + 1)) : nmin(qc, qc);
      if (r.nc2 > 1) low = nmin(low, nmin(q(c - r.s2), q(c + r.s2)));
      if (r.nc3 > 1) low = nmin(low, nmin(q(c - r.s3), q(c + r.s3)));
      double phi = 0.;
      if (qc > 0.) {
        double t = low / qc;
        phi = t != t ? t : (t < 0. ? 0. : (t > 1. ? 1. : t));
      }
      // flux_positivity_relax (flux_positivity.cpp:143-149)
      if (td < 1. && td > 0.) th = th + phi * (nmin(th / td, 1.) - th);
      theta[vo + c] = th;
    }
  }
};

template <typename TF>
__global__ void __launch_bounds__(256)
    scalar_theta_kernel(double const* __restrict__ w, double const* __restrict__ s,
                        TF const* __restrict__ h1, TF const* __restrict__ h2,
                        TF const* __restrict__ h3, double const* __restrict__ f1,
                        double const* __restrict__ f2, double const* __restrict__ f3,
                        double* __restrict__ theta, ThetaArgs a) {
  ScalarThetaBody<TF>::run(block_index(), grid_size(), w, s, h1, h2, h3, f1, f2, f3, theta, a);
}

// ---------------------------------------------------------------------------
// divergence, implicit transfer, RK combine, primitive
// ---------------------------------------------------------------------------
struct UpdArgs {
  Range3 r;
  int ns;
  int act[3];
  double area[3], vol;
  double const* dt_dev;
  double dt;
  double w0, w1, w2;
};

template <bool COMBINE>
struct ScalarUpdateBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ f1,
                                             double const* __restrict__ f2,
                                             double const* __restrict__ f3,
                                             double const* __restrict__ theta,
                                             double const* __restrict__ s0, double* __restrict__ s,
                                             double* __restrict__ ds_out, UpdArgs a) {
    Range3 const& r = a.r;
    int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int i = r.lo1 + int(idx % r.n1);
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    int64_t c = k * r.s3 + j * r.s2 + i;
    double dt = dt_at(a.dt_dev, a.dt, r, k);
    double const* fl[3] = {f1, f2, f3};
    int64_t st[3] = {1, r.s2, r.s3};
    for (int v = 0; v < a.ns; ++v) {
      int64_t vo = v * r.sv;
      // CoordinateImpl::divergence in direction order, each face flux scaled
      // by its donor's theta (flux_positivity_factor: the lower cell where
      // F > 0)
      double div = 0.;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (!a.act[d]) continue;
        double lo = fl[d][vo + c], hi = fl[d][vo + c + st[d]];
        if (theta) {
          lo = lo * (lo > 0. ? theta[vo + c - st[d]] : theta[vo + c]);
          hi = hi * (hi > 0. ? theta[vo + c] : theta[vo + c + st[d]]);
        }
        div = div + (a.area[d] * hi - a.area[d] * lo);
      }
      double ds = -dt * (div / a.vol);
      if constexpr (COMBINE)
        s[vo + c] = a.w0 * s0[vo + c] + a.w1 * s[vo + c] + a.w2 * ds;
      else
        ds_out[vo + c] = ds;
    }
  }
};

template <bool COMBINE>
__global__ void __launch_bounds__(256)
    scalar_update_kernel(double const* __restrict__ f1, double const* __restrict__ f2,
                         double const* __restrict__ f3, double const* __restrict__ theta,
                         double const* __restrict__ s0, double* __restrict__ s,
                         double* __restrict__ ds_out, UpdArgs a) {
  ScalarUpdateBody<COMBINE>::run(block_index(), grid_size(), f1, f2, f3, theta, s0, s, ds_out, a);
}

// one thread per (row, column): P = r_donor M through the face below each
// cell, the donor cap sequentially upward (donor_cap_cuda), then
// ds + (P - P_above) / V and the combine, cell by cell as the cap passes
struct ScalarTransferBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ moved,
                                             double const* __restrict__ rr,
                                             double const* __restrict__ s0, double* __restrict__ s,
                                             double const* __restrict__ ds, UpdArgs a) {
    Range3 const& r = a.r;
    int64_t ncol = int64_t(r.n2) * r.n3;
    int64_t t = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (t >= ncol * a.ns) return;
    int v = int(t / ncol);
    int64_t col = t - v * ncol;
    int j = r.lo2 + int(col % r.n2), k = r.lo3 + int(col / r.n2);
    int64_t row = k * r.s3 + j * r.s2;
    int64_t vo = v * r.sv;
    double const V = a.vol;
    int const il = r.lo1, n = r.n1;
    double avail = 0., Pprev = 0., sprev = 0., dprev = 0., s0prev = 0.;
    for (int l = 0; l < n; ++l) {
      int64_t c = row + il + l;
      double M = moved[c];
      double P = (M > 0. ? rr[vo + c - 1] : rr[vo + c]) * M;
      double sc = s[vo + c], dc = ds[vo + c], s0c = s0[vo + c];
      double x = sc + dc;
      double held = (x < 0. ? 0. : x) * V;
      if (l == 0) {
        avail = held;
      } else {
        double cap = P < avail ? P : avail;
        P = cap > -held ? cap : -held;
        avail = held + P;
        // the cell below is final: its transfer in (Pprev) and out (P)
        double d = dprev + (Pprev - P) / V;
        s[vo + c - 1] = a.w0 * s0prev + a.w1 * sprev + a.w2 * d;
      }
      Pprev = P;
      sprev = sc;
      dprev = dc;
      s0prev = s0c;
    }
    // the top cell: nothing leaves through the column's top face
    double d = dprev + (Pprev - 0.) / V;
    s[vo + row + il + n - 1] = a.w0 * s0prev + a.w1 * sprev + a.w2 * d;
  }
};

__global__ void __launch_bounds__(128)
    scalar_transfer_kernel(double const* __restrict__ moved, double const* __restrict__ rr,
                           double const* __restrict__ s0, double* __restrict__ s,
                           double const* __restrict__ ds, UpdArgs a) {
  ScalarTransferBody::run(block_index(), grid_size(), moved, rr, s0, s, ds, a);
}

template <typename TU>
struct ScalarPrimBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             TU const* __restrict__ u,
                                             float const* __restrict__ comp,
                                             double const* __restrict__ s, double* __restrict__ rr,
                                             int ns, Range3 r) {
    int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int i = r.lo1 + int(idx % r.n1);
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    int64_t c = k * r.s3 + j * r.s2 + i;
    double rho = double(u[IDN * r.sv + c]);
    if (comp) rho = rho + double(comp[IDN * r.sv + c]);
    for (int v = 0; v < ns; ++v) rr[v * r.sv + c] = s[v * r.sv + c] / rho;
  }
};

template <typename TU>
__global__ void __launch_bounds__(256)
    scalar_prim_kernel(TU const* __restrict__ u, float const* __restrict__ comp,
                       double const* __restrict__ s, double* __restrict__ rr, int ns, Range3 r) {
  ScalarPrimBody<TU>::run(block_index(), grid_size(), u, comp, s, rr, ns, r);
}

int scalar_checks(const snapy_grid* g, const snapy_physics* p, int ns, char const* what) {
  if (int e = check_grid(g)) return e;
  if (!p) return fail(-1, std::string(what) + ": NULL physics");
  if (ns < 1) return fail(-13, std::string(what) + ": ns must be >= 1");
  if (member_axis(*g)) return fail(-4, std::string(what) + ": no member axis with scalars");
  if (p->eos != SNAPY_EOS_IDEAL_GAS) return fail(-4, std::string(what) + ": ideal gas only");
  return 0;
}

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

int snapy_scalar_flux(int dir, const snapy_grid* g, const snapy_physics* p, int ns,
                      const double* w, const double* s, int fdtype, const void* hflux,
                      const double* x1tab, double* sflux, void* stream) {
  if (int e = scalar_checks(g, p, ns, "scalar_flux")) return e;
  if (!w || !s || !hflux || !sflux) return fail(-1, "scalar_flux: NULL argument");
  if (dir < 1 || dir > 3) return fail(-5, "scalar_flux: dir must be 1, 2 or 3");
  if (!axis_active(*g, dir)) return fail(-5, "scalar_flux: inactive direction");
  if (fdtype != SNAPY_F64 && fdtype != SNAPY_F32) return fail(-10, "unknown dtype");
  FluxArgs a;
  a.r = ranges(*g);
  a.ns = ns;
  int nd = dir == 1 ? g->nx1 : (dir == 2 ? g->nx2 : g->nx3);
  a.nf = nd + 1;
  a.st = dir == 1 ? 1 : (dir == 2 ? a.r.s2 : a.r.s3);
  a.cov = dir > 1 && p->flux_covariance && g->nc1 >= 3;
  if (a.cov && !x1tab) return fail(-1, "scalar_flux: the covariance needs x1tab");
  a.shock = p->shock_x23;
  a.wall_lo = p->bc[2 * (dir - 1)] == SNAPY_BC_REFLECTING;
  a.wall_hi = p->bc[2 * (dir - 1) + 1] == SNAPY_BC_REFLECTING;
  a.s2 = g->dx1 * g->dx1 / 12.;
  int64_t faces = int64_t(dir == 1 ? a.nf : g->nx1) * (dir == 2 ? a.nf : g->nx2) *
                  (dir == 3 ? a.nf : g->nx3);
  dim3 grid(blocks_for(faces, 256));
  hipStream_t st = as_stream(stream);
  auto go = [&](auto tf) {
    using TF = decltype(tf);
    TF const* h = static_cast<TF const*>(hflux);
    if (dir == 1)
      launch<ScalarFluxBody<1, TF>, 256, 1>((scalar_flux_kernel<1, TF>), grid, dim3(256), 0, st, w, s, h, x1tab, sflux, a);
    else if (dir == 2)
      launch<ScalarFluxBody<2, TF>, 256, 1>((scalar_flux_kernel<2, TF>), grid, dim3(256), 0, st, w, s, h, x1tab, sflux, a);
    else
      launch<ScalarFluxBody<3, TF>, 256, 1>((scalar_flux_kernel<3, TF>), grid, dim3(256), 0, st, w, s, h, x1tab, sflux, a);
  };
  if (fdtype == SNAPY_F64)
    go(double(0));
  else
    go(float(0));
  return launch_status("scalar_flux");
}

int snapy_scalar_theta(const snapy_grid* g, const snapy_physics* p, int ns, const double* w,
                       const double* s, int fdtype, const void* const hflux[3],
                       const double* const sflux[3], double dt, const double* dt_dev,
                       double* theta, void* stream) {
  if (int e = scalar_checks(g, p, ns, "scalar_theta")) return e;
  if (!w || !s || !hflux || !sflux || !theta) return fail(-1, "scalar_theta: NULL argument");
  if (fdtype != SNAPY_F64 && fdtype != SNAPY_F32) return fail(-10, "unknown dtype");
  ThetaArgs a;
  a.r = ranges(*g);
  a.ns = ns;
  for (int d = 0; d < 3; ++d) {
    a.act[d] = axis_active(*g, d + 1);
    if (a.act[d] && (!hflux[d] || !sflux[d])) return fail(-1, "scalar_theta: NULL flux");
  }
  a.area[0] = g->dx3 * g->dx2;
  a.area[1] = g->dx3 * g->dx1;
  a.area[2] = g->dx2 * g->dx1;
  a.vol = g->dx3 * (g->dx2 * g->dx1);
  a.margin = 4096. * 2.220446049250313e-16;
  a.dt_dev = dt_dev;
  a.dt = dt;
  int64_t tot = int64_t(g->nc1) * g->nc2 * g->nc3;
  dim3 grid(blocks_for(tot, 256));
  hipStream_t st = as_stream(stream);
  if (fdtype == SNAPY_F64)
    launch<ScalarThetaBody<double>, 256, 1>(scalar_theta_kernel<double>, grid, dim3(256), 0, st, w, s,
                       static_cast<double const*>(hflux[0]), static_cast<double const*>(hflux[1]),
                       static_cast<double const*>(hflux[2]), sflux[0], sflux[1], sflux[2], theta, a);
  else
    launch<ScalarThetaBody<float>, 256, 1>(scalar_theta_kernel<float>, grid, dim3(256), 0, st, w, s,
                       static_cast<float const*>(hflux[0]), static_cast<float const*>(hflux[1]),
                       static_cast<float const*>(hflux[2]), sflux[0], sflux[1], sflux[2], theta, a);
  if (int e = launch_status("scalar_theta")) return e;
  // the ghosts of theta filled as a scalar field (flux_positivity_sync_theta_)
  return snapy_scalar_boundaries(g, p, ns, theta, stream);
}

int snapy_scalar_update(const snapy_grid* g, const snapy_physics* p, int ns,
                        const double* const sflux[3], const double* theta, const double* r,
                        const double* moved, const double* s0, double* s, double* ds, double dt,
                        const double* dt_dev, double w0, double w1, double w2, void* stream) {
  if (int e = scalar_checks(g, p, ns, "scalar_update")) return e;
  if (!sflux || !s0 || !s) return fail(-1, "scalar_update: NULL argument");
  if (moved && (!r || !ds)) return fail(-1, "scalar_update: the implicit transfer needs r and ds");
  UpdArgs a;
  a.r = ranges(*g);
  a.ns = ns;
  for (int d = 0; d < 3; ++d) {
    a.act[d] = axis_active(*g, d + 1);
    if (a.act[d] && !sflux[d]) return fail(-1, "scalar_update: NULL flux");
  }
  a.area[0] = g->dx3 * g->dx2;
  a.area[1] = g->dx3 * g->dx1;
  a.area[2] = g->dx2 * g->dx1;
  a.vol = g->dx3 * (g->dx2 * g->dx1);
  a.dt_dev = dt_dev;
  a.dt = dt;
  a.w0 = w0;
  a.w1 = w1;
  a.w2 = w2;
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  hipStream_t st = as_stream(stream);
  if (!moved) {
    launch<ScalarUpdateBody<true>, 256, 1>(scalar_update_kernel<true>, dim3(blocks_for(tot, 256)), dim3(256), 0, st,
                       sflux[0], sflux[1], sflux[2], theta, s0, s, nullptr, a);
  } else {
    launch<ScalarUpdateBody<false>, 256, 1>(scalar_update_kernel<false>, dim3(blocks_for(tot, 256)), dim3(256), 0, st,
                       sflux[0], sflux[1], sflux[2], theta, s0, s, ds, a);
    if (int e = launch_status("scalar_update")) return e;
    int64_t cols = int64_t(g->nx2) * g->nx3 * ns;
    launch<ScalarTransferBody, 128, 1>(scalar_transfer_kernel, dim3(blocks_for(cols, 128)), dim3(128), 0, st, moved,
                       r, s0, s, ds, a);
  }
  if (int e = launch_status("scalar_update")) return e;
  return snapy_scalar_boundaries(g, p, ns, s, stream);
}

int snapy_scalar_prim(const snapy_grid* g, const snapy_physics* p, int ns, int udtype,
                      const void* u, const float* comp, const double* s, double* r,
                      void* stream) {
  if (int e = scalar_checks(g, p, ns, "scalar_prim")) return e;
  if (!u || !s || !r) return fail(-1, "scalar_prim: NULL argument");
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  dim3 grid(blocks_for(tot, 256));
  hipStream_t st = as_stream(stream);
  if (udtype == SNAPY_F64) {
    if (comp) return fail(-10, "scalar_prim: a compensation goes with the FP32 state");
    launch<ScalarPrimBody<double>, 256, 1>(scalar_prim_kernel<double>, grid, dim3(256), 0, st,
                       static_cast<double const*>(u), nullptr, s, r, ns, ranges(*g));
  } else if (udtype == SNAPY_F32) {
    launch<ScalarPrimBody<float>, 256, 1>(scalar_prim_kernel<float>, grid, dim3(256), 0, st,
                       static_cast<float const*>(u), comp, s, r, ns, ranges(*g));
  } else {
    return fail(-10, "unknown dtype");
  }
  if (int e = launch_status("scalar_prim")) return e;
  return snapy_scalar_boundaries(g, p, ns, r, stream);
}

}  // extern "C"
