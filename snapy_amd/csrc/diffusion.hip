// Isotropic Cartesian diffusion (forcing/diffusion, kinematic form, no x1
// profiles, stress-free walls), FP64 on gfx950:
//   * one thread per face of an active direction builds the viscous
//     momentum flux, its work in the energy row and the Fourier flux
//     (DiffusionImpl::cartesian_fluxes, src/forcing/diffusion.cpp:557-706),
//     keeps the momentum rows in a workspace and adds the energy row to the
//     hydro flux total of every face that is not a prescribed wall
//     (post_fluxes, diffusion_cubed_sphere.cpp:365-379, then
//     apply_prescribed, hydro_forward.cpp:930-935);
//   * one thread per interior cell turns the momentum rows into the stage
//     increment -dt div(F) (DiffusionImpl::add, diffusion.cpp:532-547),
//     which the stage increment kernels add as their du_src.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

struct DiffArgs {
  Range3 r;
  int ax;       // the face direction: 0 (x1), 1 (x2), 2 (x3)
  int nf;       // faces along it
  int act[3];   // active axes
  int64_t s[3]; // strides of the axes
  double dx[3];
  double nu, kappa, rd, cv;
  int xwall_lo, xwall_hi;  // reflecting x1 walls: extrapolated face coefficients
  int pres_lo, pres_hi;    // prescribed wall faces of this direction
};

template <typename TT>
struct DiffusionFaceBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double* __restrict__ work, TT* __restrict__ total,
                                             DiffArgs a) {
    Range3 const& r = a.r;
    int n[3] = {r.n1, r.n2, r.n3};
    int lo[3] = {r.lo1, r.lo2, r.lo3};
    n[a.ax] = a.nf;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(n[0]) * n[1] * n[2];
    if (idx >= tot) return;
    int c0 = int(idx % n[0]);
    int64_t rest = idx / n[0];
    int c1 = int(rest % n[1]);
    int c2 = int(rest / n[1]);
    int ci[3] = {lo[0] + c0, lo[1] + c1, lo[2] + c2};
    int f = ci[a.ax] - lo[a.ax];  // face index along the direction
    int64_t c = ci[2] * r.s3 + ci[1] * r.s2 + ci[0];  // the cell above the face
    int64_t sd = a.s[a.ax];
    int64_t cm = c - sd;  // the cell below
    auto W = [&](int v, int64_t q) { return w[v * r.sv + q]; };
    // centered_derivative (diffusion.cpp:73-95)
    auto cd = [&](int v, int64_t q, int ax) {
      return (W(v, q + a.s[ax]) - W(v, q - a.s[ax])) / (2. * a.dx[ax]);
    };
    auto divv = [&](int64_t q) {
      double d = 0.;
      for (int ax = 0; ax < 3; ++ax)
        if (a.act[ax]) d = d + cd(IVX + ax, q, ax);
      return d;
    };
    // face_coefficient (diffusion.cpp:164-185): the two-cell average; a
    // reflecting x1 wall face extrapolates from the two nearest active cells
    bool xwall = a.ax == 0 && a.nf >= 3 &&
                 ((f == 0 && a.xwall_lo) || (f == a.nf - 1 && a.xwall_hi));
    auto coef = [&](auto val) {
      if (!xwall) return 0.5 * (val(c) + val(cm));
      double t = 0.5;
      double near = f == 0 ? val(c) : val(cm);
      double next = f == 0 ? val(c + 1) : val(cm - 1);
      double wall = (1. + t) * near - t * next;
      return wall > 0. ? wall : near;
    };
    double const dxd = a.dx[a.ax];
    double out[4] = {0., 0., 0., 0.};  // momentum x1, x2, x3, energy
    double rho_face = coef([&](int64_t q) { return W(IDN, q); });
    if (a.nu > 0.) {
      double div_face = 0.5 * (divv(c) + divv(cm));
      for (int iv = 0; iv < 3; ++iv) {
        double stress = (W(IVX + iv, c) - W(IVX + iv, cm)) / dxd;
        if (iv == a.ax) {
          stress = 2. * stress - (2. / 3.) * div_face;
        } else if (a.act[iv]) {
          stress = stress + 0.5 * (cd(IVX + a.ax, c, iv) + cd(IVX + a.ax, cm, iv));
        }
        double mflux = -a.nu * rho_face * stress;
        out[iv] = mflux;
        out[3] = out[3] + 0.5 * (W(IVX + iv, c) + W(IVX + iv, cm)) * mflux;
      }
    }
    if (a.kappa > 0.) {
      auto temp = [&](int64_t q) { return W(IPR, q) / (W(IDN, q) * a.rd); };
      double dtdn = (temp(c) - temp(cm)) / dxd;
      double rcv = coef([&](int64_t q) { return W(IDN, q) * a.cv; });
      out[3] = out[3] - a.kappa * rcv * dtdn;
    }
    int64_t ncell = int64_t(r.nc1) * r.nc2 * r.nc3;
    for (int q = 0; q < 4; ++q) work[(a.ax * 4 + q) * ncell + c] = out[q];
    bool wall = (f == 0 && a.pres_lo) || (f == a.nf - 1 && a.pres_hi);
    if (!wall) total[IPR * r.sv + c] = TT(double(total[IPR * r.sv + c]) + out[3]);
  }
};

template <typename TT>
__global__ void __launch_bounds__(256)
    diffusion_face_kernel(double const* __restrict__ w, double* __restrict__ work,
                          TT* __restrict__ total, DiffArgs a) {
  DiffusionFaceBody<TT>::run(block_index(), grid_size(), w, work, total, a);
}

struct DiffDivArgs {
  Range3 r;
  int act[3];
  double a1, a2, a3, vol, mdt;
  double const* dt_dev;
};

struct DiffusionDuBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ work,
                                             double* __restrict__ dsrc, DiffDivArgs a) {
    Range3 const& r = a.r;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
    if (idx >= tot) return;
    int i = r.lo1 + int(idx % r.n1);
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    int64_t c = k * r.s3 + j * r.s2 + i;
    int64_t ncell = int64_t(r.nc1) * r.nc2 * r.nc3;
    double dt = dt_at(a.dt_dev, -a.mdt, r, k);
    double A[3] = {a.a1, a.a2, a.a3};
    int64_t s[3] = {1, r.s2, r.s3};
    dsrc[IDN * r.sv + c] = 0.;
    dsrc[IPR * r.sv + c] = 0.;
    for (int q = 0; q < 3; ++q) {
      // CoordinateImpl::divergence (coordinate.cpp:618-662)
      double div = 0.;
      bool first = true;
      for (int ax = 0; ax < 3; ++ax) {
        if (!a.act[ax]) continue;
        double const* F = work + (ax * 4 + q) * ncell;
        double d = A[ax] * F[c + s[ax]] - A[ax] * F[c];
        div = first ? d : div + d;
        first = false;
      }
      div = div / a.vol;
      dsrc[(IVX + q) * r.sv + c] = -(dt * div);
    }
  }
};

__global__ void __launch_bounds__(256)
    diffusion_du_kernel(double const* __restrict__ work, double* __restrict__ dsrc,
                        DiffDivArgs a) {
  DiffusionDuBody::run(block_index(), grid_size(), work, dsrc, a);
}

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

int64_t snapy_diffusion_workspace(const snapy_grid* g) {
  if (!g) return 0;
  return int64_t(12) * g->nc1 * g->nc2 * g->nc3;
}

int snapy_diffusion(const snapy_grid* g, const snapy_physics* p, int dtype,
                    const double* w, void* const totals[3], double* work,
                    double* du_src, double dt, const double* dt_dev,
                    void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !totals || !work || !du_src)
    return fail(-1, "diffusion: NULL argument");
  if (g->nvar != 5) return fail(-6, "diffusion: nvar 5");
  if (g->nghost < 2) return fail(-9, "diffusion: needs nghost >= 2");
  if (!(p->nu_iso >= 0.) || !(p->kappa_iso >= 0.))
    return fail(-8, "diffusion: coefficients must be >= 0");
  if (p->kappa_iso > 0. && !(p->gas_r > 0.))
    return fail(-8, "diffusion: heat conduction needs a gas constant > 0");
  if (dtype != SNAPY_F64 && dtype != SNAPY_F32) return fail(-10, "unknown dtype");
  Range3 r = ranges(*g);
  int act[3] = {g->nc1 > 1, g->nc2 > 1, axis_active(*g, 3)};
  int nx[3] = {g->nx1, g->nx2, g->nx3};
  double dx[3] = {g->dx1, g->dx2, g->dx3};
  for (int ax = 0; ax < 3; ++ax)
    if (act[ax] && !totals[ax]) return fail(-1, "diffusion: NULL flux total");
  hipStream_t s = as_stream(stream);
  for (int ax = 0; ax < 3; ++ax) {
    if (!act[ax]) continue;
    DiffArgs a;
    a.r = r;
    a.ax = ax;
    a.nf = nx[ax] + 1;
    for (int q = 0; q < 3; ++q) {
      a.act[q] = act[q];
      a.dx[q] = dx[q];
    }
    a.s[0] = 1;
    a.s[1] = r.s2;
    a.s[2] = r.s3;
    a.nu = p->nu_iso;
    a.kappa = p->kappa_iso;
    a.rd = p->gas_r;
    a.cv = p->gas_r / (p->gammad - 1.);
    a.xwall_lo = p->bc[0] == SNAPY_BC_REFLECTING;
    a.xwall_hi = p->bc[1] == SNAPY_BC_REFLECTING;
    a.pres_lo = p->bc[2 * ax] == SNAPY_BC_REFLECTING;
    a.pres_hi = p->bc[2 * ax + 1] == SNAPY_BC_REFLECTING;
    int64_t nfaces = int64_t(a.nf) * (ax == 0 ? 1 : nx[0]) *
                     (ax == 1 ? 1 : nx[1]) * (ax == 2 ? 1 : nx[2]);
    if (dtype == SNAPY_F64)
      launch<DiffusionFaceBody<double>, 256, 1>(diffusion_face_kernel<double>,
                         dim3(blocks_for(nfaces, 256)), dim3(256), 0, s, w,
                         work, static_cast<double*>(totals[ax]), a);
    else
      launch<DiffusionFaceBody<float>, 256, 1>(diffusion_face_kernel<float>,
                         dim3(blocks_for(nfaces, 256)), dim3(256), 0, s, w,
                         work, static_cast<float*>(totals[ax]), a);
    if (int e = launch_status("diffusion_faces")) return e;
  }
  DiffDivArgs d;
  d.r = r;
  for (int q = 0; q < 3; ++q) d.act[q] = act[q];
  d.a1 = g->dx3 * g->dx2;
  d.a2 = g->dx3 * g->dx1;
  d.a3 = g->dx2 * g->dx1;
  d.vol = g->dx3 * (g->dx2 * g->dx1);
  d.mdt = -dt;
  d.dt_dev = dt_dev;
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  launch<DiffusionDuBody, 256, 1>(diffusion_du_kernel, dim3(blocks_for(tot, 256)),
                     dim3(256), 0, s, work, du_src, d);
  return launch_status("diffusion_du");
}

}  // extern "C"
