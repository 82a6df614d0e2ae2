// The shallow-water stage on a Cartesian grid (eos shallow-water, Riemann
// shallow-roe, dir xy): rows (phi, phi v1, phi v2, phi v3), phi = g h.
//   * cons <-> prim (ShallowWaterImpl::_cons2prim/_prim2cons,
//     src/eos/shallow_water.cpp:31-55; _mom_from_vel_,
//     src/eos/equation_of_state.cpp:290-295);
//   * one thread per interior face: weno5 on phi, cp5 on the velocities
//     unless shock, then shallow_roe_impl (src/riemann/riemann_kernels.h:
//     447-495) with shallow_axes (:44-47);
//   * the CFL bound with the sound speed sqrt(phi) (shallow_water.cpp:20-22).
// The divergence, combine and boundaries are the row-generic snapy_update /
// snapy_apply_boundaries with nvar 4.
#include <hip/hip_runtime.h>

#include <cstring>

#include "common.hpp"
#include "point.hpp"

namespace snapy {
namespace {

constexpr int kSweRows = 4;

template <typename T>
__global__ void __launch_bounds__(256)
    swe_cons2prim_kernel(T const* __restrict__ u, T* __restrict__ w, int64_t n,
                         int64_t sv) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  T h = u[c];
  w[c] = h;
  for (int v = 1; v < kSweRows; ++v) w[v * sv + c] = u[v * sv + c] / h;
}

template <typename T>
__global__ void __launch_bounds__(256)
    swe_prim2cons_kernel(T const* __restrict__ w, T* __restrict__ u, int64_t n,
                         int64_t sv) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  T h = w[c];
  u[c] = h;
  for (int v = 1; v < kSweRows; ++v) u[v * sv + c] = w[v * sv + c] * h;
}

// shallow_roe_impl on one face; dim_axis 0 = x1, 1 = x2 (dir xy)
template <typename T>
__device__ __forceinline__ void shallow_roe(T const* wl, T const* wr, int ax,
                                            T* flx) {
  int ivx = ax == 0 ? 1 : 2, ivy = ax == 0 ? 2 : 1, ivz = 3;
  T sqrtdl = sqrt(wl[0]);
  T sqrtdr = sqrt(wr[0]);
  T isdlpdr = T(1) / (sqrtdl + sqrtdr);
  T ubar = (wl[ivx] * sqrtdl + wr[ivx] * sqrtdr) * isdlpdr;
  T vbar = (wl[ivy] * sqrtdl + wr[ivy] * sqrtdr) * isdlpdr;
  T cbar = sqrt(T(0.5) * (wl[0] + wr[0]));
  T del0 = wr[0] - wl[0];
  T delx = wr[ivx] - wl[ivx];
  T dely = wr[ivy] - wl[ivy];
  T hbar = sqrt(wl[0] * wr[0]);
  T a1 = T(0.5) * (cbar * del0 - hbar * delx) / cbar;
  T a2 = hbar * dely;
  T a3 = T(0.5) * (cbar * del0 + hbar * delx) / cbar;
  T wave0[4] = {0, 0, 0, 0}, wave1[4] = {0, 0, 0, 0}, wave2[4] = {0, 0, 0, 0};
  wave0[0] = a1;
  wave0[ivx] = a1 * (ubar - cbar);
  wave0[ivy] = a1 * vbar;
  wave1[ivy] = a2;
  wave2[0] = a3;
  wave2[ivx] = a3 * (ubar + cbar);
  wave2[ivy] = a3 * vbar;
  T speed0 = fabs(ubar - cbar);
  T speed1 = fabs(ubar);
  T speed2 = fabs(ubar + cbar);
  flx[0] = T(0.5) * (wl[0] * wl[ivx] + wr[0] * wr[ivx]);
  flx[ivx] = T(0.5) * (wl[0] * wl[ivx] * wl[ivx] + T(0.5) * wl[0] * wl[0] +
                       wr[0] * wr[ivx] * wr[ivx] + T(0.5) * wr[0] * wr[0]);
  flx[ivy] = T(0.5) * (wl[0] * wl[ivx] * wl[ivy] + wr[0] * wr[ivx] * wr[ivy]);
  flx[ivz] = T(0);
  for (int v = 0; v < 4; ++v)
    flx[v] -= T(0.5) * (speed0 * wave0[v] + speed1 * wave1[v] + speed2 * wave2[v]);
}

struct SweArgs {
  Range3 r;
  int ax;  // 0 = x1, 1 = x2
  int nf;  // faces along the sweep
  int shock;
  int wall_lo, wall_hi;
};

template <typename T>
__global__ void __launch_bounds__(256)
    swe_sweep_kernel(T const* __restrict__ w, T* __restrict__ flx, SweArgs a) {
  Range3 const& r = a.r;
  int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int n1 = a.ax == 0 ? a.nf : r.n1, n2 = a.ax == 1 ? a.nf : r.n2;
  int64_t tot = int64_t(n1) * n2 * r.n3;
  if (idx >= tot) return;
  int i = r.lo1 + int(idx % n1);
  int64_t rest = idx / n1;
  int j = r.lo2 + int(rest % n2);
  int k = r.lo3 + int(rest / n2);
  int64_t base = k * r.s3 + j * r.s2 + i;
  int64_t stride = a.ax == 0 ? 1 : r.s2;
  T wl[4], wr[4];
  for (int v = 0; v < kSweRows; ++v) {
    T const* p = w + v * r.sv + base;
    T c0 = p[-3 * stride], c1 = p[-2 * stride], c2 = p[-stride], c3 = p[0],
      c4 = p[stride], c5 = p[2 * stride];
    // phi always weno5; the velocities cp5 unless shock
    // (ReconstructImpl::forward, reconstruct.cpp:303-322)
    if (v == 0 || a.shock) {
      wl[v] = weno5_plus(c0, c1, c2, c3, c4);
      wr[v] = weno5_minus(c1, c2, c3, c4, c5);
    } else {
      wl[v] = cp5_plus(c0, c1, c2, c3, c4);
      wr[v] = cp5_minus(c1, c2, c3, c4, c5);
    }
  }
  T out[4];
  shallow_roe(wl, wr, a.ax, out);
  int f = a.ax == 0 ? i - r.lo1 : j - r.lo2;
  // a wall plane keeps only the momentum rows
  if ((f == 0 && a.wall_lo) || (f == a.nf - 1 && a.wall_hi)) out[0] = T(0);
  for (int v = 0; v < kSweRows; ++v) flx[v * r.sv + base] = out[v];
}

template <typename T>
__global__ void __launch_bounds__(256)
    swe_cfl_kernel(T const* __restrict__ w, unsigned long long* out, Range3 r,
                   double dx1, double dx2, int act1, int act2) {
  int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
  double best = 1.e9;
  if (idx < tot) {
    int i = r.lo1 + int(idx % r.n1);
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    int64_t c = k * r.s3 + j * r.s2 + i;
    T cs = sqrt(w[c]);
    if (act1) best = fmin(best, double(T(dx1) / (fabs(w[r.sv + c]) + cs)));
    if (act2) best = fmin(best, double(T(dx2) / (fabs(w[2 * r.sv + c]) + cs)));
  }
  // the minimum of positive doubles orders as their bit patterns
  for (int off = 32; off > 0; off >>= 1) best = fmin(best, __shfl_down(best, off, 64));
  if ((threadIdx.x & 63) == 0) {
    unsigned long long bits;
    memcpy(&bits, &best, sizeof(bits));
    atomicMin(out, bits);
  }
}

__global__ void swe_fill_kernel(unsigned long long* out, unsigned long long bits) {
  *out = bits;
}

__global__ void swe_scale_kernel(double* out, double s) { *out = *out * s; }

template <typename F>
int swe_dispatch(int dtype, F&& f) {
  if (dtype == SNAPY_F64) return f(double(0));
  if (dtype == SNAPY_F32) return f(float(0));
  return fail(-10, "unknown dtype");
}

}  // namespace

int swe_cons2prim(const snapy_grid* g, int dtype, const void* u, void* w,
                  void* stream) {
  if (g->nvar != kSweRows) return fail(-6, "shallow water: nvar 4");
  Range3 r = ranges(*g);
  return swe_dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    hipLaunchKernelGGL(swe_cons2prim_kernel<T>, dim3(blocks_for(r.sv, 256)),
                       dim3(256), 0, as_stream(stream),
                       static_cast<T const*>(u), static_cast<T*>(w), r.sv, r.sv);
    return launch_status("swe_cons2prim");
  });
}

int swe_prim2cons(const snapy_grid* g, int dtype, const void* w, void* u,
                  void* stream) {
  if (g->nvar != kSweRows) return fail(-6, "shallow water: nvar 4");
  Range3 r = ranges(*g);
  return swe_dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    hipLaunchKernelGGL(swe_prim2cons_kernel<T>, dim3(blocks_for(r.sv, 256)),
                       dim3(256), 0, as_stream(stream),
                       static_cast<T const*>(w), static_cast<T*>(u), r.sv, r.sv);
    return launch_status("swe_prim2cons");
  });
}

int swe_sweep(int dim, const snapy_grid* g, const snapy_physics* p, int dtype,
              const void* w, void* flux, void* stream) {
  if (g->nvar != kSweRows) return fail(-6, "shallow water: nvar 4");
  if (dim == 3) return fail(-8, "shallow water: dir xy sweeps x1 and x2");
  int nx[3] = {g->nx1, g->nx2, g->nx3};
  if (nx[dim - 1] < 2) return fail(-9, "sweep: inactive direction");
  SweArgs a;
  a.r = ranges(*g);
  a.ax = dim - 1;
  a.nf = nx[dim - 1] + 1;
  a.shock = dim == 1 ? p->shock_x1 : p->shock_x23;
  a.wall_lo = p->bc[2 * (dim - 1)] == SNAPY_BC_REFLECTING;
  a.wall_hi = p->bc[2 * (dim - 1) + 1] == SNAPY_BC_REFLECTING;
  int64_t tot = int64_t(a.nf) * (dim == 1 ? g->nx2 : g->nx1) * g->nx3;
  return swe_dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    hipLaunchKernelGGL(swe_sweep_kernel<T>, dim3(blocks_for(tot, 256)),
                       dim3(256), 0, as_stream(stream),
                       static_cast<T const*>(w), static_cast<T*>(flux), a);
    return launch_status("swe_sweep");
  });
}

int swe_cfl(const snapy_grid* g, int dtype, const void* w, double scale,
            double* dt_out, void* stream) {
  if (g->nvar != kSweRows) return fail(-6, "shallow water: nvar 4");
  Range3 r = ranges(*g);
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  auto out = reinterpret_cast<unsigned long long*>(dt_out);
  double init = 1.e9;
  unsigned long long bits;
  std::memcpy(&bits, &init, sizeof(bits));
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(swe_fill_kernel, dim3(1), dim3(1), 0, s, out, bits);
  return swe_dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    hipLaunchKernelGGL(swe_cfl_kernel<T>, dim3(blocks_for(tot, 256)),
                       dim3(256), 0, s, static_cast<T const*>(w), out, r,
                       g->dx1, g->dx2, g->nc1 > 1, g->nc2 > 1);
    if (scale != 1.)
      hipLaunchKernelGGL(swe_scale_kernel, dim3(1), dim3(1), 0, s, dt_out,
                         scale);
    return launch_status("swe_cfl");
  });
}

}  // namespace snapy
