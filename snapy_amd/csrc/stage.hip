// Stage kernels of the Cartesian finite-volume path on gfx950 and their C-ABI
// wrappers: EOS, the fused reconstruction + Riemann face sweep, the fused
// divergence + RK combine, the FluxRegister step field, the CFL reduction and
// the deterministic FP64 budget.
//
// Design: HBM-bound stencil work, 64-wide wavefronts, the lane index walking
// the contiguous x1 axis in every kernel so each wavefront load is one
// coalesced 512 B (FP64) row.  MFMA is not used: nothing here is a dense
// contraction.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {

namespace {
thread_local std::string g_last_error;
}

void set_error(std::string const& msg) { g_last_error = msg; }

int fail(int code, std::string const& msg) {
  set_error(msg);
  return code;
}

int launch_status(char const* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    return fail(static_cast<int>(e), std::string(what) + ": " +
                                         hipGetErrorString(e));
  }
  return 0;
}

int check_grid(snapy_grid const* g) {
  if (!g) return fail(-1, "grid descriptor is NULL");
  if (g->nghost < 3) return fail(-2, "nghost must be >= 3 for weno5");
  int nx[3] = {g->nx1, g->nx2, g->nx3};
  int nc[3] = {g->nc1, g->nc2, g->nc3};
  if (g->members3 > 1 && (g->nx3 != g->members3 || g->nc3 != g->members3))
    return fail(-4, "a member x3 axis has nx3 = nc3 = members3 (no x3 ghosts)");
  for (int d = 0; d < 3; ++d) {
    if (nx[d] < 1) return fail(-3, "nx must be >= 1");
    if (d == 2 && g->members3 > 1) continue;
    int want = nx[d] > 1 ? nx[d] + 2 * g->nghost : 1;
    if (nc[d] != want) return fail(-4, "nc does not match nx + 2*nghost");
    if (nx[d] > 1 && nx[d] < g->nghost)
      return fail(-5, "an active axis needs at least nghost cells");
  }
  if (g->nvar != 5 && g->nvar != 4) return fail(-6, "nvar must be 5 or 4");
  return 0;
}

// ---------------------------------------------------------------------------
// EOS
// ---------------------------------------------------------------------------
template <typename T>
struct Cons2PrimBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ u, T* __restrict__ w, int64_t n,
                                             int64_t sv, double gd) {
    int64_t c = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    T uu[5], ww[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) uu[v] = u[v * sv + c];
    cons2prim5(uu, ww, gd);
#pragma unroll
    for (int v = 0; v < 5; ++v) w[v * sv + c] = ww[v];
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    cons2prim_kernel(T const* __restrict__ u, T* __restrict__ w, int64_t n,
                     int64_t sv, double gd) {
  Cons2PrimBody<T>::run(block_index(), grid_size(), u, w, n, sv, gd);
}

template <typename T>
struct Prim2ConsBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ w, T* __restrict__ u, int64_t n,
                                             int64_t sv, T gm1) {
    int64_t c = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    T uu[5], ww[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) ww[v] = w[v * sv + c];
    prim2cons5(ww, uu, gm1);
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v * sv + c] = uu[v];
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    prim2cons_kernel(T const* __restrict__ w, T* __restrict__ u, int64_t n,
                     int64_t sv, T gm1) {
  Prim2ConsBody<T>::run(block_index(), grid_size(), w, u, n, sv, gm1);
}

// ---------------------------------------------------------------------------
// fused reconstruction + Riemann sweep
// ---------------------------------------------------------------------------
struct SweepArgs {
  Range3 r;
  int nf;  // faces along the sweep axis: nx + 1
  double gammad;
  int riemann;
  int shock;
  int wall_lo, wall_hi;  // the sweep's lower/upper boundary face is a wall
  int gauge;             // FP32 LMARS: subtract the line's pressure gauge
};

// ---------------------------------------------------------------------------
// the FP32 state form: the conserved state is an FP32 pair (u, c) whose sum
// in FP64 is the state (c the compensation of the RK accumulation)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pair_value(float hi, float lo) {
  return double(hi) + double(lo);
}

__device__ __forceinline__ void pair_store(float* hi, float* lo, double x) {
  float h = float(x);
  *hi = h;
  *lo = float(x - double(h));
}

struct Cons2PrimPairBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             float const* __restrict__ u,
                                             float const* __restrict__ c, double* __restrict__ w,
                                             int64_t n, int64_t sv, double gd) {
    int64_t q = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double uu[5], ww[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) uu[v] = pair_value(u[v * sv + q], c[v * sv + q]);
    cons2prim5(uu, ww, gd);
#pragma unroll
    for (int v = 0; v < 5; ++v) w[v * sv + q] = ww[v];
  }
};

__global__ void __launch_bounds__(256)
    cons2prim_pair_kernel(float const* __restrict__ u, float const* __restrict__ c,
                          double* __restrict__ w, int64_t n, int64_t sv,
                          double gd) {
  Cons2PrimPairBody::run(block_index(), grid_size(), u, c, w, n, sv, gd);
}

struct Prim2ConsPairBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w, float* __restrict__ u,
                                             float* __restrict__ c, int64_t n, int64_t sv,
                                             double gm1) {
    int64_t q = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double uu[5], ww[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) ww[v] = w[v * sv + q];
    prim2cons5(ww, uu, gm1);
#pragma unroll
    for (int v = 0; v < 5; ++v) pair_store(u + v * sv + q, c + v * sv + q, uu[v]);
  }
};

__global__ void __launch_bounds__(256)
    prim2cons_pair_kernel(double const* __restrict__ w, float* __restrict__ u,
                          float* __restrict__ c, int64_t n, int64_t sv,
                          double gm1) {
  Prim2ConsPairBody::run(block_index(), grid_size(), w, u, c, n, sv, gm1);
}

// one thread per interior face.  DIR 1: the lane walks the faces along x1;
// DIR 2/3: the lane walks x1 across the faces, so each of the six stencil
// loads of a wavefront is one coalesced row.
// The primitive state is FP64 (PrecisionPolicy keeps primitives FP64); T is
// the compute and flux precision.  With a.gauge the pressure row enters the
// reconstruction as p - g, g the pressure of the sweep line's first interior
// cell, so an FP32 reconstruction keeps the small pressure differences; the
// stored momentum flux is then gauge-free (g cancels in its divergence).
template <typename T, int DIR>
struct SweepBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w, T* __restrict__ flx,
                                             double* __restrict__ fp, SweepArgs a) {
    Range3 const& r = a.r;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int i, j, k;
    int64_t stride;
    if (DIR == 1) {
      int64_t tot = int64_t(a.nf) * r.n2 * r.n3;
      if (idx >= tot) return;
      i = r.lo1 + int(idx % a.nf);
      int64_t rest = idx / a.nf;
      j = r.lo2 + int(rest % r.n2);
      k = r.lo3 + int(rest / r.n2);
      stride = 1;
    } else if (DIR == 2) {
      int64_t tot = int64_t(r.n1) * a.nf * r.n3;
      if (idx >= tot) return;
      i = r.lo1 + int(idx % r.n1);
      int64_t rest = idx / r.n1;
      j = r.lo2 + int(rest % a.nf);
      k = r.lo3 + int(rest / a.nf);
      stride = r.s2;
    } else {
      int64_t tot = int64_t(r.n1) * r.n2 * a.nf;
      if (idx >= tot) return;
      i = r.lo1 + int(idx % r.n1);
      int64_t rest = idx / r.n1;
      j = r.lo2 + int(rest % r.n2);
      k = r.lo3 + int(rest / r.n2);
      stride = r.s3;
    }
    int64_t base = k * r.s3 + j * r.s2 + i;  // cell f: face f is its lower face

    double gauge = 0.;
    if (a.gauge) {
      int64_t first = DIR == 1 ? k * r.s3 + j * r.s2 + r.lo1
                    : DIR == 2 ? k * r.s3 + r.lo2 * r.s2 + i
                               : r.lo3 * r.s3 + j * r.s2 + i;
      gauge = w[IPR * r.sv + first];
    }
    T out[5];
    T pface;
    face_flux_line(w + base, r.sv, stride, gauge, a.gammad, a.riemann, a.shock, DIR - 1, out,
                   &pface);
    // boundary.prescribed at a wall plane (FluxRegister::apply_prescribed,
    // flux_register.cpp:707-748, planes and free rows from supply_prescribed,
    // seam_register.cpp:184-220): the momentum rows stay free, the mass and
    // energy rows are zero
    int f = DIR == 1 ? i - r.lo1 : (DIR == 2 ? j - r.lo2 : k - r.lo3);
    if ((f == 0 && a.wall_lo) || (f == a.nf - 1 && a.wall_hi)) {
      out[IDN] = T(0);
      out[IPR] = T(0);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) flx[v * r.sv + base] = out[v];
    if (DIR == 1 && fp != nullptr) fp[base] = double(pface);
  }
};

template <typename T, int DIR>
__global__ void __launch_bounds__(256)
    sweep_kernel(double const* __restrict__ w, T* __restrict__ flx,
                 double* __restrict__ fp, SweepArgs a) {
  SweepBody<T, DIR>::run(block_index(), grid_size(), w, flx, fp, a);
}

// ---------------------------------------------------------------------------
// fused divergence + RK combine
// ---------------------------------------------------------------------------
struct UpdateArgs {
  Range3 r;
  int nvar;
  double a1, a2, a3, vol;  // Cartesian face areas and cell volume
  double mdt;              // -dt (host dt)
  double const* dt_dev;    // device dt (StepDt on the device); NULL: mdt
  double w0, w1, w2;
};

template <typename T>
struct UpdateBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ f1, T const* __restrict__ f2,
                                             T const* __restrict__ f3, T const* __restrict__ u0,
                                             T* __restrict__ u, T const* __restrict__ dsrc,
                                             UpdateArgs a) {
    Range3 const& r = a.r;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
    if (idx >= tot) return;
    int i = r.lo1 + int(idx % r.n1);
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    int64_t c = k * r.s3 + j * r.s2 + i;
    T A1 = T(a.a1), A2 = T(a.a2), A3 = T(a.a3), V = T(a.vol);
    T mdt = a.dt_dev ? T(-dt_at(a.dt_dev, 0., r, k)) : T(a.mdt);
    T w0 = T(a.w0), w1 = T(a.w1), w2 = T(a.w2);
    for (int v = 0; v < a.nvar; ++v) {
      int64_t o = v * r.sv + c;
      // dflx = 0; dflx += A F(+) - A F(-) per active direction; / vol
      // (CoordinateImpl::divergence, coordinate.cpp:618-662)
      T div = T(0);
      bool first = true;
      if (f1) {
        T d = A1 * f1[o + 1] - A1 * f1[o];
        div = first ? d : div + d;
        first = false;
      }
      if (f2) {
        T d = A2 * f2[o + r.s2] - A2 * f2[o];
        div = first ? d : div + d;
        first = false;
      }
      if (f3) {
        T d = A3 * f3[o + r.s3] - A3 * f3[o];
        div = first ? d : div + d;
        first = false;
      }
      div = div / V;
      T du = mdt * div;
      if (dsrc) du = du + dsrc[o];
      u[o] = w0 * u0[o] + w1 * u[o] + w2 * du;
    }
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    update_kernel(T const* __restrict__ f1, T const* __restrict__ f2,
                  T const* __restrict__ f3, T const* __restrict__ u0,
                  T* __restrict__ u, T const* __restrict__ dsrc, UpdateArgs a) {
  UpdateBody<T>::run(block_index(), grid_size(), f1, f2, f3, u0, u, dsrc, a);
}

// The FP32 state form of update_kernel: FP32 face fluxes, the divergence and
// the combine in FP64 on the (u, c) pairs, du_src FP64.
struct UpdatePairBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             float const* __restrict__ f1,
                                             float const* __restrict__ f2,
                                             float const* __restrict__ f3,
                                             float const* __restrict__ u0,
                                             float const* __restrict__ c0, float* __restrict__ u,
                                             float* __restrict__ cc,
                                             double const* __restrict__ dsrc, UpdateArgs a) {
    Range3 const& r = a.r;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
    if (idx >= tot) return;
    int i = r.lo1 + int(idx % r.n1);
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    int64_t c = k * r.s3 + j * r.s2 + i;
    double A1 = a.a1, A2 = a.a2, A3 = a.a3, V = a.vol;
    double mdt = a.dt_dev ? -dt_at(a.dt_dev, 0., r, k) : a.mdt;
    for (int v = 0; v < a.nvar; ++v) {
      int64_t o = v * r.sv + c;
      double div = 0.;
      bool first = true;
      if (f1) {
        double d = A1 * double(f1[o + 1]) - A1 * double(f1[o]);
        div = first ? d : div + d;
        first = false;
      }
      if (f2) {
        double d = A2 * double(f2[o + r.s2]) - A2 * double(f2[o]);
        div = first ? d : div + d;
        first = false;
      }
      if (f3) {
        double d = A3 * double(f3[o + r.s3]) - A3 * double(f3[o]);
        div = first ? d : div + d;
        first = false;
      }
      div = div / V;
      double du = mdt * div;
      if (dsrc) du = du + dsrc[o];
      double x = a.w0 * pair_value(u0[o], c0[o]) + a.w1 * pair_value(u[o], cc[o]) +
                 a.w2 * du;
      pair_store(u + o, cc + o, x);
    }
  }
};

__global__ void __launch_bounds__(256)
    update_pair_kernel(float const* __restrict__ f1, float const* __restrict__ f2,
                       float const* __restrict__ f3, float const* __restrict__ u0,
                       float const* __restrict__ c0, float* __restrict__ u,
                       float* __restrict__ cc, double const* __restrict__ dsrc,
                       UpdateArgs a) {
  UpdatePairBody::run(block_index(), grid_size(), f1, f2, f3, u0, c0, u, cc, dsrc, a);
}

// S <- w1*S + w2*dt*T over every face of the total, in FP64
template <typename T>
struct StepFieldBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ t, double* __restrict__ s,
                                             int64_t n, double w1, double w2, double dt,
                                             double const* dt_dev, Range3 r) {
    int64_t c = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    // dt.scaled(w2) then axpy_ (flux_register.cpp:761-777); a member axis
    // has no x3 ghosts, so the x3 index is the member
    int k = r.mem3 > 1 ? int((c / r.s3) % r.nc3) : 0;
    double w2dt = (dt_dev ? dt_dev[k] : dt) * w2;
    s[c] = s[c] * w1 + double(t[c]) * w2dt;
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    step_field_kernel(T const* __restrict__ t, double* __restrict__ s,
                      int64_t n, double w1, double w2, double dt,
                      double const* dt_dev, Range3 r) {
  StepFieldBody<T>::run(block_index(), grid_size(), t, s, n, w1, w2, dt, dt_dev, r);
}

// ---------------------------------------------------------------------------
// CFL: min over interior cells of dx_d / (|v_d| + c), one atomic per block
// ---------------------------------------------------------------------------
struct FillU64Body {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             unsigned long long* p, unsigned long long v, int n) {
    int m = bid.x * blockDim.x + threadIdx.x;
    if (m < n) p[m] = v;
  }
};

__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v, int n) {
  FillU64Body::run(block_index(), grid_size(), p, v, n);
}

struct ScaleF64Body {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim, double* p,
                                             double scale, int n) {
    int m = bid.x * blockDim.x + threadIdx.x;
    if (m < n) p[m] = scale * p[m];
  }
};

__global__ void scale_f64_kernel(double* p, double scale, int n) {
  ScaleF64Body::run(block_index(), grid_size(), p, scale, n);
}

// blockIdx.y: the member along a member x3 axis (its own slab and output),
// else 0 (every interior cell, one output)
template <typename T>
struct CflBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ w, unsigned long long* out,
                                             Range3 r, double gd, double dx1, double dx2,
                                             double dx3, int act1, int act2, int act3, double adv1) {
    int const m = bid.y;
    int const n3 = r.mem3 > 1 ? 1 : r.n3;
    int64_t tot = int64_t(r.n1) * r.n2 * n3;
    double best = 1.e9;
    out += m;
    // grid-stride: a bounded grid keeps the final atomics few
    for (int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x; idx < tot;
         idx += int64_t(gdim.x) * blockDim.x) {
      int i = r.lo1 + int(idx % r.n1);
      int64_t rest = idx / r.n1;
      int j = r.lo2 + int(rest % r.n2);
      int k = r.lo3 + (r.mem3 > 1 ? m : int(rest / r.n2));
      int64_t c = k * r.s3 + j * r.s2 + i;
      T rho = w[c], p = w[4 * r.sv + c];
      // sound_speed: sqrt(gamma * p / rho) (equation_of_state.cpp:235-238)
      T cs = sqrt(T(gd) * p / rho);
      if (act1) {
        // adv1 > 0: the implicit x1 direction's advection bound
        T t = adv1 > 0. ? T(dx1) / (fabs(w[r.sv + c]) * T(adv1))
                        : T(dx1) / (fabs(w[r.sv + c]) + cs);
        best = fmin(best, double(t));
      }
      if (act2) {
        T t = T(dx2) / (fabs(w[2 * r.sv + c]) + cs);
        best = fmin(best, double(t));
      }
      if (act3) {
        T t = T(dx3) / (fabs(w[3 * r.sv + c]) + cs);
        best = fmin(best, double(t));
      }
    }
    // wavefront reduction (64 lanes), then across the 4 wavefronts of the block
    for (int off = 32; off > 0; off >>= 1)
      best = fmin(best, __shfl_xor(best, off, 64));
    __shared__ double part[4];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = part[0];
      for (int q = 1; q < int(blockDim.x >> 6); ++q) m = fmin(m, part[q]);
      // positive doubles order like their bit patterns
      atomicMin(out, static_cast<unsigned long long>(__double_as_longlong(m)));
    }
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    cfl_kernel(T const* __restrict__ w, unsigned long long* out, Range3 r,
               double gd, double dx1, double dx2, double dx3, int act1,
               int act2, int act3, double adv1) {
  CflBody<T>::run(block_index(), grid_size(), w, out, r, gd, dx1, dx2, dx3, act1, act2, act3, adv1);
}

// ---------------------------------------------------------------------------
// FP64 budget: fixed-shape two-pass sum (deterministic)
// ---------------------------------------------------------------------------
constexpr int kBudgetTpb = 256;

template <typename T>
__global__ void __launch_bounds__(kBudgetTpb)
    budget_partial_kernel(T const* __restrict__ u, double* __restrict__ part,
                          Range3 r, int nvar, double vol, int64_t nblk) {
  int v = blockIdx.y;
  int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
  double acc = 0.;
  for (int64_t idx = int64_t(blockIdx.x) * kBudgetTpb + threadIdx.x; idx < tot;
       idx += int64_t(gridDim.x) * kBudgetTpb) {
    int i = r.lo1 + int(idx % r.n1);
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    acc += double(u[v * r.sv + k * r.s3 + j * r.s2 + i]) * vol;
  }
  __shared__ double sm[kBudgetTpb];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBudgetTpb / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[v * nblk + blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(kBudgetTpb)
    budget_final_kernel(double const* __restrict__ part, double* out,
                        int64_t nblk) {
  int v = blockIdx.x;
  double acc = 0.;
  for (int64_t b = threadIdx.x; b < nblk; b += kBudgetTpb)
    acc += part[v * nblk + b];
  __shared__ double sm[kBudgetTpb];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBudgetTpb / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = sm[0];
}

constexpr int64_t kBudgetBlocks = 1024;

// The allocation-free budget's partial sums (snapy_budget_noalloc): one FP64
// slot per (row, first-pass workgroup) for up to kScratchRows rows (the five
// conserved rows of the ideal gas), 40 KiB.
// A file-scope device array: every GPU the library runs on holds its own
// copy until the process ends.  Host code never hands this name to a kernel
// (on the host it only stands for the symbol); it takes the address of the
// current device's copy from hipGetSymbolAddress (scratch_address below).
constexpr int kScratchRows = 5;
static __device__ __align__(256) double scratch[kScratchRows * kBudgetBlocks];

}  // namespace snapy

// ===========================================================================
// C-ABI
// ===========================================================================
using namespace snapy;

namespace {

template <typename F>
int dispatch(int dtype, F&& f) {
  if (dtype == SNAPY_F64) return f(double(0));
  if (dtype == SNAPY_F32) return f(float(0));
  return fail(-10, "unknown dtype");
}

// The device address of `scratch` on each device, looked up once per device
// with that device current (each GPU holds its own copy at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_address(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0)
    return fail(-12, "budget_noalloc: no current HIP device");
  if (dev >= kMaxDevices)
    return fail(-12, "budget_noalloc: device index " + std::to_string(dev) + " beyond " +
                         std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(snapy::scratch));
    if (e != hipSuccess || !ptr)
      return fail(-12, std::string("budget_noalloc: hipGetSymbolAddress failed: ") +
                           hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch[dev];
  return 0;
}

// the two passes of the FP64 budget, partial sums in ws
int budget_launch(snapy_grid const* g, int dtype, void const* u, double* ws, double* out,
                  hipStream_t s) {
  Range3 r = ranges(*g);
  double vol = g->dx3 * (g->dx2 * g->dx1);
  return dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    hipLaunchKernelGGL(budget_partial_kernel<T>, dim3(unsigned(kBudgetBlocks), unsigned(g->nvar)),
                       dim3(kBudgetTpb), 0, s, static_cast<T const*>(u), ws, r, g->nvar, vol,
                       kBudgetBlocks);
    hipLaunchKernelGGL(budget_final_kernel, dim3(unsigned(g->nvar)), dim3(kBudgetTpb), 0, s, ws, out,
                       kBudgetBlocks);
    return launch_status("budget");
  });
}

int check_phys(snapy_physics const* p) {
  if (!p) return fail(-1, "physics descriptor is NULL");
  if (p->riemann != SNAPY_RIEMANN_HLLC && p->riemann != SNAPY_RIEMANN_LMARS &&
      p->riemann != SNAPY_RIEMANN_SHALLOW_ROE)
    return fail(-7, "unknown riemann solver");
  return 0;
}

}  // namespace

extern "C" {

const char* snapy_last_error(void) { return g_last_error.c_str(); }

int snapy_kernel_count(void) { return 16; }

int snapy_cons2prim(const snapy_grid* g, const snapy_physics* p, int dtype,
                    const void* u, void* w, void* stream) {
  if (int e = check_grid(g)) return e;
  if (int e = check_phys(p)) return e;
  if (!u || !w) return fail(-1, "cons2prim: NULL buffer");
  if (p->eos == SNAPY_EOS_SHALLOW_WATER) return swe_cons2prim(g, dtype, u, w, stream);
  if (g->nvar != 5) return fail(-6, "cons2prim: ideal gas needs nvar 5");
  Range3 r = ranges(*g);
  return dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    unsigned tpb = tuned(SNAPY_TUNE_EOS, g, dtype, 256);
    launch<Cons2PrimBody<T>, 256, 1>(cons2prim_kernel<T>, dim3(blocks_for(r.sv, tpb)),
                       dim3(tpb), 0, as_stream(stream),
                       static_cast<T const*>(u), static_cast<T*>(w), r.sv,
                       r.sv, p->gammad);
    return launch_status("cons2prim");
  });
}

int snapy_prim2cons(const snapy_grid* g, const snapy_physics* p, int dtype,
                    const void* w, void* u, void* stream) {
  if (int e = check_grid(g)) return e;
  if (int e = check_phys(p)) return e;
  if (!u || !w) return fail(-1, "prim2cons: NULL buffer");
  if (p->eos == SNAPY_EOS_SHALLOW_WATER) return swe_prim2cons(g, dtype, w, u, stream);
  if (g->nvar != 5) return fail(-6, "prim2cons: ideal gas needs nvar 5");
  Range3 r = ranges(*g);
  return dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    unsigned tpb = tuned(SNAPY_TUNE_EOS, g, dtype, 256);
    launch<Prim2ConsBody<T>, 256, 1>(prim2cons_kernel<T>, dim3(blocks_for(r.sv, tpb)),
                       dim3(tpb), 0, as_stream(stream),
                       static_cast<T const*>(w), static_cast<T*>(u), r.sv,
                       r.sv, T(p->gammad - 1.));
    return launch_status("prim2cons");
  });
}

int snapy_cons2prim_f32(const snapy_grid* g, const snapy_physics* p,
                        const float* u, const float* c, double* w,
                        void* stream) {
  if (int e = check_grid(g)) return e;
  if (int e = check_phys(p)) return e;
  if (!u || !c || !w) return fail(-1, "cons2prim_f32: NULL buffer");
  if (g->nvar != 5) return fail(-6, "cons2prim_f32: ideal gas needs nvar 5");
  Range3 r = ranges(*g);
  unsigned tpb = tuned(SNAPY_TUNE_EOS, g, SNAPY_F32, 256);
  launch<Cons2PrimPairBody, 256, 1>(cons2prim_pair_kernel, dim3(blocks_for(r.sv, tpb)),
                     dim3(tpb), 0, as_stream(stream), u, c, w, r.sv, r.sv,
                     p->gammad);
  return launch_status("cons2prim_f32");
}

int snapy_prim2cons_f32(const snapy_grid* g, const snapy_physics* p,
                        const double* w, float* u, float* c, void* stream) {
  if (int e = check_grid(g)) return e;
  if (int e = check_phys(p)) return e;
  if (!u || !c || !w) return fail(-1, "prim2cons_f32: NULL buffer");
  if (g->nvar != 5) return fail(-6, "prim2cons_f32: ideal gas needs nvar 5");
  Range3 r = ranges(*g);
  unsigned tpb = tuned(SNAPY_TUNE_EOS, g, SNAPY_F32, 256);
  launch<Prim2ConsPairBody, 256, 1>(prim2cons_pair_kernel, dim3(blocks_for(r.sv, tpb)),
                     dim3(tpb), 0, as_stream(stream), w, u, c, r.sv, r.sv,
                     p->gammad - 1.);
  return launch_status("prim2cons_f32");
}

int snapy_update_f32(const snapy_grid* g, const void* const flux[3],
                     const float* u0, const float* c0, float* u, float* c,
                     const double* du_src, double dt, const double* dt_dev,
                     double w0, double w1, double w2, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!u0 || !c0 || !u || !c || !flux) return fail(-1, "update_f32: NULL buffer");
  UpdateArgs a;
  a.r = ranges(*g);
  a.nvar = g->nvar;
  a.a1 = g->dx3 * g->dx2;
  a.a2 = g->dx3 * g->dx1;
  a.a3 = g->dx2 * g->dx1;
  a.vol = g->dx3 * (g->dx2 * g->dx1);
  a.mdt = -dt;
  a.dt_dev = dt_dev;
  a.w0 = w0;
  a.w1 = w1;
  a.w2 = w2;
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  unsigned tpb = tuned(SNAPY_TUNE_UPDATE, g, SNAPY_F32, 256);
  launch<UpdatePairBody, 256, 1>(update_pair_kernel, dim3(blocks_for(tot, tpb)), dim3(tpb),
                     0, as_stream(stream), static_cast<float const*>(flux[0]),
                     static_cast<float const*>(flux[1]),
                     static_cast<float const*>(flux[2]), u0, c0, u, c, du_src,
                     a);
  return launch_status("update_f32");
}

int snapy_sweep(int dim, const snapy_grid* g, const snapy_physics* p,
                int dtype, const void* w, void* flux, void* face_pressure,
                void* stream) {
  if (int e = check_grid(g)) return e;
  if (int e = check_phys(p)) return e;
  if (dim < 1 || dim > 3) return fail(-8, "sweep: dim must be 1, 2 or 3");
  if (!w || !flux) return fail(-1, "sweep: NULL buffer");
  if (p->eos == SNAPY_EOS_SHALLOW_WATER || p->riemann == SNAPY_RIEMANN_SHALLOW_ROE) {
    if (p->eos != SNAPY_EOS_SHALLOW_WATER || p->riemann != SNAPY_RIEMANN_SHALLOW_ROE)
      return fail(-7, "sweep: shallow water pairs with shallow-roe");
    return swe_sweep(dim, g, p, dtype, w, flux, stream);
  }
  if (g->nvar != 5) return fail(-6, "sweep: the Euler sweep needs nvar 5");
  int nx[3] = {g->nx1, g->nx2, g->nx3};
  if (nx[dim - 1] < 2) return fail(-9, "sweep: inactive direction");
  SweepArgs a;
  a.r = ranges(*g);
  a.nf = nx[dim - 1] + 1;
  a.gammad = p->gammad;
  a.riemann = p->riemann;
  a.shock = dim == 1 ? p->shock_x1 : p->shock_x23;
  // a reflecting face is a wall (MeshBlockOptionsImpl::face_kind,
  // src/mesh/meshblock_options.cpp:188-209)
  a.wall_lo = p->bc[2 * (dim - 1)] == SNAPY_BC_REFLECTING;
  a.wall_hi = p->bc[2 * (dim - 1) + 1] == SNAPY_BC_REFLECTING;
  int64_t tot = int64_t(a.nf);
  for (int d = 0; d < 3; ++d)
    if (d != dim - 1) tot *= nx[d];
  // FP32 LMARS sweeps carry the line gauge (HLLC's clamps and shock
  // switches need total pressures: its FP32 form reconstructs them as they are)
  a.gauge = dtype == SNAPY_F32 && p->riemann == SNAPY_RIEMANN_LMARS;
  return dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    auto wp = static_cast<double const*>(w);
    auto fl = static_cast<T*>(flux);
    auto fp = static_cast<double*>(face_pressure);
    unsigned tpb = tuned(SNAPY_TUNE_SWEEP, g, dtype, 256);
    dim3 grid(blocks_for(tot, tpb)), blk(tpb);
    hipStream_t s = as_stream(stream);
    if (dim == 1)
      launch<SweepBody<T, 1>, 256, 1>((sweep_kernel<T, 1>), grid, blk, 0, s, wp, fl, fp, a);
    else if (dim == 2)
      launch<SweepBody<T, 2>, 256, 1>((sweep_kernel<T, 2>), grid, blk, 0, s, wp, fl, fp, a);
    else
      launch<SweepBody<T, 3>, 256, 1>((sweep_kernel<T, 3>), grid, blk, 0, s, wp, fl, fp, a);
    return launch_status("sweep");
  });
}

int snapy_update(const snapy_grid* g, int dtype, const void* const flux[3],
                 const void* u0, void* u, const void* du_src, double dt,
                 const double* dt_dev, double w0, double w1, double w2,
                 void* stream) {
  if (int e = check_grid(g)) return e;
  if (!u0 || !u || !flux) return fail(-1, "update: NULL buffer");
  UpdateArgs a;
  a.r = ranges(*g);
  a.nvar = g->nvar;
  a.a1 = g->dx3 * g->dx2;
  a.a2 = g->dx3 * g->dx1;
  a.a3 = g->dx2 * g->dx1;
  a.vol = g->dx3 * (g->dx2 * g->dx1);
  a.mdt = -dt;
  a.dt_dev = dt_dev;
  a.w0 = w0;
  a.w1 = w1;
  a.w2 = w2;
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  return dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    unsigned tpb = tuned(SNAPY_TUNE_UPDATE, g, dtype, 256);
    launch<UpdateBody<T>, 256, 1>(update_kernel<T>, dim3(blocks_for(tot, tpb)), dim3(tpb),
                       0, as_stream(stream), static_cast<T const*>(flux[0]),
                       static_cast<T const*>(flux[1]),
                       static_cast<T const*>(flux[2]),
                       static_cast<T const*>(u0), static_cast<T*>(u),
                       static_cast<T const*>(du_src), a);
    return launch_status("update");
  });
}

int snapy_step_field(const snapy_grid* g, int dtype, const void* total,
                     double* step, double w1, double w2, double dt,
                     const double* dt_dev, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!total || !step) return fail(-1, "step_field: NULL buffer");
  int64_t n = int64_t(g->nvar) * g->nc1 * g->nc2 * g->nc3;
  return dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    launch<StepFieldBody<T>, 256, 1>(step_field_kernel<T>, dim3(blocks_for(n, 256)),
                       dim3(256), 0, as_stream(stream),
                       static_cast<T const*>(total), step, n, w1, w2, dt,
                       dt_dev, ranges(*g));
    return launch_status("step_field");
  });
}

int snapy_cfl(const snapy_grid* g, const snapy_physics* p, int dtype,
              const void* w, double scale, double* dt_out, void* stream) {
  if (int e = check_grid(g)) return e;
  if (int e = check_phys(p)) return e;
  if (!w || !dt_out) return fail(-1, "cfl: NULL buffer");
  if (p->eos == SNAPY_EOS_SHALLOW_WATER) {
    if (member_axis(*g)) return fail(-4, "cfl: a member axis is built for the ideal gas");
    return swe_cfl(g, dtype, w, scale, dt_out, stream);
  }
  Range3 r = ranges(*g);
  int const nout = r.mem3;  // one dt per member
  int64_t tot = int64_t(g->nx1) * g->nx2 * (nout > 1 ? 1 : g->nx3);
  auto out = reinterpret_cast<unsigned long long*>(dt_out);
  double init = 1.e9;
  // the sources' bound joins before the CFL factor (hydro.cpp:396): the
  // kinematic diffusion bound dx_min^2 / (2 ndim max(nu, kappa))
  // (DiffusionImpl::max_time_step, src/forcing/diffusion.cpp:708-756)
  double coeff = p->nu_iso > p->kappa_iso ? p->nu_iso : p->kappa_iso;
  if (coeff > 0.) {
    int ndim = 0;
    double dxm = 1.e300;
    double dxs[3] = {g->dx1, g->dx2, g->dx3};
    for (int ax = 0; ax < 3; ++ax)
      if (axis_active(*g, ax + 1)) {
        ++ndim;
        dxm = dxs[ax] < dxm ? dxs[ax] : dxm;
      }
    if (ndim > 0) {
      double dtd = dxm * dxm / (2. * ndim * coeff);
      init = dtd < init ? dtd : init;
    }
  }
  unsigned long long bits;
  std::memcpy(&bits, &init, sizeof(bits));
  hipStream_t s = as_stream(stream);
  launch<FillU64Body, 1024, 1>(fill_u64_kernel, dim3(blocks_for(nout, 64)), dim3(64), 0, s, out, bits,
                     nout);
  bool act2 = axis_active(*g, 2), act3 = axis_active(*g, 3);
  return dispatch(dtype, [&](auto z) {
    using T = decltype(z);
    unsigned nb = blocks_for(tot, 256);
    unsigned cap = nout > 1 ? 64u : 2048u;
    nb = nb < cap ? nb : cap;
    launch<CflBody<T>, 256, 1>(cfl_kernel<T>, dim3(nb, unsigned(nout)), dim3(256), 0,
                       s, static_cast<T const*>(w), out, r, p->gammad, g->dx1,
                       g->dx2, g->dx3, g->nc1 > 1, act2, act3,
                       ((p->implicit_scheme & 1) && (act2 || act3))
                           ? p->advection
                           : 0.);
    if (scale != 1.)
      launch<ScaleF64Body, 1024, 1>(scale_f64_kernel, dim3(blocks_for(nout, 64)), dim3(64), 0, s, dt_out,
                         scale, nout);
    return launch_status("cfl");
  });
}

int64_t snapy_budget_workspace(const snapy_grid* g) {
  if (!g) return 0;
  return int64_t(g->nvar) * kBudgetBlocks;
}

int snapy_budget(const snapy_grid* g, int dtype, const void* u, double* ws,
                 double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!u || !ws || !out) return fail(-1, "budget: NULL buffer");
  return budget_launch(g, dtype, u, ws, out, as_stream(stream));
}

int snapy_budget_noalloc(const snapy_grid* g, int dtype, const void* u,
                         double* out, void* stream) {
  if (!g) return fail(-1, "grid descriptor is NULL");
  if (g->nvar > kScratchRows || g->nvar < 1)
    return fail(-13, "budget_noalloc: nvar " + std::to_string(g->nvar) +
                         " exceeds the " + std::to_string(kScratchRows) +
                         " rows its device scratch holds");
  if (int e = check_grid(g)) return e;
  if (!u || !out) return fail(-1, "budget_noalloc: NULL buffer");
  double* ws = nullptr;
  if (int e = scratch_address(&ws)) return e;
  // the first pass writes every slot the second pass reads: nothing of the
  // previous call's contents is read, so no clear is needed
  return budget_launch(g, dtype, u, ws, out, as_stream(stream));
}

}  // extern "C"
