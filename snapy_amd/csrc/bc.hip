// Physical-boundary ghost fills (src/bc/bc_func.cpp) as on-device gathers.
//
// One launch per active axis fills the ghost layers of both faces of that
// axis over the full extent of the other two axes, so the axes run in the
// reference's face order (x1, x2, x3) and the corners see the earlier axes'
// ghosts exactly as MeshBlockImpl::apply_boundaries leaves them
// (src/mesh/meshblock_exchange.cpp:256-286).  The copies are bit-exact.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

struct BcArgs {
  Range3 r;
  int nvar;
  int axis;    // 1 = x1, 2 = x2, 3 = x3
  int ng;
  int nc;      // cells along the axis
  int64_t st;  // stride along the axis
  int nrow;    // cells across the axis (product of the two other extents)
  int other_n; // extent of the faster other axis
  int64_t other_s_fast, other_s_slow;
  int bc_lo, bc_hi;
  int prim;    // 1: the field is primitive (kPrimitive), 0 conserved
  int vn;      // the row whose sign a reflecting face flips; -1: a kScalar
               // field (tracers, theta), which flips none
};

__device__ __forceinline__ int64_t row_offset(BcArgs const& a, int row) {
  int f = row % a.other_n, s = row / a.other_n;
  return f * a.other_s_fast + s * a.other_s_slow;
}

// U->W of the whole block into the outflow workspace (the fresh primitive
// view apply_boundaries converts before its first radiating face)
template <typename T>
struct Cons2PrimBcBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ u, T* __restrict__ w, int64_t n,
                                             int64_t sv, double gd) {
    int64_t c = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    T uu[5], ww[5];
    for (int v = 0; v < 5; ++v) uu[v] = u[v * sv + c];
    cons2prim5(uu, ww, gd);
    for (int v = 0; v < 5; ++v) w[v * sv + c] = ww[v];
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    cons2prim_kernel_bc(T const* __restrict__ u, T* __restrict__ w, int64_t n,
                        int64_t sv, double gd) {
  Cons2PrimBcBody<T>::run(block_index(), grid_size(), u, w, n, sv, gd);
}

// reflecting / periodic / extrapolation gathers; outflow is skipped here.
template <typename T>
struct BcCopyBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim, T* __restrict__ q,
                                             BcArgs a) {
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(2) * a.ng * a.nrow;
    if (idx >= tot) return;
    int row, g, side;
    if (a.st == 1) {
      // x1: the ghost layers of one column are adjacent in memory, so the
      // ghost index runs fastest (a wavefront touches ~11 columns' lines, not
      // 64)
      int gs = int(idx % (2 * a.ng));
      row = int(idx / (2 * a.ng));
      side = gs / a.ng;
      g = gs - side * a.ng;
    } else {
      row = int(idx % a.nrow);
      int64_t rest = idx / a.nrow;
      g = int(rest % a.ng);
      side = int(rest / a.ng);
    }
    int bc = side ? a.bc_hi : a.bc_lo;
    int dst, src;
    bool flip = false;
    if (bc == SNAPY_BC_REFLECTING) {
      // narrow(0, ng) = narrow(ng, ng).flip; normal velocity *= -1
      dst = side ? a.nc - a.ng + g : g;
      src = side ? a.nc - a.ng - 1 - g : 2 * a.ng - 1 - g;
      flip = true;
    } else if (bc == SNAPY_BC_PERIODIC) {
      dst = side ? a.nc - a.ng + g : g;
      src = side ? a.ng + g : a.nc - 2 * a.ng + g;
    } else if (bc == SNAPY_BC_EXTRAPOLATION) {
      dst = side ? a.nc - a.ng + g : g;
      src = side ? a.nc - a.ng - 1 : a.ng;
    } else {
      return;
    }
    int64_t off = row_offset(a, row);
    int vn = a.vn;
    for (int v = 0; v < a.nvar; ++v) {
      T x = q[v * a.r.sv + off + src * a.st];
      if (flip && v == vn) x = x * T(-1);
      q[v * a.r.sv + off + dst * a.st] = x;
    }
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    bc_copy_kernel(T* __restrict__ q, BcArgs a) {
  BcCopyBody<T>::run(block_index(), grid_size(), q, a);
}

// outflow (radiating about the saved background; bc_func.cpp:110-220) on a
// primitive field, ideal gas, no tracers, Cartesian (boundary_velocity_ is
// the identity there).  Writes the ghost cells of w for each outflow side.
template <typename T>
struct BcOutflowBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim, T* __restrict__ w,
                                             T const* __restrict__ b, BcArgs a, double gd,
                                             double dfloor, double pfloor) {
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(2) * a.ng * a.nrow;
    if (idx >= tot) return;
    int row = int(idx % a.nrow);
    int64_t rest = idx / a.nrow;
    int g = int(rest % a.ng);
    int side = int(rest / a.ng);
    if ((side ? a.bc_hi : a.bc_lo) != SNAPY_BC_OUTFLOW) return;
    int64_t off = row_offset(a, row);
    int src = side ? a.nc - a.ng - 1 : a.ng;
    int dst = (side ? a.nc - a.ng : 0) + g;
    int vn = a.axis;
    T wi[5], bi[5], bg[5], out[5];
    for (int v = 0; v < 5; ++v) {
      wi[v] = w[v * a.r.sv + off + src * a.st];
      bi[v] = b[v * a.r.sv + off + src * a.st];
      bg[v] = b[v * a.r.sv + off + dst * a.st];
    }
    radiating_ghost(wi, bi, bg, side, vn, gd, dfloor, pfloor, out);
    for (int v = 0; v < 5; ++v) w[v * a.r.sv + off + dst * a.st] = out[v];
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    bc_outflow_kernel(T* __restrict__ w, T const* __restrict__ b, BcArgs a,
                      double gd, double dfloor, double pfloor) {
  BcOutflowBody<T>::run(block_index(), grid_size(), w, b, a, gd, dfloor, pfloor);
}

// W->U of the ghost layers of the outflow sides of one axis
template <typename T>
struct BcFlushBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ w, T* __restrict__ u, BcArgs a,
                                             T gm1) {
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(2) * a.ng * a.nrow;
    if (idx >= tot) return;
    int row = int(idx % a.nrow);
    int64_t rest = idx / a.nrow;
    int g = int(rest % a.ng);
    int side = int(rest / a.ng);
    if ((side ? a.bc_hi : a.bc_lo) != SNAPY_BC_OUTFLOW) return;
    int64_t o = row_offset(a, row) + int64_t((side ? a.nc - a.ng : 0) + g) * a.st;
    T ww[5], uu[5];
    for (int v = 0; v < 5; ++v) ww[v] = w[v * a.r.sv + o];
    prim2cons5(ww, uu, gm1);
    for (int v = 0; v < 5; ++v) u[v * a.r.sv + o] = uu[v];
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    bc_flush_kernel(T const* __restrict__ w, T* __restrict__ u, BcArgs a,
                    T gm1) {
  BcFlushBody<T>::run(block_index(), grid_size(), w, u, a, gm1);
}

BcArgs make_args(snapy_grid const& g, snapy_physics const& p, int axis,
                 int prim) {
  BcArgs a;
  a.r = ranges(g);
  a.nvar = g.nvar;
  a.axis = axis;
  a.vn = axis;  // IVX for x1, IVY for x2, IVZ for x3
  a.ng = g.nghost;
  a.prim = prim;
  if (axis == 1) {
    a.nc = g.nc1;
    a.st = 1;
    a.other_n = g.nc2;
    a.other_s_fast = a.r.s2;
    a.other_s_slow = a.r.s3;
    a.nrow = g.nc2 * g.nc3;
  } else if (axis == 2) {
    a.nc = g.nc2;
    a.st = a.r.s2;
    a.other_n = g.nc1;
    a.other_s_fast = 1;
    a.other_s_slow = a.r.s3;
    a.nrow = g.nc1 * g.nc3;
  } else {
    a.nc = g.nc3;
    a.st = a.r.s3;
    a.other_n = g.nc1;
    a.other_s_fast = 1;
    a.other_s_slow = a.r.s2;
    a.nrow = g.nc1 * g.nc2;
  }
  a.bc_lo = p.bc[2 * (axis - 1)];
  a.bc_hi = p.bc[2 * (axis - 1) + 1];
  return a;
}

template <typename T>
int boundaries(snapy_grid const& g, snapy_physics const& p, T* q,
               T const* bref, T* scratch, bool prim, hipStream_t s) {
  bool any_outflow = false;
  for (int f = 0; f < 6; ++f)
    if (p.bc[f] == SNAPY_BC_OUTFLOW && axis_active(g, f / 2 + 1)) any_outflow = true;
  if (any_outflow && !bref)
    return fail(-11, "outflow: missing initial boundary reference");
  if (any_outflow && !prim && !scratch)
    return fail(-11, "outflow: missing primitive workspace");
  Range3 r = ranges(g);
  T gm1 = T(p.gammad - 1.);
  // conserved path: one primitive view for the consecutive outflow faces
  // (apply_boundaries shares the conversion between them, then flushes
  // before a non-outflow face and at the end)
  bool pending = false;
  int pending_axes[3] = {0, 0, 0};
  auto flush = [&]() {
    if (!pending) return;
    for (int ax = 1; ax <= 3; ++ax) {
      if (!pending_axes[ax - 1]) continue;
      BcArgs a = make_args(g, p, ax, 0);
      int64_t tot = int64_t(2) * a.ng * a.nrow;
      launch<BcFlushBody<T>, 256, 1>(bc_flush_kernel<T>, dim3(blocks_for(tot, 256)),
                         dim3(256), 0, s, scratch, q, a, gm1);
      pending_axes[ax - 1] = 0;
    }
    pending = false;
  };
  for (int ax = 1; ax <= 3; ++ax) {
    // a single cell or a member axis has no ghosts
    if (!axis_active(g, ax)) continue;
    BcArgs a = make_args(g, p, ax, prim ? 1 : 0);
    int64_t tot = int64_t(2) * a.ng * a.nrow;
    bool out_lo = a.bc_lo == SNAPY_BC_OUTFLOW, out_hi = a.bc_hi == SNAPY_BC_OUTFLOW;
    if (out_lo || out_hi) {
      // the outflow faces of this axis, then the non-outflow ones; within
      // one axis the two faces touch disjoint ghosts and read disjoint
      // source planes, so their order does not matter
      T* wv = q;
      if (!prim) {
        if (!pending) {
          launch<Cons2PrimBcBody<T>, 256, 1>(cons2prim_kernel_bc<T>, dim3(blocks_for(r.sv, 256)),
                             dim3(256), 0, s, q, scratch, r.sv, r.sv, p.gammad);
        }
        wv = scratch;
        pending = true;
        pending_axes[ax - 1] = 1;
      }
      launch<BcOutflowBody<T>, 256, 1>(bc_outflow_kernel<T>, dim3(blocks_for(tot, 256)),
                         dim3(256), 0, s, wv, bref, a, p.gammad,
                         p.density_floor, p.pressure_floor);
      if (!(out_lo && out_hi)) {
        // the other face of this axis is a gather on q: flush first
        flush();
        launch<BcCopyBody<T>, 256, 1>(bc_copy_kernel<T>, dim3(blocks_for(tot, 256)),
                           dim3(256), 0, s, q, a);
      }
    } else {
      flush();
      launch<BcCopyBody<T>, 256, 1>(bc_copy_kernel<T>, dim3(blocks_for(tot, 256)),
                         dim3(256), 0, s, q, a);
    }
  }
  flush();
  return launch_status("apply_boundaries");
}

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

int snapy_scalar_boundaries(const snapy_grid* g, const snapy_physics* p, int ns, double* q,
                            void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !q) return fail(-1, "scalar_boundaries: NULL argument");
  if (ns < 1) return fail(-13, "scalar_boundaries: ns must be >= 1");
  hipStream_t s = as_stream(stream);
  for (int ax = 1; ax <= 3; ++ax) {
    if (!axis_active(*g, ax)) continue;
    BcArgs a = make_args(*g, *p, ax, 0);
    if (a.bc_lo == SNAPY_BC_OUTFLOW || a.bc_hi == SNAPY_BC_OUTFLOW)
      return fail(-11, "scalar_boundaries: no radiating boundary for a scalar field");
    a.nvar = ns;
    a.vn = -1;
    int64_t tot = int64_t(2) * a.ng * a.nrow;
    launch<BcCopyBody<double>, 256, 1>(bc_copy_kernel<double>, dim3(blocks_for(tot, 256)), dim3(256), 0, s, q, a);
  }
  return launch_status("scalar_boundaries");
}

int snapy_apply_boundaries(const snapy_grid* g, const snapy_physics* p,
                           int dtype, void* u, const void* bref_w,
                           void* scratch_w, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p) return fail(-1, "physics descriptor is NULL");
  if (!u) return fail(-1, "apply_boundaries: NULL state");
  if (dtype == SNAPY_F64)
    return boundaries<double>(*g, *p, static_cast<double*>(u),
                              static_cast<double const*>(bref_w),
                              static_cast<double*>(scratch_w), false,
                              as_stream(stream));
  if (dtype == SNAPY_F32)
    return boundaries<float>(*g, *p, static_cast<float*>(u),
                             static_cast<float const*>(bref_w),
                             static_cast<float*>(scratch_w), false,
                             as_stream(stream));
  return fail(-10, "unknown dtype");
}

int snapy_apply_boundaries_prim(const snapy_grid* g, const snapy_physics* p,
                                int dtype, void* w, const void* bref_w,
                                void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p) return fail(-1, "physics descriptor is NULL");
  if (!w) return fail(-1, "apply_boundaries_prim: NULL state");
  if (dtype == SNAPY_F64)
    return boundaries<double>(*g, *p, static_cast<double*>(w),
                              static_cast<double const*>(bref_w), nullptr,
                              true, as_stream(stream));
  if (dtype == SNAPY_F32)
    return boundaries<float>(*g, *p, static_cast<float*>(w),
                             static_cast<float const*>(bref_w), nullptr, true,
                             as_stream(stream));
  return fail(-10, "unknown dtype");
}

}  // extern "C"
