// The work list: many independent x1 blocks of uneven size advanced by one
// launch per phase.
//
// A shock-tube request of the serving mix is a 1D block of 128 to 8192
// cells; alone, each of its stage kernels fills a fraction of one CU and the
// stage is launch-bound.  Here a batch of such blocks is one list of
// descriptors in device memory (each block keeps its own buffers and its own
// device dt), and every phase of the RK stage is one persistent launch over
// the concatenation of the blocks' items (cells, faces, interior cells,
// ghost cells): a grid of at most kWlGrid workgroups (or the tuned cap,
// SNAPY_TUNE_WL_GRID) strides over the items, and a thread finds its block by a binary search of the phase's
// prefix counts.
//
// The per-item arithmetic is the single-block kernels' (the same device
// functions: cons2prim5, face_flux_line, radiating_ghost, prim2cons5 and
// the divergence + combine expression of update_kernel), so a block's state
// is bit-identical alone and in any batch.
//
// Phases (MeshBlockImpl::forward for an x1 block without sources,
// src/mesh/meshblock_advance.cpp:88-105 and :314-349): backup (stage 0),
// cons2prim, the x1 sweep (FluxRegister total T_1), the divergence + RK
// combine, the boundaries (outflow fused with its own cons2prim and
// prim2cons); and the CFL, one dt per block.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "point.hpp"

namespace snapy {
namespace {

constexpr int kWlTpb = 256;
constexpr unsigned kWlGrid = 2048;  // the persistent grid's default cap (8 per CU)
enum { WL_CELLS = 0, WL_FACES = 1, WL_INTERIOR = 2, WL_GHOSTS = 3 };

struct WlArgs {
  snapy_wl_block const* blocks;
  int32_t const* prefix;  // 4 x (nb + 1)
  int nb, ng;
  int64_t total;
  double gammad, dfloor, pfloor;
  int riemann, shock;
  double w0, w1, w2;
};

// the block holding global item g of a phase: the last b with
// prefix[b] <= g
__device__ __forceinline__ int wl_block(int32_t const* pre, int nb, int64_t g) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

#define WL_LOOP(PHASE)                                                         \
  int32_t const* pre = a.prefix + (PHASE) * (a.nb + 1);                        \
  for (int64_t g = int64_t(blockIdx.x) * kWlTpb + threadIdx.x; g < a.total;    \
       g += int64_t(gridDim.x) * kWlTpb)

// u0 <- u over every cell (StateStore::backup at stage 0)
__global__ void __launch_bounds__(kWlTpb) wl_backup_kernel(WlArgs a) {
  WL_LOOP(WL_CELLS) {
    int b = wl_block(pre, a.nb, g);
    snapy_wl_block const& B = a.blocks[b];
    int c = int(g - pre[b]);
    int nc1 = B.nx1 + 2 * a.ng;
    for (int v = 0; v < 5; ++v) B.u0[v * nc1 + c] = B.u[v * nc1 + c];
  }
}

// w <- cons2prim(u) over every cell (the stage's EOS)
__global__ void __launch_bounds__(kWlTpb) wl_cons2prim_kernel(WlArgs a) {
  WL_LOOP(WL_CELLS) {
    int b = wl_block(pre, a.nb, g);
    snapy_wl_block const& B = a.blocks[b];
    int c = int(g - pre[b]);
    int nc1 = B.nx1 + 2 * a.ng;
    double uu[5], ww[5];
    for (int v = 0; v < 5; ++v) uu[v] = B.u[v * nc1 + c];
    cons2prim5(uu, ww, a.gammad);
    for (int v = 0; v < 5; ++v) B.w[v * nc1 + c] = ww[v];
  }
}

// the x1 face fluxes of every interior face (sweep_kernel<double, 1>)
__global__ void __launch_bounds__(kWlTpb) wl_sweep_kernel(WlArgs a) {
  WL_LOOP(WL_FACES) {
    int b = wl_block(pre, a.nb, g);
    snapy_wl_block const& B = a.blocks[b];
    int f = int(g - pre[b]);
    int nc1 = B.nx1 + 2 * a.ng;
    int i = a.ng + f;
    double out[5], pface;
    face_flux_line(B.w + i, nc1, 1, 0., a.gammad, a.riemann, a.shock, 0, out, &pface);
    if ((f == 0 && B.bc_lo == SNAPY_BC_REFLECTING) ||
        (f == B.nx1 && B.bc_hi == SNAPY_BC_REFLECTING)) {
      out[IDN] = 0.;
      out[IPR] = 0.;
    }
    for (int v = 0; v < 5; ++v) B.flux[v * nc1 + i] = out[v];
  }
}

// du = -dt div F; u <- w0 u0 + w1 u + w2 du on every interior cell
// (update_kernel<double> with the x1 totals only)
__global__ void __launch_bounds__(kWlTpb) wl_update_kernel(WlArgs a) {
  WL_LOOP(WL_INTERIOR) {
    int b = wl_block(pre, a.nb, g);
    snapy_wl_block const& B = a.blocks[b];
    int nc1 = B.nx1 + 2 * a.ng;
    int i = a.ng + int(g - pre[b]);
    double A1 = B.dx3 * B.dx2, V = B.dx3 * (B.dx2 * B.dx1);
    double mdt = -*B.dt;
    for (int v = 0; v < 5; ++v) {
      int o = v * nc1 + i;
      double div = A1 * B.flux[o + 1] - A1 * B.flux[o];
      div = div / V;
      double du = mdt * div;
      B.u[o] = a.w0 * B.u0[o] + a.w1 * B.u[o] + a.w2 * du;
    }
  }
}

// the x1 ghost cells of u: outflow (radiating about the block's background,
// on the primitive view of the boundary cell, flushed back to conserved),
// or the reflecting / periodic / extrapolation gathers (bc_func.cpp:10-227);
// every ghost reads interior cells only, so the ghosts are independent
__global__ void __launch_bounds__(kWlTpb) wl_boundaries_kernel(WlArgs a) {
  WL_LOOP(WL_GHOSTS) {
    int b = wl_block(pre, a.nb, g);
    snapy_wl_block const& B = a.blocks[b];
    int q = int(g - pre[b]);
    int side = q / a.ng, gg = q % a.ng;
    int nc1 = B.nx1 + 2 * a.ng, ng = a.ng;
    int bc = side ? B.bc_hi : B.bc_lo;
    int dst = side ? nc1 - ng + gg : gg;
    if (bc == SNAPY_BC_OUTFLOW) {
      int src = side ? nc1 - ng - 1 : ng;
      double uu[5], wi[5], bi[5], bg[5], out[5], uo[5];
      for (int v = 0; v < 5; ++v) {
        uu[v] = B.u[v * nc1 + src];
        bi[v] = B.bref[v * nc1 + src];
        bg[v] = B.bref[v * nc1 + dst];
      }
      cons2prim5(uu, wi, a.gammad);
      radiating_ghost(wi, bi, bg, side, IVX, a.gammad, a.dfloor, a.pfloor, out);
      prim2cons5(out, uo, a.gammad - 1.);
      for (int v = 0; v < 5; ++v) B.u[v * nc1 + dst] = uo[v];
      continue;
    }
    int src;
    bool flip = false;
    if (bc == SNAPY_BC_REFLECTING) {
      src = side ? nc1 - ng - 1 - gg : 2 * ng - 1 - gg;
      flip = true;
    } else if (bc == SNAPY_BC_PERIODIC) {
      src = side ? ng + gg : nc1 - 2 * ng + gg;
    } else if (bc == SNAPY_BC_EXTRAPOLATION) {
      src = side ? nc1 - ng - 1 : ng;
    } else {
      continue;
    }
    for (int v = 0; v < 5; ++v) {
      double x = B.u[v * nc1 + src];
      if (flip && v == IVX) x = x * -1.;
      B.u[v * nc1 + dst] = x;
    }
  }
}

// dt of each block: scale * min over its interior cells of dx1 / (|v1| + c),
// one workgroup per block (cfl_kernel<double> for an x1 block)
__global__ void __launch_bounds__(kWlTpb) wl_cfl_kernel(WlArgs a) {
  int b = blockIdx.x;
  snapy_wl_block const& B = a.blocks[b];
  int nc1 = B.nx1 + 2 * a.ng;
  double best = 1.e9;
  for (int c = threadIdx.x; c < B.nx1; c += kWlTpb) {
    int i = a.ng + c;
    double rho = B.w[i], p = B.w[IPR * nc1 + i];
    double cs = sqrt(a.gammad * p / rho);
    double t = B.dx1 / (fabs(B.w[IVX * nc1 + i]) + cs);
    best = fmin(best, t);
  }
  for (int off = 32; off > 0; off >>= 1) best = fmin(best, __shfl_xor(best, off, 64));
  __shared__ double part[kWlTpb / 64];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
    for (int q = 1; q < kWlTpb / 64; ++q) m = fmin(m, part[q]);
    *B.dt = B.scale != 1. ? B.scale * m : m;
  }
}

int wl_args(snapy_work_list const* wl, snapy_physics const* p, int phase, WlArgs& a) {
  if (!wl || !p) return fail(-1, "work list: NULL argument");
  if (!wl->blocks || !wl->prefix || wl->nblocks < 1)
    return fail(-1, "work list: no blocks");
  if (wl->nghost != 3) return fail(-2, "work list: nghost must be 3");
  a.blocks = wl->blocks;
  a.prefix = wl->prefix;
  a.nb = wl->nblocks;
  a.ng = wl->nghost;
  a.total = phase >= 0 ? wl->total[phase] : 0;
  a.gammad = p->gammad;
  a.dfloor = p->density_floor;
  a.pfloor = p->pressure_floor;
  a.riemann = p->riemann;
  a.shock = p->shock_x1;
  a.w0 = a.w1 = a.w2 = 0.;
  return 0;
}

unsigned wl_grid(int64_t total) {
  unsigned cap = tuned(SNAPY_TUNE_WL_GRID, nullptr, 0, kWlGrid);
  unsigned nb = blocks_for(total, kWlTpb);
  return nb < cap ? (nb > 0 ? nb : 1) : cap;
}

template <typename K>
int wl_launch(K kernel, WlArgs const& a, void* stream, char const* what) {
  if (a.total <= 0) return 0;
  hipLaunchKernelGGL(kernel, dim3(wl_grid(a.total)), dim3(kWlTpb), 0, as_stream(stream), a);
  return launch_status(what);
}

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

int snapy_wl_counts(const snapy_wl_block* blocks, int32_t nblocks, int32_t nghost,
                    int32_t* prefix, int64_t* total) {
  if (!blocks || !prefix || !total || nblocks < 1) return fail(-1, "wl_counts: NULL argument");
  int64_t run[4] = {0, 0, 0, 0};
  for (int ph = 0; ph < 4; ++ph) prefix[ph * (nblocks + 1)] = 0;
  for (int b = 0; b < nblocks; ++b) {
    int n = blocks[b].nx1;
    if (n < nghost) return fail(-5, "wl_counts: a block needs at least nghost cells");
    int64_t cnt[4] = {int64_t(n) + 2 * nghost, int64_t(n) + 1, int64_t(n), int64_t(2) * nghost};
    for (int ph = 0; ph < 4; ++ph) {
      run[ph] += cnt[ph];
      if (run[ph] > INT32_MAX) return fail(-9, "wl_counts: the batch exceeds 2^31 items");
      prefix[ph * (nblocks + 1) + b + 1] = int32_t(run[ph]);
    }
  }
  for (int ph = 0; ph < 4; ++ph) total[ph] = run[ph];
  return 0;
}

int snapy_wl_backup(const snapy_work_list* wl, const snapy_physics* p, void* stream) {
  WlArgs a;
  if (int e = wl_args(wl, p, WL_CELLS, a)) return e;
  return wl_launch(wl_backup_kernel, a, stream, "wl_backup");
}

int snapy_wl_cons2prim(const snapy_work_list* wl, const snapy_physics* p, void* stream) {
  WlArgs a;
  if (int e = wl_args(wl, p, WL_CELLS, a)) return e;
  return wl_launch(wl_cons2prim_kernel, a, stream, "wl_cons2prim");
}

int snapy_wl_sweep_x1(const snapy_work_list* wl, const snapy_physics* p, void* stream) {
  WlArgs a;
  if (int e = wl_args(wl, p, WL_FACES, a)) return e;
  return wl_launch(wl_sweep_kernel, a, stream, "wl_sweep_x1");
}

int snapy_wl_update(const snapy_work_list* wl, const snapy_physics* p, double w0, double w1,
                    double w2, void* stream) {
  WlArgs a;
  if (int e = wl_args(wl, p, WL_INTERIOR, a)) return e;
  a.w0 = w0;
  a.w1 = w1;
  a.w2 = w2;
  return wl_launch(wl_update_kernel, a, stream, "wl_update");
}

int snapy_wl_boundaries(const snapy_work_list* wl, const snapy_physics* p, void* stream) {
  WlArgs a;
  if (int e = wl_args(wl, p, WL_GHOSTS, a)) return e;
  return wl_launch(wl_boundaries_kernel, a, stream, "wl_boundaries");
}

int snapy_wl_cfl(const snapy_work_list* wl, const snapy_physics* p, void* stream) {
  WlArgs a;
  if (int e = wl_args(wl, p, -1, a)) return e;
  hipLaunchKernelGGL(wl_cfl_kernel, dim3(unsigned(a.nb)), dim3(kWlTpb), 0, as_stream(stream), a);
  return launch_status("wl_cfl");
}

int snapy_wl_stage(const snapy_work_list* wl, const snapy_physics* p, int stage, double w0,
                   double w1, double w2, void* stream) {
  if (stage == 0)
    if (int e = snapy_wl_backup(wl, p, stream)) return e;
  if (int e = snapy_wl_cons2prim(wl, p, stream)) return e;
  if (int e = snapy_wl_sweep_x1(wl, p, stream)) return e;
  if (int e = snapy_wl_update(wl, p, w0, w1, w2, stream)) return e;
  return snapy_wl_boundaries(wl, p, stream);
}

}  // extern "C"
