// Shared host/device helpers of the stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/snapy_amd.h"

namespace snapy {

// the thread-local message behind snapy_last_error()
void set_error(std::string const& msg);
int fail(int code, std::string const& msg);
int launch_status(char const* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// index ranges of one block: [lo, lo + n) are the interior cells of an axis
struct Range3 {
  int lo1, n1, lo2, n2, lo3, n3;
  int nc1, nc2, nc3;
  int64_t s2, s3, sv;  // strides of x2, x3 and the variable row
  int mem3;            // 1, or the members along a member x3 axis
};

// x3 is a member axis (snapy_grid.members3): no x3 ghosts, no coupling
inline bool member_axis(snapy_grid const& g) { return g.members3 > 1; }
// an axis with a stencil, fluxes and boundaries
inline bool axis_active(snapy_grid const& g, int d) {
  int nc = d == 1 ? g.nc1 : (d == 2 ? g.nc2 : g.nc3);
  return nc > 1 && !(d == 3 && member_axis(g));
}

inline Range3 ranges(snapy_grid const& g) {
  Range3 r;
  r.nc1 = g.nc1;
  r.nc2 = g.nc2;
  r.nc3 = g.nc3;
  r.lo1 = g.nc1 > 1 ? g.nghost : 0;
  r.lo2 = g.nc2 > 1 ? g.nghost : 0;
  r.lo3 = axis_active(g, 3) ? g.nghost : 0;
  r.mem3 = member_axis(g) ? g.members3 : 1;
  r.n1 = g.nx1;
  r.n2 = g.nx2;
  r.n3 = g.nx3;
  r.s2 = g.nc1;
  r.s3 = int64_t(g.nc1) * g.nc2;
  r.sv = r.s3 * g.nc3;
  return r;
}

// The stage dt of interior x3 index k: the device dt (one value, or one
// per member along a member axis), else the host dt.
__device__ __forceinline__ double dt_at(double const* dt_dev, double dt_host, Range3 const& r,
                                        int k) {
  return dt_dev ? dt_dev[r.mem3 > 1 ? k - r.lo3 : 0] : dt_host;
}

// host-side validation of a grid descriptor; returns 0 or a negative code
int check_grid(snapy_grid const* g);

inline unsigned blocks_for(int64_t n, int tpb) {
  return static_cast<unsigned>((n + tpb - 1) / tpb);
}

// the tuned workgroup size (or work-list grid) of a kernel family for a
// block of this shape and dtype, dflt when none is set (tune.hip)
unsigned tuned(int tunable, snapy_grid const* g, int dtype, unsigned dflt);

// MI355X dispatches workgroup b of a launch to XCD b mod 8, and each XCD has
// its own L2.  Kernels whose neighbouring tiles share input (stencil halos)
// take their tile index through this map, so that each XCD works on one
// contiguous range of tiles: b -> (b mod 8) * (nb / 8) + b / 8 for the
// first 8 * (nb / 8) workgroups, the identity for the rest.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  unsigned q = nb / kXcds;
  return b < q * kXcds ? (b % kXcds) * q + b / kXcds : b;
}

// The vertical implicit solve's column status word.  Bit 0: this call's
// solve of the column failed (the column keeps its explicit increment; the
// finish kernels read it).  Bit 1: a call since the caller last cleared the
// word failed (a step-level latch: the host clears the words at a step's
// first stage and check_redo reads them, as the reference ORs each stage's
// bits into its step status, src/implicit/implicit_hydro.cpp:610-617).
constexpr int kVicStageFailed = 1;
constexpr int kVicStepFailed = 2;
__device__ __forceinline__ void vic_mark(int* status, int64_t col, bool ok) {
  int const s = status[col];
  status[col] = ok ? (s & kVicStepFailed) : (kVicStageFailed | kVicStepFailed);
}

// the shallow-water stage (swe.hip), dispatched from the generic entry
// points when p->eos is SNAPY_EOS_SHALLOW_WATER
int swe_cons2prim(const snapy_grid* g, int dtype, const void* u, void* w,
                  void* stream);
int swe_prim2cons(const snapy_grid* g, int dtype, const void* w, void* u,
                  void* stream);
int swe_sweep(int dim, const snapy_grid* g, const snapy_physics* p, int dtype,
              const void* w, void* flux, void* stream);
int swe_cfl(const snapy_grid* g, int dtype, const void* w, double scale,
            double* dt_out, void* stream);

}  // namespace snapy
