// The cubed-sphere shallow-water stage: six gnomonic-equiangle panels per
// member, every member and panel resident in one tensor (B = members * 6
// blocks of (4, nc, nc, 1), block b = member * 6 + panel), each kernel one
// launch over all of them.
//
//   * cons <-> prim with the metric raise / lower
//     (ShallowWaterImpl::_cons2prim/_prim2cons, src/eos/shallow_water.cpp:
//     31-55; coord_vec_raise/lower_impl, src/coord/coord_utils_impl.h:9-25);
//   * the panel-edge ghost fill as a gather over precomputed index maps: per
//     ghost cell the source cells on the neighbour panel, the along-edge
//     weights and the pointwise frame matrices (CubedSphereLayoutImpl::
//     serialize/deserialize, src/layout/cubed_sphere_layout.cpp:563-800;
//     _interp_ghost, src/coord/gnomonic_equiangle.cpp:537-605;
//     call_cs_matrix_vec, src/coord/coord_dispatch.cpp:66-90);
//   * the x2 / x3 sweeps: weno5 / cp5, the face-local frame, shallow-roe in
//     the y-z plane, the covariant global rows (prim2local2_/3_,
//     flux2global2_/3_, src/coord/gnomonic_equiangle.cpp:342-470); each
//     panel's interior-side states at its edge faces go to an edge buffer;
//   * the seams: per seam face the owner's (lower panel's) flux from its own
//     state and the neighbour's, through the frame matrices, and the other
//     panel's copy through the global Cartesian frame (the face-state sync
//     of HydroImpl::forward, src/hydro/hydro_forward.cpp:707-757;
//     reconcile_panel_seams, src/layout/seam_register.cpp:35-186);
//   * divergence with the gnomonic areas and volumes, the shallow-water
//     metric sources, the x1 mass source, the RK3 combine
//     (CoordinateImpl::divergence, src/coord/coordinate.cpp:618-662;
//     GnomonicEquiangleImpl::forward / x1_mass_source_,
//     src/coord/gnomonic_equiangle.cpp:472-535);
//   * the CFL bound per member (HydroImpl::max_time_step,
//     src/hydro/hydro.cpp:372-397).
// Every expression keeps the reference's operation order (FP64, no
// contraction), so the result is the CPU path's bit for bit.
#include <hip/hip_runtime.h>

#include <cstring>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

constexpr int kRows = 4;  // phi, phi v1, phi v2, phi v3

// cs_local_to_cart_vel / cs_cart_to_local_vel
// (src/layout/cubed_sphere_constants.h:22-37)
__constant__ int kL2CIdx[6][3] = {{1, 2, 0}, {2, 1, 0}, {1, 2, 0}, {0, 2, 1}, {2, 1, 0}, {0, 2, 1}};
__constant__ int kL2CSgn[6][3] = {{1, 1, 1}, {1, -1, 1}, {-1, -1, 1}, {1, 1, -1}, {-1, 1, 1}, {-1, 1, 1}};
__constant__ int kC2LIdx[6][3] = {{2, 0, 1}, {2, 1, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {0, 2, 1}};
__constant__ int kC2LSgn[6][3] = {{1, 1, 1}, {1, -1, 1}, {1, -1, -1}, {1, -1, 1}, {1, 1, -1}, {-1, 1, 1}};

// the packed grid tables (snapy_cs_mesh.geom)
struct Geom {
  double const *cos_c, *sin_c, *vol, *xr, *yr, *w2, *w3;
  double const *cos_f2, *sin_f2, *area2;  // nc x (nc + 1)
  double const *cos_f3, *sin_f3, *area3;  // (nc + 1) x nc
};

struct Dims {
  int n, ng, nc, members;
  int64_t plane, block;  // nc * nc, kRows * plane
};

Geom geom_of(double const* g, int nc) {
  int64_t p = int64_t(nc) * nc, f = int64_t(nc) * (nc + 1);
  Geom o;
  o.cos_c = g;
  o.sin_c = g + p;
  o.vol = g + 2 * p;
  o.xr = g + 3 * p;
  o.yr = g + 4 * p;
  o.w2 = g + 5 * p;
  o.w3 = g + 6 * p;
  double const* h = g + 7 * p;
  o.cos_f2 = h;
  o.sin_f2 = h + f;
  o.area2 = h + 2 * f;
  o.cos_f3 = h + 3 * f;
  o.sin_f3 = h + 4 * f;
  o.area3 = h + 5 * f;
  return o;
}

Dims dims_of(snapy_cs_mesh const& m) {
  Dims d;
  d.n = m.n;
  d.ng = m.nghost;
  d.nc = m.n + 2 * m.nghost;
  d.members = m.members;
  d.plane = int64_t(d.nc) * d.nc;
  d.block = kRows * d.plane;
  return d;
}

// ---------------------------------------------------------------------------
// EOS
// ---------------------------------------------------------------------------
// The FP32 form keeps the conserved state as the FP32 pair (u, c), value
// double(u) + double(c), and the primitives, face fluxes and edge states in
// FP32; the EOS, the ghost gather, the seams and the update compute in FP64.
__device__ __forceinline__ double cs_val(double const* u, float const*, int64_t o) { return u[o]; }
__device__ __forceinline__ double cs_val(float const* u, float const* c, int64_t o) {
  return double(u[o]) + double(c[o]);
}
__device__ __forceinline__ void cs_put(double* u, float*, int64_t o, double v) { u[o] = v; }
__device__ __forceinline__ void cs_put(float* u, float* c, int64_t o, double v) {
  float h = float(v);
  u[o] = h;
  c[o] = float(v - double(h));
}

template <typename TU, typename TW>
struct CsCons2PrimBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             TU const* __restrict__ u, float const* __restrict__ uc,
                                             TW* __restrict__ w, double const* __restrict__ cos_c,
                                             Dims d, int64_t ncell) {
    int64_t t = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int64_t b = t / d.plane, c = t % d.plane;
    int64_t o = b * d.block + c;
    TW* wb = w + o;
    double h = cs_val(u, uc, o);
    double v1 = cs_val(u, uc, o + d.plane) / h, v2 = cs_val(u, uc, o + 2 * d.plane) / h,
           v3 = cs_val(u, uc, o + 3 * d.plane) / h;
    double cth = cos_c[c];
    double sth2 = 1. - cth * cth;
    wb[0] = TW(h);
    wb[d.plane] = TW(v1);
    wb[2 * d.plane] = TW(v2 / sth2 - v3 * cth / sth2);
    wb[3 * d.plane] = TW(-v2 * cth / sth2 + v3 / sth2);
  }
};

template <typename TU, typename TW>
__global__ void __launch_bounds__(256)
    cs_cons2prim_kernel(TU const* __restrict__ u, float const* __restrict__ uc,
                        TW* __restrict__ w, double const* __restrict__ cos_c, Dims d,
                        int64_t ncell) {
  CsCons2PrimBody<TU, TW>::run(block_index(), grid_size(), u, uc, w, cos_c, d, ncell);
}

template <typename TU>
struct CsPrim2ConsBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w, TU* __restrict__ u,
                                             float* __restrict__ uc,
                                             double const* __restrict__ cos_c, Dims d,
                                             int64_t ncell) {
    int64_t t = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int64_t b = t / d.plane, c = t % d.plane;
    int64_t o = b * d.block + c;
    double const* wb = w + o;
    double h = wb[0];
    double m1 = wb[d.plane] * h, m2 = wb[2 * d.plane] * h, m3 = wb[3 * d.plane] * h;
    double cth = cos_c[c];
    cs_put(u, uc, o, h);
    cs_put(u, uc, o + d.plane, m1);
    cs_put(u, uc, o + 2 * d.plane, m2 + m3 * cth);
    cs_put(u, uc, o + 3 * d.plane, m3 + m2 * cth);
  }
};

template <typename TU>
__global__ void __launch_bounds__(256)
    cs_prim2cons_kernel(double const* __restrict__ w, TU* __restrict__ u, float* __restrict__ uc,
                        double const* __restrict__ cos_c, Dims d, int64_t ncell) {
  CsPrim2ConsBody<TU>::run(block_index(), grid_size(), w, u, uc, cos_c, d, ncell);
}

// ---------------------------------------------------------------------------
// ghost fill
// ---------------------------------------------------------------------------
// out_i = m_i0 v0 + m_i1 v1 + m_i2 v2
__device__ __forceinline__ void matvec(double const* m, double const* v, double* o) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

// the members of a ghost-fill tile: all of them, up to 64 (a wavefront is
// 64 / mt entries x mt members).  Of 4, 8, 16 and 32 members per wavefront
// at 32 members, 32 (two entries per wavefront) was the fastest: 75 us
// against 104 us for 8 and 131 us for one entry per 64 lanes of a member
// (FP32 pair, 6 x 192^2 panels)
__host__ __device__ inline int ghost_tile_members(int members) { return members < 64 ? members : 64; }

// the threads of a ghost fill over nent entries and `members` members
inline int64_t ghost_threads(int64_t nent, int members) {
  int mt = ghost_tile_members(members), et = 64 / mt;
  return ((nent + et - 1) / et) * ((members + mt - 1) / mt) * 64;
}

// one thread per (ghost entry, member): npts 2 (linear), 4 (cubic) or 1 (a
// corner cell: a plain copy of the nearest interior cell)
// TU double: the FP64 state; float: the FP32 pair (u, uc), gathered and
// interpolated in FP64 (the same arithmetic on the pair's values)
template <typename TU>
struct CsGhostBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim, TU* __restrict__ u,
                                             float* __restrict__ uc,
                                             int32_t const* __restrict__ idx,
                                             double const* __restrict__ wt,
                                             double const* __restrict__ mat, int64_t nent, Dims d) {
    // a wavefront takes a tile of et consecutive entries x mt members, the
    // entries fastest: the lanes of one entry share its tables (indices,
    // weights, the 45 frame-matrix values), read once per wavefront for its
    // mt members instead of once per member
    int64_t t = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int const mt = ghost_tile_members(d.members), et = 64 / mt;
    int64_t const net = (nent + et - 1) / et;
    int64_t const tile = t >> 6;
    int const l = int(t & 63);
    int const moff = l / et, eoff = l - moff * et;
    int64_t e = (tile % net) * et + eoff;
    int64_t m = (tile / net) * mt + moff;
    if (moff >= mt || e >= nent || m >= d.members) return;
    int32_t const* ix = idx + 8 * e;
    int npts = ix[7];
    int64_t dst = (m * 6 + ix[0]) * d.block + ix[1];
    int64_t src = (m * 6 + ix[2]) * d.block;
    if (npts == 1) {
      for (int v = 0; v < kRows; ++v) {
        u[dst + v * d.plane] = u[src + v * d.plane + ix[3]];
        if constexpr (sizeof(TU) == 4) uc[dst + v * d.plane] = uc[src + v * d.plane + ix[3]];
      }
      return;
    }
    double const* M = mat + 45 * e;
    // each source point: phi, and its momentum to the geographic frame
    double b[4][kRows];
    for (int q = 0; q < npts; ++q) {
      int64_t s = src + ix[3 + q];
      double v[3] = {cs_val(u, uc, s + d.plane), cs_val(u, uc, s + 2 * d.plane),
                     cs_val(u, uc, s + 3 * d.plane)};
      b[q][0] = cs_val(u, uc, s);
      matvec(M + 9 * q, v, &b[q][1]);
    }
    double g[kRows];
    double const* wq = wt + 3 * e;
    if (npts == 2) {
      double x = wq[0];
      for (int v = 0; v < kRows; ++v) g[v] = b[0][v] + x * (b[1][v] - b[0][v]);
    } else {
      double c0 = wq[0], c2 = wq[1], c3 = wq[2];
      for (int v = 0; v < kRows; ++v)
        g[v] = b[1][v] + c0 * (b[0][v] - b[1][v]) + c2 * (b[2][v] - b[1][v]) +
               c3 * (b[3][v] - b[1][v]);
    }
    double o[3];
    matvec(M + 36, &g[1], o);
    cs_put(u, uc, dst, g[0]);
    cs_put(u, uc, dst + d.plane, o[0]);
    cs_put(u, uc, dst + 2 * d.plane, o[1]);
    cs_put(u, uc, dst + 3 * d.plane, o[2]);
  }
};

template <typename TU>
__global__ void __launch_bounds__(256)
    cs_ghost_kernel(TU* __restrict__ u, float* __restrict__ uc, int32_t const* __restrict__ idx,
                    double const* __restrict__ wt, double const* __restrict__ mat,
                    int64_t nent, Dims d) {
  CsGhostBody<TU>::run(block_index(), grid_size(), u, uc, idx, wt, mat, nent, d);
}

// ---------------------------------------------------------------------------
// the face frame and the solver
// ---------------------------------------------------------------------------
// shallow_roe_impl (riemann_kernels.h:447-495) with shallow_axes dir yz
// (:44-47): normal row ivx, tangential ivy, the radial row ivz gets no flux
template <typename T>
__device__ __forceinline__ void roe_yz(T const* wl, T const* wr, int ivx, int ivy, int ivz,
                                       T* flx) {
  T sqrtdl = qsqrt(wl[0]);
  T sqrtdr = qsqrt(wr[0]);
  T isdlpdr = qdiv(T(1), sqrtdl + sqrtdr);
  T ubar = (wl[ivx] * sqrtdl + wr[ivx] * sqrtdr) * isdlpdr;
  T vbar = (wl[ivy] * sqrtdl + wr[ivy] * sqrtdr) * isdlpdr;
  T cbar = qsqrt(T(0.5) * (wl[0] + wr[0]));
  T del0 = wr[0] - wl[0];
  T delx = wr[ivx] - wl[ivx];
  T dely = wr[ivy] - wl[ivy];
  T hbar = qsqrt(wl[0] * wr[0]);
  T a1 = qdiv(T(0.5) * (cbar * del0 - hbar * delx), cbar);
  T a2 = hbar * dely;
  T a3 = qdiv(T(0.5) * (cbar * del0 + hbar * delx), cbar);
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
    flx[v] = flx[v] - T(0.5) * (speed0 * wave0[v] + speed1 * wave1[v] + speed2 * wave2[v]);
}

// the solve at one face of axis AX (2 = x2, 3 = x3) with the face's metric
// cosine c and sine s: prim2local, the solver, flux2global
template <int AX, typename T>
__device__ __forceinline__ void face_flux(T* wl, T* wr, T c, T s, T* f) {
  T gi = qdiv(T(1), s * s);
  if (AX == 2) {
    T t22 = qdiv(T(1), qsqrt(gi));
    for (T* w : {wl, wr}) {
      T uu2 = w[2], uu3 = w[3];
      w[3] = c * uu2 + T(1) * uu3;
      w[2] = t22 * uu2;
    }
    roe_yz<T>(wl, wr, 2, 3, 1, f);
    T sg = qsqrt(gi);
    T t32 = qdiv(-sg * c, T(1));
    T txy = f[2], txz = f[3];
    f[3] = t32 * txy + T(1) * txz;
    f[2] = sg * txy;
  } else {
    T t33 = qdiv(T(1), qsqrt(gi));
    for (T* w : {wl, wr}) {
      T uu2 = w[2], uu3 = w[3];
      w[2] = T(1) * uu2 + c * uu3;
      w[3] = t33 * uu3;
    }
    roe_yz<T>(wl, wr, 3, 2, 1, f);
    T sg = qsqrt(gi);
    T t23 = qdiv(-c, T(1)) * sg;
    T txy = f[2], txz = f[3];
    f[2] = T(1) * txy + t23 * txz;
    f[3] = sg * txz;
  }
  T ty = f[2], tz = f[3];
  f[2] = ty + tz * c;
  f[3] = tz + ty * c;
}

// one thread per interior face (x2: faces ng..ng+n of the interior rows; x3:
// the transpose); edge faces also store the interior-side state
template <int AX, typename T>
struct CsSweepBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             T const* __restrict__ w, T* __restrict__ flx,
                                             T* __restrict__ edge, Geom gm, Dims d, int shock) {
    // x3: neighbouring face lines share five of their six stencil rows, so the
    // workgroups go through the XCD map (each XCD then sweeps a contiguous
    // range of lines and the shared rows stay in its L2)
    unsigned wg = AX == 3 ? xcd_tile(bid.x, gdim.x) : bid.x;
    int64_t t = int64_t(wg) * blockDim.x + threadIdx.x;
    int nf = d.n + 1;
    int64_t per = int64_t(d.n) * nf;
    if (t >= per * 6 * d.members) return;
    int64_t b = t / per;
    int r = int(t % per);
    // consecutive threads step along x2 (the contiguous axis) in both sweeps:
    // along the faces of an x2 sweep, across the lines of an x3 sweep
    int f, a;  // face along the sweep, interior line across
    if (AX == 2) {
      f = r % nf;
      a = r / nf;
    } else {
      a = r % d.n;
      f = r / d.n;
    }
    int k, j;
    int64_t stride;
    if (AX == 2) {
      k = d.ng + a;
      j = d.ng + f;
      stride = 1;
    } else {
      k = d.ng + f;
      j = d.ng + a;
      stride = d.nc;
    }
    int64_t c = int64_t(k) * d.nc + j;
    T const* wb = w + b * d.block + c;
    T wl[kRows], wr[kRows];
    for (int v = 0; v < kRows; ++v) {
      T const* p = wb + v * d.plane;
      T c0 = p[-3 * stride], c1 = p[-2 * stride], c2 = p[-stride], c3 = p[0],
        c4 = p[stride], c5 = p[2 * stride];
      if (v == 0 || shock) {
        wl[v] = weno5_plus(c0, c1, c2, c3, c4);
        wr[v] = weno5_minus(c1, c2, c3, c4, c5);
      } else {
        wl[v] = cp5_plus(c0, c1, c2, c3, c4);
        wr[v] = cp5_minus(c1, c2, c3, c4, c5);
      }
    }
    if (f == 0 || f == d.n) {
      // sides L/R (x2) or B/T (x3): the state on the panel's own side
      int side = (AX == 2 ? 0 : 2) + (f == d.n ? 1 : 0);
      T* e = edge + ((b * 4 + side) * int64_t(d.n) + a) * kRows;
      T const* own = f == d.n ? wl : wr;
      for (int v = 0; v < kRows; ++v) e[v] = own[v];
    }
    double cf, sf;
    if (AX == 2) {
      int64_t q = int64_t(k) * (d.nc + 1) + j;
      cf = gm.cos_f2[q];
      sf = gm.sin_f2[q];
    } else {
      int64_t q = int64_t(k) * d.nc + j;
      cf = gm.cos_f3[q];
      sf = gm.sin_f3[q];
    }
    T out[kRows];
    face_flux<AX, T>(wl, wr, T(cf), T(sf), out);
    T* fb = flx + b * d.block + c;
    for (int v = 0; v < kRows; ++v) fb[v * d.plane] = out[v];
  }
};

template <int AX, typename T>
__global__ void __launch_bounds__(256)
    cs_sweep_kernel(T const* __restrict__ w, T* __restrict__ flx,
                    T* __restrict__ edge, Geom gm, Dims d, int shock) {
  CsSweepBody<AX, T>::run(block_index(), grid_size(), w, flx, edge, gm, d, shock);
}

// ---------------------------------------------------------------------------
// seams
// ---------------------------------------------------------------------------
// one thread per (seam face, member): the owner's flux from its own state and
// the neighbour's (through the two frame matrices), then the neighbour's copy
// through the global Cartesian frame
template <typename T>
struct CsSeamBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim, T* __restrict__ f2,
                                             T* __restrict__ f3, T const* __restrict__ edge,
                                             int32_t const* __restrict__ idx,
                                             double const* __restrict__ sd, int64_t nseam, Geom gm,
                                             Dims d) {
    int64_t t = int64_t(bid.x) * blockDim.x + threadIdx.x;
    if (t >= nseam * d.members) return;
    // the members of one seam face on consecutive lanes (its 30 frame values
    // read once per wavefront, as the ghost fill does)
    int64_t m = t % d.members, e = t / d.members;
    int32_t const* ix = idx + 8 * e;
    int pa = ix[0], sa = ix[1], ia = ix[2], pb = ix[3], sb = ix[4], ib = ix[5];
    double sign = double(ix[6]);
    double const* D = sd + 30 * e;
    int64_t ba = m * 6 + pa, bb = m * 6 + pb;
    T const* own = edge + ((ba * 4 + sa) * int64_t(d.n) + ia) * kRows;
    T const* oth = edge + ((bb * 4 + sb) * int64_t(d.n) + ib) * kRows;
    double mine[kRows], other[kRows];
    for (int v = 0; v < kRows; ++v) mine[v] = own[v];
    {
      double v0[3] = {oth[1], oth[2], oth[3]}, g[3], o[3];
      matvec(D, v0, g);      // the neighbour's components to the geographic frame
      matvec(D + 9, g, o);   // and into the owner's
      other[0] = oth[0];
      other[1] = o[0];
      other[2] = o[1];
      other[3] = o[2];
    }
    bool up = sa == 1 || sa == 3;
    double* wl = up ? mine : other;
    double* wr = up ? other : mine;
    int plane_a = sa & 1 ? d.ng + d.n : d.ng;
    double F[kRows];
    int64_t ca;
    if (sa < 2) {
      int k = d.ng + ia;
      int64_t q = int64_t(k) * (d.nc + 1) + plane_a;
      face_flux<2, double>(wl, wr, gm.cos_f2[q], gm.sin_f2[q], F);
      ca = int64_t(k) * d.nc + plane_a;
    } else {
      int j = d.ng + ia;
      int64_t q = int64_t(plane_a) * d.nc + j;
      face_flux<3, double>(wl, wr, gm.cos_f3[q], gm.sin_f3[q], F);
      ca = int64_t(plane_a) * d.nc + j;
    }
    T* fa = (sa < 2 ? f2 : f3) + ba * d.block + ca;
    for (int v = 0; v < kRows; ++v) fa[v * d.plane] = T(F[v]);

    // the owner's rows: raise with its face cosine, to Cartesian (z, x, y)
    double const* A = D + 18;  // x, y, delta, C, D, cos at the owner's face
    double const* Bm = D + 24;
    double cth = A[5];
    double sth2 = 1. - cth * cth;
    double r1 = F[1];
    double r2 = F[2] / sth2 - F[3] * cth / sth2;
    double r3 = -F[2] * cth / sth2 + F[3] / sth2;
    double cart[3];
    {
      double x = A[0], y = A[1], delta = A[2], C = A[3], Dd = A[4];
      double vz = r1, vx = r2, vy = r3;
      cart[kL2CIdx[pa][0]] = double(kL2CSgn[pa][0]) * (vz - vx * x / Dd - vy * y / C) / delta;
      cart[kL2CIdx[pa][1]] = double(kL2CSgn[pa][1]) * (vz * x + vx * Dd - (vy * x * y) / C) / delta;
      cart[kL2CIdx[pa][2]] = double(kL2CSgn[pa][2]) * (vz * y + vy * C - (vx * x * y) / Dd) / delta;
    }
    double g0 = F[0];
    if (sign < 0.) {
      g0 = -g0;
      cart[0] = -cart[0];
      cart[1] = -cart[1];
      cart[2] = -cart[2];
    }
    // into the other panel's contravariant, then lower with its face cosine
    double o1, o2, o3;
    {
      double x = Bm[0], y = Bm[1], delta = Bm[2], C = Bm[3], Dd = Bm[4];
      double loc[3];
      loc[kC2LIdx[pb][0]] = double(kC2LSgn[pb][0]) * cart[0];
      loc[kC2LIdx[pb][1]] = double(kC2LSgn[pb][1]) * cart[1];
      loc[kC2LIdx[pb][2]] = double(kC2LSgn[pb][2]) * cart[2];
      double vz = loc[0], vx = loc[1], vy = loc[2];
      o1 = (vz + x * vx + y * vy) / delta;
      o2 = Dd / delta * (vx - x * vz);
      o3 = C / delta * (vy - y * vz);
    }
    double cb = Bm[5];
    double l2 = o2 + o3 * cb;
    double l3 = o3 + o2 * cb;
    int plane_b = sb & 1 ? d.ng + d.n : d.ng;
    int64_t cbi = sb < 2 ? int64_t(d.ng + ib) * d.nc + plane_b
                         : int64_t(plane_b) * d.nc + (d.ng + ib);
    T* fb = (sb < 2 ? f2 : f3) + bb * d.block + cbi;
    fb[0] = T(g0);
    fb[d.plane] = T(o1);
    fb[2 * d.plane] = T(l2);
    fb[3 * d.plane] = T(l3);
  }
};

template <typename T>
__global__ void __launch_bounds__(256)
    cs_seam_kernel(T* __restrict__ f2, T* __restrict__ f3,
                   T const* __restrict__ edge, int32_t const* __restrict__ idx,
                   double const* __restrict__ sd, int64_t nseam, Geom gm, Dims d) {
  CsSeamBody<T>::run(block_index(), grid_size(), f2, f3, edge, idx, sd, nseam, gm, d);
}

// ---------------------------------------------------------------------------
// divergence, sources, combine
// ---------------------------------------------------------------------------
struct CsUpdArgs {
  double w0, w1, w2;
  double dt;
  double const* dt_dev;  // per member, or null
  double radius;         // x1v
};

// TU: the state (double, or the FP32 pair with u0c / uc); TW: the
// primitives and fluxes; the divergence, sources and combine in FP64
template <typename TU, typename TW>
struct CsUpdateBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             TU const* __restrict__ u0,
                                             float const* __restrict__ u0c, TU* __restrict__ u,
                                             float* __restrict__ uc, TW const* __restrict__ w,
                                             TW const* __restrict__ f2, TW const* __restrict__ f3,
                                             Geom gm, Dims d, CsUpdArgs a) {
    int64_t t = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t per = int64_t(d.n) * d.n;
    if (t >= per * 6 * d.members) return;
    int64_t b = t / per;
    int r = int(t % per);
    int k = d.ng + r / d.n, j = d.ng + r % d.n;
    int64_t c = int64_t(k) * d.nc + j;
    int64_t c2 = int64_t(k) * (d.nc + 1) + j;  // x2 face j of row k in area2
    double dt = a.dt_dev ? a.dt_dev[b / 6] : a.dt;
    double vol = gm.vol[c];
    TW const* wb = w + b * d.block + c;
    double phi = wb[0], v1 = wb[d.plane], v2 = wb[2 * d.plane], v3 = wb[3 * d.plane];
    double cth = gm.cos_c[c];
    double sine2 = gm.sin_c[c] * gm.sin_c[c];
    double pr = 0.5 * (phi * phi);
    double mdt = (-dt) * ((phi * v1) / a.radius);
    double a2l = gm.area2[c2], a2r = gm.area2[c2 + 1];
    double a3l = gm.area3[c], a3r = gm.area3[c + d.nc];
    for (int v = 0; v < kRows; ++v) {
      int64_t o = b * d.block + v * d.plane + c;
      double dv = 0. + (a2r * double(f2[o + 1]) - a2l * double(f2[o]));
      dv = dv + (a3r * double(f3[o + d.nc]) - a3l * double(f3[o]));
      dv = dv / vol;
      if (v == 2) dv = dv - gm.xr[c] * (pr + phi * v3 * v3 * sine2);
      if (v == 3) dv = dv - gm.yr[c] * (pr + phi * v2 * v2 * sine2);
      double du = -dt * dv;
      // the x1 mass source is added to every row (zero on phi and phi v1)
      if (v == 2)
        du = du + (0. + mdt * (v2 + v3 * cth));
      else if (v == 3)
        du = du + (0. + mdt * (v3 + v2 * cth));
      else
        du = du + 0.;
      cs_put(u, uc, o, a.w0 * cs_val(u0, u0c, o) + a.w1 * cs_val(u, uc, o) + a.w2 * du);
    }
  }
};

template <typename TU, typename TW>
__global__ void __launch_bounds__(256)
    cs_update_kernel(TU const* __restrict__ u0, float const* __restrict__ u0c,
                     TU* __restrict__ u, float* __restrict__ uc,
                     TW const* __restrict__ w, TW const* __restrict__ f2,
                     TW const* __restrict__ f3, Geom gm, Dims d, CsUpdArgs a) {
  CsUpdateBody<TU, TW>::run(block_index(), grid_size(), u0, u0c, u, uc, w, f2, f3, gm, d, a);
}

// ---------------------------------------------------------------------------
// CFL, per member
// ---------------------------------------------------------------------------
struct CsCflInitBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             unsigned long long* out, int members) {
    int m = bid.x * blockDim.x + threadIdx.x;
    if (m < members) {
      double big = 1.e9;
      unsigned long long bits;
      memcpy(&bits, &big, sizeof(bits));
      out[m] = bits;
    }
  }
};

__global__ void cs_cfl_init_kernel(unsigned long long* out, int members) {
  CsCflInitBody::run(block_index(), grid_size(), out, members);
}

template <typename TW>
struct CsCflBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             TW const* __restrict__ w, Geom gm, Dims d,
                                             unsigned long long* __restrict__ out) {
    // one block row of 256 threads per (member, chunk); bid.y = member
    int m = bid.y;
    int64_t per = int64_t(d.n) * d.n * 6;
    double best = 1.e9;
    for (int64_t t = int64_t(bid.x) * blockDim.x + threadIdx.x; t < per;
         t += int64_t(gdim.x) * blockDim.x) {
      int64_t p = t / (int64_t(d.n) * d.n);
      int r = int(t % (int64_t(d.n) * d.n));
      int64_t c = int64_t(d.ng + r / d.n) * d.nc + (d.ng + r % d.n);
      TW const* wb = w + (int64_t(m) * 6 + p) * d.block + c;
      double cs = sqrt(double(wb[0]));
      best = fmin(best, gm.w2[c] / (fabs(double(wb[2 * d.plane])) + cs));
      best = fmin(best, gm.w3[c] / (fabs(double(wb[3 * d.plane])) + cs));
    }
    for (int off = 32; off > 0; off >>= 1) best = fmin(best, __shfl_down(best, off, 64));
    if ((threadIdx.x & 63) == 0) {
      unsigned long long bits;
      memcpy(&bits, &best, sizeof(bits));
      atomicMin(out + m, bits);
    }
  }
};

template <typename TW>
__global__ void __launch_bounds__(256)
    cs_cfl_kernel(TW const* __restrict__ w, Geom gm, Dims d,
                  unsigned long long* __restrict__ out) {
  CsCflBody<TW>::run(block_index(), grid_size(), w, gm, d, out);
}

struct CsCflScaleBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim, double* out,
                                             double s, int members) {
    int m = bid.x * blockDim.x + threadIdx.x;
    if (m < members) out[m] = out[m] * s;
  }
};

__global__ void cs_cfl_scale_kernel(double* out, double s, int members) {
  CsCflScaleBody::run(block_index(), grid_size(), out, s, members);
}

int check_mesh(snapy_cs_mesh const* m) {
  if (!m) return fail(-1, "null mesh descriptor");
  if (m->n < 4 || m->nghost != 3 || m->members < 1)
    return fail(-2, "cubed sphere: n >= 4, nghost 3, members >= 1");
  if (m->interp_order != 2 && m->interp_order != 4)
    return fail(-3, "cubed sphere: interp_order 2 or 4");
  if (!m->geom) return fail(-4, "cubed sphere: null grid tables");
  return 0;
}

}  // namespace
}  // namespace snapy

using namespace snapy;

namespace {

template <typename T>
int cs_sweep_launch(const snapy_cs_mesh* m, int dim, const T* w, T* flux, T* edge,
                           void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (dim != 2 && dim != 3) return fail(-8, "cubed sphere: sweeps are x2 and x3");
  if (!w || !flux || !edge) return fail(-4, "cubed sphere: null sweep buffer");
  Dims d = dims_of(*m);
  Geom gm = geom_of(m->geom, d.nc);
  int64_t n = int64_t(d.n) * (d.n + 1) * 6 * d.members;
  if (dim == 2)
    launch<CsSweepBody<2, T>, 256, 1>((cs_sweep_kernel<2, T>), dim3(blocks_for(n, 256)), dim3(256), 0,
                       as_stream(stream), w, flux, edge, gm, d, m->shock);
  else
    launch<CsSweepBody<3, T>, 256, 1>((cs_sweep_kernel<3, T>), dim3(blocks_for(n, 256)), dim3(256), 0,
                       as_stream(stream), w, flux, edge, gm, d, m->shock);
  return launch_status("cs_sweep");
}

template <typename T>
int cs_seams_launch(const snapy_cs_mesh* m, T* flux2, T* flux3, const T* edge,
                           void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (!m->seam_idx || !m->seam_d) return fail(-4, "cubed sphere: null seam tables");
  if (!flux2 || !flux3 || !edge) return fail(-4, "cubed sphere: null seam buffer");
  Dims d = dims_of(*m);
  Geom gm = geom_of(m->geom, d.nc);
  int64_t nseam = int64_t(12) * d.n;
  int64_t n = nseam * d.members;
  launch<CsSeamBody<T>, 256, 1>(cs_seam_kernel<T>, dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), flux2, flux3, edge, m->seam_idx, m->seam_d, nseam,
                     gm, d);
  return launch_status("cs_seams");
}

template <typename TW>
int cs_cfl_launch(const snapy_cs_mesh* m, const TW* w, double scale, double* dt_out,
                         void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (!dt_out || !w) return fail(-4, "cubed sphere: null dt output or primitives");
  Dims d = dims_of(*m);
  Geom gm = geom_of(m->geom, d.nc);
  auto* bits = reinterpret_cast<unsigned long long*>(dt_out);
  unsigned mb = blocks_for(d.members, 64);
  launch<CsCflInitBody, 1024, 1>(cs_cfl_init_kernel, dim3(mb), dim3(64), 0, as_stream(stream), bits,
                     d.members);
  int64_t per = int64_t(d.n) * d.n * 6;
  unsigned gx = blocks_for(per, 256);
  if (gx > 64) gx = 64;
  launch<CsCflBody<TW>, 256, 1>(cs_cfl_kernel<TW>, dim3(gx, d.members), dim3(256), 0, as_stream(stream), w,
                     gm, d, bits);
  launch<CsCflScaleBody, 1024, 1>(cs_cfl_scale_kernel, dim3(mb), dim3(64), 0, as_stream(stream), dt_out,
                     scale, d.members);
  return launch_status("cs_cfl");
}

}  // namespace

extern "C" {

int64_t snapy_cs_geom_size(int32_t n, int32_t nghost) {
  int64_t nc = n + 2 * nghost;
  return 7 * nc * nc + 6 * nc * (nc + 1);
}

int64_t snapy_cs_edge_size(const snapy_cs_mesh* m) {
  if (!m) return -1;
  return int64_t(m->members) * 6 * 4 * m->n * kRows;
}

int snapy_cs_cons2prim(const snapy_cs_mesh* m, const double* u, double* w, void* stream) {
  if (int rc = check_mesh(m)) return rc;
  Dims d = dims_of(*m);
  int64_t n = d.plane * 6 * d.members;
  launch<CsCons2PrimBody<double, double>, 256, 1>((cs_cons2prim_kernel<double, double>), dim3(blocks_for(n, 256)), dim3(256),
                     0, as_stream(stream), u, nullptr, w, m->geom, d, n);
  return launch_status("cs_cons2prim");
}

int snapy_cs_prim2cons(const snapy_cs_mesh* m, const double* w, double* u, void* stream) {
  if (int rc = check_mesh(m)) return rc;
  Dims d = dims_of(*m);
  int64_t n = d.plane * 6 * d.members;
  launch<CsPrim2ConsBody<double>, 256, 1>(cs_prim2cons_kernel<double>, dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), w, u, nullptr, m->geom, d, n);
  return launch_status("cs_prim2cons");
}

int snapy_cs_ghosts(const snapy_cs_mesh* m, int conserved, double* q, void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (!m->ghost_idx || !m->ghost_w || !m->ghost_mat_c || !m->ghost_mat_p)
    return fail(-4, "cubed sphere: null ghost tables");
  Dims d = dims_of(*m);
  int64_t n = ghost_threads(m->nghost_entries, d.members);
  launch<CsGhostBody<double>, 256, 1>(cs_ghost_kernel<double>, dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), q, nullptr, m->ghost_idx, m->ghost_w,
                     conserved ? m->ghost_mat_c : m->ghost_mat_p, m->nghost_entries, d);
  return launch_status("cs_ghosts");
}


int snapy_cs_sweep(const snapy_cs_mesh* m, int dim, const double* w, double* flux,
                   double* edge, void* stream) {
  return cs_sweep_launch<double>(m, dim, w, flux, edge, stream);
}


int snapy_cs_seams(const snapy_cs_mesh* m, double* flux2, double* flux3,
                   const double* edge, void* stream) {
  return cs_seams_launch<double>(m, flux2, flux3, edge, stream);
}

int snapy_cs_update(const snapy_cs_mesh* m, const double* u0, double* u, const double* w,
                    const double* flux2, const double* flux3, double dt,
                    const double* dt_dev, double w0, double w1, double w2, void* stream) {
  if (int rc = check_mesh(m)) return rc;
  Dims d = dims_of(*m);
  Geom gm = geom_of(m->geom, d.nc);
  CsUpdArgs a{w0, w1, w2, dt, dt_dev, m->radius};
  int64_t n = int64_t(d.n) * d.n * 6 * d.members;
  launch<CsUpdateBody<double, double>, 256, 1>((cs_update_kernel<double, double>), dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), u0, nullptr, u, nullptr, w, flux2, flux3, gm, d, a);
  return launch_status("cs_update");
}


int snapy_cs_cfl(const snapy_cs_mesh* m, const double* w, double scale, double* dt_out,
                 void* stream) {
  return cs_cfl_launch<double>(m, w, scale, dt_out, stream);
}

// ---- the FP32 form ----------------------------------------------------------
int snapy_cs_cons2prim_f32(const snapy_cs_mesh* m, const float* u, const float* c, float* w,
                           void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (!u || !c || !w) return fail(-4, "cubed sphere: null buffer");
  Dims d = dims_of(*m);
  int64_t n = d.plane * 6 * d.members;
  launch<CsCons2PrimBody<float, float>, 256, 1>((cs_cons2prim_kernel<float, float>), dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), u, c, w, m->geom, d, n);
  return launch_status("cs_cons2prim_f32");
}

int snapy_cs_prim2cons_f32(const snapy_cs_mesh* m, const double* w, float* u, float* c,
                           void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (!u || !c || !w) return fail(-4, "cubed sphere: null buffer");
  Dims d = dims_of(*m);
  int64_t n = d.plane * 6 * d.members;
  launch<CsPrim2ConsBody<float>, 256, 1>(cs_prim2cons_kernel<float>, dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), w, u, c, m->geom, d, n);
  return launch_status("cs_prim2cons_f32");
}

int snapy_cs_ghosts_f32(const snapy_cs_mesh* m, float* u, float* c, void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (!m->ghost_idx || !m->ghost_w || !m->ghost_mat_c)
    return fail(-4, "cubed sphere: null ghost tables");
  if (!u || !c) return fail(-4, "cubed sphere: null buffer");
  Dims d = dims_of(*m);
  int64_t n = ghost_threads(m->nghost_entries, d.members);
  launch<CsGhostBody<float>, 256, 1>(cs_ghost_kernel<float>, dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), u, c, m->ghost_idx, m->ghost_w, m->ghost_mat_c,
                     m->nghost_entries, d);
  return launch_status("cs_ghosts_f32");
}

int snapy_cs_sweep_f32(const snapy_cs_mesh* m, int dim, const float* w, float* flux,
                       float* edge, void* stream) {
  return cs_sweep_launch<float>(m, dim, w, flux, edge, stream);
}

int snapy_cs_seams_f32(const snapy_cs_mesh* m, float* flux2, float* flux3, const float* edge,
                       void* stream) {
  return cs_seams_launch<float>(m, flux2, flux3, edge, stream);
}

int snapy_cs_update_f32(const snapy_cs_mesh* m, const float* u0, const float* c0, float* u,
                        float* c, const float* w, const float* flux2, const float* flux3,
                        double dt, const double* dt_dev, double w0, double w1, double w2,
                        void* stream) {
  if (int rc = check_mesh(m)) return rc;
  if (!u0 || !c0 || !u || !c || !w || !flux2 || !flux3)
    return fail(-4, "cubed sphere: null buffer");
  Dims d = dims_of(*m);
  Geom gm = geom_of(m->geom, d.nc);
  CsUpdArgs a{w0, w1, w2, dt, dt_dev, m->radius};
  int64_t n = int64_t(d.n) * d.n * 6 * d.members;
  launch<CsUpdateBody<float, float>, 256, 1>((cs_update_kernel<float, float>), dim3(blocks_for(n, 256)), dim3(256), 0,
                     as_stream(stream), u0, c0, u, c, w, flux2, flux3, gm, d, a);
  return launch_status("cs_update_f32");
}

int snapy_cs_cfl_f32(const snapy_cs_mesh* m, const float* w, double scale, double* dt_out,
                     void* stream) {
  return cs_cfl_launch<float>(m, w, scale, dt_out, stream);
}

}  // extern "C"
