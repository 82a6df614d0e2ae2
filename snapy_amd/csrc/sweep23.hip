// x2/x3 face sweep with the x1 flux covariance (dynamics/flux-covariance,
// the reference default): the O(dx1^2) covariance of the enthalpy flux is
// added to the energy row of every x2/x3 face flux
// (HydroImpl::_flux_covariance, src/hydro/hydro_forward.cpp:58-231, the
// Cartesian branch :196-213; add_cov :782-793).
//
// The covariance of a face needs the face states of its two x1 neighbours:
// every lane reconstructs its own face states, the enthalpy h and the normal
// velocity u_n go through LDS, and the inner lanes solve the Riemann problem
// and add the covariance.  When a whole x1 line (nx1 + 2 cells with the two
// halo cells) fits, a workgroup packs several whole lines; otherwise a
// workgroup of 64 lanes walks 62 consecutive x1 cells plus one halo lane on
// each side.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

constexpr int kLanes = 64;
constexpr int kInner = kLanes - 2;

struct CovArgs {
  Range3 r;
  int nf;  // faces along the sweep axis
  double gammad;
  int riemann, shock;
  int wall_lo, wall_hi;
  int cov;    // add the covariance
  double s2;  // dx1^2 / 12 (face_moment2_x1, uniform x1)
  int gauge;  // FP32 LMARS: subtract the line's pressure gauge
};

// The primitive state is FP64; T is the compute and flux precision.  With
// a.gauge the pressure row is reconstructed as p - g, g the pressure of the
// line's first interior cell along the sweep (the momentum flux stays
// gauge-free; the enthalpy and the sound speed take p - g + g).
//
// The face states of the face below cell (i, j, k) along the sweep, and the
// h and u_n the covariance of the neighbouring x1 faces reads.
template <typename T, int DIR>
__device__ __forceinline__ void cov_states(double const* __restrict__ w, CovArgs const& a, int i,
                                           int j, int k, bool live, Prim5<T>& L, Prim5<T>& R,
                                           double& h, double& un, double& gauge) {
  Range3 const& r = a.r;
  int64_t stride = DIR == 2 ? r.s2 : r.s3;
  h = 0.;
  un = 0.;
  gauge = 0.;
  if (live && a.gauge)
    gauge = w[IPR * r.sv + (DIR == 2 ? k * r.s3 + r.lo2 * r.s2 + i
                                     : r.lo3 * r.s3 + j * r.s2 + i)];
  if (!live) return;
  int64_t base = k * r.s3 + j * r.s2 + i;
  T* lv[5] = {&L.r, &L.v[0], &L.v[1], &L.v[2], &L.p};
  T* rv[5] = {&R.r, &R.v[0], &R.v[1], &R.v[2], &R.p};
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double const* p = w + v * r.sv + base;
    double gv = v == IPR ? gauge : 0.;
    T c0 = T(p[-3 * stride] - gv), c1 = T(p[-2 * stride] - gv),
      c2 = T(p[-stride] - gv), c3 = T(p[0] - gv), c4 = T(p[stride] - gv),
      c5 = T(p[2 * stride] - gv);
    if (v == IDN || a.shock) {
      *lv[v] = weno5_plus(c0, c1, c2, c3, c4);
      *rv[v] = weno5_minus(c1, c2, c3, c4, c5);
    } else {
      *lv[v] = cp5_plus(c0, c1, c2, c3, c4);
      *rv[v] = cp5_minus(c1, c2, c3, c4, c5);
    }
  }
  if (a.cov) {
    // wbar = 0.5 (wl + wr); h = (W->I(wbar) + p) / rho; u_n.  FP64 in the
    // FP64 form; in the FP32 form the FP32 quotients of qdiv (the term is a
    // small correction to the energy flux)
    double rho = 0.5 * (double(L.r) + double(R.r));
    double p = 0.5 * (double(L.p) + double(R.p)) + gauge;
    if constexpr (sizeof(T) == 8) {
      double enth = p / (a.gammad - 1.) + p;
      h = enth / rho;
    } else {
      float enth = qdiv(float(p), float(a.gammad - 1.)) + float(p);
      h = double(qdiv(enth, float(rho)));
    }
    un = 0.5 * (double(L.v[DIR - 1]) + double(R.v[DIR - 1]));
  }
}

// cov_states with the stencil from a staged tile: st(v, m) is row v of the
// m-th cell of the face's stencil (m = 0..5 for the cells f - 3 .. f + 2
// along the sweep), already shifted by the gauge and in T.
template <typename T, int DIR, typename ST>
__device__ __forceinline__ void cov_states_tile(CovArgs const& a, ST st, double gauge,
                                                Prim5<T>& L, Prim5<T>& R, double& h,
                                                double& un) {
  T* lv[5] = {&L.r, &L.v[0], &L.v[1], &L.v[2], &L.p};
  T* rv[5] = {&R.r, &R.v[0], &R.v[1], &R.v[2], &R.p};
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    T c0 = st(v, 0), c1 = st(v, 1), c2 = st(v, 2), c3 = st(v, 3), c4 = st(v, 4), c5 = st(v, 5);
    if (v == IDN || a.shock) {
      *lv[v] = weno5_plus(c0, c1, c2, c3, c4);
      *rv[v] = weno5_minus(c1, c2, c3, c4, c5);
    } else {
      *lv[v] = cp5_plus(c0, c1, c2, c3, c4);
      *rv[v] = cp5_minus(c1, c2, c3, c4, c5);
    }
  }
  h = 0.;
  un = 0.;
  if (a.cov) {
    double rho = 0.5 * (double(L.r) + double(R.r));
    double p = 0.5 * (double(L.p) + double(R.p)) + gauge;
    if constexpr (sizeof(T) == 8) {
      double enth = p / (a.gammad - 1.) + p;
      h = enth / rho;
    } else {
      float enth = qdiv(float(p), float(a.gammad - 1.)) + float(p);
      h = double(qdiv(enth, float(rho)));
    }
    un = 0.5 * (double(L.v[DIR - 1]) + double(R.v[DIR - 1]));
  }
}

// The flux of an interior face from its states, the covariance from the
// x1 neighbours' h and u_n (hm, hp, um, up), the wall prescription, stored.
template <typename T, int DIR>
__device__ __forceinline__ void cov_flux(double const* __restrict__ x1v, T* __restrict__ flx,
                                         CovArgs const& a, int i, int j, int k,
                                         Prim5<T> const& L, Prim5<T> const& R, double gauge,
                                         double hm, double hp, double um, double up) {
  Range3 const& r = a.r;
  T g = T(a.gammad);
  T out[5];
  T pface;
  if (a.riemann == SNAPY_RIEMANN_LMARS)
    lmars(L, R, g, g, DIR - 1, out, &pface, T(gauge));
  else
    hllc(L, R, g, g, DIR - 1, out, &pface);
  if (a.cov) {
    double rho = 0.5 * (double(L.r) + double(R.r));
    double dx1 = x1v[i + 1] - x1v[i - 1];
    double ce;
    if constexpr (sizeof(T) == 8) {
      double d1h = (hp - hm) / dx1;
      double dun = (up - um) / dx1;
      ce = ((a.s2 * rho) * d1h) * dun;
    } else {
      float idx = qdiv(1.f, float(dx1));
      float d1h = float(hp - hm) * idx, dun = float(up - um) * idx;
      ce = double((float(a.s2 * rho) * d1h) * dun);
    }
    out[IPR] = T(double(out[IPR]) + ce);
  }
  int f = DIR == 2 ? j - r.lo2 : k - r.lo3;
  if ((f == 0 && a.wall_lo) || (f == a.nf - 1 && a.wall_hi)) {
    out[IDN] = T(0);
    out[IPR] = T(0);
  }
  int64_t base = k * r.s3 + j * r.s2 + i;
#pragma unroll
  for (int v = 0; v < 5; ++v) flx[v * r.sv + base] = out[v];
}

// Segmented lines: a workgroup of 64 lanes walks 62 consecutive x1 cells
// plus one halo lane on each side (any nx1).
template <typename T, int DIR>
struct Sweep23Body {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ x1v, T* __restrict__ flx,
                                             CovArgs a) {
    Range3 const& r = a.r;
    __shared__ double sh_h[kLanes], sh_un[kLanes];
    int t = threadIdx.x;
    int i = r.lo1 - 1 + int(bid.y) * kInner + t;
    // bid.x: (other axis, face) -- DIR 2: (k, j face); DIR 3: (k face, j).
    // DIR 2 goes through the XCD map, so the neighbouring faces that share
    // stencil lines share an L2; in DIR 3 the faces that share a line are n2
    // workgroups apart (the same XCD when 8 divides n2), and the natural order
    // keeps each XCD on an eighth of the lines, whose reuse distance then fits
    // its L2
    int64_t y = DIR == 2 ? xcd_tile(bid.x, gdim.x) : bid.x;
    int j, k;
    if (DIR == 2) {
      j = r.lo2 + int(y % a.nf);
      k = r.lo3 + int(y / a.nf);
    } else {
      j = r.lo2 + int(y % r.n2);
      k = r.lo3 + int(y / r.n2);
    }
    bool live = i >= r.lo1 - 1 && i <= r.lo1 + r.n1;
    Prim5<T> L, R;
    double h, un, gauge;
    cov_states<T, DIR>(w, a, i, j, k, live, L, R, h, un, gauge);
    sh_h[t] = h;
    sh_un[t] = un;
    __syncthreads();
    if (t == 0 || t == kLanes - 1) return;
    if (!(i >= r.lo1 && i < r.lo1 + r.n1)) return;
    cov_flux<T, DIR>(x1v, flx, a, i, j, k, L, R, gauge, sh_h[t - 1], sh_h[t + 1], sh_un[t - 1],
                     sh_un[t + 1]);
  }
};

template <typename T, int DIR>
__global__ void __launch_bounds__(kLanes)
    sweep23_kernel(double const* __restrict__ w, double const* __restrict__ x1v,
                   T* __restrict__ flx, CovArgs a) {
  Sweep23Body<T, DIR>::run(block_index(), grid_size(), w, x1v, flx, a);
}

// Packed lines (nx1 + 2 <= kPackMax): a workgroup holds `lines` whole x1
// lines of nx1 + 2 lanes each (the two halo cells included), so few lanes
// idle; its lines are consecutive faces along the sweep (the face index
// fastest), which share five of their six stencil rows, and the workgroups
// go through the XCD map so neighbouring workgroups share an L2.
constexpr int kPackThreads = 1024;
constexpr int kPackMax = 512;

template <typename T, int DIR>
struct Sweep23PackedBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ x1v, T* __restrict__ flx,
                                             CovArgs a, int lines) {
    Range3 const& r = a.r;
    __shared__ double sh_h[kPackThreads], sh_un[kPackThreads];
    int t = threadIdx.x;
    int len = r.n1 + 2;
    int q = t / len, p = t - q * len;
    int64_t nline = int64_t(a.nf) * (DIR == 2 ? r.n3 : r.n2);
    int64_t y = int64_t(xcd_tile(bid.x, gdim.x)) * lines + q;
    bool in = q < lines && y < nline;
    int i = r.lo1 - 1 + p;
    int j, k;
    if (DIR == 2) {
      j = r.lo2 + int(y % a.nf);
      k = r.lo3 + int(y / a.nf);
    } else {
      k = r.lo3 + int(y % a.nf);
      j = r.lo2 + int(y / a.nf);
    }
    Prim5<T> L, R;
    double h, un, gauge;
    cov_states<T, DIR>(w, a, i, j, k, in, L, R, h, un, gauge);
    sh_h[t] = h;
    sh_un[t] = un;
    __syncthreads();
    if (!in || p == 0 || p == len - 1) return;
    cov_flux<T, DIR>(x1v, flx, a, i, j, k, L, R, gauge, sh_h[t - 1], sh_h[t + 1], sh_un[t - 1],
                     sh_un[t + 1]);
  }
};

template <typename T, int DIR>
__global__ void __launch_bounds__(kPackThreads)
    sweep23_packed_kernel(double const* __restrict__ w, double const* __restrict__ x1v,
                          T* __restrict__ flx, CovArgs a, int lines) {
  Sweep23PackedBody<T, DIR>::run(block_index(), grid_size(), w, x1v, flx, a, lines);
}

// Tiled lines (nx1 + 2 <= kPackMax): as the packed kernel, a workgroup
// holds `lines` whole x1 lines of nx1 + 2 lanes, the faces f0 .. f0 +
// lines - 1 of one line along the sweep (a tile never straddles two lines,
// so the tile shares one pressure gauge per x1 cell).  The tile's cells
// f0 - 3 .. f0 + lines + 1 along the sweep are staged once in LDS, every
// row already shifted by the gauge and rounded to T exactly as the lanes of
// the packed kernel round them (so the fluxes are the same bits), with the
// loads coalesced along x1; each lane then reads its six-cell stencils from
// LDS instead of global memory.  Dynamic LDS: 5 (lines + 5) (nx1 + 2) T.
template <typename T, int DIR>
struct Sweep23TileBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ x1v, T* __restrict__ flx,
                                             CovArgs a, int lines, int tiles_per_line) {
    extern __shared__ char smem_raw[];
    T* st = reinterpret_cast<T*>(smem_raw);
    __shared__ double sh_h[kPackThreads], sh_un[kPackThreads];
    Range3 const& r = a.r;
    int const t = threadIdx.x;
    int const len = r.n1 + 2;
    int const ncell = lines + 5;
    // the tile: line o (k for DIR 2, j for DIR 3) and its first face f0
    int64_t tile = xcd_tile(bid.x, gdim.x);
    int o = int(tile / tiles_per_line);
    int f0 = int(tile - int64_t(o) * tiles_per_line) * lines;
    int nl = a.nf - f0 < lines ? a.nf - f0 : lines;
    int64_t sweep = DIR == 2 ? r.s2 : r.s3;
    int lo = DIR == 2 ? r.lo2 : r.lo3;
    int64_t obase = DIR == 2 ? int64_t(r.lo3 + o) * r.s3 : int64_t(r.lo2 + o) * r.s2;
    // the line's gauge row: the first interior cell along the sweep
    double const* gp = w + IPR * r.sv + obase + int64_t(lo) * sweep + (r.lo1 - 1);
    // stage cells lo + f0 - 3 .. lo + f0 + nl + 1 (the rest of the tile area
    // is left unread)
    {
      // rows of len cells, `per` rows at a time: thread t takes position sp
      // of rows sr, sr + per, ...; each row's loads run along x1
      int const per = int(blockDim.x) / len;
      int const sr = t / len, sp = t - sr * len;
      if (sr < per) {
        double const g = a.gauge ? gp[sp] : 0.;
        double const* src = w + obase + int64_t(lo + f0 - 3) * sweep + (r.lo1 - 1 + sp);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          double const gv = v == IPR ? g : 0.;
          for (int c = sr; c < nl + 5; c += per)
            st[(v * ncell + c) * len + sp] = T(src[v * r.sv + c * sweep] - gv);
        }
      }
    }
    __syncthreads();
    int q = t / len, p = t - q * len;
    bool in = q < nl;
    int i = r.lo1 - 1 + p;
    int f = f0 + q;
    int j = DIR == 2 ? lo + f : r.lo2 + o;
    int k = DIR == 2 ? r.lo3 + o : lo + f;
    Prim5<T> L, R;
    double h = 0., un = 0., gauge = 0.;
    if (in) {
      gauge = a.gauge ? gp[p] : 0.;
      cov_states_tile<T, DIR>(
          a, [&](int v, int m) { return st[(v * ncell + q + m) * len + p]; }, gauge, L, R, h, un);
    }
    sh_h[t] = h;
    sh_un[t] = un;
    __syncthreads();
    if (!in || p == 0 || p == len - 1) return;
    cov_flux<T, DIR>(x1v, flx, a, i, j, k, L, R, gauge, sh_h[t - 1], sh_h[t + 1], sh_un[t - 1],
                     sh_un[t + 1]);
  }
};

template <typename T, int DIR>
__global__ void __launch_bounds__(kPackThreads)
    sweep23_tile_kernel(double const* __restrict__ w, double const* __restrict__ x1v,
                        T* __restrict__ flx, CovArgs a, int lines, int tiles_per_line) {
  Sweep23TileBody<T, DIR>::run(block_index(), grid_size(), w, x1v, flx, a, lines, tiles_per_line);
}

// lines per packed workgroup: the count whose lanes waste the least of
// their whole wavefronts (ties: fewer lines)
int packed_lines(int len) {
  int best = 1;
  double eff = 0.;
  for (int P = 1; P * len <= kPackThreads; ++P) {
    int lanes = P * len;
    double e = double(lanes) / double((lanes + 63) / 64 * 64);
    if (e > eff + 1e-12) {
      eff = e;
      best = P;
    }
  }
  return best;
}

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

int snapy_sweep_cov(int dim, const snapy_grid* g, const snapy_physics* p,
                    int dtype, const void* w, const double* x1tab, void* flux,
                    void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !x1tab || !flux) return fail(-1, "sweep_cov: NULL argument");
  if (dim != 2 && dim != 3) return fail(-8, "sweep_cov: dim must be 2 or 3");
  if (g->nvar != 5) return fail(-6, "sweep_cov: nvar 5");
  if (g->nx1 < 3) return fail(-9, "sweep_cov: the covariance needs nx1 >= 3");
  int nx[3] = {g->nx1, g->nx2, g->nx3};
  if (nx[dim - 1] < 2) return fail(-9, "sweep_cov: inactive direction");
  CovArgs a;
  a.r = ranges(*g);
  a.nf = nx[dim - 1] + 1;
  a.gammad = p->gammad;
  a.riemann = p->riemann;
  a.shock = p->shock_x23;
  a.wall_lo = p->bc[2 * (dim - 1)] == SNAPY_BC_REFLECTING;
  a.wall_hi = p->bc[2 * (dim - 1) + 1] == SNAPY_BC_REFLECTING;
  a.cov = p->flux_covariance;
  a.s2 = g->dx1 * g->dx1 / 12.;
  a.gauge = dtype == SNAPY_F32 && p->riemann == SNAPY_RIEMANN_LMARS;
  unsigned nseg = blocks_for(g->nx1, kInner);
  int64_t ny = dim == 2 ? int64_t(a.nf) * g->nx3 : int64_t(g->nx2) * a.nf;
  dim3 grid(unsigned(ny), nseg);
  hipStream_t s = as_stream(stream);
  int len = g->nx1 + 2;
  if (len <= kPackMax) {
    int lines = packed_lines(len);
    size_t tsize = dtype == SNAPY_F64 ? sizeof(double) : sizeof(float);
    // the tiled kernel: fewer lines per workgroup when the staged tile
    // (5 (lines + 5) len values) would pass 48 KiB
    // (FP32 only: in FP64 the tile halves the lines per workgroup, and the
    // packed kernel is faster: 0.33 against 0.49 ms at 256x256x128)
    int tl = lines;
    while (tl > 1 && size_t(5) * (tl + 5) * len * tsize > 48 * 1024) --tl;
    size_t tlds = size_t(5) * (tl + 5) * len * tsize;
    if (tlds <= 48 * 1024 && dtype == SNAPY_F32) {
      int per = (a.nf + tl - 1) / tl;
      int64_t nlines = dim == 2 ? g->nx3 : g->nx2;
      unsigned threads = unsigned((tl * len + 63) / 64 * 64);
      unsigned blocks = unsigned(nlines * per);
      if (dtype == SNAPY_F64) {
        auto go = [&](auto body, auto k) {
          launch_flex<decltype(body), kPackThreads, 1>(k, dim3(blocks), dim3(threads), tlds, s,
                                                  static_cast<double const*>(w), x1tab,
                                                  static_cast<double*>(flux), a, tl, per);
        };
        if (dim == 2)
          go(Sweep23TileBody<double, 2>{}, sweep23_tile_kernel<double, 2>);
        else
          go(Sweep23TileBody<double, 3>{}, sweep23_tile_kernel<double, 3>);
      } else {
        auto go = [&](auto body, auto k) {
          launch_flex<decltype(body), kPackThreads, 1>(k, dim3(blocks), dim3(threads), tlds, s,
                                                  static_cast<double const*>(w), x1tab,
                                                  static_cast<float*>(flux), a, tl, per);
        };
        if (dim == 2)
          go(Sweep23TileBody<float, 2>{}, sweep23_tile_kernel<float, 2>);
        else
          go(Sweep23TileBody<float, 3>{}, sweep23_tile_kernel<float, 3>);
      }
      return launch_status("sweep_cov");
    }
    unsigned threads = unsigned((lines * len + 63) / 64 * 64);
    unsigned blocks = unsigned((ny + lines - 1) / lines);
    if (dtype == SNAPY_F64) {
      auto go = [&](auto body, auto k) {
        launch_flex<decltype(body), kPackThreads, 1>(k, dim3(blocks), dim3(threads), 0, s,
                                                static_cast<double const*>(w), x1tab,
                                                static_cast<double*>(flux), a, lines);
      };
      if (dim == 2)
        go(Sweep23PackedBody<double, 2>{}, sweep23_packed_kernel<double, 2>);
      else
        go(Sweep23PackedBody<double, 3>{}, sweep23_packed_kernel<double, 3>);
    } else if (dtype == SNAPY_F32) {
      auto go = [&](auto body, auto k) {
        launch_flex<decltype(body), kPackThreads, 1>(k, dim3(blocks), dim3(threads), 0, s,
                                                static_cast<double const*>(w), x1tab,
                                                static_cast<float*>(flux), a, lines);
      };
      if (dim == 2)
        go(Sweep23PackedBody<float, 2>{}, sweep23_packed_kernel<float, 2>);
      else
        go(Sweep23PackedBody<float, 3>{}, sweep23_packed_kernel<float, 3>);
    } else {
      return fail(-10, "unknown dtype");
    }
    return launch_status("sweep_cov");
  }
  if (dtype == SNAPY_F64) {
    auto go = [&](auto body, auto k) {
      launch<decltype(body), kLanes, 1>(k, grid, dim3(kLanes), 0, s,
                                        static_cast<double const*>(w), x1tab,
                                        static_cast<double*>(flux), a);
    };
    if (dim == 2)
      go(Sweep23Body<double, 2>{}, sweep23_kernel<double, 2>);
    else
      go(Sweep23Body<double, 3>{}, sweep23_kernel<double, 3>);
  } else if (dtype == SNAPY_F32) {
    auto go = [&](auto body, auto k) {
      launch<decltype(body), kLanes, 1>(k, grid, dim3(kLanes), 0, s,
                                        static_cast<double const*>(w), x1tab,
                                        static_cast<float*>(flux), a);
    };
    if (dim == 2)
      go(Sweep23Body<float, 2>{}, sweep23_kernel<float, 2>);
    else
      go(Sweep23Body<float, 3>{}, sweep23_kernel<float, 3>);
  } else {
    return fail(-10, "unknown dtype");
  }
  return launch_status("sweep_cov");
}

}  // extern "C"
