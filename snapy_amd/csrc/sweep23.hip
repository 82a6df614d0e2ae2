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
// halo cells) fits, a workgroup packs several whole lines (FP64: one face
// per lane; FP32: each lane walks a run of faces along the sweep and
// reconstructs every cell once); otherwise a workgroup of 64 lanes walks 62
// consecutive x1 cells plus one halo lane on each side.
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

// Runs of faces (FP32, nx1 + 2 <= kPackMax): a workgroup holds `rows`
// rows of nx1 + 2 lanes (one whole x1 line each, the two halo cells
// included), and each row walks kRunFaces consecutive faces of one line
// along the sweep, the rows taking the consecutive runs of the workgroup's
// rows x kRunFaces faces; the workgroups go through the XCD map, so the
// neighbouring runs of a line, which share stencil cells, share an L2.  A
// lane keeps a five-cell window of its x1 position in registers (each
// cell's FP64 primitives read once, along x1 across the lanes, one cell
// ahead of their use, shifted by the gauge and rounded to FP32) and
// reconstructs every cell of its run once, both edges together: the right
// edge of cell f - 1 and the left edge of cell f are the states of face f.
// Per face that is one weno5 smoothness evaluation instead of two and one
// load of the primitives instead of six.  The covariance's h and u_n go
// through LDS one face at a time, double-buffered (one barrier per face).
// Every value is rounded as a one-face-per-lane reconstruction of the face's
// six-cell stencil rounds it (weno5_plus/minus, cp5_plus/minus), so the
// fluxes are the same bits.  Dynamic LDS: 2 rows (nx1 + 2) (8 + 4) bytes.
constexpr int kRunFaces = 8;

template <int DIR>
struct Sweep23RunBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w,
                                             double const* __restrict__ x1v,
                                             float* __restrict__ flx, CovArgs a, int rows,
                                             int tiles_per_line) {
    extern __shared__ char smem_raw[];
    Range3 const& r = a.r;
    int const t = threadIdx.x;
    int const len = r.n1 + 2;
    int const nt = rows * len;
    double* sh_un = reinterpret_cast<double*>(smem_raw);  // [2][nt]
    float* sh_h = reinterpret_cast<float*>(sh_un + 2 * nt);  // [2][nt]
    int64_t tile = xcd_tile(bid.x, gdim.x);
    int o = int(tile / tiles_per_line);
    int ftile = int(tile - int64_t(o) * tiles_per_line) * rows * kRunFaces;
    // a thread past the rows (a group launch may run more) walks clamped
    // cells and stores nothing
    bool const act = t < nt;
    int const tc = act ? t : 0;
    int q = tc / len, p = tc - q * len;
    int fs = ftile + q * kRunFaces;
    int const nf = a.nf;
    int64_t sweep = DIR == 2 ? r.s2 : r.s3;
    int lo = DIR == 2 ? r.lo2 : r.lo3;
    int64_t obase = DIR == 2 ? int64_t(r.lo3 + o) * r.s3 : int64_t(r.lo2 + o) * r.s2;
    int i = r.lo1 - 1 + p;
    double const* col = w + obase + i;
    // the line's gauge at this x1 position: its first interior cell along
    // the sweep
    double const gauge = a.gauge ? col[IPR * r.sv + lo * sweep] : 0.;
    // the last ghost cell along the sweep; a run past the last face reads
    // clamped cells
    int const cmax = lo + nf + 1;
    auto ld = [&](int v, int cell) -> float {
      cell = cell > cmax ? cmax : cell;
      double gv = v == IPR ? gauge : 0.;
      return float(col[v * r.sv + cell * sweep] - gv);
    };
    float win[5][5];  // [row][window position]
#pragma unroll
    for (int v = 0; v < 5; ++v)
#pragma unroll
      for (int m = 0; m < 4; ++m) win[v][m] = ld(v, lo + fs - 3 + m);
    // the raw FP64 row values of the next cell, loaded one cell ahead
    double nxt[5];
    auto fetch = [&](int cell) {
      cell = cell > cmax ? cmax : cell;
#pragma unroll
      for (int v = 0; v < 5; ++v) nxt[v] = col[v * r.sv + cell * sweep];
    };
    fetch(lo + fs + 1);
    float ep_prev[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    bool const inner = act && p != 0 && p != len - 1;
    int j = DIR == 2 ? 0 : r.lo2 + o;
    int k = DIR == 2 ? r.lo3 + o : 0;
    // cell fs - 1 + m has the window lo + fs - 3 + m .. lo + fs + 1 + m
#pragma unroll 1
    for (int m = 0; m <= kRunFaces; ++m) {
#pragma unroll
      for (int v = 0; v < 5; ++v) win[v][4] = float(nxt[v] - (v == IPR ? gauge : 0.));
      if (m < kRunFaces) fetch(lo + fs + 2 + m);
      float em[5], ep[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (v == IDN || a.shock) {
          weno5_edges_f32(win[v][0], win[v][1], win[v][2], win[v][3], win[v][4], em[v], ep[v]);
        } else {
          em[v] = cp5_minus(win[v][0], win[v][1], win[v][2], win[v][3], win[v][4]);
          ep[v] = cp5_plus(win[v][0], win[v][1], win[v][2], win[v][3], win[v][4]);
        }
      }
      if (m >= 1) {
        // face f = fs + m - 1: the left state is the right edge of cell
        // f - 1, the right state the left edge of cell f
        int f = fs + m - 1;
        Prim5<float> L{ep_prev[IDN], {ep_prev[IVX], ep_prev[IVY], ep_prev[IVZ]}, ep_prev[IPR]};
        Prim5<float> R{em[IDN], {em[IVX], em[IVY], em[IVZ]}, em[IPR]};
        double h = 0., un = 0.;
        if (a.cov) {
          double rho = 0.5 * (double(L.r) + double(R.r));
          double pr = 0.5 * (double(L.p) + double(R.p)) + gauge;
          float enth = qdiv(float(pr), float(a.gammad - 1.)) + float(pr);
          h = double(qdiv(enth, float(rho)));
          un = 0.5 * (double(L.v[DIR - 1]) + double(R.v[DIR - 1]));
        }
        int const slot = (m - 1) & 1;
        if (act) {
          sh_h[slot * nt + t] = float(h);
          sh_un[slot * nt + t] = un;
        }
        __syncthreads();
        if (inner && f < nf) {
          if (DIR == 2)
            j = lo + f;
          else
            k = lo + f;
          cov_flux<float, DIR>(x1v, flx, a, i, j, k, L, R, gauge, double(sh_h[slot * nt + t - 1]),
                               double(sh_h[slot * nt + t + 1]), sh_un[slot * nt + t - 1],
                               sh_un[slot * nt + t + 1]);
        }
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        ep_prev[v] = ep[v];
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) win[v][mm] = win[v][mm + 1];
      }
    }
  }
};

template <int DIR>
__global__ void __launch_bounds__(kPackThreads)
    sweep23_run_kernel(double const* __restrict__ w, double const* __restrict__ x1v,
                       float* __restrict__ flx, CovArgs a, int rows, int tiles_per_line) {
  Sweep23RunBody<DIR>::run(block_index(), grid_size(), w, x1v, flx, a, rows, tiles_per_line);
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
  if (dtype != SNAPY_F64 && dtype != SNAPY_F32) return fail(-10, "unknown dtype");
  unsigned nseg = blocks_for(g->nx1, kInner);
  int64_t ny = dim == 2 ? int64_t(a.nf) * g->nx3 : int64_t(g->nx2) * a.nf;
  dim3 grid(unsigned(ny), nseg);
  hipStream_t s = as_stream(stream);
  int len = g->nx1 + 2;
  if (len <= kPackMax && dtype == SNAPY_F32) {
    int rows = packed_lines(len);
    int per = (a.nf + rows * kRunFaces - 1) / (rows * kRunFaces);
    int64_t nlines = dim == 2 ? g->nx3 : g->nx2;
    unsigned threads = unsigned((rows * len + 63) / 64 * 64);
    unsigned blocks = unsigned(nlines * per);
    size_t lds = size_t(2) * rows * len * (sizeof(double) + sizeof(float));
    auto go = [&](auto body, auto k) {
      launch_flex<decltype(body), kPackThreads, 1>(k, dim3(blocks), dim3(threads), lds, s,
                                              static_cast<double const*>(w), x1tab,
                                              static_cast<float*>(flux), a, rows, per);
    };
    if (dim == 2)
      go(Sweep23RunBody<2>{}, sweep23_run_kernel<2>);
    else
      go(Sweep23RunBody<3>{}, sweep23_run_kernel<3>);
    return launch_status("sweep_cov");
  }
  if (len <= kPackMax) {
    // FP64: the packed kernel, one face per lane
    int lines = packed_lines(len);
    unsigned threads = unsigned((lines * len + 63) / 64 * 64);
    unsigned blocks = unsigned((ny + lines - 1) / lines);
    auto go = [&](auto body, auto k) {
      launch_flex<decltype(body), kPackThreads, 1>(k, dim3(blocks), dim3(threads), 0, s,
                                              static_cast<double const*>(w), x1tab,
                                              static_cast<double*>(flux), a, lines);
    };
    if (dim == 2)
      go(Sweep23PackedBody<double, 2>{}, sweep23_packed_kernel<double, 2>);
    else
      go(Sweep23PackedBody<double, 3>{}, sweep23_packed_kernel<double, 3>);
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
