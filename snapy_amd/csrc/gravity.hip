// The gravity path of the x1 sweep on gfx950: the hydrostatic x1 reference
// scan, the fused well-balanced x1 reconstruction + Riemann sweep (alone,
// and with the reference fused in, the stage's launch), and the fused
// divergence + const gravity + face gravity work + RK combine.
//
// Uniform Cartesian block owning its whole x1 column (nb1 = 1), conversion
// order 2, dry ideal gas, non-hydrostatic factor 1 (the const-gravity
// default, src/forcing/forcing.hpp:41).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cfloat>

#include "common.hpp"
#include "group.hpp"
#include "point.hpp"

namespace snapy {
namespace {

// ---------------------------------------------------------------------------
// hydrostatic x1 reference, its scans one thread per column
// (hydro_ref_x1_impl, src/hydro/hydro_ref_x1_impl.h:24-272, uniform branch)
// ---------------------------------------------------------------------------
// ref layout: 5 planes of (nc3, nc2, nc1): psf_lo, psf_hi, pref, dsf, dref
enum { R_PSF_LO = 0, R_PSF_HI = 1, R_PREF = 2, R_DSF = 3, R_DREF = 4 };

__constant__ double kW6[6] = {11. / 1440., -31. / 480., 401. / 720.,
                              401. / 720., -31. / 480., 11. / 1440.};
__constant__ double kW6E[2][6] = {
    {95. / 288., 1427. / 1440., -133. / 240., 241. / 720., -173. / 1440.,
     3. / 160.},
    {-3. / 160., 637. / 1440., 511. / 720., -43. / 240., 77. / 1440.,
     -11. / 1440.}};

__device__ __forceinline__ double wall_rop(double r0, double r1, int k) {
  double q = r0 / r1, qk = q;
  for (int n = 1; n < k; ++n) qk *= q;
  double e = r1 > r0 ? r0 * qk : r0 + double(k) * (r0 - r1);
  return e > 0. ? e : r0;
}

struct RefArgs {
  Range3 r;
  int iu;        // last interior x1 cell
  double grav;   // downward magnitude (> 0)
  double dx1;    // uniform x1 width
  int wall_clamp;
};

// ro = rho / p per cell
__device__ __forceinline__ double rop_smooth(double const* ro, int i, int jlo,
                                             int jhi, bool ext_lo,
                                             bool ext_hi) {
  auto rop = [&](int j) { return ro[j]; };
  if (!(ext_lo && i - 2 < jlo) && !(ext_hi && i + 2 > jhi))
    return rop(i < jlo ? jlo : (i > jhi ? jhi : i));
  double v[5];
  for (int m = -2; m <= 2; ++m) {
    int j = i + m;
    if (j < jlo)
      v[m + 2] = ext_lo ? wall_rop(rop(jlo), rop(jlo + 1), jlo - j) : rop(jlo);
    else if (j > jhi)
      v[m + 2] = ext_hi ? wall_rop(rop(jhi), rop(jhi - 1), j - jhi) : rop(jhi);
    else
      v[m + 2] = rop(j);
  }
  return (v[0] + 4. * v[1] + 6. * v[2] + 4. * v[3] + v[4]) / 16.;
}

// kRefCols columns per workgroup of kRefThreads threads.  Every cell's
// pressure step g rho dx and rho / p are formed across the threads
// (coalesced along x1) into LDS; one thread per column then runs the two
// face scans in their serial order (hydro_ref_x1_scan_impl) while the
// others smooth rho / p; then the threads share the cell pass
// (hydro_ref_x1_cell_impl), whose cells are independent given the faces.
// The scans keep the raw face values F (F[iu + 1] = the top anchor, F[i] =
// F[i + 1] + dp_i below it and F[i + 1] = F[i] - dp_i above it, as the
// reference's running face); cell i's faces are F[i] and F[i + 1], floored
// at DBL_MIN where read, as the reference stores them.
// LDS per column: dp, rho / p, F (nc1 + 1), smoothed rho / p.
constexpr int kRefCols = 4;
constexpr int kRefThreads = 256;

__global__ void __launch_bounds__(kRefThreads)
    hydro_ref_x1_kernel(double const* __restrict__ w, double* __restrict__ ref,
                        RefArgs a, int nc) {
  extern __shared__ double smem[];
  Range3 const& r = a.r;
  int64_t ncol = int64_t(r.nc2) * r.nc3;
  int64_t col0 = int64_t(blockIdx.x) * nc;
  int const tid = threadIdx.x;
  int nc1 = r.nc1, iu = a.iu;
  int const stride = 4 * nc1 + 1;  // doubles per column
  int ncl = ncol - col0 < nc ? int(ncol - col0) : nc;
  double const g = a.grav, dx = a.dx1, mp = DBL_MIN;
  auto floor_mp = [&](double v) { return v > mp ? v : mp; };
  int const cells = ncl * nc1;
  for (int e = tid; e < cells; e += kRefThreads) {
    int c = e / nc1, i = e - c * nc1;
    int64_t base = (col0 + c) * nc1;
    double rho = w[base + i], pr = w[4 * r.sv + base + i];
    double* sc = smem + c * stride;
    sc[i] = g * rho * dx;
    sc[nc1 + i] = rho / pr;
  }
  __syncthreads();
  int il = nc1 - 1 - iu;
  bool wc = a.wall_clamp;
  bool thin = (iu - il + 1) < 5, fits = iu >= 4;
  int jlo = wc ? il : 0, jhi = wc ? iu : nc1 - 1;
  bool ext_lo = wc && il + 1 <= iu, ext_hi = wc && iu - 1 >= il;
  if (tid < ncl) {
    // the scans from the top anchor, one column per thread
    double* sc = smem + tid * stride;
    double const* s_dp = sc;
    double* s_F = sc + 2 * nc1;
    int64_t base = (col0 + tid) * nc1;
    double rho_u = w[base + iu], pr_u = w[4 * r.sv + base + iu];
    double rdt_top = pr_u / rho_u;
    double top = pr_u * exp(-g * 0.5 * dx / rdt_top);
    s_F[iu + 1] = top;
    double face = top;
    constexpr int kB = 8;
    int i = iu;
    for (; i - kB + 1 >= 0; i -= kB) {
      double t[kB];
#pragma unroll
      for (int m = 0; m < kB; ++m) t[m] = s_dp[i - m];
#pragma unroll
      for (int m = 0; m < kB; ++m) {
        face = face + t[m];
        s_F[i - m] = face;
      }
    }
    for (; i >= 0; --i) {
      face = face + s_dp[i];
      s_F[i] = face;
    }
    face = top;
    for (int k = iu + 1; k < nc1; ++k) {
      face = face - s_dp[k];
      s_F[k + 1] = face;
    }
  }
  // the smoothed rho / p, meanwhile on the other threads
  for (int e = tid; e < cells; e += kRefThreads) {
    int c = e / nc1, i = e - c * nc1;
    double* sc = smem + c * stride;
    sc[3 * nc1 + 1 + i] = rop_smooth(sc + nc1, i, jlo, jhi, ext_lo, ext_hi);
  }
  __syncthreads();
  for (int e = tid; e < cells; e += kRefThreads) {
    int c = e / nc1, i = e - c * nc1;
    double const* sc = smem + c * stride;
    double const* s_F = sc + 2 * nc1;
    double const* s_rs = sc + 3 * nc1 + 1;
    int64_t base = (col0 + c) * nc1;
    bool wall_in = wc && i >= il && i < il + 2;
    bool wall_out = wc && i > iu - 2 && i <= iu;
    bool clamp_in = wall_in && fits && !thin;
    bool clamp_out = wall_out && fits && !thin;
    double lo = floor_mp(s_F[i]), hi = floor_mp(s_F[i + 1]);
    double cell = 0.5 * (lo + hi);
    double lower = lo < hi ? lo : hi, upper = lo > hi ? lo : hi;
    auto six = [&](int start, double const* wt, bool rev) {
      double value = 0.;
      for (int m = 0; m < 6; ++m)
        value += wt[rev ? 5 - m : m] * floor_mp(s_F[start + m]);
      if (value >= lower && value <= upper) cell = value;
    };
    if (i >= 2 && i < nc1 - 2 && !wall_in && !wall_out) six(i - 2, kW6, false);
    if (clamp_in) six(il, kW6E[i - il], false);
    if (clamp_out) {
      int s = iu + 1 - 5;
      six(s, kW6E[4 - (i - s)], true);
    }
    double rs = s_rs[i];
    double rf = i > 0 ? 0.5 * (s_rs[i - 1] + rs) : rs;
    ref[R_PSF_LO * r.sv + base + i] = lo;
    ref[R_PSF_HI * r.sv + base + i] = hi;
    ref[R_PREF * r.sv + base + i] = cell;
    ref[R_DREF * r.sv + base + i] = cell * rs;
    ref[R_DSF * r.sv + base + i] = lo * rf;
  }
}

// ---------------------------------------------------------------------------
// fused well-balanced x1 reconstruction + Riemann sweep
// (WellBalancedX1::forward, src/recon/well_balanced_x1.cpp:16-154, with the
// face densities of X1Reference::evaluate, src/recon/x1_reference.cpp:
// 177-243), one thread per interior x1 face
// ---------------------------------------------------------------------------
struct WbArgs {
  Range3 r;
  int nf;
  double gammad;
  int riemann, shock;
  int mirror_lo, mirror_hi;  // even perturbation ghosts at a physical x1 face
  int refl_lo, refl_hi;      // reflecting wall: one face density both sides
  int wall_lo, wall_hi;      // prescribed wall planes
};

// T is the compute precision of the face (the primitives w, the reference
// and the restored face densities are FP64); TF is the stored flux
// precision.  An FP32 flux keeps the x1 momentum row as the anomaly about
// the reference face pressure psf (the FP64 update adds psf back), so the
// FP32 word holds the small non-hydrostatic part of the flux.  With FP32
// compute (LMARS only) the face states carry the pressure perturbation
// itself, p - psf, and LMARS takes psf as its gauge: the enthalpy and the
// sound speed see the total pressure, the momentum flux and the face
// pressure come out as anomalies about psf, so FP32 keeps the precision of
// the non-hydrostatic part.
// One x1 face of the well-balanced sweep: face fi (x1 index i) of the
// column at row; pref(m) and dref(m) give the reference pressure and
// density of cell m of the column, psf and dsf the scan's face pressure and
// density at the face.
template <typename T, typename TF, typename PREF, typename DREF>
__device__ __forceinline__ void wb_face(double const* __restrict__ w, int64_t row, int fi, double psf,
                                        double dsf, PREF pref, DREF dref,
                                        TF* __restrict__ flx, double* __restrict__ fp,
                                        WbArgs const& a) {
  Range3 const& r = a.r;
  int i = r.lo1 + fi;
  int il = r.lo1, iu = r.lo1 + r.n1 - 1;

  // perturbation of row c at cell m, with the even mirror at the walls
  auto pert = [&](int c, int m) -> T {
    int src = m;
    if (m < il && a.mirror_lo) src = 2 * il - 1 - m;
    if (m > iu && a.mirror_hi) src = 2 * iu + 1 - m;
    double rf = c == IPR ? pref(src) : dref(src);
    return T(double(w[c * r.sv + row + src]) - rf);
  };

  Prim5<T> L, R;
  // velocities: the plain face states
  T* lv[3] = {&L.v[0], &L.v[1], &L.v[2]};
  T* rv[3] = {&R.v[0], &R.v[1], &R.v[2]};
  for (int v = IVX; v <= IVZ; ++v) {
    double const* p = w + v * r.sv + row + i;
    T c0 = T(p[-3]), c1 = T(p[-2]), c2 = T(p[-1]), c3 = T(p[0]), c4 = T(p[1]), c5 = T(p[2]);
    if (a.shock) {
      *lv[v - 1] = weno5_plus(c0, c1, c2, c3, c4);
      *rv[v - 1] = weno5_minus(c1, c2, c3, c4, c5);
    } else {
      *lv[v - 1] = cp5_plus(c0, c1, c2, c3, c4);
      *rv[v - 1] = cp5_minus(c1, c2, c3, c4, c5);
    }
  }
  // pressure perturbation (cp5 unless shock), density perturbation (weno5)
  {
    T c0 = pert(IPR, i - 3), c1 = pert(IPR, i - 2), c2 = pert(IPR, i - 1),
      c3 = pert(IPR, i), c4 = pert(IPR, i + 1), c5 = pert(IPR, i + 2);
    T pl, pr;
    if (a.shock) {
      pl = weno5_plus(c0, c1, c2, c3, c4);
      pr = weno5_minus(c1, c2, c3, c4, c5);
    } else {
      pl = cp5_plus(c0, c1, c2, c3, c4);
      pr = cp5_minus(c1, c2, c3, c4, c5);
    }
    double tl = double(pl) + psf, tr = double(pr) + psf;
    if constexpr (sizeof(T) == 4) {
      // the perturbation itself (zero where the total is floored to psf)
      L.p = tl > 0. ? pl : T(0);
      R.p = tr > 0. ? pr : T(0);
    } else {
      L.p = T(tl > 0. ? tl : psf);
      R.p = T(tr > 0. ? tr : psf);
    }
  }
  {
    T c0 = pert(IDN, i - 3), c1 = pert(IDN, i - 2), c2 = pert(IDN, i - 1),
      c3 = pert(IDN, i), c4 = pert(IDN, i + 1), c5 = pert(IDN, i + 2);
    T dlp = weno5_plus(c0, c1, c2, c3, c4);
    T drp = weno5_minus(c1, c2, c3, c4, c5);
    // face densities of the reference: cp5 rows on dref guarded by the two
    // cells' range, else the scan's dsf; one mean value on a reflecting wall
    double d0 = dref(i - 3), d1 = dref(i - 2), d2 = dref(i - 1), d3 = dref(i),
           d4 = dref(i + 1), d5 = dref(i + 2);
    double fl = cp5_plus(d0, d1, d2, d3, d4);
    double fr = cp5_minus(d1, d2, d3, d4, d5);
    double lo = d2 < d3 ? d2 : d3, hi = d2 > d3 ? d2 : d3;
    double sl = (fl >= lo && fl <= hi) ? fl : dsf;
    double sr = (fr >= lo && fr <= hi) ? fr : dsf;
    if ((fi == 0 && a.refl_lo) || (fi == a.nf - 1 && a.refl_hi)) {
      double mean = 0.5 * (sl + sr);
      sl = mean;
      sr = mean;
    }
    double tl = double(dlp) + sl, tr = double(drp) + sr;
    L.r = T(tl > 0. ? tl : double(w[row + i - 1]));
    R.r = T(tr > 0. ? tr : double(w[row + i]));
  }
  T g = T(a.gammad);
  T out[5];
  T pface;
  if constexpr (sizeof(T) == 4) {
    // anomaly form: the momentum flux and pface are about psf already
    lmars(L, R, g, g, 0, out, &pface, T(psf));
  } else {
    if (a.riemann == SNAPY_RIEMANN_LMARS)
      lmars(L, R, g, g, 0, out, &pface);
    else
      hllc(L, R, g, g, 0, out, &pface);
  }
  if ((fi == 0 && a.wall_lo) || (fi == a.nf - 1 && a.wall_hi)) {
    out[IDN] = T(0);
    out[IPR] = T(0);
  }
  if constexpr (sizeof(TF) == 4 && sizeof(T) == 8) out[IVX] = out[IVX] - T(psf);
  for (int v = 0; v < 5; ++v) flx[v * r.sv + row + i] = TF(out[v]);
  if (fp) fp[row + i] = sizeof(T) == 4 ? double(pface) + psf : double(pface);
}

template <typename T, typename TF>
__global__ void __launch_bounds__(256)
    sweep_x1_wb_kernel(double const* __restrict__ w, double const* __restrict__ ref,
                       TF* __restrict__ flx, double* __restrict__ fp, WbArgs a) {
  Range3 const& r = a.r;
  int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t tot = int64_t(a.nf) * r.n2 * r.n3;
  if (idx >= tot) return;
  int fi = int(idx % a.nf);
  int i = r.lo1 + fi;
  int64_t rest = idx / a.nf;
  int j = r.lo2 + int(rest % r.n2);
  int k = r.lo3 + int(rest / r.n2);
  int64_t row = k * r.s3 + j * r.s2;
  double const* pref = ref + R_PREF * r.sv + row;
  double const* dref = ref + R_DREF * r.sv + row;
  double psf = ref[R_PSF_LO * r.sv + row + i];
  double dsf = ref[R_DSF * r.sv + row + i];
  wb_face<T, TF>(
      w, row, fi, psf, dsf, [&](int m) { return pref[m]; }, [&](int m) { return dref[m]; },
      flx, fp, a);
}

// ---------------------------------------------------------------------------
// fused divergence + const gravity + gravity work + RK combine
// ---------------------------------------------------------------------------
// x1 table: 8 rows of nc1 doubles
enum { T_X1V = 0, T_GW_HI = 1, T_GW_LO = 2, T_GVAR = 3, T_WT0 = 4, T_WT1 = 5,
       T_WT2 = 6 };

struct GUpdArgs {
  Range3 r;
  double a1, a2, a3, vol;
  double mdt;
  double const* dt_dev;
  double w0, w1, w2;
  double grav1;
  double heat_flux, heat_dzd;  // bot-heat flux and dz * depth
  int heat_depth;               // 0 = no bot-heat
};

// The stage increment of cell (i, row) from its face fluxes (the x1 rows
// through f1v(v, s), s = 0 the face below and 1 the face above, the
// reference face pressure restored; the x1 mass flux of face q through
// m1(q)) and the x2/x3 flux arrays (NULL when inactive): the divergence, the
// constant gravity, the diffusion source dsrc (may be NULL), bot-heat and
// the corrected-PE gravity work.  ii is the cell's interior x1 index, c its
// offset in a row, dt the stage dt.
template <typename F1V, typename M1, typename F2V, typename F3V>
__device__ __forceinline__ void gravity_increment(GUpdArgs const& a, int ii, int i, double dt,
                                                  double rho, double vx, F1V f1v, M1 m1, F2V f2v,
                                                  F3V f3v, bool has2, bool has3,
                                                  double const* __restrict__ tab, int,
                                                  double const* __restrict__ dsrc, int64_t c,
                                                  double (&du)[5]) {
  using T = double;
  Range3 const& r = a.r;
  T A1 = T(a.a1), A2 = T(a.a2), A3 = T(a.a3), V = T(a.vol);
  T mdt = T(-dt), tdt = T(dt);
  T g1 = T(a.grav1);
  for (int v = 0; v < 5; ++v) {
    T div = A1 * f1v(v, 1) - A1 * f1v(v, 0);
    if (has2) div = div + (A2 * T(f2v(v, 1)) - A2 * T(f2v(v, 0)));
    if (has3) div = div + (A3 * T(f3v(v, 1)) - A3 * T(f3v(v, 0)));
    div = div / V;
    du[v] = mdt * div;
  }
  // ConstGravityImpl::forward_split (const_gravity.cpp:35-43), nh = 1
  du[IVX] = du[IVX] + tdt * rho * g1 * T(1.);
  du[IPR] = du[IPR] + tdt * rho * vx * g1 * T(1.);
  if (dsrc)
    for (int v = 0; v < 5; ++v) du[v] = du[v] + dsrc[v * r.sv + c];
  // LayerHeat bot-heat, an integer depth of bottom cells
  // (src/forcing/layer_heat.cpp:78-85): du[IPR] += dt * flux / (dz * depth)
  if (ii < a.heat_depth) du[IPR] = du[IPR] + T(dt * a.heat_flux / a.heat_dzd);
  // gravity_energy_correction (hydro_forward.cpp:1049-1124), Cartesian
  auto drho = [&](int m) {  // -dt * x1_face_difference / V at cell lo1 + m
    int q = r.lo1 + m;
    return mdt * ((A1 * m1(q + 1) - A1 * m1(q)) / V);
  };
  T gw = (T(tab[T_GW_HI * r.nc1 + i]) * (A1 * m1(i + 1)) - T(tab[T_GW_LO * r.nc1 + i]) * (A1 * m1(i))) /
         V;
  T fgw = tdt * gw;
  int jj = (ii < 1 ? 1 : (ii > r.n1 - 2 ? r.n1 - 2 : ii)) - 1;
  T s = T(tab[T_WT0 * r.nc1 + i]) * drho(jj);
  s = s + T(tab[T_WT1 * r.nc1 + i]) * drho(jj + 1);
  s = s + T(tab[T_WT2 * r.nc1 + i]) * drho(jj + 2);
  fgw = fgw + T(tab[T_GVAR * r.nc1 + i]) * s;
  T ogw = tdt * rho * vx * g1 * T(1.);
  du[IPR] = du[IPR] + (fgw - ogw);
}

// The stage increment fused into the FP32 x1 reference sweep: its
// arguments (w0..w2 unused), the x2/x3 flux totals (NULL when inactive), the
// x1 table and du, (nvar, nc3, nc2, nc1) FP64.
struct X1IncArgs {
  GUpdArgs g;
  float const* f2;
  float const* f3;
  double const* tab;
  double* du;
};

// The reference and the well-balanced sweep fused, for the interior
// columns: kRefCols columns per workgroup run hydro_ref_x1_kernel's phases
// (the pressure steps and rho / p, the scans one thread per column, the
// smoothing, the cell pass), keeping the reference pressure and density of
// every cell and the raw faces in LDS; only the face pressure psf_lo goes to
// HBM (the divergence kernel's face gravity reads it).  Then the threads
// share the column's faces, each as wb_face with the reference read from
// LDS.  The arithmetic is that of the two kernels, so the fluxes are the
// same bits.
template <typename T, typename TF>
struct SweepX1WbRefBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w, double* __restrict__ ref,
                                             TF* __restrict__ flx, double* __restrict__ fp,
                                             RefArgs ra, WbArgs a, int nc) {
    extern __shared__ double smem[];
    __shared__ int64_t s_row[kRefCols];
    int ncl = ref_phases(bid, w, ref, ra, a, nc, smem, s_row);
    Range3 const& r = a.r;
    int const tid = threadIdx.x;
    int const nc1 = r.nc1;
    int const stride = 4 * nc1 + 1;
    double const mp = DBL_MIN;
    auto floor_mp = [&](double v) { return v > mp ? v : mp; };
    if constexpr (sizeof(T) == 4) {
      wb_faces_f32(w, flx, fp, a, smem, s_row, ncl, stride, floor_mp, nullptr, 0);
    } else {
      int const nf = a.nf;
      int const faces = ncl * nf;
      for (int e = tid; e < faces; e += kRefThreads) {
        int c = e / nf, fi = e - c * nf;
        int i = r.lo1 + fi;
        double const* sc = smem + c * stride;
        double const* s_F = sc + 2 * nc1;
        double const* s_rs = sc + 3 * nc1 + 1;
        double lo = floor_mp(s_F[i]);
        double rf = i > 0 ? 0.5 * (s_rs[i - 1] + s_rs[i]) : s_rs[i];
        wb_face<T, TF>(
            w, s_row[c], fi, lo, lo * rf, [&](int m) { return sc[m]; },
            [&](int m) { return sc[nc1 + m]; }, flx, fp, a);
      }
    }
  }

  // hydro_ref_x1_kernel's phases for the workgroup's columns, into LDS (4
  // nc1 + 1 doubles per column: the reference pressure and density of every
  // cell, the raw faces, the smoothed rho / p), the face pressure psf_lo to
  // ref; returns the workgroup's column count
  static __device__ __forceinline__ int ref_phases(dim3 const& bid, double const* __restrict__ w,
                                                   double* __restrict__ ref, RefArgs const& ra,
                                                   WbArgs const& a, int nc, double* smem,
                                                   int64_t* s_row) {
    Range3 const& r = a.r;
    int64_t ncol = int64_t(r.n2) * r.n3;
    int64_t col0 = int64_t(bid.x) * nc;
    int const tid = threadIdx.x;
    int nc1 = r.nc1, iu = ra.iu;
    int const stride = 4 * nc1 + 1;  // doubles per column
    int ncl = ncol - col0 < nc ? int(ncol - col0) : nc;
    double const g = ra.grav, dx = ra.dx1, mp = DBL_MIN;
    auto floor_mp = [&](double v) { return v > mp ? v : mp; };
    if (tid < ncl) {
      int64_t col = col0 + tid;
      int j = r.lo2 + int(col % r.n2), k = r.lo3 + int(col / r.n2);
      s_row[tid] = k * r.s3 + j * r.s2;
    }
    __syncthreads();
    int const cells = ncl * nc1;
    for (int e = tid; e < cells; e += kRefThreads) {
      int c = e / nc1, i = e - c * nc1;
      int64_t base = s_row[c];
      double rho = w[base + i], pr = w[4 * r.sv + base + i];
      double* sc = smem + c * stride;
      sc[i] = g * rho * dx;
      sc[nc1 + i] = rho / pr;
    }
    __syncthreads();
    int il = nc1 - 1 - iu;
    bool wc = ra.wall_clamp;
    bool thin = (iu - il + 1) < 5, fits = iu >= 4;
    int jlo = wc ? il : 0, jhi = wc ? iu : nc1 - 1;
    bool ext_lo = wc && il + 1 <= iu, ext_hi = wc && iu - 1 >= il;
    if (tid < ncl) {
      double* sc = smem + tid * stride;
      double const* s_dp = sc;
      double* s_F = sc + 2 * nc1;
      int64_t base = s_row[tid];
      double rho_u = w[base + iu], pr_u = w[4 * r.sv + base + iu];
      double rdt_top = pr_u / rho_u;
      double top = pr_u * exp(-g * 0.5 * dx / rdt_top);
      s_F[iu + 1] = top;
      double face = top;
      constexpr int kB = 8;
      int i = iu;
      for (; i - kB + 1 >= 0; i -= kB) {
        double t[kB];
#pragma unroll
        for (int m = 0; m < kB; ++m) t[m] = s_dp[i - m];
#pragma unroll
        for (int m = 0; m < kB; ++m) {
          face = face + t[m];
          s_F[i - m] = face;
        }
      }
      for (; i >= 0; --i) {
        face = face + s_dp[i];
        s_F[i] = face;
      }
      face = top;
      for (int k = iu + 1; k < nc1; ++k) {
        face = face - s_dp[k];
        s_F[k + 1] = face;
      }
    }
    for (int e = tid; e < cells; e += kRefThreads) {
      int c = e / nc1, i = e - c * nc1;
      double* sc = smem + c * stride;
      sc[3 * nc1 + 1 + i] = rop_smooth(sc + nc1, i, jlo, jhi, ext_lo, ext_hi);
    }
    __syncthreads();
    // the cell pass: pref into the pressure-step slot, dref into the rho / p
    // slot (neither is read again)
    for (int e = tid; e < cells; e += kRefThreads) {
      int c = e / nc1, i = e - c * nc1;
      double* sc = smem + c * stride;
      double const* s_F = sc + 2 * nc1;
      double const* s_rs = sc + 3 * nc1 + 1;
      bool wall_in = wc && i >= il && i < il + 2;
      bool wall_out = wc && i > iu - 2 && i <= iu;
      bool clamp_in = wall_in && fits && !thin;
      bool clamp_out = wall_out && fits && !thin;
      double lo = floor_mp(s_F[i]), hi = floor_mp(s_F[i + 1]);
      double cell = 0.5 * (lo + hi);
      double lower = lo < hi ? lo : hi, upper = lo > hi ? lo : hi;
      auto six = [&](int start, double const* wt, bool rev) {
        double value = 0.;
        for (int m = 0; m < 6; ++m)
          value += wt[rev ? 5 - m : m] * floor_mp(s_F[start + m]);
        if (value >= lower && value <= upper) cell = value;
      };
      if (i >= 2 && i < nc1 - 2 && !wall_in && !wall_out) six(i - 2, kW6, false);
      if (clamp_in) six(il, kW6E[i - il], false);
      if (clamp_out) {
        int s = iu + 1 - 5;
        six(s, kW6E[4 - (i - s)], true);
      }
      sc[i] = cell;
      sc[nc1 + i] = cell * s_rs[i];
      ref[R_PSF_LO * r.sv + s_row[c] + i] = lo;
    }
    __syncthreads();
    return ncl;
  }

  // The faces of the FP32 form (wb_face<float, float> on every face of the
  // workgroup's columns), cell by cell.  The workgroup's cells are taken in
  // column order, 64 consecutive cells per wavefront pass, and each lane
  // forms its cell's five face-state inputs once (the velocities; the
  // pressure and density perturbations about the column's reference,
  // mirrored at the walls); the five-cell window comes from the neighbouring
  // lanes by wavefront shuffles, each lane reconstructs both edges of its
  // cell (weno5 edges sharing their smoothness indicators, or cp5), and the
  // face below the cell takes the right edge of the cell below from the
  // lane below.  Lanes 0-1 and 62-63 are the window's halo and lane 2 only
  // feeds lane 3, so faces come out on lanes 3-61; a pass owns (stores) the
  // faces of lanes 3-61, or with inc of lanes 5-58, and passes overlap.  The
  // passes need no LDS and no barrier.  Each value is rounded as
  // wb_face<float, float> rounds it, so the fluxes and face pressures are
  // its bits.
  //
  // With inc (non-NULL), each owned interior cell also forms its stage
  // increment (gravity_increment, as update_gravity_kernel<float, true,
  // true> forms it from the stored fluxes): the x1 rows of the face above
  // and the mass fluxes of the faces two below to three above come from the
  // neighbouring lanes, the x2/x3 fluxes from inc's arrays; col0 is the
  // workgroup's first column.
  template <typename FLOOR>
  static __device__ __forceinline__ void wb_faces_f32(double const* __restrict__ w,
                                                      float* __restrict__ flx,
                                                      double* __restrict__ fp, WbArgs const& a,
                                                      double const* smem, int64_t const* s_row,
                                                      int ncl, int stride, FLOOR floor_mp,
                                                      X1IncArgs const* inc, int64_t col0) {
    Range3 const& r = a.r;
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const nc1 = r.nc1;
    int const il = r.lo1, iu = r.lo1 + r.n1 - 1;
    int const off = inc ? 5 : 3, kOwn = inc ? 54 : 59;
    int const ncell = ncl * nc1;
    int const npass = (ncell + kOwn - 1) / kOwn;
    float const g = float(a.gammad);
    bool const own = lane >= off && lane < off + kOwn;
    for (int ps = wave; ps < npass; ps += kRefThreads / 64) {
      int const L = ps * kOwn - off + lane;
      bool const inb = L >= 0 && L < ncell;
      int const c = inb ? L / nc1 : 0;
      int const m = inb ? L - c * nc1 : 0;
      double const* sc = smem + c * stride;
      int64_t const row = s_row[c];
      // the increment's inputs of an owned interior cell, loaded before the
      // faces so their latency hides behind the reconstruction
      bool const inc_cell = inc && own && inb && m >= il && m <= iu;
      float x2lo[5], x2hi[5], x3lo[5], x3hi[5];
      double rho_c = 0., vx_c = 0.;
      if (inc_cell) {
        int64_t cc = row + m;
        rho_c = w[cc];
        vx_c = w[r.sv + cc];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          int64_t o = v * r.sv + cc;
          x2lo[v] = inc->f2 ? inc->f2[o] : 0.f;
          x2hi[v] = inc->f2 ? inc->f2[o + r.s2] : 0.f;
          x3lo[v] = inc->f3 ? inc->f3[o] : 0.f;
          x3hi[v] = inc->f3 ? inc->f3[o + r.s3] : 0.f;
        }
      }
      float q[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
      if (inb) {
        int src = m;
        if (m < il && a.mirror_lo) src = 2 * il - 1 - m;
        if (m > iu && a.mirror_hi) src = 2 * iu + 1 - m;
        q[IDN] = float(double(w[IDN * r.sv + row + src]) - sc[nc1 + src]);
        q[IVX] = float(w[IVX * r.sv + row + m]);
        q[IVY] = float(w[IVY * r.sv + row + m]);
        q[IVZ] = float(w[IVZ * r.sv + row + m]);
        q[IPR] = float(double(w[IPR * r.sv + row + src]) - sc[src]);
      }
      // the edge cells lo1 - 1 .. lo1 + n1 (their windows stay in the column)
      bool const edge = inb && lane >= 2 && lane <= 61 && m >= r.lo1 - 1 && m <= r.lo1 + r.n1;
      float em[5], ep[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        float v0 = __shfl_up(q[v], 2), v1 = __shfl_up(q[v], 1), v2 = q[v];
        float v3 = __shfl_down(q[v], 1), v4 = __shfl_down(q[v], 2);
        em[v] = 0.f;
        ep[v] = 0.f;
        if (edge) {
          if (v == IDN || a.shock) {
            weno5_edges_f32(v0, v1, v2, v3, v4, em[v], ep[v]);
          } else {
            em[v] = cp5_minus(v0, v1, v2, v3, v4);
            ep[v] = cp5_plus(v0, v1, v2, v3, v4);
          }
        }
      }
      float lq[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) lq[v] = __shfl_up(ep[v], 1);
      bool const face = edge && lane >= 3 && m >= r.lo1;
      int const i = m, fi = m - r.lo1;
      double const* s_F = sc + 2 * nc1;
      float out[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
      if (face) {
        double const* s_rs = sc + 3 * nc1 + 1;
        double const psf = floor_mp(s_F[i]);
        double const rf = i > 0 ? 0.5 * (s_rs[i - 1] + s_rs[i]) : s_rs[i];
        double const dsf = psf * rf;
        Prim5<float> Lf, Rf;
        Lf.v[0] = lq[IVX];
        Lf.v[1] = lq[IVY];
        Lf.v[2] = lq[IVZ];
        Rf.v[0] = em[IVX];
        Rf.v[1] = em[IVY];
        Rf.v[2] = em[IVZ];
        {
          float pl = lq[IPR], pr = em[IPR];
          double tl = double(pl) + psf, tr = double(pr) + psf;
          Lf.p = tl > 0. ? pl : 0.f;
          Rf.p = tr > 0. ? pr : 0.f;
        }
        {
          float dlp = lq[IDN], drp = em[IDN];
          double const* dref = sc + nc1;
          double d0 = dref[i - 3], d1 = dref[i - 2], d2 = dref[i - 1], d3 = dref[i],
                 d4 = dref[i + 1], d5 = dref[i + 2];
          double fl = cp5_plus(d0, d1, d2, d3, d4);
          double fr = cp5_minus(d1, d2, d3, d4, d5);
          double lo = d2 < d3 ? d2 : d3, hi = d2 > d3 ? d2 : d3;
          double sl = (fl >= lo && fl <= hi) ? fl : dsf;
          double sr = (fr >= lo && fr <= hi) ? fr : dsf;
          if ((fi == 0 && a.refl_lo) || (fi == a.nf - 1 && a.refl_hi)) {
            double mean = 0.5 * (sl + sr);
            sl = mean;
            sr = mean;
          }
          double tl = double(dlp) + sl, tr = double(drp) + sr;
          Lf.r = float(tl > 0. ? tl : double(w[row + i - 1]));
          Rf.r = float(tr > 0. ? tr : double(w[row + i]));
        }
        float pface;
        lmars(Lf, Rf, g, g, 0, out, &pface, float(psf));
        if ((fi == 0 && a.wall_lo) || (fi == a.nf - 1 && a.wall_hi)) {
          out[IDN] = 0.f;
          out[IPR] = 0.f;
        }
        if (own) {
          for (int v = 0; v < 5; ++v) flx[v * r.sv + row + i] = out[v];
          if (fp) fp[row + i] = double(pface) + psf;
        }
      }
      if (!inc) continue;
      // the stage increment of the owned interior cells
      float up[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) up[v] = __shfl_down(out[v], 1);
      float const m1w[6] = {__shfl_up(out[IDN], 2), __shfl_up(out[IDN], 1), out[IDN],
                            __shfl_down(out[IDN], 1), __shfl_down(out[IDN], 2),
                            __shfl_down(out[IDN], 3)};
      if (!inc_cell) continue;
      int const ii = m - il;
      int64_t const cc = row + i;
      int const k = r.lo3 + int((col0 + c) / r.n2);
      double const dt = dt_at(inc->g.dt_dev, -inc->g.mdt, r, k);
      double const psf0 = floor_mp(s_F[i]), psf1 = floor_mp(s_F[i + 1]);
      auto f1v = [&](int v, int sgn) {
        double x = double(sgn ? up[v] : out[v]);
        if (v == IVX) x = x + (sgn ? psf1 : psf0);
        return x;
      };
      auto m1 = [&](int qf) {
        int d = qf - i;
        float x = d == -2 ? m1w[0] : (d == -1 ? m1w[1] : (d == 0 ? m1w[2] : (d == 1 ? m1w[3] : (d == 2 ? m1w[4] : m1w[5]))));
        return double(x);
      };
      double du[5];
      gravity_increment(inc->g, ii, i, dt, rho_c, vx_c, f1v, m1,
                        [&](int v, int sgn) { return sgn ? x2hi[v] : x2lo[v]; },
                        [&](int v, int sgn) { return sgn ? x3hi[v] : x3lo[v]; }, inc->f2 != nullptr,
                        inc->f3 != nullptr, inc->tab, i, static_cast<double const*>(nullptr), cc, du);
#pragma unroll
      for (int v = 0; v < 5; ++v) inc->du[v * r.sv + cc] = du[v];
    }
  }
};

// The FP32 reference sweep with the stage increment fused (the VIC path's
// du): the reference phases, then wb_faces_f32 with X1IncArgs.
struct SweepX1WbIncBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             double const* __restrict__ w, double* __restrict__ ref,
                                             float* __restrict__ flx, double* __restrict__ fp,
                                             RefArgs ra, WbArgs a, int nc, X1IncArgs ia) {
    extern __shared__ double smem[];
    __shared__ int64_t s_row[kRefCols];
    int ncl = SweepX1WbRefBody<float, float>::ref_phases(bid, w, ref, ra, a, nc, smem, s_row);
    double const mp = DBL_MIN;
    auto floor_mp = [&](double v) { return v > mp ? v : mp; };
    SweepX1WbRefBody<float, float>::wb_faces_f32(w, flx, fp, a, smem, s_row, ncl, 4 * a.r.nc1 + 1,
                                                 floor_mp, &ia, int64_t(bid.x) * nc);
  }
};

__global__ void __launch_bounds__(kRefThreads)
    sweep_x1_wb_inc_kernel(double const* __restrict__ w, double* __restrict__ ref,
                           float* __restrict__ flx, double* __restrict__ fp, RefArgs ra, WbArgs a,
                           int nc, X1IncArgs ia) {
  SweepX1WbIncBody::run(block_index(), grid_size(), w, ref, flx, fp, ra, a, nc, ia);
}

template <typename T, typename TF>
__global__ void __launch_bounds__(kRefThreads)
    sweep_x1_wb_ref_kernel(double const* __restrict__ w, double* __restrict__ ref,
                           TF* __restrict__ flx, double* __restrict__ fp, RefArgs ra,
                           WbArgs a, int nc) {
  SweepX1WbRefBody<T, TF>::run(block_index(), grid_size(), w, ref, flx, fp, ra, a, nc);
}

// WRITE_DU: store the stage increment du (for the implicit solve to finish)
// instead of combining it into u.  TF is the flux precision; PAIR: the state
// is the FP32 pair (u, c) (the FP32 state form), combined in FP64.  psf (may
// be NULL) is the reference face pressure that an FP32 x1 momentum flux
// carries as an anomaly.
template <typename TF, bool WRITE_DU, bool PAIR>
struct UpdateGravityBody {
  static __device__ __forceinline__ void run(dim3 const& bid, dim3 const& gdim,
                                             TF const* __restrict__ f1, TF const* __restrict__ f2,
                                             TF const* __restrict__ f3,
                                             void const* __restrict__ u0v,
                                             float const* __restrict__ c0, void* __restrict__ uv,
                                             float* __restrict__ cc, double const* __restrict__ w,
                                             double const* __restrict__ tab,
                                             double const* __restrict__ dsrc,
                                             double const* __restrict__ psf, GUpdArgs a) {
    using T = double;
    Range3 const& r = a.r;
    int64_t idx = int64_t(bid.x) * blockDim.x + threadIdx.x;
    int64_t tot = int64_t(r.n1) * r.n2 * r.n3;
    if (idx >= tot) return;
    int ii = int(idx % r.n1);
    int i = r.lo1 + ii;
    int64_t rest = idx / r.n1;
    int j = r.lo2 + int(rest % r.n2);
    int k = r.lo3 + int(rest / r.n2);
    int64_t row = k * r.s3 + j * r.s2;
    int64_t c = row + i;
    double dt = dt_at(a.dt_dev, -a.mdt, r, k);
    T w0 = T(a.w0), w1 = T(a.w1), w2 = T(a.w2);
    T rho = w[c], vx = w[r.sv + c];
    // x1 face flux of row v at face q (the reference face pressure restored
    // on an FP32 momentum anomaly)
    auto F1 = [&](int v, int s) {
      int64_t q = c + s;
      T x = T(f1[v * r.sv + q]);
      if (psf && v == IVX) x = x + psf[q];
      return x;
    };
    TF const* m1 = f1 + row;  // the IDN row of F1 along this column
    T du[5];
    gravity_increment(
        a, ii, i, dt, rho, vx, F1, [&](int q) { return T(m1[q]); },
        [&](int v, int sgn) { return f2[v * r.sv + c + (sgn ? r.s2 : 0)]; },
        [&](int v, int sgn) { return f3[v * r.sv + c + (sgn ? r.s3 : 0)]; }, f2 != nullptr,
        f3 != nullptr, tab, i, dsrc, c, du);
    for (int v = 0; v < 5; ++v) {
      int64_t o = v * r.sv + c;
      if (WRITE_DU) {
        static_cast<double*>(uv)[o] = du[v];
      } else if (PAIR) {
        float* uf = static_cast<float*>(uv);
        float const* u0f = static_cast<float const*>(u0v);
        double x = w0 * (double(u0f[o]) + double(c0[o])) +
                   w1 * (double(uf[o]) + double(cc[o])) + w2 * du[v];
        float h = float(x);
        uf[o] = h;
        cc[o] = float(x - double(h));
      } else {
        double* ud = static_cast<double*>(uv);
        double const* u0d = static_cast<double const*>(u0v);
        ud[o] = w0 * u0d[o] + w1 * ud[o] + w2 * du[v];
      }
    }
  }
};

template <typename TF, bool WRITE_DU, bool PAIR>
__global__ void __launch_bounds__(256)
    update_gravity_kernel(TF const* __restrict__ f1, TF const* __restrict__ f2,
                          TF const* __restrict__ f3, void const* __restrict__ u0v,
                          float const* __restrict__ c0, void* __restrict__ uv,
                          float* __restrict__ cc, double const* __restrict__ w,
                          double const* __restrict__ tab,
                          double const* __restrict__ dsrc,
                          double const* __restrict__ psf, GUpdArgs a) {
  UpdateGravityBody<TF, WRITE_DU, PAIR>::run(block_index(), grid_size(), f1, f2, f3, u0v, c0, uv,
                                             cc, w, tab, dsrc, psf, a);
}

}  // namespace
}  // namespace snapy

using namespace snapy;

static WbArgs wb_args(const snapy_grid* g, const snapy_physics* p) {
  WbArgs a;
  a.r = ranges(*g);
  a.nf = g->nx1 + 1;
  a.gammad = p->gammad;
  a.riemann = p->riemann;
  a.shock = p->shock_x1;
  a.mirror_lo = p->bc[0] != SNAPY_BC_OUTFLOW;
  a.mirror_hi = p->bc[1] != SNAPY_BC_OUTFLOW;
  a.refl_lo = p->bc[0] == SNAPY_BC_REFLECTING;
  a.refl_hi = p->bc[1] == SNAPY_BC_REFLECTING;
  a.wall_lo = a.refl_lo;
  a.wall_hi = a.refl_hi;
  return a;
}

static RefArgs ref_args(const snapy_grid* g, const snapy_physics* p) {
  RefArgs a;
  a.r = ranges(*g);
  a.iu = a.r.lo1 + g->nx1 - 1;
  a.grav = -p->grav1;
  a.dx1 = g->dx1;
  a.wall_clamp = 1;
  return a;
}

extern "C" {

int snapy_hydro_ref_x1(const snapy_grid* g, const snapy_physics* p,
                       const double* w, double* ref, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !ref) return fail(-1, "hydro_ref_x1: NULL argument");
  if (g->nc1 < 6) return fail(-9, "hydro_ref_x1: x1 is not resolved");
  if (p->grav1 == 0.) return fail(-12, "hydro_ref_x1: no x1 gravity");
  RefArgs a = ref_args(g, p);
  int64_t ncol = int64_t(g->nc2) * g->nc3;
  // columns per workgroup: kRefCols, fewer for a tall column
  int nc = kRefCols;
  auto lds_of = [&](int c) { return size_t(c) * (4 * g->nc1 + 1) * sizeof(double); };
  while (nc > 1 && lds_of(nc) > 64 * 1024) nc /= 2;
  if (lds_of(nc) > 64 * 1024) return fail(-9, "hydro_ref_x1: an x1 column above 2047 cells");
  hipLaunchKernelGGL(hydro_ref_x1_kernel, dim3(unsigned((ncol + nc - 1) / nc)),
                     dim3(kRefThreads), lds_of(nc), as_stream(stream), w, ref, a, nc);
  return launch_status("hydro_ref_x1");
}

int snapy_sweep_x1_wb(const snapy_grid* g, const snapy_physics* p, int dtype,
                      const void* w, const double* ref, void* flux,
                      void* face_pressure, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !ref || !flux) return fail(-1, "sweep_x1_wb: NULL argument");
  if (g->nx1 < 2) return fail(-9, "sweep_x1_wb: x1 is not resolved");
  if (dtype != SNAPY_F64 && dtype != SNAPY_F32) return fail(-10, "unknown dtype");
  WbArgs a = wb_args(g, p);
  int64_t tot = int64_t(a.nf) * g->nx2 * g->nx3;
  unsigned tpb = tuned(SNAPY_TUNE_SWEEP, g, dtype, 256);
  if (dtype == SNAPY_F64)
    hipLaunchKernelGGL((sweep_x1_wb_kernel<double, double>),
                       dim3(blocks_for(tot, tpb)), dim3(tpb), 0,
                       as_stream(stream), static_cast<double const*>(w), ref,
                       static_cast<double*>(flux),
                       static_cast<double*>(face_pressure), a);
  else if (p->riemann == SNAPY_RIEMANN_LMARS)
    // FP32 fluxes with LMARS: FP32 compute in the anomaly form
    hipLaunchKernelGGL((sweep_x1_wb_kernel<float, float>),
                       dim3(blocks_for(tot, tpb)), dim3(tpb), 0,
                       as_stream(stream), static_cast<double const*>(w), ref,
                       static_cast<float*>(flux),
                       static_cast<double*>(face_pressure), a);
  else
    hipLaunchKernelGGL((sweep_x1_wb_kernel<double, float>),
                       dim3(blocks_for(tot, tpb)), dim3(tpb), 0,
                       as_stream(stream), static_cast<double const*>(w), ref,
                       static_cast<float*>(flux),
                       static_cast<double*>(face_pressure), a);
  return launch_status("sweep_x1_wb");
}

int snapy_sweep_x1_wb_ref(const snapy_grid* g, const snapy_physics* p, int dtype,
                          const double* w, double* ref, void* flux, void* face_pressure,
                          void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !ref || !flux) return fail(-1, "sweep_x1_wb_ref: NULL argument");
  if (g->nc1 < 6) return fail(-9, "sweep_x1_wb_ref: x1 is not resolved");
  if (p->grav1 == 0.) return fail(-12, "sweep_x1_wb_ref: no x1 gravity");
  if (dtype != SNAPY_F64 && dtype != SNAPY_F32) return fail(-10, "unknown dtype");
  WbArgs a = wb_args(g, p);
  RefArgs ra = ref_args(g, p);
  int64_t ncol = int64_t(g->nx2) * g->nx3;
  // columns per workgroup: the tuned value (default kRefCols), fewer for a
  // tall column
  int nc = int(tuned(SNAPY_TUNE_REF_COLS, g, dtype, kRefCols));
  if (nc > kRefCols) nc = kRefCols;
  auto lds_of = [&](int c) { return size_t(c) * (4 * g->nc1 + 1) * sizeof(double); };
  while (nc > 1 && lds_of(nc) > 64 * 1024) nc /= 2;
  if (lds_of(nc) > 64 * 1024) return fail(-9, "sweep_x1_wb_ref: an x1 column too tall for LDS");
  dim3 grid(unsigned((ncol + nc - 1) / nc));
  if (dtype == SNAPY_F64)
    launch<SweepX1WbRefBody<double, double>, kRefThreads, 1>((sweep_x1_wb_ref_kernel<double, double>), grid, dim3(kRefThreads),
                       lds_of(nc), as_stream(stream), w, ref, static_cast<double*>(flux),
                       static_cast<double*>(face_pressure), ra, a, nc);
  else if (p->riemann == SNAPY_RIEMANN_LMARS)
    // FP32 fluxes with LMARS: FP32 compute in the anomaly form
    launch<SweepX1WbRefBody<float, float>, kRefThreads, 1>((sweep_x1_wb_ref_kernel<float, float>), grid, dim3(kRefThreads),
                       lds_of(nc), as_stream(stream), w, ref, static_cast<float*>(flux),
                       static_cast<double*>(face_pressure), ra, a, nc);
  else
    launch<SweepX1WbRefBody<double, float>, kRefThreads, 1>((sweep_x1_wb_ref_kernel<double, float>), grid, dim3(kRefThreads),
                       lds_of(nc), as_stream(stream), w, ref, static_cast<float*>(flux),
                       static_cast<double*>(face_pressure), ra, a, nc);
  return launch_status("sweep_x1_wb_ref");
}

int snapy_sweep_x1_wb_inc_f32(const snapy_grid* g, const snapy_physics* p, const double* w,
                              double* ref, float* flux1, double* face_pressure,
                              const float* flux2, const float* flux3, const double* x1tab,
                              double* du_out, double dt, const double* dt_dev, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !w || !ref || !flux1 || !x1tab || !du_out)
    return fail(-1, "sweep_x1_wb_inc_f32: NULL argument");
  if (g->nvar != 5) return fail(-6, "sweep_x1_wb_inc_f32: ideal gas needs nvar 5");
  if (g->nc1 < 6 || g->nx1 < 3) return fail(-9, "sweep_x1_wb_inc_f32: x1 is not resolved");
  if (p->grav1 == 0.) return fail(-12, "sweep_x1_wb_inc_f32: no x1 gravity");
  if (p->riemann != SNAPY_RIEMANN_LMARS)
    return fail(-13, "sweep_x1_wb_inc_f32: the FP32 faces take LMARS");
  WbArgs a = wb_args(g, p);
  RefArgs ra = ref_args(g, p);
  int64_t ncol = int64_t(g->nx2) * g->nx3;
  int nc = int(tuned(SNAPY_TUNE_REF_COLS, g, SNAPY_F32, kRefCols));
  if (nc > kRefCols) nc = kRefCols;
  auto lds_of = [&](int c) { return size_t(c) * (4 * g->nc1 + 1) * sizeof(double); };
  while (nc > 1 && lds_of(nc) > 64 * 1024) nc /= 2;
  if (lds_of(nc) > 64 * 1024) return fail(-9, "sweep_x1_wb_inc_f32: an x1 column too tall for LDS");
  X1IncArgs ia;
  GUpdArgs& ga = ia.g;
  ga.r = ranges(*g);
  ga.a1 = g->dx3 * g->dx2;
  ga.a2 = g->dx3 * g->dx1;
  ga.a3 = g->dx2 * g->dx1;
  ga.vol = g->dx3 * (g->dx2 * g->dx1);
  ga.mdt = -dt;
  ga.dt_dev = dt_dev;
  ga.w0 = ga.w1 = ga.w2 = 0.;
  ga.grav1 = p->grav1;
  ga.heat_depth = p->bot_heat_flux != 0. ? p->bot_heat_depth : 0;
  ga.heat_flux = p->bot_heat_flux;
  ga.heat_dzd = g->dx1 * double(p->bot_heat_depth);
  ia.f2 = flux2;
  ia.f3 = flux3;
  ia.tab = x1tab;
  ia.du = du_out;
  dim3 grid(unsigned((ncol + nc - 1) / nc));
  launch<SweepX1WbIncBody, kRefThreads, 1>(sweep_x1_wb_inc_kernel, grid, dim3(kRefThreads), lds_of(nc),
                                           as_stream(stream), w, ref, flux1, face_pressure, ra, a, nc,
                                           ia);
  return launch_status("sweep_x1_wb_inc_f32");
}

int snapy_update_gravity(const snapy_grid* g, const snapy_physics* p,
                         int dtype, const void* const flux[3], const void* u0,
                         void* u, const void* w, const double* x1tab,
                         const void* du_src, double dt, const double* dt_dev,
                         double w0, double w1, double w2, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !flux || !flux[0] || !u || !w || !x1tab)
    return fail(-1, "update_gravity: NULL argument");
  if (g->nx1 < 3) return fail(-9, "update_gravity: needs >= 3 x1 cells");
  if (dtype != SNAPY_F64) return fail(-10, "update_gravity: FP64 state");
  GUpdArgs a;
  a.r = ranges(*g);
  a.a1 = g->dx3 * g->dx2;
  a.a2 = g->dx3 * g->dx1;
  a.a3 = g->dx2 * g->dx1;
  a.vol = g->dx3 * (g->dx2 * g->dx1);
  a.mdt = -dt;
  a.dt_dev = dt_dev;
  a.w0 = w0;
  a.w1 = w1;
  a.w2 = w2;
  a.grav1 = p->grav1;
  a.heat_depth = p->bot_heat_flux != 0. ? p->bot_heat_depth : 0;
  a.heat_flux = p->bot_heat_flux;
  a.heat_dzd = g->dx1 * double(p->bot_heat_depth);
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  unsigned tpb = tuned(SNAPY_TUNE_UPDATE, g, SNAPY_F64, 256);
  auto go = [&](auto body, auto kern) {
    launch<decltype(body), 256, 1>(kern, dim3(blocks_for(tot, tpb)), dim3(tpb), 0,
                                   as_stream(stream), static_cast<double const*>(flux[0]),
                                   static_cast<double const*>(flux[1]),
                                   static_cast<double const*>(flux[2]), u0, nullptr, u,
                                   nullptr, static_cast<double const*>(w), x1tab,
                                   static_cast<double const*>(du_src), nullptr, a);
  };
  if (u0)
    go(UpdateGravityBody<double, false, false>{}, update_gravity_kernel<double, false, false>);
  else
    go(UpdateGravityBody<double, true, false>{}, update_gravity_kernel<double, true, false>);
  return launch_status("update_gravity");
}

int snapy_update_gravity_f32(const snapy_grid* g, const snapy_physics* p,
                             const void* const flux[3], const double* ref,
                             const float* u0, const float* c0, float* u,
                             float* c, double* du_out, const double* w,
                             const double* x1tab, const double* du_src,
                             double dt, const double* dt_dev, double w0,
                             double w1, double w2, void* stream) {
  if (int e = check_grid(g)) return e;
  if (!p || !flux || !flux[0] || !ref || !w || !x1tab)
    return fail(-1, "update_gravity_f32: NULL argument");
  if (!du_out && (!u0 || !c0 || !u || !c))
    return fail(-1, "update_gravity_f32: NULL state");
  if (g->nx1 < 3) return fail(-9, "update_gravity_f32: needs >= 3 x1 cells");
  GUpdArgs a;
  a.r = ranges(*g);
  a.a1 = g->dx3 * g->dx2;
  a.a2 = g->dx3 * g->dx1;
  a.a3 = g->dx2 * g->dx1;
  a.vol = g->dx3 * (g->dx2 * g->dx1);
  a.mdt = -dt;
  a.dt_dev = dt_dev;
  a.w0 = w0;
  a.w1 = w1;
  a.w2 = w2;
  a.grav1 = p->grav1;
  a.heat_depth = p->bot_heat_flux != 0. ? p->bot_heat_depth : 0;
  a.heat_flux = p->bot_heat_flux;
  a.heat_dzd = g->dx1 * double(p->bot_heat_depth);
  int64_t tot = int64_t(g->nx1) * g->nx2 * g->nx3;
  double const* psf = ref + R_PSF_LO * ranges(*g).sv;
  void* uo = du_out ? static_cast<void*>(du_out) : static_cast<void*>(u);
  unsigned tpb = tuned(SNAPY_TUNE_UPDATE, g, SNAPY_F32, 256);
  auto go = [&](auto body, auto kern) {
    launch<decltype(body), 256, 1>(kern, dim3(blocks_for(tot, tpb)), dim3(tpb), 0,
                                   as_stream(stream), static_cast<float const*>(flux[0]),
                                   static_cast<float const*>(flux[1]),
                                   static_cast<float const*>(flux[2]), u0, c0, uo, c, w,
                                   x1tab, du_src, psf, a);
  };
  if (du_out)
    go(UpdateGravityBody<float, true, true>{}, update_gravity_kernel<float, true, true>);
  else
    go(UpdateGravityBody<float, false, true>{}, update_gravity_kernel<float, false, true>);
  return launch_status("update_gravity_f32");
}

}  // extern "C"
