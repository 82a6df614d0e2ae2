// Device point functions of the finite-volume stage: one face value of the
// reconstruction, one face flux of the Riemann solvers, one cell of the EOS.
// Every function restates the reference's arithmetic order term by term so the
// FP64 form tracks the reference CPU path to round-off; the library is built
// with -ffp-contract=off so no product/sum pair is fused behind our back.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace snapy {

enum { IDN = 0, IVX = 1, IVY = 2, IVZ = 3, IPR = 4 };

template <typename T>
__device__ __forceinline__ T sqr(T x) {
  return x * x;
}

// Quotient and square root of the reconstruction and the Riemann solvers:
// correctly rounded in FP64 (the form held to the reference's round-off);
// in FP32 the hardware reciprocal and square root (1 ulp), two instructions
// instead of the ten of a correctly rounded FP32 quotient.  The FP32 form is
// held to 1e-5 relative L2 of the FP64 result, far above these ulps.
template <typename T>
__device__ __forceinline__ T qdiv(T a, T b) {
  return a / b;
}
template <>
__device__ __forceinline__ float qdiv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
template <typename T>
__device__ __forceinline__ T qsqrt(T a) {
  return sqrt(a);
}
template <>
__device__ __forceinline__ float qsqrt<float>(float a) {
  return __builtin_amdgcn_sqrtf(a);
}

// ---------------------------------------------------------------------------
// Reconstruction (src/recon/recon_point.h, tables src/recon/reconstruct.cpp)
// ---------------------------------------------------------------------------
//
// The window is the 5 cells v[0..4] centred on the reconstructed cell.  The
// "minus" table (cm) gives the LEFT edge value of the centre cell, the "plus"
// table (cp = cm with every row reversed) its RIGHT edge value.  Zero
// coefficients of the reference tables are dropped: 0*x adds an exact zero.

// weno5 with the uniform table and scale = false (recon_point.h:114-147;
// rows reconstruct.cpp:74-86, d = .3,.6,.1 at :84, eps = 1e-6 at :65-67).
// The weight-averaged candidate split into its numerator sum(a_k p_k) and
// denominator sum(a_k), which the scaled form needs apart.
template <typename T>
__device__ __forceinline__ void weno5_minus_parts(T v0, T v1, T v2, T v3, T v4, T& num, T& den) {
  // candidate rows
  T p0 = T(-1. / 6.) * v0;
  p0 = p0 + T(5. / 6.) * v1;
  p0 = p0 + T(1. / 3.) * v2;
  T p1 = T(1. / 3.) * v1;
  p1 = p1 + T(5. / 6.) * v2;
  p1 = p1 + T(-1. / 6.) * v3;
  T p2 = T(11. / 6.) * v2;
  p2 = p2 + T(-7. / 6.) * v3;
  p2 = p2 + T(1. / 3.) * v4;
  // smoothness rows
  T s4 = v0 + T(-2.) * v1;
  s4 = s4 + v2;
  T s5 = v0 + T(-4.) * v1;
  s5 = s5 + T(3.) * v2;
  T s6 = v1 + T(-2.) * v2;
  s6 = s6 + v3;
  T s7 = T(-1.) * v1;
  s7 = s7 + v3;
  T s8 = v2 + T(-2.) * v3;
  s8 = s8 + v4;
  T s9 = T(3.) * v2;
  s9 = s9 + T(-4.) * v3;
  s9 = s9 + v4;
  T beta0 = T(13. / 12.) * sqr(s4) + T(.25) * sqr(s5);
  T beta1 = T(13. / 12.) * sqr(s6) + T(.25) * sqr(s7);
  T beta2 = T(13. / 12.) * sqr(s8) + T(.25) * sqr(s9);
  T a0 = qdiv(T(.3), sqr(beta0 + T(1e-6)));
  T a1 = qdiv(T(.6), sqr(beta1 + T(1e-6)));
  T a2 = qdiv(T(.1), sqr(beta2 + T(1e-6)));
  num = a0 * p0 + a1 * p1 + a2 * p2;
  den = a0 + a1 + a2;
}

template <typename T>
__device__ __forceinline__ T weno5_minus(T v0, T v1, T v2, T v3, T v4) {
  T num, den;
  weno5_minus_parts(v0, v1, v2, v3, v4, num, den);
  return qdiv(num, den);
}

template <typename T>
__device__ __forceinline__ void weno5_plus_parts(T v0, T v1, T v2, T v3, T v4, T& num, T& den) {
  // the reversed rows, accumulated in window order as the reference does
  T p0 = T(1. / 3.) * v2;
  p0 = p0 + T(5. / 6.) * v3;
  p0 = p0 + T(-1. / 6.) * v4;
  T p1 = T(-1. / 6.) * v1;
  p1 = p1 + T(5. / 6.) * v2;
  p1 = p1 + T(1. / 3.) * v3;
  T p2 = T(1. / 3.) * v0;
  p2 = p2 + T(-7. / 6.) * v1;
  p2 = p2 + T(11. / 6.) * v2;
  T s4 = v2 + T(-2.) * v3;
  s4 = s4 + v4;
  T s5 = T(3.) * v2;
  s5 = s5 + T(-4.) * v3;
  s5 = s5 + v4;
  T s6 = v1 + T(-2.) * v2;
  s6 = s6 + v3;
  T s7 = v1 + T(-1.) * v3;
  T s8 = v0 + T(-2.) * v1;
  s8 = s8 + v2;
  T s9 = v0 + T(-4.) * v1;
  s9 = s9 + T(3.) * v2;
  T beta0 = T(13. / 12.) * sqr(s4) + T(.25) * sqr(s5);
  T beta1 = T(13. / 12.) * sqr(s6) + T(.25) * sqr(s7);
  T beta2 = T(13. / 12.) * sqr(s8) + T(.25) * sqr(s9);
  T a0 = qdiv(T(.3), sqr(beta0 + T(1e-6)));
  T a1 = qdiv(T(.6), sqr(beta1 + T(1e-6)));
  T a2 = qdiv(T(.1), sqr(beta2 + T(1e-6)));
  num = a0 * p0 + a1 * p1 + a2 * p2;
  den = a0 + a1 + a2;
}

template <typename T>
__device__ __forceinline__ T weno5_plus(T v0, T v1, T v2, T v3, T v4) {
  T num, den;
  weno5_plus_parts(v0, v1, v2, v3, v4, num, den);
  return qdiv(num, den);
}

// Both edges of the centre cell of the window v0..v4 with weno5 (the left
// edge em = weno5_minus(v0..v4), the right edge ep = weno5_plus(v0..v4)),
// each term in the order of weno5_minus_parts / weno5_plus_parts, so the
// values are those two functions' bits.  The two share their smoothness
// indicators (the right edge's beta0 and beta2 are the left edge's beta2
// and beta0, beta1 the same, and its s7 is the left edge's with the sign
// flipped, which squares to the same value) and the reciprocals of the
// squared indicators, which qdiv<float> forms as its quotients' divisors.
__device__ __forceinline__ void weno5_edges_f32(float v0, float v1, float v2, float v3, float v4,
                                                float& em, float& ep) {
  float s4 = v0 + -2.f * v1;
  s4 = s4 + v2;
  float s5 = v0 + -4.f * v1;
  s5 = s5 + 3.f * v2;
  float s6 = v1 + -2.f * v2;
  s6 = s6 + v3;
  float s7 = -1.f * v1;
  s7 = s7 + v3;
  float s8 = v2 + -2.f * v3;
  s8 = s8 + v4;
  float s9 = 3.f * v2;
  s9 = s9 + -4.f * v3;
  s9 = s9 + v4;
  float const c13 = float(13. / 12.);
  float beta0 = c13 * sqr(s4) + .25f * sqr(s5);
  float beta1 = c13 * sqr(s6) + .25f * sqr(s7);
  float beta2 = c13 * sqr(s8) + .25f * sqr(s9);
  float r0 = __builtin_amdgcn_rcpf(sqr(beta0 + 1e-6f));
  float r1 = __builtin_amdgcn_rcpf(sqr(beta1 + 1e-6f));
  float r2 = __builtin_amdgcn_rcpf(sqr(beta2 + 1e-6f));
  float const d0 = .3f, d1 = .6f, d2 = .1f;
  float const m6 = float(-1. / 6.), p56 = float(5. / 6.), p13 = float(1. / 3.),
              p116 = float(11. / 6.), m76 = float(-7. / 6.);
  {
    float p0 = m6 * v0;
    p0 = p0 + p56 * v1;
    p0 = p0 + p13 * v2;
    float p1 = p13 * v1;
    p1 = p1 + p56 * v2;
    p1 = p1 + m6 * v3;
    float p2 = p116 * v2;
    p2 = p2 + m76 * v3;
    p2 = p2 + p13 * v4;
    float a0 = d0 * r0, a1 = d1 * r1, a2 = d2 * r2;
    float num = a0 * p0 + a1 * p1 + a2 * p2;
    float den = a0 + a1 + a2;
    em = qdiv(num, den);
  }
  {
    float p0 = p13 * v2;
    p0 = p0 + p56 * v3;
    p0 = p0 + m6 * v4;
    float p1 = m6 * v1;
    p1 = p1 + p56 * v2;
    p1 = p1 + p13 * v3;
    float p2 = p13 * v0;
    p2 = p2 + m76 * v1;
    p2 = p2 + p116 * v2;
    float a0 = d0 * r2, a1 = d1 * r1, a2 = d2 * r0;
    float num = a0 * p0 + a1 * p1 + a2 * p2;
    float den = a0 + a1 + a2;
    ep = qdiv(num, den);
  }
}

// cp5 (poly_point<T,5>, recon_point.h:54-61; table reconstruct.cpp:65-67)
template <typename T>
__device__ __forceinline__ T cp5_minus(T v0, T v1, T v2, T v3, T v4) {
  T o = T(-1. / 20.) * v0;
  o = o + T(9. / 20.) * v1;
  o = o + T(47. / 60.) * v2;
  o = o + T(-13. / 60.) * v3;
  o = o + T(1. / 30.) * v4;
  return o;
}

template <typename T>
__device__ __forceinline__ T cp5_plus(T v0, T v1, T v2, T v3, T v4) {
  T o = T(1. / 30.) * v0;
  o = o + T(-13. / 60.) * v1;
  o = o + T(47. / 60.) * v2;
  o = o + T(9. / 20.) * v3;
  o = o + T(-1. / 20.) * v4;
  return o;
}

// ---------------------------------------------------------------------------
// Riemann solvers (src/riemann/riemann_kernels.h)
// ---------------------------------------------------------------------------
// Face states are in the face frame: n = normal row, t1/t2 the tangential
// rows in cyclic order (face_axes, riemann_kernels.h:37-40).  A state is
// {rho, v[IVX], v[IVY], v[IVZ], p} indexed by the global rows; the solver
// reads the normal component through n.

template <typename T>
struct Prim5 {
  T r, v[3], p;  // v[0] = IVX row, v[1] = IVY row, v[2] = IVZ row
};

// HLLC, FP64/FP32 forms of hllc_impl (riemann_kernels.h:56-194), dry
// (ny = 0: the dry fraction is 1) and without the face background.  e is the
// internal energy per volume p/(gamma-1), c the sound speed
// (src/riemann/riemann.cpp:232-260).
template <typename T>
__device__ __forceinline__ void hllc(const Prim5<T>& L, const Prim5<T>& R,
                                     T gl, T gr, int n, T out[5], T* fp) {
  T el = L.p / (gl - T(1));
  T er = R.p / (gr - T(1));
  T cl = sqrt(gl * L.p / L.r);
  T cr = sqrt(gr * R.p / R.r);
  const int t1 = (n + 1) % 3, t2 = (n + 2) % 3;
  T uL = L.v[n], uR = R.v[n];

  el = el + T(0.5) * L.r * (sqr(L.v[0]) + sqr(L.v[1]) + sqr(L.v[2]));
  er = er + T(0.5) * R.r * (sqr(R.v[0]) + sqr(R.v[1]) + sqr(R.v[2]));

  T rhoa = T(0.5) * (L.r + R.r);
  T ca = T(0.5) * (cl + cr);
  T pmid = T(0.5) * (L.p + R.p + (uL - uR) * rhoa * ca);

  T shock_l = pmid / L.p - T(1.0);
  T shock_r = pmid / R.p - T(1.0);
  T ql = (pmid <= L.p) ? T(1.0)
                       : sqrt(T(1.0) + (gl + 1) / (2 * gl) * shock_l);
  T qr = (pmid <= R.p) ? T(1.0)
                       : sqrt(T(1.0) + (gr + 1) / (2 * gr) * shock_r);

  T al = uL - cl * ql;
  T ar = uR + cr * qr;
  T bp = ar > T(0.0) ? ar : T(1.0e-10);
  T bm = al < T(0.0) ? al : -T(1.0e-10);

  T vxl, vxr;
  if constexpr (sizeof(T) == 4) {
    vxl = cl * ql;
    vxr = -cr * qr;
  } else {
    vxl = uL - al;
    vxr = uR - ar;
  }
  T tl = L.p + vxl * L.r * uL;
  T tr = R.p + vxr * R.r * uR;
  T ml = L.r * vxl;
  T mr = -(R.r * vxr);
  T am = (tl - tr) / (ml + mr);
  T cp = (ml * tr + mr * tl) / (ml + mr);
  cp = cp > T(0.0) ? cp : T(0.0);
  if (fp) *fp = cp;

  if constexpr (sizeof(T) == 4) {
    vxl = al < T(0) ? cl * ql : uL + T(1.0e-10);
    vxr = ar > T(0) ? -cr * qr : uR - T(1.0e-10);
  } else {
    vxl = uL - bm;
    vxr = uR - bp;
  }

  T flI = L.r * vxl, frI = R.r * vxr;
  T fln = L.r * uL * vxl + L.p, frn = R.r * uR * vxr + R.p;
  T fl1 = L.r * L.v[t1] * vxl, fr1 = R.r * R.v[t1] * vxr;
  T fl2 = L.r * L.v[t2] * vxl, fr2 = R.r * R.v[t2] * vxr;
  T flE = el * vxl + L.p * uL, frE = er * vxr + R.p * uR;

  T sl, sr, sm;
  if (am >= T(0.0)) {
    sl = am / (am - bm);
    sr = T(0.0);
    sm = -bm / (am - bm);
  } else {
    sl = T(0.0);
    sr = -am / (bp - am);
    sm = bp / (bp - am);
  }
  out[IDN] = sl * flI + sr * frI;
  out[IVX + n] = sl * fln + sr * frn + sm * cp;
  out[IVX + t1] = sl * fl1 + sr * fr1;
  out[IVX + t2] = sl * fl2 + sr * fr2;
  out[IPR] = sl * flE + sr * frE + sm * cp * am;
}

// LMARS (lmars_impl, riemann_kernels.h:197-255), dry.  The specific internal
// energy e/rho = p/((gamma-1) rho) enters the enthalpy
// (src/riemann/riemann.cpp:262-268).
//
// With a gauge g the face states carry p - g (a constant along the sweep
// line): the enthalpy and the sound speed take the total p' + g, the
// momentum flux and the face pressure stay gauge-free (their constant g
// cancels in the flux difference).  g = 0 is the plain solver exactly.
template <typename T>
__device__ __forceinline__ void lmars(const Prim5<T>& L, const Prim5<T>& R,
                                      T gl, T gr, int n, T out[5], T* fp,
                                      T gauge = T(0)) {
  const int t1 = (n + 1) % 3, t2 = (n + 2) % 3;
  T pl = L.p + gauge, pr = R.p + gauge;  // total pressures
  T hl = qdiv(qdiv(pl, gl - T(1)), L.r);
  T hr = qdiv(qdiv(pr, gr - T(1)), R.r);
  // wli = {rho, v_n, v_t1, v_t2, p}
  T lx = L.v[n], ly = L.v[t1], lz = L.v[t2];
  T rx = R.v[n], ry = R.v[t1], rz = R.v[t2];
  // hl += a + b is hl + (a + b)
  hl = hl + (T(0.5) * (sqr(lx) + sqr(ly) + sqr(lz)) + qdiv(pl, L.r));
  hr = hr + (T(0.5) * (sqr(rx) + sqr(ry) + sqr(rz)) + qdiv(pr, R.r));

  T rhobar = T(0.5) * (L.r + R.r);
  T gamma_bar = T(0.5) * (gl + gr);
  T cbar = qsqrt(qdiv(T(0.5) * gamma_bar * (pl + pr), rhobar));
  T pbar = T(0.5) * (L.p + R.p) + T(0.5) * (rhobar * cbar) * (lx - rx);
  if (fp) *fp = pbar;
  T ubar = T(0.5) * (lx + rx) + qdiv(T(0.5), rhobar * cbar) * (L.p - R.p);
  if (ubar > T(0)) {
    out[IDN] = ubar * L.r;  // * rd (= 1, dry)
    out[IVX + n] = ubar * L.r * lx + pbar;
    out[IVX + t1] = ubar * L.r * ly;
    out[IVX + t2] = ubar * L.r * lz;
    out[IPR] = ubar * L.r * hl;
  } else {
    out[IDN] = ubar * R.r;
    out[IVX + n] = ubar * R.r * rx + pbar;
    out[IVX + t1] = ubar * R.r * ry;
    out[IVX + t2] = ubar * R.r * rz;
    out[IPR] = ubar * R.r * hr;
  }
}

// ---------------------------------------------------------------------------
// one face of a sweep: reconstruction + Riemann solver
// ---------------------------------------------------------------------------
// The face below cell f of a sweep line: p points at row IDN of cell f, sv
// is the row stride and stride the cell stride along the sweep.  Density is
// always weno5; velocity and pressure cp5 unless shock
// (reconstruct.cpp:303-322).  With a gauge g the pressure row enters the
// reconstruction as p - g (LMARS only).  n is the sweep's normal row
// (0, 1, 2 for x1, x2, x3).
template <typename T>
__device__ __forceinline__ void face_flux_line(double const* p, int64_t sv, int64_t stride,
                                               double gauge, double gammad, int riemann,
                                               int shock, int n, T out[5], T* pface) {
  Prim5<T> L, R;
  T* lv[5] = {&L.r, &L.v[0], &L.v[1], &L.v[2], &L.p};
  T* rv[5] = {&R.r, &R.v[0], &R.v[1], &R.v[2], &R.p};
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double const* q = p + v * sv;
    double gv = v == IPR ? gauge : 0.;
    T c0 = T(q[-3 * stride] - gv), c1 = T(q[-2 * stride] - gv),
      c2 = T(q[-stride] - gv), c3 = T(q[0] - gv), c4 = T(q[stride] - gv),
      c5 = T(q[2 * stride] - gv);
    // left state: right edge of cell f-1 (window c0..c4); right state: left
    // edge of cell f (window c1..c5)
    if (v == IDN || shock) {
      *lv[v] = weno5_plus(c0, c1, c2, c3, c4);
      *rv[v] = weno5_minus(c1, c2, c3, c4, c5);
    } else {
      *lv[v] = cp5_plus(c0, c1, c2, c3, c4);
      *rv[v] = cp5_minus(c1, c2, c3, c4, c5);
    }
  }
  T g = T(gammad);
  if (riemann == SNAPY_RIEMANN_LMARS)
    lmars(L, R, g, g, n, out, pface, T(gauge));
  else
    hllc(L, R, g, g, n, out, pface);
}

// ---------------------------------------------------------------------------
// the outflow (radiating) ghost about the saved background (bc_func.cpp:
// 110-220), ideal gas, no tracers, Cartesian: wi and bi the interior
// boundary cell's state and background, bg the ghost cell's background,
// side 0 / 1 the lower / upper face, vn the normal velocity row.  Writes
// the ghost's primitive state.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void radiating_ghost(T wi[5], T const bi[5], T const bg[5], int side,
                                                int vn, double gd, double dfloor,
                                                double pfloor, T out[5]) {
  double sign = side ? 1. : -1.;
  T d[5];
  T gm = T(gd);
  T c0 = sqrt(gm * bi[IPR] / bi[IDN]);
  T c = sqrt(gm * wi[IPR] / wi[IDN]);
  bool ok = isfinite(c) && c > T(0);
  if (!ok) {
    for (int v = 0; v < 5; ++v) wi[v] = bi[v];
    c = c0;
  }
  for (int v = 0; v < 5; ++v) d[v] = wi[v] - bi[v];
  T un = T(sign) * wi[vn];
  bool advect = un >= T(0);
  T impedance = bi[IDN] * c0;
  T plus = (un + c >= T(0)) ? T(sign) * d[vn] + d[IPR] / impedance : T(0);
  T minus = (un - c >= T(0)) ? T(sign) * d[vn] - d[IPR] / impedance : T(0);
  T c02 = c0 * c0;
  T entropy = advect ? d[IDN] - d[IPR] / c02 : T(0);
  for (int v = 0; v < 5; ++v) d[v] = d[v] * T(advect ? 1 : 0);
  d[vn] = T(sign) * T(0.5) * (plus + minus);
  d[IPR] = T(0.5) * impedance * (plus - minus);
  d[IDN] = entropy + d[IPR] / c02;
  T alpha = T(1);
  auto lower = [&](T base, T delta, T floor) {
    T bound = delta < T(0) ? (base - floor) / (-delta) : T(1);
    T cand = bound < T(1) ? T(0.99) * bound : bound;
    alpha = cand < alpha ? cand : alpha;
  };
  lower(bg[IDN], d[IDN], T(dfloor));
  lower(bg[IPR], d[IPR], T(pfloor));
  alpha = alpha < T(0) ? T(0) : (alpha > T(1) ? T(1) : alpha);
  for (int v = 0; v < 5; ++v) out[v] = bg[v] + alpha * d[v];
}

// ---------------------------------------------------------------------------
// Equation of state (src/eos/ideal_gas_impl.h, src/eos/ideal_gas.cpp)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cons2prim5(const T u[5], T w[5], double gd) {
  // ideal_gas_cons2prim (ideal_gas_impl.h:16-35); Cartesian: cos(theta) = 0
  w[IDN] = u[IDN];
  w[IVX] = u[IVX] / w[IDN];
  w[IVY] = u[IVY] / w[IDN];
  w[IVZ] = u[IVZ] / w[IDN];
  // the kernel mixes T with the double gammad: ke and ie promote to double
  double ke = 0.5 * (w[IVX] * u[IVX] + w[IVY] * u[IVY] + w[IVZ] * u[IVZ]);
  double ie = u[IPR] - ke;
  w[IPR] = T((gd - 1.) * ie);
}

template <typename T>
__device__ __forceinline__ void prim2cons5(const T w[5], T u[5], T gm1) {
  // IdealGasImpl::_prim2cons (ideal_gas.cpp:49-65): mom = v*rho,
  // E = p/(gamma-1) + 0.5*sum(v*m)
  u[IDN] = w[IDN];
  u[IVX] = w[IVX] * w[IDN];
  u[IVY] = w[IVY] * w[IDN];
  u[IVZ] = w[IVZ] * w[IDN];
  T ke = T(0.5) * (w[IVX] * u[IVX] + w[IVY] * u[IVY] + w[IVZ] * u[IVZ]);
  u[IPR] = w[IPR] / gm1;
  u[IPR] = u[IPR] + ke;
}

}  // namespace snapy
