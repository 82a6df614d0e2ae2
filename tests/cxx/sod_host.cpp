// A C++ host with no Python and no torch: it advances the Sod deck
// (snapy_amd/decks/sod.yaml: 256 cells on x1, weno5 with shock, HLLC,
// outflow, RK3, CFL 0.9) through the C-ABI of include/snapy_amd.h alone,
// owning every device buffer, and writes the final conserved state.  It
// makes the launches MeshBlock makes for this deck, in the same order
// (initialize: the outflow background, the primitive ghost fill, prim2cons;
// each step: the primitives refreshed from the state after the first step,
// the CFL dt on the device; each stage: the RK backup in the first, EOS, the
// x1 sweep, the divergence + RK combine, the boundaries), so its result is
// MeshBlock's bit for bit (tests/test_gpu_capi_host.py).
//
// usage: sod_host STEPS OUT.bin  (OUT: the (5, 262) FP64 state, row-major)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "snapy_amd.h"

#define CALL(x)                                                                  \
  do {                                                                           \
    int rc_ = (x);                                                               \
    if (rc_ != 0) {                                                              \
      std::fprintf(stderr, "%s failed (%d): %s\n", #x, rc_, snapy_last_error()); \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

#define HIP(x)                                                                      \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));           \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s STEPS OUT.bin\n", argv[0]);
    return 2;
  }
  int const steps = std::atoi(argv[1]);
  int const ng = 3, nx = 256, nc = nx + 2 * ng;
  // the deck's grid: x1 in [0, 1], x2 and x3 single cells of unit width
  snapy_grid g{};
  g.nvar = 5;
  g.nghost = ng;
  g.nx1 = nx;
  g.nx2 = g.nx3 = 1;
  g.nc1 = nc;
  g.nc2 = g.nc3 = 1;
  g.dx1 = 1.0 / nx;
  g.dx2 = g.dx3 = 1.0;
  g.x1min = 0.0;
  g.x2min = g.x3min = -0.5;
  snapy_physics p{};
  p.gammad = 1.4;
  p.density_floor = p.pressure_floor = 1.e-10;
  p.riemann = SNAPY_RIEMANN_HLLC;
  p.shock_x1 = p.shock_x23 = 1;
  for (int f = 0; f < 6; ++f) p.bc[f] = SNAPY_BC_OUTFLOW;
  p.advection = 0.9;
  p.eos = SNAPY_EOS_IDEAL_GAS;
  double const cfl = 0.9;
  // RK3 (SSP) stage weights: u <- w0 u0 + w1 u + w2 dt du
  double const wt[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};

  // the Sod state on every cell, ghosts included: (1, 0, 1) left of x = 0.5,
  // (0.125, 0, 0.1) right of it
  size_t const n = size_t(5) * nc;
  std::vector<double> w0(n, 0.0);
  for (int i = 0; i < nc; ++i) {
    double x = g.x1min + (i - ng + 0.5) * g.dx1;
    bool left = x < 0.5;
    w0[0 * nc + i] = left ? 1.0 : 0.125;
    w0[4 * nc + i] = left ? 1.0 : 0.1;
  }
  double *u, *u0, *w, *bref, *scratch, *flux, *fp, *dt;
  HIP(hipMalloc(&u, n * sizeof(double)));
  HIP(hipMalloc(&u0, n * sizeof(double)));
  HIP(hipMalloc(&w, n * sizeof(double)));
  HIP(hipMalloc(&bref, n * sizeof(double)));
  HIP(hipMalloc(&scratch, n * sizeof(double)));
  HIP(hipMalloc(&flux, n * sizeof(double)));
  HIP(hipMalloc(&fp, nc * sizeof(double)));
  HIP(hipMalloc(&dt, sizeof(double)));
  HIP(hipMemset(flux, 0, n * sizeof(double)));
  HIP(hipMemset(fp, 0, nc * sizeof(double)));
  hipStream_t s;
  HIP(hipStreamCreate(&s));
  HIP(hipMemcpy(w, w0.data(), n * sizeof(double), hipMemcpyHostToDevice));
  HIP(hipMemcpy(bref, w0.data(), n * sizeof(double), hipMemcpyHostToDevice));
  // initialize: the primitive ghost fill, then the conserved state
  CALL(snapy_apply_boundaries_prim(&g, &p, SNAPY_F64, w, bref, s));
  CALL(snapy_prim2cons(&g, &p, SNAPY_F64, w, u, s));
  void const* fl[3] = {flux, nullptr, nullptr};
  for (int step = 0; step < steps; ++step) {
    // the time step reads the primitives of the current state
    if (step > 0) CALL(snapy_cons2prim(&g, &p, SNAPY_F64, u, w, s));
    CALL(snapy_cfl(&g, &p, SNAPY_F64, w, cfl, dt, s));
    for (int k = 0; k < 3; ++k) {
      if (k == 0) HIP(hipMemcpyAsync(u0, u, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      CALL(snapy_cons2prim(&g, &p, SNAPY_F64, u, w, s));
      CALL(snapy_sweep(1, &g, &p, SNAPY_F64, w, flux, fp, s));
      CALL(snapy_update(&g, SNAPY_F64, fl, u0, u, nullptr, 0.0, dt, wt[k][0], wt[k][1], wt[k][2], s));
      CALL(snapy_apply_boundaries(&g, &p, SNAPY_F64, u, bref, scratch, s));
    }
  }
  std::vector<double> out(n);
  HIP(hipMemcpyAsync(out.data(), u, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP(hipStreamSynchronize(s));
  std::FILE* f = std::fopen(argv[2], "wb");
  if (!f || std::fwrite(out.data(), sizeof(double), n, f) != n) {
    std::fprintf(stderr, "cannot write %s\n", argv[2]);
    return 1;
  }
  std::fclose(f);
  double mass = 0.0;
  for (int i = ng; i < ng + nx; ++i) mass += out[i] * g.dx1;
  std::printf("sod_host: %d steps, interior mass %.15e\n", steps, mass);
  HIP(hipStreamDestroy(s));
  for (double* b : {u, u0, w, bref, scratch, flux, fp, dt}) HIP(hipFree(b));
  return 0;
}
