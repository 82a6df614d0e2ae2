/*
 * snapy_amd.h -- the C-ABI of the MI355X (gfx950) finite-volume stage.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t
 * passed as `void *` (NULL = the null stream), launches asynchronously on
 * that stream and returns 0 on success.  A negative return is an argument
 * error detected on the host before any launch; a positive return is the
 * hipError_t of a failed launch.  snapy_last_error() gives the message of the
 * last failure on the calling thread.  No entry point allocates device
 * memory: the caller passes every buffer (state, face fluxes, workspace).
 *
 * Layout of every cell field: (nvar, nc3, nc2, nc1), x1 contiguous, with
 * nc_d = nx_d + 2*nghost when nx_d > 1 and nc_d = 1 otherwise
 * (reference: src/coord/coordinate.hpp:45-47).  A face field of direction d
 * has the same shape; face i is the lower face of cell i
 * (src/recon/reconstruct.cpp:217-229).  Rows are IDN=0, IVX=1, IVY=2, IVZ=3,
 * IPR=4 (src/snap.h:34-45); shallow water uses rows 0..3.
 *
 * The reference seam each entry replaces is cited beside it.  The reference
 * dispatches one ATen DispatchStub per operator; this ABI exposes fused
 * stage-level kernels instead.
 */
#ifndef SNAPY_AMD_H
#define SNAPY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums ------------------------------------------------------------ */
enum snapy_dtype { SNAPY_F64 = 0, SNAPY_F32 = 1 };

/* src/riemann/riemann.hpp RiemannKind (hllc, lmars, shallow-roe subset) */
enum snapy_riemann {
  SNAPY_RIEMANN_HLLC = 0,
  SNAPY_RIEMANN_LMARS = 1,
  SNAPY_RIEMANN_SHALLOW_ROE = 2
};

/* equation-of-state/type (src/eos/equation_of_state.cpp) */
enum snapy_eos { SNAPY_EOS_IDEAL_GAS = 0, SNAPY_EOS_SHALLOW_WATER = 1 };

/* src/bc/bc_func.cpp: boundary functions by name */
enum snapy_bc {
  SNAPY_BC_NONE = 0, /* custom / no-op */
  SNAPY_BC_REFLECTING = 1,
  SNAPY_BC_PERIODIC = 2,
  SNAPY_BC_EXTRAPOLATION = 3,
  SNAPY_BC_OUTFLOW = 4 /* radiating about the initial background */
};

/* ---- descriptors ------------------------------------------------------ */

/* One block's uniform Cartesian grid (src/coord/coordinate.cpp:340-420).
 *
 * members3 > 1 makes x3 a member axis: the tensor holds members3
 * independent blocks of a deck whose x3 is a single cell (nx3 = nc3 =
 * members3, no x3 ghosts), advanced together by every entry point with no
 * coupling across x3 (no x3 flux, boundary or diffusion), each member with
 * its own dt: a device dt (dt_dev) then holds members3 values, and
 * snapy_cfl writes members3 values.  0 or 1: an ordinary block. */
typedef struct snapy_grid {
  int32_t nvar;   /* stored rows: 5 Euler, 4 shallow water */
  int32_t nghost; /* 3 for weno5 */
  int32_t nx1, nx2, nx3;
  int32_t nc1, nc2, nc3;
  double dx1, dx2, dx3; /* cell widths (global extent / global cells) */
  double x1min, x2min, x3min; /* lower edge of the first interior cell */
  int32_t members3; /* independent members along x3 (see above) */
  int32_t pad3;
} snapy_grid;

/* Deck physics the stage kernels read (src/hydro/hydro.hpp options). */
typedef struct snapy_physics {
  double gammad;         /* ideal-gas adiabatic index */
  double grav1;          /* const-gravity along x1 (0 = none) */
  double density_floor;  /* eos density-floor (outflow bounds) */
  double pressure_floor; /* eos pressure-floor (outflow bounds) */
  int32_t riemann;       /* enum snapy_riemann */
  int32_t shock_x1;      /* reconstruct/vertical shock: weno5 on every row */
  int32_t shock_x23;     /* reconstruct/horizontal shock */
  int32_t bc[6];         /* x1-inner, x1-outer, x2-inner, ... x3-outer */
  int32_t implicit_scheme; /* integration/implicit-scheme (0, 1, 9) */
  double advection;      /* cfl / implicit-advection-cfl (CFL of an implicit
                            direction, src/hydro/hydro.cpp:326-370) */
  double bot_heat_flux;  /* forcing/bot-heat flux (W/m^2), 0 = none */
  int32_t bot_heat_depth; /* forcing/bot-heat depth (cells) */
  int32_t flux_covariance; /* dynamics/flux-covariance on the x2/x3 faces */
  double nu_iso;         /* forcing/diffusion kinematic viscosity (m^2/s) */
  double kappa_iso;      /* forcing/diffusion thermal diffusivity (m^2/s) */
  double gas_r;          /* dry gas constant Rgas / weight (J/kg/K) */
  int32_t eos;           /* enum snapy_eos; shallow water: nvar 4, rows
                            (phi, phi v1, phi v2, phi v3), Riemann
                            shallow-roe, dir xy (x1 and x2 sweeps) */
} snapy_physics;

/* ---- entry points ----------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char *snapy_last_error(void);

/* Number of gfx950 kernels compiled into the library (load check). */
int snapy_kernel_count(void);

/* Conserved -> primitive on every cell of the block.
 * Replaces: ideal_gas_cons2prim stub (src/eos/eos_dispatch.hpp:9-16,
 * kernel src/eos/ideal_gas_impl.h:16-35). */
int snapy_cons2prim(const snapy_grid *g, const snapy_physics *p, int dtype,
                    const void *u, void *w, void *stream);

/* Primitive -> conserved on every cell (IdealGasImpl::_prim2cons,
 * src/eos/ideal_gas.cpp:49-65). */
int snapy_prim2cons(const snapy_grid *g, const snapy_physics *p, int dtype,
                    const void *w, void *u, void *stream);

/* Fused face sweep of direction dim (1 = x1, 2 = x2, 3 = x3): the weno5/cp5
 * face states of every interior face straight from w, then the Riemann flux,
 * written into the FluxRegister total `flux` (faces lo..hi+1 along dim,
 * interior cells across it).  face_pressure (x1 only, may be NULL, FP64)
 * receives the solver's face pressure.  w is the FP64 primitive state (the
 * PrecisionPolicy keeps primitives FP64); dtype is the compute and flux
 * precision.  The FP32 LMARS form reconstructs the pressure about the
 * pressure g of each sweep line's first interior cell and stores a
 * gauge-free momentum flux (g cancels in the divergence).  At a reflecting (wall) boundary face the mass
 * and energy rows are written as zero (FluxRegister boundary.prescribed,
 * src/layout/seam_register.cpp:184-220).
 * Replaces: call_recon (src/recon/recon_dispatch.hpp:13-16) twice per row
 * group, and call_riemann (src/riemann/riemann.hpp:165-171) with its EOS face
 * pass (src/riemann/riemann.cpp:219-346). */
int snapy_sweep(int dim, const snapy_grid *g, const snapy_physics *p,
                int dtype, const void *w, void *flux, void *face_pressure,
                void *stream);

/* x2/x3 sweep (dim 2 or 3) with the x1 flux covariance of the energy flux
 * when p->flux_covariance is set (HydroImpl::_flux_covariance, Cartesian,
 * src/hydro/hydro_forward.cpp:58-231, added as add_cov :782-793), else the
 * plain sweep.  x1tab row 0 holds x1v (nc1 doubles).  Interior faces only;
 * wall planes prescribed as in snapy_sweep; w FP64, dtype and the FP32
 * gauge as in snapy_sweep. */
int snapy_sweep_cov(int dim, const snapy_grid *g, const snapy_physics *p,
                    int dtype, const void *w, const double *x1tab, void *flux,
                    void *stream);

/* Fused flux divergence + RK stage combine on the interior cells:
 *   du = -dt * div(F) + du_src      (du_src may be NULL)
 *   u  = w0*u0 + w1*u + w2*du
 * flux[d] may be NULL for an inactive direction.  dt_dev (may be NULL) is a
 * device FP64 dt read by the kernel in place of dt (StepDt on the device,
 * src/utils/step_dt.hpp:17-63), so a captured stage replays with a new dt.
 * Replaces: CoordinateImpl::divergence (src/coord/coordinate.cpp:618-662),
 * the du assembly (src/hydro/hydro_forward.cpp:983-985) and the harp combine
 * pintg->forward(stage, u0, u, du, u) (src/mesh/meshblock_advance.cpp:329). */
int snapy_update(const snapy_grid *g, int dtype, const void *const flux[3],
                 const void *u0, void *u, const void *du_src, double dt,
                 const double *dt_dev, double w0, double w1, double w2,
                 void *stream);

/* FluxRegister step field S_d <- w1*S_d + (w2*dt)*T_d in FP64 over every
 * face of the total (FluxRegister::end_stage, src/hydro/flux_register.cpp:
 * 761-777).  step is FP64; total is dtype; dt_dev as in snapy_update. */
int snapy_step_field(const snapy_grid *g, int dtype, const void *total,
                     double *step, double w1, double w2, double dt,
                     const double *dt_dev, void *stream);

/* CFL reduction: *dt_out (device, FP64) = scale * min over interior cells
 * and active directions of dx_d / (|v_d| + c), and of the kinematic
 * diffusion bound when p->nu_iso or p->kappa_iso is set (HydroImpl::max_time_step,
 * explicit branch, src/hydro/hydro.cpp:372-392; scale = 2^-redo * cfl as in
 * global_time_step, src/mesh/meshblock_advance.cpp:65-72).  dt_out is
 * overwritten, not merged; no host read. */
int snapy_cfl(const snapy_grid *g, const snapy_physics *p, int dtype,
              const void *w, double scale, double *dt_out, void *stream);

/* Physical-boundary ghost fill of the conserved state, faces in the order
 * x1-inner, x1-outer, x2-..., x3-... (MeshBlockImpl::apply_boundaries,
 * src/mesh/meshblock_exchange.cpp:214-287; functions src/bc/bc_func.cpp).
 * bref_w is the initial primitive background for outflow faces (may be NULL
 * without outflow); scratch_w is a primitive-sized workspace (may be NULL
 * without outflow). */
int snapy_apply_boundaries(const snapy_grid *g, const snapy_physics *p,
                           int dtype, void *u, const void *bref_w,
                           void *scratch_w, void *stream);

/* Same ghost fill on a primitive state (apply_boundaries(..., primitive=true),
 * used by initialize). */
int snapy_apply_boundaries_prim(const snapy_grid *g, const snapy_physics *p,
                                int dtype, void *w, const void *bref_w,
                                void *stream);

/* Hydrostatic x1 reference of every column (the well-balanced x1 path that
 * runs whenever grav1 != 0, src/hydro/hydro_forward.cpp:479-546):
 * ref = 5 FP64 planes (psf_lo, psf_hi, pref, dsf, dref), each
 * (nc3, nc2, nc1), from the FP64 primitive w.  One block owns its whole x1
 * column (nb1 = 1), uniform x1, wall clamp on.
 * Replaces: call_hydro_ref_x1 (src/hydro/hydro_dispatch.hpp:26-42; kernel
 * src/hydro/hydro_ref_x1_impl.h:24-272). */
int snapy_hydro_ref_x1(const snapy_grid *g, const snapy_physics *p,
                       const double *w, double *ref, void *stream);

/* Fused well-balanced x1 sweep: perturbation reconstruction about ref with
 * even wall ghosts, the reference face pressure/density restored with their
 * positivity fallbacks, then the Riemann flux into the x1 total, computed
 * in FP64 from the FP64 w.  With dtype FP32 the stored x1 momentum row is
 * the anomaly flux - psf_lo (the FP32 update restores psf_lo).
 * Replaces: WellBalancedX1::forward (src/recon/well_balanced_x1.cpp:16-154),
 * X1Reference::evaluate's face densities (src/recon/x1_reference.cpp:
 * 177-243) and the x1 call_riemann. */
int snapy_sweep_x1_wb(const snapy_grid *g, const snapy_physics *p, int dtype,
                      const void *w, const double *ref, void *flux,
                      void *face_pressure, void *stream);

/* snapy_hydro_ref_x1 and snapy_sweep_x1_wb in one launch, for the interior
 * columns: the reference stays on chip (LDS) between the two, and only its
 * psf_lo plane is written to ref (the other four planes are left as they
 * are; snapy_update_gravity reads psf_lo).  Fluxes and face pressure are
 * the same bits as the two entries' (the same arithmetic).
 * Replaces: the x1 reference evaluation and WellBalancedX1::forward of one
 * stage (src/hydro/hydro_forward.cpp:479-546, src/hydro/hydro_ref_x1_impl.h:
 * 24-272, src/recon/well_balanced_x1.cpp:16-154). */
int snapy_sweep_x1_wb_ref(const snapy_grid *g, const snapy_physics *p, int dtype,
                          const double *w, double *ref, void *flux,
                          void *face_pressure, void *stream);

/* snapy_sweep_x1_wb_ref in the FP32 form (LMARS, dtype FP32, nvar 5) with
 * the stage increment fused: each interior cell also forms its du, FP64, as
 * snapy_update_gravity_f32 with du_out forms it from the stored fluxes (the
 * x1 rows of this launch, the x2/x3 totals flux2/flux3, NULL when
 * inactive, the x1 table; no diffusion source).  flux2/flux3 must hold
 * this stage's x2/x3 sweeps.  flux1 and face_pressure (may be NULL) are
 * written as snapy_sweep_x1_wb_ref writes them, du_out the same bits as
 * snapy_update_gravity_f32's.
 * Replaces: the x1 sweep (as snapy_sweep_x1_wb_ref) followed by the
 * divergence, const gravity, bot-heat and gravity work of the stage
 * (src/hydro/hydro_forward.cpp:479-546, 1049-1124;
 * src/forcing/const_gravity.cpp:35-43; src/forcing/layer_heat.cpp:78-85). */
int snapy_sweep_x1_wb_inc_f32(const snapy_grid *g, const snapy_physics *p,
                              const double *w, double *ref, float *flux1,
                              double *face_pressure, const float *flux2,
                              const float *flux3, const double *x1tab,
                              double *du_out, double dt, const double *dt_dev,
                              void *stream);

/* snapy_update with x1 gravity: du also receives the const-gravity source
 * (src/forcing/const_gravity.cpp:30-55) and the face-based gravity work plus
 * the corrected potential-energy work in place of the cell-centred work
 * (src/hydro/hydro_forward.cpp:1049-1124, src/hydro/gravity_work.hpp).
 * x1tab is an FP64 device table of 8 rows of nc1: x1v, the face-work
 * factors grav1*(x1f[i+1]-x1v[i]) and grav1*(x1f[i]-x1v[i]), grav1*dx1^2/12,
 * the three centroid-slope weights (interior cells), and a spare row.
 * With u0 NULL the kernel writes the stage increment du into u instead of
 * combining it (the input of snapy_vic_full). */
int snapy_update_gravity(const snapy_grid *g, const snapy_physics *p,
                         int dtype, const void *const flux[3], const void *u0,
                         void *u, const void *w, const double *x1tab,
                         const void *du_src, double dt, const double *dt_dev,
                         double w0, double w1, double w2, void *stream);

/* Vertical implicit correction, full 5x5 matrix (implicit-scheme 9), FP64,
 * one column per 8-lane group (lane p owns row p of every 5x5 block; the
 * blocks are eliminated by Gauss-Jordan with scaled partial pivoting across
 * the row lanes), then the RK combine u = w0*u0 + w1*u + w2*du'
 * with du' the corrected increment (interior cells).  du is the explicit
 * stage increment (snapy_update_gravity with u0 NULL).  The solve runs with
 * dt_c = wstage * dt (the stage weight inside the solve,
 * src/hydro/hydro_forward.cpp:1137-1175).  vtab: FP64 rows of nc1 (face
 * work weights lo/hi, corrected-PE bands lo/mid/hi, cell potential,
 * face-minus-cell potentials lo/hi); t0, t1 the wall third points; x1tab as
 * for snapy_update_gravity.  ws: snapy_vic_workspace(g) doubles; status:
 * nx2*nx3 ints, one word per column: bit 0 is set when this call's solve of
 * the column failed (singular pivot or a non-finite solution; the column
 * keeps its explicit increment) and cleared when it succeeded; bit 1 is set
 * with bit 0 and kept by later calls, a step-level latch the caller clears
 * at the step's first stage (the reference ORs every stage's bits into the
 * step status, src/implicit/implicit_hydro.cpp:610-617, and resets it at
 * step start, src/mesh/meshblock_advance.cpp:186-188).  moved: NULL, or an
 * FP64 (nc3, nc2, nc1) array that receives at each interior cell the
 * clamped dry-gas transfer through the face below it (0 at a column's
 * bottom cell and in a failed column), the mass_fix row the passive
 * scalars' implicit transfer reads (src/implicit/vic_redistribute_impl.h:
 * 80-90, src/mesh/meshblock_advance.cpp:215-233).
 * Replaces: vic_stage assemble/solve/redistribute (src/implicit/
 * implicit_dispatch.hpp:50-57, src/implicit/vic_kernels.h:66-297) and
 * ImplicitHydroImpl::forward_masked (src/implicit/implicit_hydro.cpp:
 * 301-623), radial-exact coupling with Woodbury, no solids.  A grid of
 * at most 1536 columns and 144 layers is solved by cyclic reduction (a
 * workgroup per column), a larger one by the five-lane forward sweep: the
 * same system, solutions apart by rounding (about 1e-13 of the increment). */
int64_t snapy_vic_workspace(const snapy_grid *g);
int snapy_vic_full(const snapy_grid *g, const snapy_physics *p, const double *w,
                   const double *du, const double *u0, double *u,
                   const double *vtab, const double *x1tab, double t0,
                   double t1, double dt, const double *dt_dev, double wstage,
                   double w0, double w1, double w2, double *ws, int *status,
                   double *moved, void *stream);
/* The same solve with one column per thread and textbook Gaussian
 * elimination with scaled partial pivoting on each block: the same
 * arguments and outputs as snapy_vic_full, kept as the independent
 * cross-check of the lane-group kernel (no counterpart of its own in the
 * reference; it replaces the same vic_stage seam). */
int snapy_vic_full_serial(const snapy_grid *g, const snapy_physics *p,
                          const double *w, const double *du, const double *u0,
                          double *u, const double *vtab, const double *x1tab,
                          double t0, double t1, double dt,
                          const double *dt_dev, double wstage, double w0,
                          double w1, double w2, double *ws, int *status,
                          double *moved, void *stream);

/* Isotropic Cartesian diffusion, kinematic form (forcing/diffusion with
 * nu_iso, kappa_iso; no x1 profiles; stress-free walls).  w is the FP64
 * primitive state with ghosts filled.  For every active direction d the
 * viscous momentum flux, its work and the Fourier flux -kappa (rho cv) dT/dn
 * are formed on the faces (a reflecting x1 wall face extrapolates the face
 * density from the two nearest active cells); the energy row is added to
 * totals[d] (dtype) on every face that is not a reflecting wall, and the
 * momentum rows become du_src = -dt div(F) (rows IDN and IPR zero) on the
 * interior cells, FP64, the du_src of snapy_update / snapy_update_gravity.
 * work: snapy_diffusion_workspace(g) FP64 doubles.
 * Replaces: DiffusionImpl::cartesian_fluxes + post_fluxes + add
 * (src/forcing/diffusion.cpp:532-706, diffusion_cubed_sphere.cpp:365-379). */
int64_t snapy_diffusion_workspace(const snapy_grid *g);
int snapy_diffusion(const snapy_grid *g, const snapy_physics *p, int dtype,
                    const double *w, void *const totals[3], double *work,
                    double *du_src, double dt, const double *dt_dev,
                    void *stream);

/* ---- the FP32 state form ----------------------------------------------
 * Under dynamics/precision float32 the conserved state is an FP32 pair
 * (u, c): the state is double(u) + double(c), c the compensation that keeps
 * the RK3 accumulation exact to FP64 rounding (Kahan-compensated storage).
 * The RK backup is the pair (u0, c0).  Primitives, the x1 reference, the
 * implicit solve and the budgets stay FP64; face fluxes are FP32.  Replaces:
 * PrecisionPolicy's float32 rows (src/mesh/precision_policy.hpp:31-71) with
 * the compensated accumulation of the GPU form. */

/* w (FP64) = prim(u + c) on every cell (ghosts included). */
int snapy_cons2prim_f32(const snapy_grid *g, const snapy_physics *p,
                        const float *u, const float *c, double *w,
                        void *stream);
/* (u, c) = the FP32 pair of cons(w) on every cell. */
int snapy_prim2cons_f32(const snapy_grid *g, const snapy_physics *p,
                        const double *w, float *u, float *c, void *stream);
/* snapy_update on the pair: FP32 fluxes, FP64 divergence, du_src FP64 (may
 * be NULL), (u, c) <- w0 (u0 + c0) + w1 (u + c) + w2 du in FP64. */
int snapy_update_f32(const snapy_grid *g, const void *const flux[3],
                     const float *u0, const float *c0, float *u, float *c,
                     const double *du_src, double dt, const double *dt_dev,
                     double w0, double w1, double w2, void *stream);
/* snapy_update_gravity on the pair: FP32 fluxes whose x1 momentum row is
 * the anomaly about the reference face pressure in ref (snapy_hydro_ref_x1);
 * with du_out non-NULL the FP64 stage increment is written there instead of
 * the combine (the input of snapy_vic_full_f32). */
int snapy_update_gravity_f32(const snapy_grid *g, const snapy_physics *p,
                             const void *const flux[3], const double *ref,
                             const float *u0, const float *c0, float *u,
                             float *c, double *du_out, const double *w,
                             const double *x1tab, const double *du_src,
                             double dt, const double *dt_dev, double w0,
                             double w1, double w2, void *stream);
/* The vertical implicit correction with the partial matrix
 * (implicit-scheme 1): 3x3 blocks in (rho, m1, E) with the tangential
 * momenta folded in as dm_t = v_t drho + q_t, the corrected-PE coupling and
 * its Woodbury wall points, the mass redistribution with the tangential
 * momentum each moved mass carries (the donor cell's velocity), the
 * post-solve remainder and the RK combine; arguments as snapy_vic_full.
 * One thread per column assembles, sweeps, solves the Woodbury 2x2 and
 * back-substitutes (records of the factored rows and the partial solutions
 * in ws, snapy_vic_partial_workspace(g) doubles); a few columns per
 * workgroup then redistribute (the serial chains one thread per column)
 * and combine.
 * Replaces: vic_stage for VicMatrix::partial (vic_kernels.h:66-297 with
 * N = 3; vic_assemble_impl.h:45-218; forward_sweep_impl.h:104-121;
 * vic_redistribute_impl.h:177-205) and ImplicitHydroImpl::forward_masked
 * (implicit_hydro.cpp:301-623). */
int64_t snapy_vic_partial_workspace(const snapy_grid *g);
int snapy_vic_partial(const snapy_grid *g, const snapy_physics *p,
                      const double *w, const double *du, const double *u0,
                      double *u, const double *vtab, const double *x1tab,
                      double t0, double t1, double dt, const double *dt_dev,
                      double wstage, double w0, double w1, double w2,
                      double *ws, int *status, double *moved, void *stream);
/* snapy_vic_partial finishing into the FP32 pair (u, c) from (u0, c0). */
int snapy_vic_partial_f32(const snapy_grid *g, const snapy_physics *p,
                          const double *w, const double *du, const float *u0,
                          const float *c0, float *u, float *c,
                          const double *vtab, const double *x1tab, double t0,
                          double t1, double dt, const double *dt_dev,
                          double wstage, double w0, double w1, double w2,
                          double *ws, int *status, double *moved,
                          void *stream);

/* snapy_vic_partial_f32 that also writes the primitives of the combined
 * pair into the interior of w (w is read first, as in
 * snapy_vic_partial_f32; nvar 5): the values snapy_cons2prim_f32 would form
 * from (u, c) there, so a caller's next stage needs only the ghost fill on
 * the primitives.
 * Replaces: the same solve, followed by the next stage's
 * peos->forward(u, w) at the start of HydroImpl::forward
 * (src/hydro/hydro_forward.cpp:340) on the interior. */
int snapy_vic_partial_prim_f32(const snapy_grid *g, const snapy_physics *p,
                               double *w, const double *du, const float *u0,
                               const float *c0, float *u, float *c,
                               const double *vtab, const double *x1tab,
                               double t0, double t1, double dt,
                               const double *dt_dev, double wstage, double w0,
                               double w1, double w2, double *ws, int *status,
                               double *moved, void *stream);

/* snapy_vic_full finishing into the pair (u, c) from (u0, c0). */
int snapy_vic_full_f32(const snapy_grid *g, const snapy_physics *p,
                       const double *w, const double *du, const float *u0,
                       const float *c0, float *u, float *c,
                       const double *vtab, const double *x1tab, double t0,
                       double t1, double dt, const double *dt_dev,
                       double wstage, double w0, double w1, double w2,
                       double *ws, int *status, double *moved,
                       void *stream);

/* ---- passive scalars (tracers) ----------------------------------------
 * ns tracer rows s = rho r (FP64 in both state forms), (ns, nc3, nc2, nc1)
 * like the hydro rows; the stage's dry density is the primitive density
 * w[IDN].  Dry ideal gas, no member axis.  Per stage, after the hydro
 * sweeps: snapy_scalar_flux per active direction, snapy_scalar_theta when
 * the covariance is on, then (after the hydro update, whose implicit solve
 * writes the dry-gas transfer) snapy_scalar_update and, after the hydro
 * boundaries, snapy_scalar_prim. */

/* The face flux of direction dir (1, 2, 3) at the faces of the interior
 * cells (ScalarImpl::forward, src/scalar/scalar.cpp:104-174): the scaled
 * WENO5 faces of rho and s clamped at 0 (recon_point.h:69-147,
 * reconstruct.cpp:262-266), r = s_face / rho_face or the donor cell's r,
 * the upwind flux on row IDN of hflux (the hydro's stage flux total of
 * dir, dtype fdtype; riemann.cpp:164-171), on x2/x3 with
 * p->flux_covariance the covariance row sigma1^2 rho D1[u_n] D1[q]
 * (hydro_forward.cpp:233-297; x1tab row 0 = x1v), and 0 on reflecting
 * faces.  sflux: FP64, the face below each cell index. */
int snapy_scalar_flux(int dir, const snapy_grid *g, const snapy_physics *p,
                      int ns, const double *w, const double *s, int fdtype,
                      const void *hflux, const double *x1tab, double *sflux,
                      void *stream);
/* The scalars' positivity limiter factors (flux_positivity_theta, _share,
 * _relax, src/hydro/flux_positivity.cpp:39-149, as scalar.cpp:193-205
 * applies them): theta of every row on the interior cells relaxed by the
 * dry air's theta of the mass fluxes, 1 on the ghosts, then the ghosts
 * filled as a scalar field. */
int snapy_scalar_theta(const snapy_grid *g, const snapy_physics *p, int ns,
                       const double *w, const double *s, int fdtype,
                       const void *const hflux[3], const double *const sflux[3],
                       double dt, const double *dt_dev, double *theta,
                       void *stream);
/* ds = -dt div(F) on the interior, each face flux scaled by its donor's
 * theta (NULL: none); with moved (the dry-gas transfer of the implicit
 * solve, snapy_vic_*) the transfer P = r_donor M capped up each column by
 * what the donors hold (meshblock_advance.cpp:215-233, donor_cap_cuda,
 * implicit_dispatch.cu:69-87; r the stage's scalar primitives, ds an FP64
 * workspace like s); then s = w0 s0 + w1 s + w2 ds and the ghosts of s. */
int snapy_scalar_update(const snapy_grid *g, const snapy_physics *p, int ns,
                        const double *const sflux[3], const double *theta,
                        const double *r, const double *moved, const double *s0,
                        double *s, double *ds, double dt, const double *dt_dev,
                        double w0, double w1, double w2, void *stream);
/* r = s / rho of the state u (dtype udtype; comp the FP32 pair's
 * compensation or NULL) on the interior, then the ghosts of r
 * (set_scalar_primitive, src/mesh/meshblock.cpp:67-79). */
int snapy_scalar_prim(const snapy_grid *g, const snapy_physics *p, int ns,
                      int udtype, const void *u, const float *comp,
                      const double *s, double *r, void *stream);
/* The boundary functions on a kScalar field of ns FP64 rows: reflecting
 * mirrors without a sign, periodic, extrapolation (bc_func.cpp:10-65); a
 * radiating face is refused. */
int snapy_scalar_boundaries(const snapy_grid *g, const snapy_physics *p,
                            int ns, double *q, void *stream);

/* FP64 budget: out[v] = sum over interior cells of u[v] * volume, for
 * v < nvar, deterministic (fixed reduction tree).  out is device FP64 of
 * length nvar; ws is FP64 device workspace of at least
 * snapy_budget_workspace(g) doubles.
 * Replaces: BudgetConsumer domain totals (src/diagnostics/budget.cpp:267). */
int64_t snapy_budget_workspace(const snapy_grid *g);
int snapy_budget(const snapy_grid *g, int dtype, const void *u, double *ws,
                 double *out, void *stream);
/* snapy_budget without a workspace argument and without any allocation: the
 * partial sums live in a static device array of the library (40 KiB per GPU,
 * held until the process ends), whose address on the current device is
 * looked up once per device.  The same kernels, so the same bits as
 * snapy_budget.  The array holds 5 rows: a larger nvar returns an error
 * before any launch.  Calls on one GPU share that array: calls that overlap in
 * time on one GPU (different streams or threads) race and must be
 * serialised by the caller.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc). */
int snapy_budget_noalloc(const snapy_grid *g, int dtype, const void *u,
                         double *out, void *stream);

/* ---- grouped launches: many blocks of different grids, one launch -------
 *
 * Replaces the reference's worker threads, one per block of a process, each
 * launching its block's kernels on its own stream (src/mesh/mesh.cpp:
 * 74-110): here the blocks that share a deck share each launch.  Between
 * snapy_record_begin and snapy_record_end the calling thread's entry points
 * make no launch: they record it (its kernel, arguments and grid) in a
 * plan.  Plans of blocks that made the same launches in the same order (one
 * deck on grids that differ only where the launches allow: the x2/x3 extents
 * and spacings, with the same x1 cells and launch shapes) become a group,
 * whose launch makes each recorded launch once over all the blocks'
 * workgroups; every workgroup runs its own block's arguments, so the group
 * step is bit-for-bit the blocks' own steps.  A plan and a group keep the
 * buffer addresses and scalars they were recorded with (a stage's dt is read
 * from each block's device dt).  Errors: -20 recording state, -21 plans that
 * do not match, -22 no device memory. */
int snapy_record_begin(void);
int snapy_record_end(void **plan);
/* the launches a plan holds */
int snapy_plan_launches(const void *plan);
void snapy_plan_free(void *plan);
/* one group of nplans plans (its argument records in device memory) */
int snapy_group_build(void *const *plans, int nplans, void **group);
/* the group's launches, in the recorded order, on the stream */
int snapy_group_launch(const void *group, void *stream);
int snapy_group_launches(const void *group);
/* frees the group (after the work launched from it has finished) */
void snapy_group_free(void *group);

/* ---- launch-shape tuning --------------------------------------------------
 *
 * The workgroup size of a kernel family for blocks of one shape (nx1, nx2,
 * nx3) and dtype, the columns per workgroup of the column kernels, or the
 * persistent grid of the work list (any shape): set by the project's
 * autotuner, read at every launch.  A tuned shape only changes how items are
 * grouped into workgroups, so results are bit-identical for every value.
 * Workgroup sizes: 64, 128 or 256 (default 256); columns per workgroup: 1,
 * 2, 4 or 8 for the scheme-1 finish (default 8), 1, 2 or 4 for the fused x1
 * reference sweep (default 4), fewer where a tall column's LDS needs it; the
 * work list's grid: 64 .. 65536 workgroups (default 2048).  value 0 restores
 * the default.  No counterpart in the reference (its CUDA launches use fixed
 * shapes). */
enum snapy_tunable {
  SNAPY_TUNE_EOS = 0,     /* cons2prim / prim2cons (FP64 and the FP32 pair) */
  SNAPY_TUNE_SWEEP = 1,   /* snapy_sweep (x1, x2, x3) and snapy_sweep_x1_wb */
  SNAPY_TUNE_UPDATE = 2,  /* snapy_update(_f32), snapy_update_gravity(_f32) */
  SNAPY_TUNE_WL_GRID = 3, /* the work list's persistent grid */
  SNAPY_TUNE_VIC_COLS = 4, /* columns per workgroup, scheme-1 finish */
  SNAPY_TUNE_REF_COLS = 5, /* columns per workgroup, snapy_sweep_x1_wb_ref */
  SNAPY_TUNE_COUNT = 6
};
int snapy_tune_set(int tunable, const snapy_grid *g, int dtype, int value);
int snapy_tune_get(int tunable, const snapy_grid *g, int dtype);

/* ---- the work list: many x1 blocks of uneven size per launch ------------
 *
 * A batch of independent 1D (x1) ideal-gas blocks without sources (the
 * serving mix's shock tubes), FP64.  Each block keeps its own buffers and its
 * own device dt; every phase of the RK stage is one persistent launch over
 * the concatenated items of all blocks (cells, faces, interior cells, ghost
 * cells), a grid of at most 2048 workgroups striding over them.  The
 * per-item arithmetic is the single-block kernels', so a block's state is
 * bit-identical alone and in any batch.  The descriptor array and the prefix
 * counts live in device memory; snapy_wl_counts fills the prefix counts on
 * the host.
 * Replaces: the per-block launches of MeshBlockImpl::forward
 * (meshblock_advance.cpp:88-105) for many blocks, as the reference's Mesh
 * runs its blocks one after another (mesh.cpp:74-110). */
typedef struct snapy_wl_block {
  int32_t nx1;          /* interior cells; nc1 = nx1 + 2 nghost */
  int32_t bc_lo, bc_hi; /* SNAPY_BC_* of the two x1 faces */
  int32_t pad;
  double dx1, dx2, dx3; /* cell widths (x2, x3: the block's unit extents) */
  double scale;         /* CFL factor 2^-redo * cfl */
  double *u, *u0, *w;   /* (5, nc1) FP64: state, RK backup, primitives */
  double *flux;         /* (5, nc1) FP64: the x1 face totals */
  const double *bref;   /* outflow background w (5, nc1), or NULL */
  double *dt;           /* the block's device dt */
} snapy_wl_block;

typedef struct snapy_work_list {
  const snapy_wl_block *blocks; /* device array of nblocks descriptors */
  const int32_t *prefix;        /* device: 4 x (nblocks + 1) prefix counts of
                                   cells, faces, interior cells, ghost cells */
  int32_t nblocks;
  int32_t nghost;               /* 3 */
  int64_t total[4];             /* the four item totals */
} snapy_work_list;

/* Host: the prefix counts (4 x (nblocks + 1)) and totals of a batch. */
int snapy_wl_counts(const snapy_wl_block *blocks, int32_t nblocks,
                    int32_t nghost, int32_t *prefix, int64_t *total);
/* u0 <- u for every block (StateStore::backup at stage 0). */
int snapy_wl_backup(const snapy_work_list *wl, const snapy_physics *p,
                    void *stream);
/* w <- cons2prim(u) over every cell of every block. */
int snapy_wl_cons2prim(const snapy_work_list *wl, const snapy_physics *p,
                       void *stream);
/* The x1 face totals of every block (weno5/cp5 + HLLC/LMARS, walls).
 * Replaces: call_recon x2 + call_riemann on x1 (sweep.cpp:105-157). */
int snapy_wl_sweep_x1(const snapy_work_list *wl, const snapy_physics *p,
                      void *stream);
/* u <- w0 u0 + w1 u + w2 (-dt div F) on the interior cells, each block with
 * its own dt.  Replaces: divergence + harp's combine for each block. */
int snapy_wl_update(const snapy_work_list *wl, const snapy_physics *p,
                    double w0, double w1, double w2, void *stream);
/* The x1 ghost cells of u (outflow, reflecting, periodic, extrapolation).
 * Replaces: apply_boundaries on x1 (bc_func.cpp:10-227). */
int snapy_wl_boundaries(const snapy_work_list *wl, const snapy_physics *p,
                        void *stream);
/* Each block's dt = scale * min dx1 / (|v1| + c), from w.
 * Replaces: HydroImpl::max_time_step + global_time_step per block. */
int snapy_wl_cfl(const snapy_work_list *wl, const snapy_physics *p,
                 void *stream);
/* One RK stage of every block: backup (stage 0), cons2prim, sweep, update,
 * boundaries. */
int snapy_wl_stage(const snapy_work_list *wl, const snapy_physics *p,
                   int stage, double w0, double w1, double w2, void *stream);

/* ---- the cubed-sphere shallow-water stage -------------------------------
 *
 * Six gnomonic-equiangle panels per member (the reference's cubed-sphere
 * layout with one block per panel, all on one device), shallow water with
 * shallow-roe in the y-z plane.  A state tensor holds members * 6 blocks of
 * (4, nc, nc, 1) FP64, x2 fastest, block b = member * 6 + panel, nc = n + 6.
 * Faces of axis d are stored at the index of the cell above them (as the
 * flux registers).  The descriptor's tables are grid-only, built once on the
 * host (snapy_amd/cubed.py) and kept on the device:
 *   geom         7 planes (nc x nc): cos_c, sin_c, volume, x_ov_rD, y_ov_rC,
 *                width2, width3; then nc x (nc + 1): cos_f2, sin_f2, area2;
 *                then (nc + 1) x nc: cos_f3, sin_f3, area3
 *                (snapy_cs_geom_size doubles);
 *   ghost_*      per ghost entry (4 sides x 3 layers x n per panel, plus the
 *                corner cells): int32 [dst panel, dst cell, src panel,
 *                src cells 0..3, npts], FP64 weights [3], FP64 matrices
 *                [5][3][3] (the sender's frame matrix at each source cell,
 *                then the receiver's), for conserved (_c) and primitive (_p)
 *                variables;
 *   seam_*       per seam face (12 n): int32 [owner panel, side, index,
 *                other panel, side, index, sign, 0], FP64 [30]: the other
 *                panel's to-geographic matrix, the owner's from-geographic
 *                matrix, then (x, y, delta, C, D, face cosine) at the owner's
 *                face and at the other panel's. */
typedef struct snapy_cs_mesh {
  int32_t n;            /* cells along a panel edge (nx2 = nx3) */
  int32_t nghost;       /* 3 */
  int32_t members;      /* independent six-panel meshes in the tensor */
  int32_t interp_order; /* geometry/cells/interp_order: 2 or 4 */
  int32_t shock;        /* reconstruct/horizontal shock */
  int32_t pad;
  double radius;        /* x1v of the shell (the x1 mass source) */
  const double *geom;
  const int32_t *ghost_idx;
  const double *ghost_w;
  const double *ghost_mat_c;
  const double *ghost_mat_p;
  int64_t nghost_entries;
  const int32_t *seam_idx;
  const double *seam_d;
} snapy_cs_mesh;

/* Doubles of the packed grid tables for n cells per edge. */
int64_t snapy_cs_geom_size(int32_t n, int32_t nghost);
/* Doubles of the edge-state buffer the sweeps fill and the seams read. */
int64_t snapy_cs_edge_size(const snapy_cs_mesh *m);

/* u -> w over every cell: v = m / phi, raised with the cell cosine.
 * Replaces: ShallowWaterImpl::_cons2prim (src/eos/shallow_water.cpp:31-41). */
int snapy_cs_cons2prim(const snapy_cs_mesh *m, const double *u, double *w,
                       void *stream);
/* w -> u: m = phi v, lowered with the cell cosine.
 * Replaces: ShallowWaterImpl::_prim2cons + _mom_from_vel_
 * (shallow_water.cpp:43-55, equation_of_state.cpp:290-295). */
int snapy_cs_prim2cons(const snapy_cs_mesh *m, const double *w, double *u,
                       void *stream);
/* The panel-edge ghost fill of q (conserved != 0: hydro_u, else hydro_w),
 * in place: each edge ghost from the neighbour panel's interior,
 * interpolated along the edge in the geographic frame; corner cells copy the
 * nearest interior cell.  Bit-exact with the host restatement.
 * Replaces: MeshBlockImpl::exchange_state over the cubed-sphere layout
 * (meshblock_exchange.cpp:103-117; cubed_sphere_layout.cpp:563-800;
 * gnomonic_equiangle.cpp:286-305,537-605). */
int snapy_cs_ghosts(const snapy_cs_mesh *m, int conserved, double *q,
                    void *stream);
/* The x2 (dim 2) or x3 (dim 3) sweep of every panel: weno5/cp5 face states,
 * the face-local frame, shallow-roe (dir yz), covariant global rows into
 * flux; the interior-side states at the panel's edge faces into edge.
 * Replaces: reconstruct + riemann_solve + to_face_frame_ / to_global_frame_
 * on a gnomonic grid (sweep.cpp:105-157, riemann.cpp:16-51,219-346). */
int snapy_cs_sweep(const snapy_cs_mesh *m, int dim, const double *w,
                   double *flux, double *edge, void *stream);
/* The twelve panel seams: the owner's face flux from the two panels' face
 * states (the face-state sync), and the other panel's copy through the
 * global Cartesian frame, written into flux2 / flux3.
 * Replaces: the cross-panel face-state exchange (hydro_forward.cpp:707-757)
 * and reconcile_panel_seams (seam_register.cpp:35-186). */
int snapy_cs_seams(const snapy_cs_mesh *m, double *flux2, double *flux3,
                   const double *edge, void *stream);
/* Divergence with the gnomonic areas and volume, the shallow-water metric
 * sources and the x1 mass source, du = -dt div, and the RK combine
 * u = w0 u0 + w1 u + w2 du on the interior cells.  dt_dev (per member) if
 * non-null, else dt.
 * Replaces: CoordinateImpl::divergence + GnomonicEquiangleImpl::forward
 * (coordinate.cpp:618-662, gnomonic_equiangle.cpp:472-535) + harp's
 * combine. */
int snapy_cs_update(const snapy_cs_mesh *m, const double *u0, double *u,
                    const double *w, const double *flux2, const double *flux3,
                    double dt, const double *dt_dev, double w0, double w1,
                    double w2, void *stream);
/* dt_out[member] = scale * min over the member's six panels of
 * width_d / (|v_d| + sqrt(phi)), d = x2, x3.
 * Replaces: HydroImpl::max_time_step (hydro.cpp:372-397) on each block and
 * the mesh minimum. */
int snapy_cs_cfl(const snapy_cs_mesh *m, const double *w, double scale,
                 double *dt_out, void *stream);

/* ---- the FP32 form of the cubed-sphere stage ----------------------------
 *
 * The conserved state is the FP32 pair (u, c), whose value is
 * double(u) + double(c) (the compensated RK accumulation); the primitives,
 * the face fluxes and the edge states are FP32 and the sweeps compute in
 * FP32.  The EOS, the ghost gather, the seam fluxes, the divergence, the
 * sources and the RK combine compute in FP64 from the FP32 operands and
 * round the result into the pair, so mass is conserved to FP64 rounding.
 * Each entry replaces the same reference seam as its FP64 counterpart
 * above (the reference has no FP32 form of the shallow-water or gnomonic
 * path, src/hydro/hydro.cpp:120-139). */
int snapy_cs_cons2prim_f32(const snapy_cs_mesh *m, const float *u,
                           const float *c, float *w, void *stream);
/* FP64 primitives w -> the pair (u, c). */
int snapy_cs_prim2cons_f32(const snapy_cs_mesh *m, const double *w, float *u,
                           float *c, void *stream);
/* The panel-edge ghost fill of the conserved pair, gathered and
 * interpolated in FP64. */
int snapy_cs_ghosts_f32(const snapy_cs_mesh *m, float *u, float *c,
                        void *stream);
int snapy_cs_sweep_f32(const snapy_cs_mesh *m, int dim, const float *w,
                       float *flux, float *edge, void *stream);
int snapy_cs_seams_f32(const snapy_cs_mesh *m, float *flux2, float *flux3,
                       const float *edge, void *stream);
int snapy_cs_update_f32(const snapy_cs_mesh *m, const float *u0,
                        const float *c0, float *u, float *c, const float *w,
                        const float *flux2, const float *flux3, double dt,
                        const double *dt_dev, double w0, double w1, double w2,
                        void *stream);
int snapy_cs_cfl_f32(const snapy_cs_mesh *m, const float *w, double scale,
                     double *dt_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SNAPY_AMD_H */
