/*
 * cueva.h -- C ABI of the MI355X (gfx950) implementation of one CUEVA 2D
 * resistive relativistic MHD time step.
 *
 * The reference drives one step through zero-argument Fortran subroutines
 * that share module globals (main_cueva_code.f95:47-356).  Each entry point
 * below replaces one of those subroutines; the state they share lives on the
 * GPU inside a cueva_ctx.  A Fortran host binds them with iso_c_binding
 * (mhd-kh-instability_amd/fortran/cueva_gpu.f95); Python binds them with
 * ctypes (mhd-kh-instability_amd/cueva/abi.py).
 *
 * Conventions
 *  - Every function returns int: CUEVA_OK (0) or a CUEVA_E_* status.
 *  - Stage and substep indices are 1-based, as in the reference (l, m).
 *  - Host field arrays use the reference layout (-6:imax+6, -6:jmax+6), x
 *    fastest, ld = imax+13 doubles between rows; a set of 14 "prims" planes
 *    is in the restart order psi phi Bx By Bz Vx Vy Vz Ex Ey Ez q rho p
 *    (22_writer.f95:1116-1172).
 *  - Calls on one ctx are ordered on the ctx's HIP stream; only
 *    cueva_step, cueva_c2p, cueva_download*, the counter, record and error
 *    queries, cueva_reduce_globals and cueva_sync wait for the device.  A
 *    ctx is not thread safe.
 *  - Failures: where the reference prints and stops, the device records the
 *    first failure in the reference's own order (the launch that failed
 *    first; in it, the first cell of the reference's loop, j outer, i inner)
 *    with its step, stage and cell, and every later launch of the step path
 *    leaves the data alone.  The status is sticky: cueva_step and cueva_c2p
 *    return it from then on, until cueva_upload_prims loads a new state.
 *    It surfaces at the next call that waits (cueva_c2p or cueva_step at the
 *    latest, for a face failure of the per-subroutine sequence).
 */
#ifndef CUEVA_H
#define CUEVA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUEVA_ABI_VERSION 1

enum cueva_status {
  CUEVA_OK = 0,
  CUEVA_E_ARG = 1,           /* invalid argument or configuration           */
  CUEVA_E_HIP = 2,           /* HIP runtime error                           */
  CUEVA_E_NOMEM = 3,         /* device allocation failed                    */
  CUEVA_E_HLLC_VEL = 7,      /* get_hllc_vel did not converge / NaN contact  */
  CUEVA_E_HLLC_STAR = 8,     /* "Revise Vxstar velocity" (NaN contact speed) */
  CUEVA_E_HLLE_NAN = 9,      /* HLLE wave speed is NaN                      */
  CUEVA_E_C2P_OLDV = 20,     /* not returned: the reference clamps a superluminal
                                previous velocity (12_varprimitive.f95:236-241) */
  CUEVA_E_C2P_IMAG = 21,     /* no admissible resolvent root                 */
  CUEVA_E_C2P_DET = 22,      /* det < 0 in the Cardano quadratic             */
  CUEVA_E_C2P_NOROOT = 23,   /* both quartic roots below 1                   */
  CUEVA_E_C2P_NAN = 24,      /* no root selected (NaN quartic roots) with
                                finite quartic coefficients; the reference does
                                not stop here but goes on with the previous
                                cell's W (12_varprimitive.f95:463-500).  With a
                                NaN coefficient the point goes on to NaN
                                results, as in the reference                  */
  CUEVA_E_C2P_SUPERLUM = 25, /* superluminal v and magnetisation < 10        */
  CUEVA_E_C2P_DRIFT = 26     /* drift-velocity fix still superluminal        */
};

/* Kinds of fix-up records (the reference's files on units 301, 302, 303). */
enum { CUEVA_FIX_RHO = 0, CUEVA_FIX_P = 1, CUEVA_FIX_DRIFT = 2 };

/* Boundary map per axis (the BOUND modes of 04_boundary.f95 are pairs of these:
 * BOUND 1 = copy/copy, BOUND 2 = periodic/periodic, BOUND 6 = copy x / periodic y;
 * the Kelvin-Helmholtz setup uses periodic x / copy y). */
enum { CUEVA_BC_COPY = 0, CUEVA_BC_PERIODIC = 1 };

/* Riemann solver + reconstruction (parameters.f95:21): 5/6/7 = HLL/HLLE/HLLC
 * with MPx reconstruction; 2 = HLL with the second-order monotonised-central
 * slope limiter (16_hll_flow_mcl.f95); 4 = HLLC with the same limiter
 * (16_hllc_flow_mcl.f95); 1 = Lax-Friedrichs flux splitting with the same
 * limiter (15_lax_flow.f95). */
enum { CUEVA_LAX = 1, CUEVA_HLL_MCL = 2, CUEVA_HLLE_MCL = 3, CUEVA_HLLC_MCL = 4, CUEVA_HLL = 5, CUEVA_HLLE = 6, CUEVA_HLLC = 7 };

/* Arithmetic of the fluxes: plain FP64, or FP32 reconstruction/Riemann/flux
 * difference with FP64 primitive recovery, E solve and sums. */
enum { CUEVA_FP64 = 0, CUEVA_MIXED = 1 };

/* Primitive recovery: Cardano quartic (12_varprimitive.f95), or a learned
 * initial guess refined by <= 2 FP64 Newton steps with Cardano fallback. */
enum { CUEVA_C2P_CARDANO = 0, CUEVA_C2P_LEARNED = 1 };

/* Ghost-fill targets of cueva_boundary (04_boundary.f95, 24_boundary_conserved.f95,
 * 25_boundary_electric.f95, 24_boundary_flux_conserved.f95). */
enum { CUEVA_BND_PRIM = 0, CUEVA_BND_CONS = 1, CUEVA_BND_ELEC = 2, CUEVA_BND_FLUX = 3 };

/* Plain-old-data mirror of the compile-time parameters the path reads
 * (parameters.f95:10-334).  Fill with cueva_params_default, then override. */
typedef struct cueva_params {
  int32_t imax, jmax;     /* cells per axis (parameters.f95:110-111)           */
  int32_t order;          /* 1 = first-order scheme (MIRK1), 2 = IMEX SSP2(2,2,2)
                             or MIRK2, 3 = SSP2(3,3,2)-LUM, 4 = SSP3(4,3,3)
                             (:10); stages lmax = order                        */
  int32_t flux_solver;    /* CUEVA_HLL/_HLLE/_HLLC/_HLL_MCL/_HLLE_MCL/_HLLC_MCL/_LAX (:21) */
  int32_t rec_mp;         /* 0 MP5, 1 MP7, 2 MP9 (:33)                          */
  int32_t bc_x, bc_y;     /* CUEVA_BC_* per axis                                */
  int32_t precision;      /* CUEVA_FP64 / CUEVA_MIXED                           */
  int32_t c2p_mode;       /* CUEVA_C2P_*                                        */
  int32_t hllc_seed;      /* 1: first contact-speed Newton point as the flang
                             build of the reference; 0: the clamped seed       */
  int32_t crmath;         /* 1: correctly rounded acos/cos/pow in the Cardano
                             step (double-double); 0: device math library      */
  int32_t primitive_rec;  /* 1: recons_var_primitive before each step (:26)    */
  int32_t slop;           /* slope limiter of the second-order face states
                             (:17): 1 MinMod, 2 monotonised central, 3 Superbee;
                             0 = 2                                               */
  int32_t limit;          /* the same for recons_var_primitive (:18); 0 = 2    */
  int32_t mirk;           /* electric field integration (:11): 0 IMEX implicit
                             stage solve, 1 MIRK first order, 2 MIRK second
                             order (11_elecvarint.f95, 20_electricfield.f95);
                             order 2 with MIRK 1 is refused, as the reference
                             stops on it (imex_scheme.f95:51-56)              */
  int32_t rec_prim;       /* 1: primitive recovery of every stage state and the
                             stage's fluxes from those primitives (:23,
                             main_cueva_code.f95:220-228, 13_intercon_var.f95);
                             refused with the second-order HLL/HLLE/HLLC,
                             whose REC_PRIM branch in the reference reads an
                             unset z-velocity (16_hll_flow_mcl.f95:911-958)  */
  double longx, longy;    /* domain lengths Lx, Ly (:174-175)                   */
  double cfl;             /* CFL number (:309)                                  */
  double sigma;           /* conductivity sigma_0 (:196)                        */
  double gamma;           /* adiabatic index (:148)                             */
  double dt;              /* time step; <= 0 selects CFL*min(Delx,Dely,1) (:312) */
  double reserved_d[6];
} cueva_params;

typedef struct cueva_ctx cueva_ctx;

/* version / build information */
int cueva_abi_version(void);

/* Defaults: 128x256, SSP2(2,2,2), HLL+MP5, copy/periodic, FP64, Cardano. */
int cueva_params_default(cueva_params *p);

/* Grid spacings and time step as the reference derives them
 * (parameters.f95:306-312, 325-334). out = {Delx, Dely, Delt, kappa_psi, kappa_phi}. */
int cueva_derived(const cueva_params *p, double out[5]);

/* Allocate device state for one grid on the current HIP device.
 * Replaces the module arrays of scalar.f95 / threevectors.f95 / fourvectors.f95. */
int cueva_create(const cueva_params *p, cueva_ctx **ctx);
int cueva_destroy(cueva_ctx *ctx);
/* Wait for all work queued on ctx. */
int cueva_sync(cueva_ctx *ctx);
/* Device bytes held by ctx. */
int cueva_device_bytes(const cueva_ctx *ctx, int64_t *bytes);

/* Host <-> device.  prims14 holds 14 planes of (imax+13)*(jmax+13) doubles.
 * Upload replaces the RESTART read (03_restart_primitive.f95:50-61). */
int cueva_upload_prims(cueva_ctx *ctx, const double *prims14);
/* Download the 14 planes and, if cons5 != NULL, the evolved D tau Sx Sy Sz. */
int cueva_download(cueva_ctx *ctx, double *prims14, double *cons5);

/* ---- one entry point per reference subroutine of the step ---------------- */
/* boundary / boundary_conserved / boundary_electric / boundary_flux_conserved
 * (04_boundary.f95:1, 24_boundary_conserved.f95:2, 25_boundary_electric.f95:1,
 *  24_boundary_flux_conserved.f95:2); kind = CUEVA_BND_*, l = stage (1-based). */
int cueva_boundary(cueva_ctx *ctx, int kind, int l);
/* Second-order diagonal MC reconstruction and re-averaging of the 14
 * primitives on the interior (recons_var_primitive, 05_reconstruct.f95:6;
 * the reference calls it, then boundary, before startconvar when
 * primitive_rec = 1, main_cueva_code.f95:170-175).  Ghosts are not filled. */
int cueva_recons_var_primitive(cueva_ctx *ctx);
int cueva_startconvar(cueva_ctx *ctx);             /* 06_conserved_var.f95:5   */
int cueva_cleanvarint(cueva_ctx *ctx);             /* 07_clean_var.f95:5       */
int cueva_swap1(cueva_ctx *ctx);                   /* 08_swap1.f95:2           */
int cueva_sumint(cueva_ctx *ctx, int l, int m);    /* 09_sumint.f95:5          */
int cueva_varint(cueva_ctx *ctx, int l);           /* 10_varint.f95:5          */
int cueva_electricvarint(cueva_ctx *ctx, int l);   /* 11_elecvarint.f95:3      */
/* REC_PRIM = 1 only (main_cueva_code.f95:220-228): varprimitivecardano with
 * l_swap = 0 on stage l's state (12_varprimitive.f95:111-135,792-806),
 * started from that stage's primitives as they stand (the previous step's;
 * zero before the first step); its fix-ups count towards the step's
 * counters.  Does not wait: a failure surfaces at the next call that waits. */
int cueva_c2p_stage(cueva_ctx *ctx, int l);
/* REC_PRIM = 1 only: interconvar (13_intercon_var.f95:6), stage l's
 * conserved variables from its recovered primitives, then the ghosts of the
 * stage's conserved and E planes. */
int cueva_interconvar(cueva_ctx *ctx, int l);
int cueva_flowinit(cueva_ctx *ctx, int l);         /* 14_flowinit.f95          */
/* hll_flow_mpx / hlle_flow_mpx / hllc_flow_mpx / hll_flow / hlle_flow /
 * hllc_flow / lax_flow per params.flux_solver (16_hll_flow_mpx.f95,
 * 16_hlle_flow_mpx.f95, 16_hllc_flow_mpx.f95, 16_hll_flow_mcl.f95:5,
 * 16_hlle_flow_mcl.f95:1638, 16_hllc_flow_mcl.f95, 15_lax_flow.f95:5). */
int cueva_flux(cueva_ctx *ctx, int l);
int cueva_fluxandsource(cueva_ctx *ctx, int l);    /* 17_fluxandsource.f95:6   */
int cueva_sumfin(cueva_ctx *ctx, int l);           /* 18_sumfin.f95:5          */
int cueva_varfin(cueva_ctx *ctx);                  /* 19_varfin.f95:5          */
int cueva_electricfield(cueva_ctx *ctx);           /* 20_electricfield.f95     */
/* varprimitivecardano with l_swap = 1, or with REC_PRIM = 1 l_swap = 2 (from
 * the last stage's primitives; 12_varprimitive.f95:6, main_cueva_code.f95:
 * 301-309).  Waits for the device; counters = {contador_r, contador_p,
 * contador_v} of this call (with REC_PRIM = 1: of the step since
 * cueva_cleanvarint, stage recoveries included), whose fix-up records
 * cueva_fixup_records returns; fallback = cells that fell back
 * to Cardano in the learned mode (else 0).  A superluminal previous velocity
 * is clamped with the reference's eps1 (1e-6 once an HLL+MPx face solve ran
 * in this ctx, else 0; 12_varprimitive.f95:236-241, 16_hll_flow_mpx.f95:1375).
 * Returns the first failure of the step so far (face solves included).
 * Ends a step of the per-subroutine sequence (the step index advances). */
int cueva_c2p(cueva_ctx *ctx, int l_swap, int32_t counters[3], int32_t *fallback);

/* nsteps whole steps (main_cueva_code.f95:176-332 per step), fused into a
 * few kernels: same results bit for bit as the sequence boundary,
 * startconvar, ..., varprimitivecardano, boundary.  Per step: one marching
 * face launch per stage (reconstruction, Riemann solve, flux difference and
 * the next stage's state, or U^{n+1} on the last stage), one primitive-
 * recovery launch (which within one call also forms the next step's first
 * stage state) and ghost fills. */
int cueva_step(cueva_ctx *ctx, int nsteps);
/* Fix-up counters and Cardano fallbacks accumulated by cueva_step since the
 * last call (waits for the device); also clears the fix-up records. */
int cueva_step_counters(cueva_ctx *ctx, int64_t counters[3], int64_t *fallback);

/* The first failure (0 if none; waits for the device) and where it arose:
 * where = {step index h, stage l (1-based; lmax for the primitive recovery),
 * i, j} -- the reference's (h, l, i, j) of its stop message
 * (12_varprimitive.f95:374-425,482-486,684-693; funciones.f95:238-244).
 * After a failure inside cueva_step the primitives are those at the start
 * of the failing step (the step's later launches change no data, and its
 * recovery writes to scratch planes that take the primitives' place only
 * once it has succeeded).  No reference counterpart beyond the stop message
 * itself. */
int cueva_last_error(cueva_ctx *ctx, int32_t where[4]);

/* The step index the next step carries in failure locations and fix-up
 * records (the reference's h; 1 after cueva_create, then +1 per fused step
 * and per cueva_c2p call).  A driver that resumes at step hini sets it. */
int cueva_set_step_index(cueva_ctx *ctx, int64_t h);

/* Fix-up records of kind CUEVA_FIX_* since the counters were last cleared
 * (cueva_cleanvarint, cueva_c2p, cueva_step_counters), in the order the
 * reference writes them to its unit 301/302/303 file
 * (12_varprimitive.f95:544,601,663): by step, then j, then i.  Per record
 * hij = {h, i, j} and vals = the two values the reference prints (rho* and
 * rho^n; p* and p^n; v*^2 and v_drift^2); the running count and x, y of its
 * record follow from these.  *n = number of records; at most cap are
 * copied.  CUEVA_E_NOMEM if more than 65536 records of the kind arose. */
int cueva_fixup_records(cueva_ctx *ctx, int kind, int32_t *hij, double *vals, int64_t cap, int64_t *n);

/* The 37 global diagnostics the reference writes every hglb steps
 * (22_writer.f95:818-991), in the order of its output units 80..217:
 * out = {sum_psi, sum_phi, sum_ex, sum_ey, sum_ez, sum_exey, sum_exez,
 * sum_eyez, sum_bx, sum_by, sum_bz, sum_bxby, sum_bxbz, sum_bybz, sum_vxvyvz,
 * sum_bxbybz, sum_vx, sum_vy, sum_vz, sum_vxvy, sum_vxvz, sum_vyvz, sum_q,
 * sum_p, sum_rho, ekin_c, ekin_r, etot, efluid_r, emag, eint_c, eint_r,
 * lorentz, sum_D, sum_D_med, max_val_bx, max_val_phi} over i=0..imax,
 * j=0..jmax with each cell term formed as the reference forms it.  FP64,
 * deterministic; differs from the reference's sequential sum only by the
 * summation order.  Waits for the device. */
#define CUEVA_N_GLOBALS 37
int cueva_reduce_globals(cueva_ctx *ctx, double out[CUEVA_N_GLOBALS]);

/* cueva_reduce_globals for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): no context, no allocation and no wait.  No reference
 * counterpart beyond cueva_reduce_globals itself.  prims14 / cons5: device
 * memory holding the 14 primitive planes (restart order) and D tau Sx Sy Sz
 * in the host layout of cueva_upload_prims ((imax+13)*(jmax+13) doubles a
 * plane, x fastest); out: device memory for the 37 results; stream: the
 * hipStream_t to queue on (NULL: the null stream).  The block partial sums
 * live in a fixed device array of the library, 1024 x 37 doubles per GPU
 * (every grid fits: the reduction strides over it), held until the process
 * ends; calls that overlap on one GPU share it and race. */
int cueva_reduce_globals_noalloc(const cueva_params *p, const double *prims14, const double *cons5, double *out,
                                 void *stream);

/* Largest relativistic acoustic signal speed |lambda| over the interior and
 * both directions.  No reference counterpart: the reference takes the fixed
 * dt = CFL*min(Delx, Dely, 1) of parameters.f95:312 (light speed 1, which
 * bounds every wave of the augmented system); this lets a caller check that
 * bound or choose an adaptive dt.  Waits for the device. */
int cueva_max_speed(cueva_ctx *ctx, double *vmax);

/* Replace the conserved D tau Sx Sy Sz (5 planes, host layout) that the
 * primitive recovery reads; with cueva_upload_prims this loads a restart
 * state given in conserved form.  No reference counterpart (the reference
 * restarts from primitives only). */
int cueva_upload_cons(cueva_ctx *ctx, const double *cons5);

/* Minimum (out[0]) and maximum (out[1]) of one device plane over the
 * interior points; group and index as cueva_download_plane.  No reference
 * counterpart: the KH transverse-velocity amplitude (max vy - min vy)/2 of a
 * growth-rate series without copying the plane to the host. */
int cueva_plane_range(cueva_ctx *ctx, int group, int index, double out[2]);

/* Copy one device plane to the host (host layout).  group: 0 primitives
 * (14, restart order), 1 D tau Sx Sy Sz W, 2 stage planes, 3 x-face fluxes,
 * 4 y-face fluxes, 5+l stage-l right-hand sides (17).  For tests. */
int cueva_download_plane(cueva_ctx *ctx, int group, int index, double *out);

/* Kernel timing on the ctx stream: while on, every launch is bracketed by
 * HIP events; cueva_kernel_times sums them per kernel class
 * {stage, faces, fluxsource, final, ghosts, other} in ms, with launch counts. */
int cueva_set_profiling(cueva_ctx *ctx, int on);
int cueva_kernel_times(cueva_ctx *ctx, double ms[6], int64_t launches[6]);
/* Milliseconds of nsteps fused steps timed with HIP events on the ctx stream. */
int cueva_time_steps(cueva_ctx *ctx, int nsteps, double *ms);

/* ---- learned start point of the primitive recovery ----------------------
 * No reference counterpart.  A 3 -> 32 -> 32 -> 2 MLP (tanh, BF16 weights on
 * the matrix cores, FP32 accumulation) maps the normalised features
 * (log(D/C2), |S - E x B|/C2, log C2), C2 = tau - (E^2+B^2)/2, of a conserved
 * state to (log W, log(p/C2)).  In c2p_mode CUEVA_C2P_LEARNED its W starts at
 * most two FP64 Newton steps on the quartic of 12_varprimitive.f95:245-296;
 * a result with relative quartic residual > 1e-10 or W < 1 falls back to the
 * Cardano path (counted in the fallback counter), and the reference's
 * pressure loop, fix-ups and velocities follow either way.
 * Packed parameters, CUEVA_MLP_NPACK floats: W1[32][3] b1[32] W2[32][32]
 * b2[32] W3[2][32] b3[2] (the CUEVA_MLP_NPARAM trainable ones), then
 * in_mean[3] in_std[3] out_mean[2] out_std[2]. */
#define CUEVA_MLP_NPARAM 1250
#define CUEVA_MLP_NPACK 1260
typedef struct cueva_mlp cueva_mlp;

/* Attach trained parameters to a solver (required before a step in
 * CUEVA_C2P_LEARNED mode). */
int cueva_set_c2p_params(cueva_ctx *ctx, const float *packed);

/* Training on the device: FP32 master weights, Adam, mean squared error on
 * normalised targets.  Forward on the matrix cores, backward through LDS. */
int cueva_mlp_create(cueva_mlp **m);
int cueva_mlp_destroy(cueva_mlp *m);
int cueva_mlp_set_params(cueva_mlp *m, const float *packed);  /* also resets Adam */
int cueva_mlp_get_params(cueva_mlp *m, float *packed);
/* n samples: x[n][3] normalised features, y[n][2] normalised targets */
int cueva_mlp_set_data(cueva_mlp *m, const float *x, const float *y, int64_t n);
/* gradient (CUEVA_MLP_NPARAM floats) and mean loss over samples [off, off+cnt) */
int cueva_mlp_grad(cueva_mlp *m, int64_t off, int64_t cnt, float *grad, double *loss);
/* steps Adam updates over consecutive batches; loss = mean batch loss */
int cueva_mlp_train(cueva_mlp *m, int64_t batch, int steps, float lr, double *loss);
/* normalised outputs y[n][2] of normalised inputs x[n][3] */
int cueva_mlp_forward(cueva_mlp *m, const float *x, int64_t n, float *y);

/* Seeded Kelvin-Helmholtz initial data on the host (no reference counterpart:
 * the reference has no such setup).  Periodic x in [-Lx/2, Lx/2], outflow y in
 * [-Ly/2, Ly/2]; rho = p = 1, vx = v0 tanh(y/0.01) + U(-1e-4,1e-4) noise,
 * vy = 0.01 sin(2 pi x / Lx) exp(-y^2/0.01), B = (bx0, 0, bz0), E = -v x B.
 * rotated != 0 writes the same data in the frame x' = y, y' = -x (so that
 * the reference's BOUND 6 applies) with imax' = jmax, jmax' = imax, i.e. the
 * caller's imax/jmax are the rotated grid's. */
int cueva_kh_init(int imax, int jmax, double lx, double ly, int rotated, double v0,
                  double bx0, double bz0, uint64_t seed, double *prims14);

/* Circularly polarised Alfven wave along x, uniform in y, for a periodic
 * box (03_init_var_primitive.f95:129-160 extended to 2D): rho0, p0, guide
 * field b0, amplitude eta; *va receives the exact Alfven speed (one period
 * is Lx / va).  Ghosts are filled periodically. */
int cueva_cpaw_init(int imax, int jmax, double lx, double gamma, double rho0, double p0, double b0, double eta,
                    double *prims14, double *va);

#ifdef __cplusplus
}
#endif
#endif
