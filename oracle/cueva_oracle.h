/*
 * CPU restatement (FP64) of one CUEVA 2D resistive-RMHD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; tests, the smoke check and the CPU-baseline leg of the benchmark
 * use it as the checker.  Every function cites the reference file:line it
 * follows (paths relative to the reference's cueva_2D_base/ directory).
 *
 * Layout: every field is one plane of (imax+13) x (jmax+13) doubles, x
 * fastest, holding the reference index range (-6:imax+6, -6:jmax+6).
 */
#ifndef CUEVA_ORACLE_H
#define CUEVA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order of the 14 evolved primitive/EM planes: the reference restart
 * column order (22_writer.f95:1116-1172). */
enum {
  OR_PSI = 0, OR_PHI, OR_BX, OR_BY, OR_BZ, OR_VX, OR_VY, OR_VZ,
  OR_EX, OR_EY, OR_EZ, OR_Q, OR_RHO, OR_P, OR_NPRIM
};

/* Boundary kinds per axis, as index maps of the reference ghost fills. */
enum { OR_BC_COPY = 0, OR_BC_PERIODIC = 1 };

/* Riemann solvers (parameters.f95:21 flux_solver 5/6/7). */
enum { OR_LAX = 1, OR_HLL_MCL = 2, OR_HLLE_MCL = 3, OR_HLLC_MCL = 4, OR_HLL = 5, OR_HLLE = 6, OR_HLLC = 7 };

typedef struct {
  int32_t imax, jmax;
  int32_t lmax;            /* stages = order: 1 (MIRK1 or first-order IMEX), 2 = SSP2(2,2,2) or MIRK2,
                              3 = SSP2(3,3,2)-LUM, 4 = SSP3(4,3,3) */
  int32_t flux_solver;     /* 5 HLL, 6 HLLE, 7 HLLC */
  int32_t rec_order;       /* 5, 7 or 9 (MP5/MP7/MP9) */
  int32_t bc_x, bc_y;      /* OR_BC_COPY / OR_BC_PERIODIC */
  int32_t pad0;
  double dt, dx, dy;
  double sigma, kappa_psi, kappa_phi, gamma;
  double wb[8], wbt[8], Ab[8][8], Abt[8][8];
  /* modes of this instance (oracle_step installs them for its own run) */
  int32_t primitive_rec;   /* 1: recons_var_primitive before each step (parameters.f95:26) */
  int32_t hllc_seed;       /* 1: first contact-speed Newton point as the reference build; 0: clamped seed */
  int32_t libm_mode;       /* 0: acos/cos/pow of the C library (as the reference build); 1: correctly rounded */
  int32_t slop;            /* second-order slope limiter (parameters.f95:17): 1 MinMod, 2 MC, 3 Superbee */
  int32_t limit;           /* the same for recons_var_primitive (parameters.f95:18) */
  int32_t mirk;            /* 0 IMEX, 1 MIRK first order, 2 MIRK second order (parameters.f95:11) */
  int32_t rec_prim;        /* 1: primitive recovery of every stage state (parameters.f95:23) */
  int32_t pad1;
} oracle_cfg;

typedef struct oracle_state oracle_state;

/* Fill cfg with the reference's constants for a grid and scheme
 * (parameters.f95:303-334, imex_scheme.f95:13-49,77-86,280-291), MIRK = 0. */
int oracle_default_cfg(oracle_cfg *c, int imax, int jmax, double lx, double ly,
                       int lmax, int flux_solver, int rec_order, int bc_x, int bc_y,
                       double sigma, double cfl);
/* Refill the Butcher tables of cfg for its lmax (= order) and mirk
 * (imex_scheme.f95:13-58); nonzero for the combination the reference stops
 * on (order 2 with MIRK = 1). */
int oracle_scheme_tables(oracle_cfg *c);

oracle_state *oracle_create(const oracle_cfg *c);
void oracle_destroy(oracle_state *s);
size_t oracle_plane_len(const oracle_state *s);

/* Copy the 14 planes in (restart order) into the state, then fill ghosts. */
void oracle_set_prims(oracle_state *s, const double *prims14);
/* Copy out the 14 planes and the 5 evolved conserved planes D tau Sx Sy Sz. */
void oracle_get_fields(const oracle_state *s, double *prims14, double *cons5);
void oracle_get_counters(const oracle_state *s, int64_t counters3[3]);
/* Replace the conserved D tau Sx Sy Sz (5 planes) the recovery reads. */
void oracle_set_cons(oracle_state *s, const double *cons5);
/* varprimitivecardano alone with l_swap = 1 (12_varprimitive.f95:6-824) on
 * the state's primitives (previous velocity, rho, p; E, B) and conserved
 * planes; counters and fix-up records restart from zero.  Returns 0 or the
 * status of the reference's first stop in its loop order. */
int oracle_recover(oracle_state *s);
/* Where the last nonzero status arose: {step (1-based within the
 * oracle_step call), stage l (1-based; lmax for the recovery), i, j}. */
void oracle_last_stop(const oracle_state *s, int32_t out[4]);
/* How often the MPx face solves took their fallbacks since create: {the
 * superluminal face-velocity fallback (16_hll_flow_mpx.f95:1392-1461), the
 * HLLC face-density fix (16_hllc_flow_mpx.f95:2663-2693)}. */
void oracle_face_events(const oracle_state *s, int64_t out[2]);
/* The same events one by one, in the order the reference prints them:
 * {kind, stage l, i, j} and a value per event; kind 0 = fallback from an
 * x face ("FIXED velocity ... V2_X"), 1 = from a y face only ("V2_Y"),
 * 2..5 = density fix of y_L, y_R, x_L, x_R ("Negative density"; value =
 * the negative density).  Returns the count; copies at most cap. */
int oracle_get_face_events(const oracle_state *s, int32_t *kind_l_i_j, double *vals, int cap);

/* Advance nsteps steps; returns 0, or a nonzero status naming the failure
 * the reference would stop on (see cueva_oracle.c). */
int oracle_step(oracle_state *s, int nsteps);

/* Debug view of the last completed stage l (0-based) of the last step:
 * which = 0..16 stage RHS (14 flux divergences then 3 sources). */
void oracle_get_rhs(const oracle_state *s, int l, int which, double *out);
void oracle_get_face(const oracle_state *s, int dir, int slot, double *out);

/* The 37 global diagnostics of 22_writer.f95:818-991 in the reference's
 * order and summation order (i outer, j inner). */
void oracle_globals(const oracle_state *s, double out[37]);

/* Fix-up records of the last step's primitive recovery, as the reference
 * writes them to units 301 (rho), 302 (p) and 303 (drift) in
 * 12_varprimitive.f95:544,601,663: per record i, j and the two values it
 * prints (rho* and rho^n; p* and p^n; v*^2 and v_drift^2), in the
 * reference's loop order (j outer, i inner).  Returns the number of
 * records; copies at most cap of them. */
int oracle_get_fixups(const oracle_state *s, int kind, int32_t *ij, double *vals, int cap);

/* Pointwise/linewise pieces, exported for unit tests. */
void oracle_mp_line(int order, const double *u, int n, double *up, double *um, int positivity);
double oracle_hllc_vel(double a, double b, double c, double rt, double v0);
/* One cell of varprimitivecardano with l_swap = 1 (12_varprimitive.f95:48-813).
 * in = {D, tau, Sx, Sy, Sz, Ex, Ey, Ez, Bx, By, Bz, Vx, Vy, Vz} (V: the
 * previous velocity); out[3], out[4] hold the previous rho and p on entry;
 * on return out = {Vx, Vy, Vz, rho, p}, flags = fix-ups {rho, p, drift},
 * rec = the two values of each fix-up record {rho*, rho^n, p*, p^n, v*^2,
 * v_drift^2}.  eps1: the reference's module variable eps1, which clamps a
 * superluminal previous velocity (:238): 1e-6 once hll_flow_mpx has run in
 * the process (16_hll_flow_mpx.f95:1375), else 0.  Returns 0 or the status
 * of the reference's stop. */
int oracle_c2p_point(double gamma, double eps1, const double in[14], double out[5], int flags[3], double rec[6]);

#ifdef __cplusplus
}
#endif
#endif
