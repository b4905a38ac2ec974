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
  int32_t lmax;            /* IMEX stages: 2 = SSP2(2,2,2), 4 = SSP3(4,3,3) */
  int32_t flux_solver;     /* 5 HLL, 6 HLLE, 7 HLLC */
  int32_t rec_order;       /* 5, 7 or 9 (MP5/MP7/MP9) */
  int32_t bc_x, bc_y;      /* OR_BC_COPY / OR_BC_PERIODIC */
  int32_t pad0;
  double dt, dx, dy;
  double sigma, kappa_psi, kappa_phi, gamma;
  double wb[8], wbt[8], Ab[8][8], Abt[8][8];
} oracle_cfg;

typedef struct oracle_state oracle_state;

/* Fill cfg with the reference's constants for a grid and scheme
 * (parameters.f95:303-334, imex_scheme.f95:40-49,280-291). */
int oracle_default_cfg(oracle_cfg *c, int imax, int jmax, double lx, double ly,
                       int lmax, int flux_solver, int rec_order, int bc_x, int bc_y,
                       double sigma, double cfl);

oracle_state *oracle_create(const oracle_cfg *c);
void oracle_destroy(oracle_state *s);
size_t oracle_plane_len(const oracle_state *s);

/* Copy the 14 planes in (restart order) into the state, then fill ghosts. */
void oracle_set_prims(oracle_state *s, const double *prims14);
/* Copy out the 14 planes and the 5 evolved conserved planes D tau Sx Sy Sz. */
void oracle_get_fields(const oracle_state *s, double *prims14, double *cons5);
void oracle_get_counters(const oracle_state *s, int64_t counters3[3]);

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

/* Pointwise/linewise pieces, exported for unit tests. */
void oracle_mp_line(int order, const double *u, int n, double *up, double *um, int positivity);
double oracle_hllc_vel(double a, double b, double c, double rt, double v0);
/* 1: recons_var_primitive before each step (primitive_rec = 1) */
void oracle_set_primitive_rec(int on);
/* 1: first contact-speed Newton point as in the reference build; 0: clamped seed */
void oracle_set_hllc_seed_mode(int mode);
/* 0: acos/cos/pow from the C library (as the reference build); 1: correctly rounded */
void oracle_set_libm_mode(int mode);
int oracle_c2p_point(double gamma, const double in[14], double out[5], int flags[3]);

#ifdef __cplusplus
}
#endif
#endif
