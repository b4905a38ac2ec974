/*
 * CPU restatement (FP64) of one CUEVA 2D resistive-RMHD time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 *
 * Each routine follows the reference routine it names, file:line relative to
 * the reference's cueva_2D_base/ directory, with the same floating-point
 * operation order (Fortran evaluates left to right; x**n is x*(x*(...x)),
 * as flang lowers it) so that, compiled with -ffp-contract=off, results match
 * the reference binary bit for bit except where noted:
 *   - get_hllc_vel (funciones.f95:188-259) reads its Newton iterate before
 *     setting it; this restatement starts Newton from the clamped seed.
 *   - acos/cos/pow come from the C library, as in the reference build.
 *   - z-direction flux differences (X(i,j)-X(i,j))/Delz are the constant +0.
 *   - the EGLM terms (FAC_EGLM = 0, parameters.f95:44) are left out; they
 *     only add a signed zero.
 *
 * Flags assumed (parameters.f95 defaults kept by every configuration):
 * MIRK=0, DIM=2, REC_PRIM=0, REC_ELEC=0, VDRIFT=0, primitive_rec=0,
 * source_rec(_mpx)=0, flux_rec_mp5=0, REC_4VECTOR=0, SLC=0, low_vel=0.
 */
#include "cueva_oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Elementary functions of the Cardano step (12_varprimitive.f95:342-434).
 * Mode 0 calls the C library, as the reference build does (and so matches it
 * bit for bit on the machine the fixtures were made on).  Mode 1 returns the
 * correctly rounded result, computed in quad precision (libquadmath) and
 * rounded once to double. */
/* The mode switches below are set from the running instance's cfg at the
 * start of every oracle_step / recovery call (install_modes). */
static int g_libm_mode = 0;
static double o_acos(double x) { return g_libm_mode ? (double)acosq((__float128)x) : acos(x); }
static double o_cos(double x) { return g_libm_mode ? (double)cosq((__float128)x) : cos(x); }
static double o_pow(double x, double y) {
  return g_libm_mode ? (double)powq((__float128)x, (__float128)y) : pow(x, y);
}

#define G 6
#define NXS 27 /* x-sweep slots 1..27   (16_hll_flow_mpx.f95:784-917)  */
#define NYS 26 /* y-sweep slots 38..63  (16_hll_flow_mpx.f95:1086-1211) */

/* x-sweep slot numbers minus one */
enum { XS_PSI, XS_PHI, XS_EX, XS_EY, XS_EZ, XS_EXM, XS_EYM, XS_EZM, XS_BX, XS_BY, XS_BZ,
       XS_BXM, XS_BYM, XS_BZM, XS_Q, XS_D, XS_TAU, XS_SX, XS_SY, XS_SZ, XS_VX, XS_VY, XS_VZ,
       XS_P, XS_RHO, XS_JX, XS_JZ };
/* y-sweep slot numbers minus 38 */
enum { YS_PSI, YS_PHI, YS_EX, YS_EY, YS_EZ, YS_EXM, YS_EYM, YS_EZM, YS_BX, YS_BY, YS_BZ,
       YS_BXM, YS_BYM, YS_BZM, YS_Q, YS_D, YS_TAU, YS_SX, YS_SY, YS_SZ, YS_VX, YS_VY, YS_VZ,
       YS_P, YS_RHO, YS_JY };

/* stage state planes (the l-indexed arrays of fourvectors.f95:12-59) */
enum { ST_PSI, ST_PHI, ST_BX, ST_BY, ST_BZ, ST_EXA, ST_EYA, ST_EZA, ST_Q, ST_D, ST_TAU,
       ST_SX, ST_SY, ST_SZ, ST_EX, ST_EY, ST_EZ, ST_JX, ST_JY, ST_JZ, ST_N };

/* 14 accumulated fields in the order of 09_sumint.f95 / 18_sumfin.f95 */
enum { AC_PSI, AC_PHI, AC_BX, AC_BY, AC_BZ, AC_EX, AC_EY, AC_EZ, AC_Q, AC_D, AC_TAU,
       AC_SX, AC_SY, AC_SZ, AC_N };

/* face fluxes used by 17_fluxandsource.f95:433-637 */
enum { FX_EXPSI, FX_BXPHI, FX_PHIBX, FX_EZBY, FX_EYBZ, FX_PSIEX, FX_BZEY, FX_BYEZ, FX_J, FX_FD,
       FX_FTAU, FX_FSX, FX_FSY, FX_FSZ, FX_N };
enum { FY_EYPSI, FY_BYPHI, FY_EZBX, FY_PHIBY, FY_EXBZ, FY_BZEX, FY_PSIEY, FY_BXEZ, FY_J, FY_FD,
       FY_FTAU, FY_FSX, FY_FSY, FY_FSZ, FY_N };

#define NRHS 17
#define MAXL 8

struct oracle_state {
  oracle_cfg c;
  int nx, ny;
  size_t n;
  double *prim[OR_NPRIM];
  double *D, *tau, *Sx, *Sy, *Sz, *W;
  double *intsum[AC_N], *finsum[AC_N];
  double *st[ST_N];
  double *xL[NXS], *xR[NXS], *yL[NYS], *yR[NYS];
  double *fx[FX_N], *fy[FY_N];
  double *rhs[MAXL][NRHS];
  int64_t cnt[3];
  double eps1; /* the reference's module eps1: 1e-6 once hll_flow_mpx ran (16_hll_flow_mpx.f95:1375) */
  /* fix-up records of the last recovery, per kind: (i, j) and two values */
  int nfix[3], capfix[3];
  int32_t stop[4]; /* step (1-based within oracle_step), stage (1-based), i, j of the last stop */
  int64_t face_events[2]; /* superluminal face fallbacks, HLLC face-density fixes (since create) */
  /* the same events as the reference prints them, in its order: kind, stage, i, j, and a value */
  int nfe, capfe;
  int32_t *fe_kind;
  double *fe_val;
  int cur_l;
  int32_t *fix_ij[3];
  double *fix_val[3];
  /* B^n and E^n kept by varfin for the MIRK electric field (19_varfin.f95:24-30) */
  double *bn[3], *en[3];
  /* stage primitives Vxint..Vzint, Wint, rhoint, pint, enthpyint (fourvectors.f95:12-13):
   * static arrays of the reference, zero until written, kept from step to step */
  double *vint[MAXL][3], *wint[MAXL], *rhoint[MAXL], *pint[MAXL], *hint[MAXL];
  /* the velocity, pressure and density the face solves reconstruct and the
   * stage right-hand side reads (K_VX..K_RHO): the step's primitives, or
   * with REC_PRIM = 1 the current stage's (set per stage by oracle_step) */
  const double *kin[5];
};
enum { K_VX, K_VY, K_VZ, K_P, K_RHO };

static void set_kin(oracle_state *s, int l) {
  const int rp = s->c.rec_prim && l >= 0;
  s->kin[K_VX] = rp ? s->vint[l][0] : s->prim[OR_VX];
  s->kin[K_VY] = rp ? s->vint[l][1] : s->prim[OR_VY];
  s->kin[K_VZ] = rp ? s->vint[l][2] : s->prim[OR_VZ];
  s->kin[K_P] = rp ? s->pint[l] : s->prim[OR_P];
  s->kin[K_RHO] = rp ? s->rhoint[l] : s->prim[OR_RHO];
}

#define IX(s, i, j) (((size_t)((j) + G)) * (size_t)(s)->nx + (size_t)((i) + G))
#define AT(s, a, i, j) ((a)[IX(s, i, j)])

static inline double sq(double x) { return x * x; }
static inline double pw3(double x) { return x * (x * x); }
static inline double pw4(double x) { return x * (x * (x * x)); }
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
/* Fortran min / max as flang evaluates them (probed with the flang of this
 * image): the first argument when it compares strictly less (greater), else
 * the second -- so min(NaN, x) = x, min(x, NaN) = NaN, min(-0, +0) = +0;
 * min(a, b, c) = min(min(a, b), c) */
static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmax2(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* configuration                                                       */
/* ------------------------------------------------------------------ */

int oracle_default_cfg(oracle_cfg *c, int imax, int jmax, double lx, double ly, int lmax,
                       int flux_solver, int rec_order, int bc_x, int bc_y, double sigma,
                       double cfl) {
  memset(c, 0, sizeof(*c));
  if (imax < 8 || jmax < 8) return 1;
  if (lmax < 1 || lmax > 4) return 2;
  if (flux_solver != OR_HLL && flux_solver != OR_HLLE && flux_solver != OR_HLLC && flux_solver != OR_HLL_MCL &&
      flux_solver != OR_LAX && flux_solver != OR_HLLC_MCL && flux_solver != OR_HLLE_MCL)
    return 3;
  if (rec_order != 5 && rec_order != 7 && rec_order != 9) return 4;
  c->imax = imax;
  c->jmax = jmax;
  c->lmax = lmax;
  c->flux_solver = flux_solver;
  c->rec_order = rec_order;
  c->bc_x = bc_x;
  c->bc_y = bc_y;
  /* parameters.f95:306-312: Delx = Longx/real(imax) (real(imax) is exact), Delz = 1 */
  c->dx = lx / (double)(float)imax;
  c->dy = ly / (double)(float)jmax;
  double m = fmin2(fmin2(c->dx, c->dy), 1.0);
  c->dt = cfl * m;
  c->sigma = sigma;
  c->gamma = 4.0 / 3.0; /* parameters.f95:148 */
  c->primitive_rec = 0;
  c->hllc_seed = 1;
  c->libm_mode = 0;
  c->slop = 2;
  c->limit = 2;
  /* parameters.f95:325-334: kappa = alpha_ch**2 / cp**2, cp = sqrt(min(Delx,Dely)/alpha) */
  double cp = sqrt(fmin2(c->dx, c->dy) / 1.0);
  c->kappa_psi = 1.0 / (cp * cp);
  c->kappa_phi = 1.0 / (cp * cp);
  c->mirk = 0;
  c->rec_prim = 0;
  return oracle_scheme_tables(c) ? 2 : 0;
}

int oracle_scheme_tables(oracle_cfg *c) {
  const int lmax = c->lmax;
  memset(c->wb, 0, sizeof(c->wb));
  memset(c->wbt, 0, sizeof(c->wbt));
  memset(c->Ab, 0, sizeof(c->Ab));
  memset(c->Abt, 0, sizeof(c->Abt));
  if (c->mirk < 0 || c->mirk > 2) return 1;
  if (lmax == 1) { /* imex_scheme.f95:13-20: the first-order scheme of MIRK = 1 (and IMEX order 1) */
    c->wb[0] = 1.0;
    c->Ab[0][0] = 1.0;
  } else if (lmax == 2 && c->mirk == 2) { /* MIRK2, imex_scheme.f95:24-36 */
    c->wb[0] = 0.5; c->wb[1] = 0.5;
    c->Ab[1][0] = 1.0;
  } else if (lmax == 2) { /* imex_scheme.f95:40-49, const = 1 - 1/sqrt(2) (parameters.f95:240) */
    if (c->mirk != 0) return 1; /* imex_scheme.f95:51-56: the reference stops */
    double g = 1.0 - 1.0 / sqrt(2.0);
    c->wb[0] = 0.5; c->wb[1] = 0.5;
    c->wbt[0] = 0.5; c->wbt[1] = 0.5;
    c->Ab[1][0] = 1.0;
    c->Abt[0][0] = g;
    c->Abt[1][0] = 1.0 - 2.0 * g;
    c->Abt[1][1] = g;
  } else if (lmax == 3) { /* SSP2(3,3,2)-LUM, imex_scheme.f95:77-86 */
    for (int l = 0; l < 3; l++) c->wb[l] = c->wbt[l] = 1.0 / 3.0;
    c->Ab[1][0] = 0.5;
    c->Ab[2][0] = 0.5; c->Ab[2][1] = 0.5;
    c->Abt[0][0] = 1.0 / 5.0;
    c->Abt[1][0] = 1.0 / 10.0; c->Abt[1][1] = 1.0 / 5.0;
    c->Abt[2][0] = 1.0 / 3.0; c->Abt[2][1] = 1.0 / 3.0; c->Abt[2][2] = 1.0 / 3.0;
  } else if (lmax == 4) { /* imex_scheme.f95:280-291 with parameters.f95:242-244 */
    double a1 = 0.24169426078821, b1 = 0.06042356519705, e1 = 0.12915286960590;
    c->wb[1] = 1.0 / 6.0; c->wb[2] = 1.0 / 6.0; c->wb[3] = 2.0 / 3.0;
    c->wbt[1] = 1.0 / 6.0; c->wbt[2] = 1.0 / 6.0; c->wbt[3] = 2.0 / 3.0;
    c->Ab[2][1] = 1.0;
    c->Ab[3][1] = 0.25; c->Ab[3][2] = 0.25;
    c->Abt[0][0] = a1;
    c->Abt[1][0] = -1.0 * a1; c->Abt[1][1] = a1;
    c->Abt[2][1] = 1.0 - a1; c->Abt[2][2] = a1;
    c->Abt[3][0] = b1; c->Abt[3][1] = e1; c->Abt[3][2] = ((0.5 - b1) - e1) - a1; c->Abt[3][3] = a1;
  } else {
    return 1;
  }
  return 0;
}

static double *plane(const oracle_state *s) { return (double *)calloc(s->n, sizeof(double)); }

/* scratch of the line sweeps: the oracle stops the process rather than go
 * on without it */
static double *scratch_doubles(size_t n) {
  double *a = (double *)malloc(sizeof(double) * n);
  if (!a) {
    fprintf(stderr, "oracle: out of memory (%zu doubles)\n", n);
    abort();
  }
  return a;
}

oracle_state *oracle_create(const oracle_cfg *c) {
  oracle_state *s = (oracle_state *)calloc(1, sizeof(*s));
  if (!s) return NULL;
  s->c = *c;
  s->nx = c->imax + 2 * G + 1;
  s->ny = c->jmax + 2 * G + 1;
  s->n = (size_t)s->nx * (size_t)s->ny;
  for (int k = 0; k < OR_NPRIM; k++) s->prim[k] = plane(s);
  s->D = plane(s); s->tau = plane(s); s->Sx = plane(s); s->Sy = plane(s); s->Sz = plane(s);
  s->W = plane(s);
  for (int k = 0; k < AC_N; k++) { s->intsum[k] = plane(s); s->finsum[k] = plane(s); }
  for (int k = 0; k < ST_N; k++) s->st[k] = plane(s);
  for (int k = 0; k < NXS; k++) { s->xL[k] = plane(s); s->xR[k] = plane(s); }
  for (int k = 0; k < NYS; k++) { s->yL[k] = plane(s); s->yR[k] = plane(s); }
  for (int k = 0; k < FX_N; k++) s->fx[k] = plane(s);
  for (int k = 0; k < FY_N; k++) s->fy[k] = plane(s);
  for (int l = 0; l < c->lmax; l++)
    for (int k = 0; k < NRHS; k++) s->rhs[l][k] = plane(s);
  set_kin(s, -1);
  if (c->mirk)
    for (int d = 0; d < 3; d++) { s->bn[d] = plane(s); s->en[d] = plane(s); }
  if (c->mirk == 2 || c->rec_prim)
    for (int l = 0; l < c->lmax; l++) {
      for (int d = 0; d < 3; d++) s->vint[l][d] = plane(s);
      s->wint[l] = plane(s); s->rhoint[l] = plane(s); s->pint[l] = plane(s); s->hint[l] = plane(s);
    }
  return s;
}

void oracle_destroy(oracle_state *s) {
  if (!s) return;
  for (int k = 0; k < OR_NPRIM; k++) free(s->prim[k]);
  free(s->D); free(s->tau); free(s->Sx); free(s->Sy); free(s->Sz); free(s->W);
  for (int k = 0; k < AC_N; k++) { free(s->intsum[k]); free(s->finsum[k]); }
  for (int k = 0; k < ST_N; k++) free(s->st[k]);
  for (int k = 0; k < NXS; k++) { free(s->xL[k]); free(s->xR[k]); }
  for (int k = 0; k < NYS; k++) { free(s->yL[k]); free(s->yR[k]); }
  for (int k = 0; k < FX_N; k++) free(s->fx[k]);
  for (int k = 0; k < FY_N; k++) free(s->fy[k]);
  for (int l = 0; l < s->c.lmax; l++)
    for (int k = 0; k < NRHS; k++) free(s->rhs[l][k]);
  for (int k = 0; k < 3; k++) { free(s->fix_ij[k]); free(s->fix_val[k]); free(s->bn[k]); free(s->en[k]); }
  for (int l = 0; l < MAXL; l++) {
    for (int d = 0; d < 3; d++) free(s->vint[l][d]);
    free(s->wint[l]); free(s->rhoint[l]); free(s->pint[l]); free(s->hint[l]);
  }
  free(s->fe_kind);
  free(s->fe_val);
  free(s);
}

size_t oracle_plane_len(const oracle_state *s) { return s->n; }

/* ------------------------------------------------------------------ */
/* boundaries: 04_boundary.f95:922-1388 (BOUND 1), 1390-1862 (BOUND 2), */
/* 2867-3033 (BOUND 6) and the same index maps in                       */
/* 24_boundary_conserved.f95, 25_boundary_electric.f95,                 */
/* 24_boundary_flux_conserved.f95.  Every ghost is a copy of one        */
/* interior point; the map is per axis: copy clamps to [0,N], periodic  */
/* wraps with the duplicated edge (ghost -k <- N-k, ghost N+k <- k).    */
/* ------------------------------------------------------------------ */

static inline int src_index(int i, int n, int kind) {
  if (i >= 0 && i <= n) return i;
  if (kind == OR_BC_PERIODIC) return i < 0 ? n + i : i - n;
  return i < 0 ? 0 : n;
}

static void fill_ghosts(const oracle_state *s, double *a) {
  const int im = s->c.imax, jm = s->c.jmax;
  for (int j = -G; j <= jm + G; j++)
    for (int i = -G; i <= im + G; i++) {
      if (i >= 0 && i <= im && j >= 0 && j <= jm) continue;
      AT(s, a, i, j) = AT(s, a, src_index(i, im, s->c.bc_x), src_index(j, jm, s->c.bc_y));
    }
}

void oracle_set_prims(oracle_state *s, const double *prims14) {
  for (int k = 0; k < OR_NPRIM; k++) {
    memcpy(s->prim[k], prims14 + (size_t)k * s->n, s->n * sizeof(double));
    fill_ghosts(s, s->prim[k]);
  }
}

void oracle_get_fields(const oracle_state *s, double *prims14, double *cons5) {
  if (prims14)
    for (int k = 0; k < OR_NPRIM; k++) memcpy(prims14 + (size_t)k * s->n, s->prim[k], s->n * sizeof(double));
  if (cons5) {
    const double *c5[5] = {s->D, s->tau, s->Sx, s->Sy, s->Sz};
    for (int k = 0; k < 5; k++) memcpy(cons5 + (size_t)k * s->n, c5[k], s->n * sizeof(double));
  }
}

void oracle_set_cons(oracle_state *s, const double *cons5) {
  double *c5[5] = {s->D, s->tau, s->Sx, s->Sy, s->Sz};
  for (int k = 0; k < 5; k++) memcpy(c5[k], cons5 + (size_t)k * s->n, s->n * sizeof(double));
}

void oracle_get_counters(const oracle_state *s, int64_t counters3[3]) {
  for (int k = 0; k < 3; k++) counters3[k] = s->cnt[k];
}

void oracle_get_rhs(const oracle_state *s, int l, int which, double *out) {
  memcpy(out, s->rhs[l][which], s->n * sizeof(double));
}

/* the face fluxes of the last face pass: dir 0 = x (FX order), 1 = y */
void oracle_get_face(const oracle_state *s, int dir, int slot, double *out) {
  memcpy(out, dir == 0 ? s->fx[slot] : s->fy[slot], s->n * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* startconvar: 06_conserved_var.f95:77-140                            */
/* ------------------------------------------------------------------ */

static void startconvar(oracle_state *s) {
  const double gam = s->c.gamma;
  double **P = s->prim;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      double vx = P[OR_VX][k], vy = P[OR_VY][k], vz = P[OR_VZ][k];
      double ex = P[OR_EX][k], ey = P[OR_EY][k], ez = P[OR_EZ][k];
      double bx = P[OR_BX][k], by = P[OR_BY][k], bz = P[OR_BZ][k];
      double rho = P[OR_RHO][k], p = P[OR_P][k];
      double v2 = (sq(vx) + sq(vy)) + sq(vz);
      double w = 1.0 / sqrt(1.0 - v2);
      double e2 = (sq(ex) + sq(ey)) + sq(ez);
      double b2 = (sq(bx) + sq(by)) + sq(bz);
      double eps = p / ((gam - 1.0) * rho);
      double h = rho * (1.0 + eps) + p;
      s->W[k] = w;
      s->D[k] = rho * w;
      s->tau[k] = (0.5 * (e2 + b2) + h * sq(w)) - p;
      s->Sx[k] = (bz * ey - by * ez) + (h * sq(w)) * vx;
      s->Sy[k] = (bx * ez - bz * ex) + (h * sq(w)) * vy;
      s->Sz[k] = (by * ex - bx * ey) + (h * sq(w)) * vz;
    }
}

/* U^n of the 14 accumulated fields, in AC_* order */
static const double *un_plane(const oracle_state *s, int a) {
  switch (a) {
    case AC_PSI: return s->prim[OR_PSI];
    case AC_PHI: return s->prim[OR_PHI];
    case AC_BX: return s->prim[OR_BX];
    case AC_BY: return s->prim[OR_BY];
    case AC_BZ: return s->prim[OR_BZ];
    case AC_EX: return s->prim[OR_EX];
    case AC_EY: return s->prim[OR_EY];
    case AC_EZ: return s->prim[OR_EZ];
    case AC_Q: return s->prim[OR_Q];
    case AC_D: return s->D;
    case AC_TAU: return s->tau;
    case AC_SX: return s->Sx;
    case AC_SY: return s->Sy;
    default: return s->Sz;
  }
}

/* stage plane that receives U^n + dt*intsum (10_varint.f95:23-63) */
static const int ac_to_st[AC_N] = {ST_PSI, ST_PHI, ST_BX, ST_BY, ST_BZ, ST_EXA, ST_EYA, ST_EZA,
                                   ST_Q, ST_D, ST_TAU, ST_SX, ST_SY, ST_SZ};

/* ------------------------------------------------------------------ */
/* cleanvarint (07_clean_var.f95:18-60), sumint (09_sumint.f95:18-61),  */
/* varint (10_varint.f95:18-63), sumfin (18_sumfin.f95:20-60),          */
/* varfin (19_varfin.f95:35-50), electricfield (20_electricfield.f95:157-163) */
/* ------------------------------------------------------------------ */

static void cleanvarint(oracle_state *s) {
  for (int a = 0; a < AC_N; a++)
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) {
        AT(s, s->intsum[a], i, j) = 0.0;
        AT(s, s->finsum[a], i, j) = 0.0;
      }
}

static void sumint(oracle_state *s, int l, int m) {
  const double ab = s->c.Ab[l][m], abt = s->c.Abt[l][m];
  for (int a = 0; a < AC_N; a++) {
    double *acc = s->intsum[a];
    const double *f = s->rhs[m][a];
    const double *src = (a >= AC_EX && a <= AC_EZ) ? s->rhs[m][14 + (a - AC_EX)] : NULL;
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) {
        size_t k = IX(s, i, j);
        if (src)
          acc[k] = acc[k] + (ab * f[k] - abt * src[k]);
        else
          acc[k] = acc[k] + ab * f[k];
      }
  }
}

static void varint(oracle_state *s) {
  const double dt = s->c.dt;
  for (int a = 0; a < AC_N; a++) {
    const double *u = un_plane(s, a);
    double *dst = s->st[ac_to_st[a]];
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) {
        size_t k = IX(s, i, j);
        dst[k] = u[k] + dt * s->intsum[a][k];
      }
  }
  /* boundary_conserved: psi phi B q D tau S (24_boundary_conserved.f95) */
  static const int cons[11] = {ST_PSI, ST_PHI, ST_BX, ST_BY, ST_BZ, ST_Q, ST_D, ST_TAU, ST_SX, ST_SY, ST_SZ};
  for (int k = 0; k < 11; k++) fill_ghosts(s, s->st[cons[k]]);
}

/* MIRK coefficients, parameters.f95:233-235: cm_1 = -0.05, cm_2 = 1,
 * cm_4 = 0.5 (1 - cm_1)**2 / cm_1 */
static const double CM1 = -0.05, CM2 = 1.0;
static double cm4(void) { return (0.5 * sq(1.0 - CM1)) / CM1; }

/* electricvarint, MIRK branches (DIM = 2): 11_elecvarint.f95:210-277.
 * MIRK = 1: the stage E is E^n.  MIRK = 2: stage 1 takes the explicit
 * value E*; later stages the linearly implicit update from E^n, v^n, B^n,
 * W^n and the stage sum of the E equation. */
static void electricvarint_mirk(oracle_state *s, int l) {
  const double sig = s->c.sigma, dt = s->c.dt;
  double **P = s->prim, **S = s->st;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      if (s->c.mirk == 1) {
        S[ST_EX][k] = P[OR_EX][k]; S[ST_EY][k] = P[OR_EY][k]; S[ST_EZ][k] = P[OR_EZ][k];
        continue;
      }
      if (l == 0) {
        S[ST_EX][k] = S[ST_EXA][k]; S[ST_EY][k] = S[ST_EYA][k]; S[ST_EZ][k] = S[ST_EZA][k];
        continue;
      }
      double w = s->W[k], vx = P[OR_VX][k], vy = P[OR_VY][k], vz = P[OR_VZ][k];
      double ex = P[OR_EX][k], ey = P[OR_EY][k], ez = P[OR_EZ][k];
      double bx = P[OR_BX][k], by = P[OR_BY][k], bz = P[OR_BZ][k];
      double sb = (dt * sig) * w;
      double edv = (ex * vx + ey * vy) + ez * vz;
      double rx = bz * vy - by * vz, ry = bx * vz - bz * vx, rz = by * vx - bx * vy;
      double inv = 1.0 / (1.0 + sb * (1.0 - CM1));
      S[ST_EX][k] = inv * ((((ex * (1.0 - sb * CM1)) + (sb * vx) * edv) - sb * rx) + dt * s->intsum[AC_EX][k]);
      S[ST_EY][k] = inv * ((((ey * (1.0 - sb * CM1)) + (sb * vy) * edv) - sb * ry) + dt * s->intsum[AC_EY][k]);
      S[ST_EZ][k] = inv * ((((ez * (1.0 - sb * CM1)) + (sb * vz) * edv) - sb * rz) + dt * s->intsum[AC_EZ][k]);
    }
  fill_ghosts(s, S[ST_EX]);
  fill_ghosts(s, S[ST_EY]);
  fill_ghosts(s, S[ST_EZ]);
}

/* electricvarint, IMEX branch: 11_elecvarint.f95:279-328 (REC_PRIM = 1: :331-378) */
static void electricvarint(oracle_state *s, int l) {
  if (s->c.mirk) {
    electricvarint_mirk(s, l);
    return;
  }
  const double abt = s->c.Abt[l][l], sig = s->c.sigma, dt = s->c.dt;
  double **P = s->prim, **S = s->st;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      double w = s->W[k], vx = P[OR_VX][k], vy = P[OR_VY][k], vz = P[OR_VZ][k];
      if (s->c.rec_prim) { /* :331-378: the stage's primitives before its recovery */
        w = s->wint[l][k];
        vx = s->vint[l][0][k]; vy = s->vint[l][1][k]; vz = s->vint[l][2][k];
      }
      double bx = S[ST_BX][k], by = S[ST_BY][k], bz = S[ST_BZ][k];
      double eax = S[ST_EXA][k], eay = S[ST_EYA][k], eaz = S[ST_EZA][k];
      double a = (abt * sig) * dt;
      double matx = ((sq(w) * a + w * sq(a)) + w) + a;
      double rx = bz * vy - by * vz;
      double ry = bx * vz - bz * vx;
      double rz = by * vx - bx * vy;
      double aw2 = a * sq(w);
      double kk = ((abt * sig) * w) * dt;
      double txx = (a + w) + aw2 * sq(vx);
      double tyy = (a + w) + aw2 * sq(vy);
      double tzz = (a + w) + aw2 * sq(vz);
      double inv = 1.0 / matx;
      S[ST_EX][k] = inv * (((txx * eax + ((aw2 * vx) * vy) * eay) + ((aw2 * vx) * vz) * eaz) -
                           kk * ((txx * rx + ((aw2 * vx) * vy) * ry) + ((aw2 * vx) * vz) * rz));
      S[ST_EY][k] = inv * (((((aw2 * vx) * vy) * eax + tyy * eay) + ((aw2 * vy) * vz) * eaz) -
                           kk * ((((aw2 * vx) * vy) * rx + tyy * ry) + ((aw2 * vy) * vz) * rz));
      S[ST_EZ][k] = inv * (((((aw2 * vx) * vz) * eax + ((aw2 * vy) * vz) * eay) + tzz * eaz) -
                           kk * ((((aw2 * vx) * vz) * rx + ((aw2 * vy) * vz) * ry) + tzz * rz));
    }
  fill_ghosts(s, S[ST_EX]);
  fill_ghosts(s, S[ST_EY]);
  fill_ghosts(s, S[ST_EZ]);
}

/* flowinit, current density only: 14_flowinit.f95:258-301 (REC_PRIM = 1:
 * :376-420, from the stage's primitives) */
static void flowinit(oracle_state *s, int l) {
  const double sig = s->c.sigma;
  double **P = s->prim, **S = s->st;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      double w = s->W[k], vx = P[OR_VX][k], vy = P[OR_VY][k], vz = P[OR_VZ][k];
      if (s->c.rec_prim) {
        w = s->wint[l][k];
        vx = s->vint[l][0][k]; vy = s->vint[l][1][k]; vz = s->vint[l][2][k];
      }
      double ex = S[ST_EX][k], ey = S[ST_EY][k], ez = S[ST_EZ][k];
      double bx = S[ST_BX][k], by = S[ST_BY][k], bz = S[ST_BZ][k], q = S[ST_Q][k];
      double edv = (ex * vx + ey * vy) + ez * vz;
      double rx = bz * vy - by * vz, ry = bx * vz - bz * vx, rz = by * vx - bx * vy;
      S[ST_JX][k] = (sig * w) * ((ex + rx) - edv * vx) + q * vx;
      S[ST_JY][k] = (sig * w) * ((ey + ry) - edv * vy) + q * vy;
      S[ST_JZ][k] = (sig * w) * ((ez + rz) - edv * vz) + q * vz;
    }
  /* 24_boundary_flux_conserved.f95:13: ghosts only for HLL and HLLC; with HLLE
   * the J ghosts keep their zero initial value. */
  if (s->c.flux_solver == OR_HLL || s->c.flux_solver == OR_HLLC) {
    fill_ghosts(s, S[ST_JX]);
    fill_ghosts(s, S[ST_JY]);
    fill_ghosts(s, S[ST_JZ]);
    /* and the stage velocity (24_boundary_flux_conserved.f95:1426-1560 and
     * the like); the stage rho and p keep zero ghosts, which no routine of
     * the reference fills */
    if (s->c.rec_prim)
      for (int d = 0; d < 3; d++) fill_ghosts(s, s->vint[l][d]);
  }
}

/* ------------------------------------------------------------------ */
/* MP5 / MP7 / MP9 along one line: mp5.f95:40-264, mp7.f95, mp9.f95.   */
/* u, up, um are indexed -6..n+6 (pointers to element 0); up/um are    */
/* written on [-2, n+2].                                                */
/* ------------------------------------------------------------------ */

static inline double dmm(double x, double y) { /* funciones.f95:148-162 */
  return (0.5 * (copysign(1.0, x) + copysign(1.0, y))) * fmin2(fabs(x), fabs(y));
}

static inline double dm4(double w, double x, double y, double z) { /* funciones.f95:165-185 */
  double sw = copysign(1.0, w);
  double m = fmin2(fmin2(fmin2(fabs(w), fabs(x)), fabs(y)), fabs(z));
  return ((0.125 * (sw + copysign(1.0, x))) * fabs((sw + copysign(1.0, y)) * (sw + copysign(1.0, z)))) * m;
}

static inline double vor_plus(int order, const double *u, int i) {
  if (order == 5) /* mp5.f95:79-80 */
    return (1.0 / 60.0) * ((((2.0 * u[i - 2] - 13.0 * u[i - 1]) + 47.0 * u[i]) + 27.0 * u[i + 1]) - 3.0 * u[i + 2]);
  if (order == 7) /* mp7.f95:111-113 */
    return (1.0 / 420.0) * ((((((-3.0 * u[i - 3] + 25.0 * u[i - 2]) - 101.0 * u[i - 1]) + 319.0 * u[i]) +
                              214.0 * u[i + 1]) - 38.0 * u[i + 2]) + 4.0 * u[i + 3]);
  /* mp9.f95:116-118 */
  return (1.0 / 2520.0) * ((((((((4.0 * u[i - 4] - 41.0 * u[i - 3]) + 199.0 * u[i - 2]) - 641.0 * u[i - 1]) +
                               1879.0 * u[i]) + 1375.0 * u[i + 1]) - 305.0 * u[i + 2]) + 55.0 * u[i + 3]) -
                           5.0 * u[i + 4]);
}

static inline double vor_minus(int order, const double *u, int i) {
  if (order == 5) /* mp5.f95:82-83 */
    return (1.0 / 60.0) * ((((-3.0 * u[i - 2] + 27.0 * u[i - 1]) + 47.0 * u[i]) - 13.0 * u[i + 1]) + 2.0 * u[i + 2]);
  if (order == 7) /* mp7.f95:115-117 */
    return (1.0 / 420.0) * ((((((-3.0 * u[i + 3] + 25.0 * u[i + 2]) - 101.0 * u[i + 1]) + 319.0 * u[i]) +
                              214.0 * u[i - 1]) - 38.0 * u[i - 2]) + 4.0 * u[i - 3]);
  /* mp9.f95:121-123 */
  return (1.0 / 2520.0) * ((((((((4.0 * u[i + 4] - 41.0 * u[i + 3]) + 199.0 * u[i + 2]) - 641.0 * u[i + 1]) +
                               1879.0 * u[i]) + 1375.0 * u[i - 1]) - 305.0 * u[i - 2]) + 55.0 * u[i - 3]) -
                           5.0 * u[i - 4]);
}

void oracle_mp_line(int order, const double *u, int n, double *up, double *um, int positivity) {
  const double b2 = 4.0 / 3.0, epsm = 1.0e-10, alpha = 2.0;
  const int ext = order == 5 ? 3 : (order == 7 ? 4 : 5); /* du/d2 range mp5.f95:64, mp7.f95:69, mp9.f95:72 */
  double *dub = scratch_doubles((size_t)(n + 32));
  double *d2b = scratch_doubles((size_t)(n + 32));
  double *du = dub + 8, *d2 = d2b + 8;
  for (int i = -ext; i <= n + ext; i++) {
    du[i] = u[i] - u[i - 1];
    d2[i] = (u[i - 1] - 2.0 * u[i]) + u[i + 1];
  }
  for (int i = -2; i <= n + 2; i++) {
    double u0 = u[i];
    double vorp = vor_plus(order, u, i), vorm = vor_minus(order, u, i);
    double vmpp = u0 + dmm(du[i + 1], alpha * du[i]);
    double vmpm = u0 + dmm(-du[i], -(alpha * du[i + 1]));
    /* U^L_{i+1/2}: mp5.f95:104-136 */
    if ((vorp - u0) * (vorp - vmpp) <= epsm) {
      up[i] = vorp;
    } else {
      double dm4jph = dm4(4.0 * d2[i] - d2[i + 1], 4.0 * d2[i + 1] - d2[i], d2[i], d2[i + 1]);
      double dm4jmh = dm4(4.0 * d2[i] - d2[i - 1], 4.0 * d2[i - 1] - d2[i], d2[i], d2[i - 1]);
      double vul = u0 + alpha * du[i];
      double vav = 0.5 * (u0 + u[i + 1]);
      double vmd = vav - 0.5 * dm4jph;
      double vlc = (u0 + 0.5 * du[i]) + b2 * dm4jmh;
      double vmin = fmax2(fmin2(fmin2(u0, u[i + 1]), vmd), fmin2(fmin2(u0, vul), vlc));
      double vmax = fmin2(fmax2(fmax2(u0, u[i + 1]), vmd), fmax2(fmax2(u0, vul), vlc));
      up[i] = vorp + dmm(vmin - vorp, vmax - vorp);
    }
    /* U^R_{i-1/2}: mp5.f95:142-167 */
    if ((vorm - u0) * (vorm - vmpm) <= epsm) {
      um[i] = vorm;
    } else {
      double dm4jph = dm4(4.0 * d2[i] - d2[i + 1], 4.0 * d2[i + 1] - d2[i], d2[i], d2[i + 1]);
      double dm4jmh = dm4(4.0 * d2[i] - d2[i - 1], 4.0 * d2[i - 1] - d2[i], d2[i], d2[i - 1]);
      double vul = u0 - alpha * du[i + 1];
      double vav = 0.5 * (u0 + u[i - 1]);
      double vmd = vav - 0.5 * dm4jmh;
      double vlc = (u0 - 0.5 * du[i + 1]) + b2 * dm4jph;
      double vmin = fmax2(fmin2(fmin2(u0, u[i - 1]), vmd), fmin2(fmin2(u0, vul), vlc));
      double vmax = fmin2(fmax2(fmax2(u0, u[i - 1]), vmd), fmax2(fmax2(u0, vul), vlc));
      um[i] = vorm + dmm(vmin - vorm, vmax - vorm);
    }
  }
  if (positivity) { /* mp5.f95:177-216 (slots 24/25 and 61/62: p and rho) */
    for (int i = -2; i <= n + 2; i++)
      if (up[i] < 0.0) {
        up[i] = u[i] + 0.5 * dmm(du[i], du[i + 1]);
        um[i] = u[i];
      }
    for (int i = -2; i <= n + 2; i++)
      if (um[i] < 0.0) {
        um[i] = u[i] - 0.5 * dmm(du[i], du[i + 1]);
        up[i] = u[i];
      }
  }
  free(dub);
  free(d2b);
}

/* ------------------------------------------------------------------ */
/* pack / MPx / unpack: 16_hll_flow_mpx.f95:781-1352 (same in hllc/hlle) */
/* ------------------------------------------------------------------ */

static double slot_x(const oracle_state *s, int slot, size_t k) {
  double *const *S = s->st;
  switch (slot) {
    case XS_PSI: return S[ST_PSI][k];
    case XS_PHI: return S[ST_PHI][k];
    case XS_EX: return S[ST_EX][k];
    case XS_EY: return S[ST_EY][k];
    case XS_EZ: return S[ST_EZ][k];
    case XS_EXM: return -S[ST_EX][k];
    case XS_EYM: return -S[ST_EY][k];
    case XS_EZM: return -S[ST_EZ][k];
    case XS_BX: return S[ST_BX][k];
    case XS_BY: return S[ST_BY][k];
    case XS_BZ: return S[ST_BZ][k];
    case XS_BXM: return -S[ST_BX][k];
    case XS_BYM: return -S[ST_BY][k];
    case XS_BZM: return -S[ST_BZ][k];
    case XS_Q: return S[ST_Q][k];
    case XS_D: return S[ST_D][k];
    case XS_TAU: return S[ST_TAU][k];
    case XS_SX: return S[ST_SX][k];
    case XS_SY: return S[ST_SY][k];
    case XS_SZ: return S[ST_SZ][k];
    case XS_VX: return s->kin[K_VX][k];
    case XS_VY: return s->kin[K_VY][k];
    case XS_VZ: return s->kin[K_VZ][k];
    case XS_P: return s->kin[K_P][k];
    case XS_RHO: return s->kin[K_RHO][k];
    case XS_JX: return S[ST_JX][k];
    default: return S[ST_JZ][k];
  }
}

static double slot_y(const oracle_state *s, int slot, size_t k) {
  if (slot == YS_JY) return s->st[ST_JY][k];
  return slot_x(s, slot, k); /* slots 38..62 repeat the x order (1..25) */
}

static void reconstruct(oracle_state *s) {
  const int im = s->c.imax, jm = s->c.jmax, ord = s->c.rec_order;
  const int nmax = (im > jm ? im : jm) + 2 * G + 1;
  double *ub = scratch_doubles((size_t)nmax);
  double *upb = scratch_doubles((size_t)nmax);
  double *umb = scratch_doubles((size_t)nmax);
  double *u = ub + G, *up = upb + G, *um = umb + G;
  /* x sweep, rows -1..jmax (the row jmax+1 of the reference is never read) */
  for (int j = -1; j <= jm; j++)
    for (int sl = 0; sl < NXS; sl++) {
      for (int i = -G; i <= im + G; i++) u[i] = slot_x(s, sl, IX(s, i, j));
      oracle_mp_line(ord, u, im, up, um, sl == XS_P || sl == XS_RHO);
      for (int i = -1; i <= im + 1; i++) {
        AT(s, s->xL[sl], i, j) = up[i];
        AT(s, s->xR[sl], i, j) = um[i + 1];
      }
    }
  /* y sweep, columns -1..imax */
  for (int i = -1; i <= im; i++)
    for (int sl = 0; sl < NYS; sl++) {
      for (int j = -G; j <= jm + G; j++) u[j] = slot_y(s, sl, IX(s, i, j));
      oracle_mp_line(ord, u, jm, up, um, sl == YS_P || sl == YS_RHO);
      for (int j = -1; j <= jm + 1; j++) {
        AT(s, s->yL[sl], i, j) = up[j];
        AT(s, s->yR[sl], i, j) = um[j + 1];
      }
    }
  free(ub); free(upb); free(umb);
}

/* ------------------------------------------------------------------ */
/* face loop helpers                                                    */
/* ------------------------------------------------------------------ */

/* one reconstructed face side */
typedef struct {
  double psi, phi, ex, ey, ez, exm, eym, ezm, bx, by, bz, bxm, bym, bzm, q, d, tau, sx, sy, sz;
  double vx, vy, vz, p, rho, j1, j3; /* j1 = normal J, j3 = Jz (x faces only) */
  double w, e2b2, h;
  double fd, ftau, fs1, fs2, fs3; /* normal-direction physical fluxes */
} side_t;

static void load_x(const oracle_state *s, double *const *arr, size_t k, side_t *o) {
  o->psi = arr[XS_PSI][k]; o->phi = arr[XS_PHI][k];
  o->ex = arr[XS_EX][k]; o->ey = arr[XS_EY][k]; o->ez = arr[XS_EZ][k];
  o->exm = arr[XS_EXM][k]; o->eym = arr[XS_EYM][k]; o->ezm = arr[XS_EZM][k];
  o->bx = arr[XS_BX][k]; o->by = arr[XS_BY][k]; o->bz = arr[XS_BZ][k];
  o->bxm = arr[XS_BXM][k]; o->bym = arr[XS_BYM][k]; o->bzm = arr[XS_BZM][k];
  o->q = arr[XS_Q][k]; o->d = arr[XS_D][k]; o->tau = arr[XS_TAU][k];
  o->sx = arr[XS_SX][k]; o->sy = arr[XS_SY][k]; o->sz = arr[XS_SZ][k];
  o->vx = arr[XS_VX][k]; o->vy = arr[XS_VY][k]; o->vz = arr[XS_VZ][k];
  o->p = arr[XS_P][k]; o->rho = arr[XS_RHO][k];
  o->j1 = arr[XS_JX][k]; o->j3 = arr[XS_JZ][k];
  (void)s;
}

static void load_y(const oracle_state *s, double *const *arr, size_t k, side_t *o) {
  o->psi = arr[YS_PSI][k]; o->phi = arr[YS_PHI][k];
  o->ex = arr[YS_EX][k]; o->ey = arr[YS_EY][k]; o->ez = arr[YS_EZ][k];
  o->exm = arr[YS_EXM][k]; o->eym = arr[YS_EYM][k]; o->ezm = arr[YS_EZM][k];
  o->bx = arr[YS_BX][k]; o->by = arr[YS_BY][k]; o->bz = arr[YS_BZ][k];
  o->bxm = arr[YS_BXM][k]; o->bym = arr[YS_BYM][k]; o->bzm = arr[YS_BZM][k];
  o->q = arr[YS_Q][k]; o->d = arr[YS_D][k]; o->tau = arr[YS_TAU][k];
  o->sx = arr[YS_SX][k]; o->sy = arr[YS_SY][k]; o->sz = arr[YS_SZ][k];
  o->vx = arr[YS_VX][k]; o->vy = arr[YS_VY][k]; o->vz = arr[YS_VZ][k];
  o->p = arr[YS_P][k]; o->rho = arr[YS_RHO][k];
  o->j1 = arr[YS_JY][k]; o->j3 = 0.0;
  (void)s;
}

/* W, E^2+B^2, enthalpy and the normal physical fluxes of one face side:
 * 16_hll_flow_mpx.f95:1463-1466 (W), 1522-1532 (E2B2), 1603-1617 (eps, h),
 * 1627-1715 (FD, Ftau, FS) */
static void side_physics(double gam, int dir, int hlle, side_t *o) {
  double v2 = (sq(o->vx) + sq(o->vy)) + sq(o->vz);
  /* HLLE subtracts the squares one by one (16_hlle_flow_mpx.f95:2302-2305) */
  o->w = hlle ? 1.0 / sqrt(((1.0 - sq(o->vx)) - sq(o->vy)) - sq(o->vz)) : 1.0 / sqrt(1.0 - v2);
  o->e2b2 = ((((sq(o->ex) + sq(o->ey)) + sq(o->ez)) + sq(o->bx)) + sq(o->by)) + sq(o->bz);
  double eps = o->p / ((gam - 1.0) * o->rho);
  o->h = o->rho * (1.0 + eps) + o->p;
  double hw2 = o->h * sq(o->w);
  if (dir == 0) {
    o->fd = o->d * o->vx;
    o->ftau = (o->bz * o->ey - o->by * o->ez) + hw2 * o->vx;
    o->fs1 = (((-sq(o->ex)) - sq(o->bx)) + hw2 * sq(o->vx)) + 0.5 * o->e2b2 + o->p;
    o->fs2 = ((-(o->ex * o->ey)) - o->bx * o->by) + (hw2 * o->vx) * o->vy;
    o->fs3 = ((-(o->ex * o->ez)) - o->bx * o->bz) + (hw2 * o->vx) * o->vz;
  } else {
    o->fd = o->d * o->vy;
    o->ftau = (o->bx * o->ez - o->bz * o->ex) + hw2 * o->vy;
    o->fs1 = ((-(o->ex * o->ey)) - o->bx * o->by) + (hw2 * o->vx) * o->vy; /* FSyx */
    o->fs2 = (((-sq(o->ey)) - sq(o->by)) + hw2 * sq(o->vy)) + 0.5 * o->e2b2 + o->p; /* FSyy */
    o->fs3 = ((-(o->ey * o->ez)) - o->by * o->bz) + (hw2 * o->vy) * o->vz; /* FSyz */
  }
}

/* (s2*FL - s1*FR + s1*s2*(UR - UL))/(s2 - s1) with s1 = -1, s2 = 1
 * (parameters.f95:283-284; 16_hll_flow_mpx.f95:1733-1958) */
static const double S1 = -1.0, S2 = 1.0;
static inline double hllf(double fl, double fr, double ul, double ur) {
  return ((S2 * fl - S1 * fr) + (S1 * S2) * (ur - ul)) / (S2 - S1);
}

/* HLL face fluxes: 16_hll_flow_mpx.f95:1738-1958 */
static void hll_x(const side_t *L, const side_t *R, double f[FX_N]) {
  f[FX_EXPSI] = hllf(L->ex, R->ex, L->psi, R->psi);
  f[FX_BXPHI] = hllf(L->bx, R->bx, L->phi, R->phi);
  f[FX_PHIBX] = hllf(L->phi, R->phi, L->bx, R->bx);
  f[FX_EZBY] = -hllf(L->ezm, R->ezm, L->by, R->by);
  f[FX_EYBZ] = hllf(L->ey, R->ey, L->bz, R->bz);
  f[FX_PSIEX] = hllf(L->psi, R->psi, L->ex, R->ex);
  f[FX_BZEY] = hllf(L->bz, R->bz, L->ey, R->ey);
  f[FX_BYEZ] = -hllf(L->bym, R->bym, L->ez, R->ez);
  f[FX_J] = hllf(L->j1, R->j1, L->q, R->q);
  f[FX_FD] = hllf(L->fd, R->fd, L->d, R->d);
  f[FX_FTAU] = hllf(L->ftau, R->ftau, L->tau, R->tau);
  f[FX_FSX] = hllf(L->fs1, R->fs1, L->sx, R->sx);
  f[FX_FSY] = hllf(L->fs2, R->fs2, L->sy, R->sy);
  f[FX_FSZ] = hllf(L->fs3, R->fs3, L->sz, R->sz);
}

static void hll_y(const side_t *L, const side_t *R, double f[FY_N]) {
  f[FY_EYPSI] = hllf(L->ey, R->ey, L->psi, R->psi);
  f[FY_BYPHI] = hllf(L->by, R->by, L->phi, R->phi);
  f[FY_EZBX] = hllf(L->ez, R->ez, L->bx, R->bx);
  f[FY_PHIBY] = hllf(L->phi, R->phi, L->by, R->by);
  f[FY_EXBZ] = -hllf(L->exm, R->exm, L->bz, R->bz);
  f[FY_BZEX] = -hllf(L->bzm, R->bzm, L->ex, R->ex);
  f[FY_PSIEY] = hllf(L->psi, R->psi, L->ey, R->ey);
  f[FY_BXEZ] = hllf(L->bx, R->bx, L->ez, R->ez);
  f[FY_J] = hllf(L->j1, R->j1, L->q, R->q);
  f[FY_FD] = hllf(L->fd, R->fd, L->d, R->d);
  f[FY_FTAU] = hllf(L->ftau, R->ftau, L->tau, R->tau);
  f[FY_FSX] = hllf(L->fs1, R->fs1, L->sx, R->sx);
  f[FY_FSY] = hllf(L->fs2, R->fs2, L->sy, R->sy);
  f[FY_FSZ] = hllf(L->fs3, R->fs3, L->sz, R->sz);
}

/* get_hllc_vel, funciones.f95:188-259: Newton on a*v^2 - b*v + c = 0 in
 * [-1,1].  The reference evaluates the first Newton step at its iterate v
 * before assigning it (funciones.f95:222-232); v0 is that point.  The seed
 * mode passes v0 = clamp(rt); the reference-build mode passes the value the
 * flang -O2 build leaves in that register (see hllc_v0_x/hllc_v0_y). */
/* max(lo, min(hi, x)) as the reference build evaluates it: a NaN passes through */
static inline double clampf(double x, double lo, double hi) {
  double t = x > hi ? hi : x;
  return t < lo ? lo : t;
}

/* last Newton iterate of the latest hllc_vel call, before the final clamp */
static double g_hllc_last_iterate = 0.0;
/* HLLC with the MC limiter: first Newton points of its x and y solves (see
 * hllc_mcl_faces); NAN when not in use */
static double g_mcl_seed_x = NAN;
static int g_mcl_seed_y = 0;

static int g_hllc_maxiter = 100; /* 200 in 16_hlle_flow_mcl.f95:4385 */

/* set when get_hllc_vel reaches its iteration limit: the reference prints
 * "[GET_HLLC_VEL]: Non convergence" and stops (funciones.f95:238-244).  A
 * NaN iterate ends the loop instead (err > minerr is false) and NaN is
 * returned, which the callers' branches on v* then stop on. */
static int g_hllc_fail = 0;

double oracle_hllc_vel(double a, double b, double c, double rt, double v0) {
  double oldv = clampf(rt, -1.0, 1.0);
  double v = v0;
  int iter = 0;
  for (;;) {
    double fv = (a * v * v - b * v) + c;
    double dfdv = (2.0 * a) * v - b;
    v = clampf(oldv - fv / dfdv, -1.0, 1.0);
    double err = fabs(oldv - v);
    oldv = v;
    iter = iter + 1;
    if (iter >= g_hllc_maxiter) { /* the reference stops here */
      g_hllc_fail = 1;
      return NAN;
    }
    if (!(err > 1.0e-12)) break;
  }
  g_hllc_last_iterate = v;
  return clampf(v, -1.0 + 1.0e-3, 1.0 - 1.0e-3);
}

/* state (0.5-weighted) HLL average and flux, 16_hllc_flow_mpx.f95:2811-3095 */
static inline double hll_state(double ul, double ur, double fl, double fr) {
  return (((S2 * ur - S1 * ul) + fl) - fr) / (S2 - S1);
}
static inline double hll_flux(double fl, double fr, double ul, double ur) {
  return 0.5 * ((fl + fr) - (ur - ul));
}

/* |v*| at or below which the contact speed is set to 0: 1e-9 in
 * 16_hllc_flow_mpx.f95:3322, 1e-6 in 16_hllc_flow_mcl.f95:2596 */
static double g_hllc_zero = 1.0e-9;

/* contact speed of 16_hllc_flow_mpx.f95:3185-3334.  v0 is the point of the
 * first Newton evaluation (see oracle_hllc_vel). */
static double hllc_contact(double a, double b, double c, double v0) {
  g_hllc_fail = 0;
  double det = sq(b) - (4.0 * a) * c;
  double qr = -(0.5 * (b + copysign(1.0, b) * fabs(sqrt(det))));
  double rt2 = c / qr;
  if (det < 0.0) rt2 = c / b;
  double v = oracle_hllc_vel(a, -b, c, rt2, v0 == v0 ? v0 : clampf(rt2, -1.0, 1.0));
  if (fabs(v) <= g_hllc_zero)
    v = 0.0;
  else if (fabs(v) > 1.0)
    v = clampf(v, -1.0 + 1.0e-6, 1.0 - 1.0e-6);
  return v;
}

/* HLLC x face: 16_hllc_flow_mpx.f95:2811-4295. Returns 0 or a stop status. */
/* First Newton point of the x-face contact solve in the flang -O2 reference
 * build: the register still holds Bzint_mp5_x_R(i,j).  For the y face it
 * holds the negated b coefficient, i.e. the b argument itself. */
static int g_hllc_seed_mode = 1; /* 1 reference build, 0 clamped seed */

static int hllc_x(const side_t *L, const side_t *R, double f[FX_N]) {
  double psi_h = hll_state(L->psi, R->psi, L->ex, R->ex);
  double phi_h = hll_state(L->phi, R->phi, L->bx, R->bx);
  double ex_h = hll_state(L->ex, R->ex, L->psi, R->psi);
  double ey_h = hll_state(L->ey, R->ey, L->bz, R->bz);
  double ez_h = hll_state(L->ez, R->ez, L->bym, R->bym);
  double bx_h = hll_state(L->bx, R->bx, L->phi, R->phi);
  double by_h = hll_state(L->by, R->by, L->ezm, R->ezm);
  double bz_h = hll_state(L->bz, R->bz, L->ey, R->ey);
  double dd_h = hll_state(L->d, R->d, L->fd, R->fd);
  double tau_h = hll_state(L->tau, R->tau, L->ftau, R->ftau);
  double sx_h = hll_state(L->sx, R->sx, L->fs1, R->fs1);
  double sy_h = hll_state(L->sy, R->sy, L->fs2, R->fs2);
  double sz_h = hll_state(L->sz, R->sz, L->fs3, R->fs3);

  double psi_f = hll_flux(L->ex, R->ex, L->psi, R->psi);
  double phi_f = hll_flux(L->bx, R->bx, L->phi, R->phi);
  double bx_f = hll_flux(L->phi, R->phi, L->bx, R->bx);
  double by_f = hll_flux(L->ezm, R->ezm, L->by, R->by);
  double bz_f = hll_flux(L->ey, R->ey, L->bz, R->bz);
  double ex_f = hll_flux(L->psi, R->psi, L->ex, R->ex);
  double ey_f = hll_flux(L->bz, R->bz, L->ey, R->ey);
  double ez_f = hll_flux(L->bym, R->bym, L->ez, R->ez);
  double q_f = hll_flux(L->j1, R->j1, L->q, R->q);
  double dd_f = hll_flux(L->fd, R->fd, L->d, R->d);
  double tau_f = hll_flux(L->ftau, R->ftau, L->tau, R->tau);
  double sx_f = hll_flux(L->fs1, R->fs1, L->sx, R->sx);
  double sy_f = hll_flux(L->fs2, R->fs2, L->sy, R->sy);
  double sz_f = hll_flux(L->fs3, R->fs3, L->sz, R->sz);

  /* star E, B, q are the HLL states on both sides (:3104-3118) */
  double erx = bz_h * ey_h - by_h * ez_h;
  double ery = bx_h * ez_h - bz_h * ex_h;
  double erz = by_h * ex_h - bx_h * ey_h;
  double e2p = sq(ey_h) + sq(ez_h);
  double b2p = sq(by_h) + sq(bz_h);
  double e2b2s = 0.5 * (((sq(ex_h) + sq(ey_h)) + sq(ez_h)) + ((sq(bx_h) + sq(by_h)) + sq(bz_h)));
  double a = tau_f - erx;
  double b = ((e2p + b2p) - tau_h) - sx_f;
  double c = sx_h - erx;
  double v = hllc_contact(a, b, c, g_mcl_seed_x == g_mcl_seed_x ? g_mcl_seed_x : (g_hllc_seed_mode ? R->bz : NAN));
  if (g_hllc_fail) return 7;

  double pstr = (sx_f + (sq(ex_h) + sq(bx_h))) - (tau_f - erx) * v;
  double psiR = (S2 * psi_h - v * ex_f) / (S2 - v);
  double psiL = (-(S1 * psi_h) + v * ex_f) / (v - S1);
  double phiR = (S2 * phi_h - v * bx_f) / (S2 - v);
  double phiL = (-(S1 * phi_h) + v * bx_f) / (v - S1);
  double sxR = (S2 * sx_h - tau_f * v) / (S2 - v);
  double sxL = (-(S1 * sx_h) + tau_f * v) / (v - S1);
  double syR = ((((S2 * sy_h - sy_f) - ery * v) - ex_h * ey_h) - bx_h * by_h) / (S2 - v);
  double syL = ((((-(S1 * sy_h) + sy_f) + ery * v) + ex_h * ey_h) + bx_h * by_h) / (v - S1);
  double szR = ((((S2 * sz_h - sz_f) - erz * v) - ex_h * ez_h) - bx_h * bz_h) / (S2 - v);
  double szL = ((((-(S1 * sz_h) + sz_f) + erz * v) + ex_h * ez_h) + bx_h * bz_h) / (v - S1);
  double tauR = ((sxR - erx) / v - pstr) + e2b2s;
  double tauL = ((sxL - erx) / v - pstr) + e2b2s;
  double ddR = (S2 * dd_h - dd_f) / (S2 - v);
  double ddL = (-(S1 * dd_h) + dd_f) / (v - S1);

  double psifR, psifL, phifR, phifL, fdR, fdL, ftR, ftL, fsxR, fsxL, fsyR, fsyL, fszR, fszL;
  if (v == 0.0) {
    psifR = psifL = psi_f; phifR = phifL = phi_f;
    fdR = fdL = dd_f; ftR = ftL = tau_f;
    fsxR = fsxL = sx_f; fsyR = fsyL = sy_f; fszR = fszL = sz_f;
  } else {
    double cr = S2 * (v - S1), cl = S1 * (S2 - v), den = S2 - S1;
    psifR = psi_f + (cr * (psiR - psiL)) / den;
    psifL = psi_f - (cl * (psiR - psiL)) / den;
    phifR = phi_f + (cr * (phiR - phiL)) / den;
    phifL = phi_f - (cl * (phiR - phiL)) / den;
    fdR = dd_f + (cr * (ddR - ddL)) / den;
    fdL = dd_f - (cl * (ddR - ddL)) / den;
    ftR = tau_f + (cr * (tauR - tauL)) / den;
    ftL = tau_f - (cl * (tauR - tauL)) / den;
    fsxR = sx_f + (cr * (sxR - sxL)) / den;
    fsxL = sx_f - (cl * (sxR - sxL)) / den;
    fsyR = sy_f + (cr * (syR - syL)) / den;
    fsyL = sy_f - (cl * (syR - syL)) / den;
    fszR = sz_f + (cr * (szR - szL)) / den;
    fszL = sz_f - (cl * (szR - szL)) / den;
  }
  /* selection :3908-4295; E/B star fluxes are the HLL fluxes */
  if (S1 < 0.0 && 0.0 < v) {
    f[FX_EXPSI] = psifL; f[FX_BXPHI] = phifL; f[FX_J] = q_f;
    f[FX_FD] = fdL; f[FX_FTAU] = ftL; f[FX_FSX] = fsxL; f[FX_FSY] = fsyL; f[FX_FSZ] = fszL;
    f[FX_PHIBX] = bx_f; f[FX_EZBY] = -by_f; f[FX_EYBZ] = bz_f;
    f[FX_PSIEX] = ex_f; f[FX_BZEY] = ey_f; f[FX_BYEZ] = -ez_f;
  } else if (v < 0.0 && 0.0 < S2) {
    f[FX_EXPSI] = psifR; f[FX_BXPHI] = phifR; f[FX_J] = q_f;
    f[FX_FD] = fdR; f[FX_FTAU] = ftR; f[FX_FSX] = fsxR; f[FX_FSY] = fsyR; f[FX_FSZ] = fszR;
    f[FX_PHIBX] = bx_f; f[FX_EZBY] = -by_f; f[FX_EYBZ] = bz_f;
    f[FX_PSIEX] = ex_f; f[FX_BZEY] = ey_f; f[FX_BYEZ] = -ez_f;
  } else if (v == 0.0) {
    f[FX_EXPSI] = 0.5 * (psifR + psifL); f[FX_BXPHI] = 0.5 * (phifR + phifL);
    f[FX_PHIBX] = 0.5 * (bx_f + bx_f); f[FX_EZBY] = -(0.5 * (by_f + by_f));
    f[FX_EYBZ] = 0.5 * (bz_f + bz_f); f[FX_PSIEX] = 0.5 * (ex_f + ex_f);
    f[FX_BZEY] = 0.5 * (ey_f + ey_f); f[FX_BYEZ] = -(0.5 * (ez_f + ez_f));
    f[FX_J] = 0.5 * (q_f + q_f); f[FX_FD] = 0.5 * (fdR + fdL); f[FX_FTAU] = 0.5 * (ftR + ftL);
    f[FX_FSX] = 0.5 * (fsxR + fsxL); f[FX_FSY] = 0.5 * (fsyR + fsyL); f[FX_FSZ] = 0.5 * (fszR + fszL);
  } else {
    return 8; /* "Revise Vxstar velocity" stop */
  }
  return 0;
}

/* HLLC y face: 16_hllc_flow_mpx.f95:2950-3095 (HLL), 3338-3489, 3536-3574,
 * 3742-3901 (star), 4305-4606 (selection) */
static int hllc_y(const side_t *L, const side_t *R, double f[FY_N]) {
  double psi_h = hll_state(L->psi, R->psi, L->ey, R->ey);
  double phi_h = hll_state(L->phi, R->phi, L->by, R->by);
  double ey_h = hll_state(L->ey, R->ey, L->psi, R->psi);
  double ez_h = hll_state(L->ez, R->ez, L->bx, R->bx);
  double ex_h = hll_state(L->ex, R->ex, L->bzm, R->bzm);
  double by_h = hll_state(L->by, R->by, L->phi, R->phi);
  double bz_h = hll_state(L->bz, R->bz, L->exm, R->exm);
  double bx_h = hll_state(L->bx, R->bx, L->ez, R->ez);
  double dd_h = hll_state(L->d, R->d, L->fd, R->fd);
  double tau_h = hll_state(L->tau, R->tau, L->ftau, R->ftau);
  double sy_h = hll_state(L->sy, R->sy, L->fs2, R->fs2);
  double sz_h = hll_state(L->sz, R->sz, L->fs3, R->fs3);
  double sx_h = hll_state(L->sx, R->sx, L->fs1, R->fs1);

  double psi_f = hll_flux(L->ey, R->ey, L->psi, R->psi);
  double phi_f = hll_flux(L->by, R->by, L->phi, R->phi);
  double by_f = hll_flux(L->phi, R->phi, L->by, R->by);
  double bz_f = hll_flux(L->exm, R->exm, L->bz, R->bz);
  double bx_f = hll_flux(L->ez, R->ez, L->bx, R->bx);
  double ey_f = hll_flux(L->psi, R->psi, L->ey, R->ey);
  double ez_f = hll_flux(L->bx, R->bx, L->ez, R->ez);
  double ex_f = hll_flux(L->bzm, R->bzm, L->ex, R->ex);
  double q_f = hll_flux(L->j1, R->j1, L->q, R->q);
  double dd_f = hll_flux(L->fd, R->fd, L->d, R->d);
  double tau_f = hll_flux(L->ftau, R->ftau, L->tau, R->tau);
  double sy_f = hll_flux(L->fs2, R->fs2, L->sy, R->sy);
  double sz_f = hll_flux(L->fs3, R->fs3, L->sz, R->sz);
  double sx_f = hll_flux(L->fs1, R->fs1, L->sx, R->sx);

  double erx = bz_h * ey_h - by_h * ez_h;
  double ery = bx_h * ez_h - bz_h * ex_h;
  double erz = by_h * ex_h - bx_h * ey_h;
  double e2p = sq(ex_h) + sq(ez_h);
  double b2p = sq(bx_h) + sq(bz_h);
  double e2b2s = 0.5 * (((sq(ex_h) + sq(ey_h)) + sq(ez_h)) + ((sq(bx_h) + sq(by_h)) + sq(bz_h)));
  double a = tau_f - ery;
  double b = ((e2p + b2p) - tau_h) - sy_f;
  double c = sy_h - ery;
  double v = hllc_contact(a, b, c, g_mcl_seed_y ? g_hllc_last_iterate : (g_hllc_seed_mode ? -b : NAN));
  if (g_hllc_fail) return 7;

  double pstr = (sy_f + (sq(ey_h) + sq(by_h))) - (tau_f - ery) * v;
  double psiR = (S2 * psi_h - v * ey_f) / (S2 - v);
  double psiL = (-(S1 * psi_h) + v * ey_f) / (v - S1);
  double phiR = (S2 * phi_h - v * by_f) / (S2 - v);
  double phiL = (-(S1 * phi_h) + v * by_f) / (v - S1);
  double syR = (S2 * sy_h - tau_f * v) / (S2 - v);
  double syL = (-(S1 * sy_h) + tau_f * v) / (v - S1);
  double szR = ((((S2 * sz_h - sz_f) - erz * v) - ey_h * ez_h) - by_h * bz_h) / (S2 - v);
  double szL = ((((-(S1 * sz_h) + sz_f) + erz * v) + ey_h * ez_h) + by_h * bz_h) / (v - S1);
  double sxR = ((((S2 * sx_h - sx_f) - erx * v) - ey_h * ex_h) - by_h * bx_h) / (S2 - v);
  double sxL = ((((-(S1 * sx_h) + sx_f) + erx * v) + ey_h * ex_h) + by_h * bx_h) / (v - S1);
  double tauR = ((syR - ery) / v - pstr) + e2b2s;
  double tauL = ((syL - ery) / v - pstr) + e2b2s;
  double ddR = (S2 * dd_h - dd_f) / (S2 - v);
  double ddL = (-(S1 * dd_h) + dd_f) / (v - S1);

  double psifR, psifL, phifR, phifL, fdR, fdL, ftR, ftL, fsxR, fsxL, fsyR, fsyL, fszR, fszL;
  if (v == 0.0) {
    psifR = psifL = psi_f; phifR = phifL = phi_f;
    fdR = fdL = dd_f; ftR = ftL = tau_f;
    fsxR = fsxL = sx_f; fsyR = fsyL = sy_f; fszR = fszL = sz_f;
  } else {
    double cr = S2 * (v - S1), cl = S1 * (S2 - v), den = S2 - S1;
    psifR = psi_f + (cr * (psiR - psiL)) / den;
    psifL = psi_f - (cl * (psiR - psiL)) / den;
    phifR = phi_f + (cr * (phiR - phiL)) / den;
    phifL = phi_f - (cl * (phiR - phiL)) / den;
    fdR = dd_f + (cr * (ddR - ddL)) / den;
    fdL = dd_f - (cl * (ddR - ddL)) / den;
    ftR = tau_f + (cr * (tauR - tauL)) / den;
    ftL = tau_f - (cl * (tauR - tauL)) / den;
    fsxR = sx_f + (cr * (sxR - sxL)) / den;
    fsxL = sx_f - (cl * (sxR - sxL)) / den;
    fsyR = sy_f + (cr * (syR - syL)) / den;
    fsyL = sy_f - (cl * (syR - syL)) / den;
    fszR = sz_f + (cr * (szR - szL)) / den;
    fszL = sz_f - (cl * (szR - szL)) / den;
  }
  if (S1 < 0.0 && 0.0 < v) {
    f[FY_EYPSI] = psifL; f[FY_BYPHI] = phifL; f[FY_J] = q_f;
    f[FY_FD] = fdL; f[FY_FTAU] = ftL; f[FY_FSX] = fsxL; f[FY_FSY] = fsyL; f[FY_FSZ] = fszL;
    f[FY_EZBX] = bx_f; f[FY_PHIBY] = by_f; f[FY_EXBZ] = -bz_f;
    f[FY_BZEX] = -ex_f; f[FY_PSIEY] = ey_f; f[FY_BXEZ] = ez_f;
  } else if (v < 0.0 && 0.0 < S2) {
    f[FY_EYPSI] = psifR; f[FY_BYPHI] = phifR; f[FY_J] = q_f;
    f[FY_FD] = fdR; f[FY_FTAU] = ftR; f[FY_FSX] = fsxR; f[FY_FSY] = fsyR; f[FY_FSZ] = fszR;
    f[FY_EZBX] = bx_f; f[FY_PHIBY] = by_f; f[FY_EXBZ] = -bz_f;
    f[FY_BZEX] = -ex_f; f[FY_PSIEY] = ey_f; f[FY_BXEZ] = ez_f;
  } else if (v == 0.0) {
    f[FY_EYPSI] = 0.5 * (psifR + psifL); f[FY_BYPHI] = 0.5 * (phifR + phifL);
    f[FY_EZBX] = 0.5 * (bx_f + bx_f); f[FY_PHIBY] = 0.5 * (by_f + by_f);
    f[FY_EXBZ] = -(0.5 * (bz_f + bz_f)); f[FY_BZEX] = -(0.5 * (ex_f + ex_f));
    f[FY_PSIEY] = 0.5 * (ey_f + ey_f); f[FY_BXEZ] = 0.5 * (ez_f + ez_f);
    f[FY_J] = 0.5 * (q_f + q_f); f[FY_FD] = 0.5 * (fdR + fdL); f[FY_FTAU] = 0.5 * (ftR + ftL);
    f[FY_FSX] = 0.5 * (fsxR + fsxL); f[FY_FSY] = 0.5 * (fsyR + fsyL); f[FY_FSZ] = 0.5 * (fszR + fszL);
  } else {
    return 8;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* HLLE: 16_hlle_flow_mpx.f95:2258-4618.  The electromagnetic subsystem */
/* is HLL with s = -1, +1; q, D, tau, S use relativistic sound-speed    */
/* wave estimates and an HLLC-structured contact for D, tau, S.         */
/* ------------------------------------------------------------------ */

/* relativistic sound-speed wave estimates, 16_hlle_flow_mpx.f95:2715-2818
 * (x) and :2893-2996 (y), including the reference's -vn*(1-cs^2) term
 * under the root (a first power of the normal speed). */
static void hlle_lambda(double gam, const side_t *o, double vn, double *lm, double *lp) {
  double cs = sqrt((gam * o->p) / o->h);
  double cs2 = sq(cs);
  double v2 = (sq(o->vx) + sq(o->vy)) + sq(o->vz);
  double root = sqrt((1.0 - v2) * ((1.0 - v2 * cs2) - vn * (1.0 - cs2)));
  double den = 1.0 - cs2 * v2;
  *lp = (vn * (1.0 - cs2) + cs * root) / den;
  *lm = (vn * (1.0 - cs2) - cs * root) / den;
}

/* min/max of the side estimates and the clamp to |s| < 1; 9 = the NaN stop */
static int hlle_speeds(double gam, const side_t *L, const side_t *R, int dir, double *s1, double *s2) {
  double lmL, lpL, lmR, lpR;
  hlle_lambda(gam, L, dir == 0 ? L->vx : L->vy, &lmL, &lpL);
  hlle_lambda(gam, R, dir == 0 ? R->vx : R->vy, &lmR, &lpR);
  double a = lmL < lmR ? lmL : lmR; /* min */
  double b = lpL > lpR ? lpL : lpR; /* max */
  if (a != a || b != b) return 9;
  const double e2 = 1.0e-9;
  if (fabs(a) > 1.0) a = clampf(a, -1.0 + e2, 1.0 - e2);
  if (fabs(b) > 1.0) b = clampf(b, -1.0 + e2, 1.0 - e2);
  *s1 = a;
  *s2 = b;
  return 0;
}

/* HLL state and flux with wave speeds a < b (16_hlle_flow_mpx.f95:2826-2862) */
static inline double hlle_state(double a, double b, double ul, double ur, double fl, double fr) {
  return (((b * ur - a * ul) + fl) - fr) / (b - a);
}
static inline double hlle_flux(double a, double b, double fl, double fr, double ul, double ur) {
  return ((b * fl - a * fr) + (a * b) * (ur - ul)) / (b - a);
}

/* fluid fluxes of one HLLE face from its wave speeds (a, b), contact speed v,
 * HLL states and fluxes and transverse terms.  n = normal index (0 x,
 * 1 y) selects which momentum is normal.  16_hlle_flow_mpx.f95:3336-3412
 * (x star states), 3459-3540 (star fluxes), 4021-4175 (x selection); y at
 * :3413-3458, 3573-3654, 4413-4500. */
typedef struct {
  double q_f, dd_h, dd_f, tau_h, tau_f, s_h[3], s_f[3];
  double er[3], e[3], b[3], e2b2s;
} hlle_face_t;

static int hlle_fluid(const hlle_face_t *F, int n, double a, double b, double v, double out[6]) {
  int t1 = n == 0 ? 1 : 2, t2 = n == 0 ? 2 : 0; /* transverse momenta in reference order */
  double pstr = (F->s_f[n] + (sq(F->e[n]) + sq(F->b[n]))) - (F->tau_f - F->er[n]) * v;
  double snR = (b * F->s_h[n] - F->tau_f * v) / (b - v);
  double snL = (-(a * F->s_h[n]) + F->tau_f * v) / (v - a);
  double s1R = ((((b * F->s_h[t1] - F->s_f[t1]) - F->er[t1] * v) - F->e[n] * F->e[t1]) - F->b[n] * F->b[t1]) / (b - v);
  double s1L = ((((-(a * F->s_h[t1]) + F->s_f[t1]) + F->er[t1] * v) + F->e[n] * F->e[t1]) + F->b[n] * F->b[t1]) / (v - a);
  double s2R = ((((b * F->s_h[t2] - F->s_f[t2]) - F->er[t2] * v) - F->e[n] * F->e[t2]) - F->b[n] * F->b[t2]) / (b - v);
  double s2L = ((((-(a * F->s_h[t2]) + F->s_f[t2]) + F->er[t2] * v) + F->e[n] * F->e[t2]) + F->b[n] * F->b[t2]) / (v - a);
  double tauR = ((snR - F->er[n]) / v - pstr) + F->e2b2s;
  double tauL = ((snL - F->er[n]) / v - pstr) + F->e2b2s;
  double ddR = (b * F->dd_h - F->dd_f) / (b - v);
  double ddL = (-(a * F->dd_h) + F->dd_f) / (v - a);
  double sR[3], sL[3];
  sR[n] = snR; sL[n] = snL; sR[t1] = s1R; sL[t1] = s1L; sR[t2] = s2R; sL[t2] = s2L;
  double fdR, fdL, ftR, ftL, fsR[3], fsL[3];
  if (v == 0.0) {
    fdR = fdL = F->dd_f; ftR = ftL = F->tau_f;
    for (int k = 0; k < 3; k++) fsR[k] = fsL[k] = F->s_f[k];
  } else {
    double cr = b * (v - a), cl = a * (b - v), den = b - a;
    fdR = F->dd_f + (cr * (ddR - ddL)) / den;
    fdL = F->dd_f - (cl * (ddR - ddL)) / den;
    ftR = F->tau_f + (cr * (tauR - tauL)) / den;
    ftL = F->tau_f - (cl * (tauR - tauL)) / den;
    for (int k = 0; k < 3; k++) {
      fsR[k] = F->s_f[k] + (cr * (sR[k] - sL[k])) / den;
      fsL[k] = F->s_f[k] - (cl * (sR[k] - sL[k])) / den;
    }
  }
  /* out = {J, FD, Ftau, FS x, FS y, FS z}.  For a face with both waves on
   * one side (a >= 0 or b < 0) the reference selects the scalars Jxint_L,
   * FDxint_L, ... (:4021-4033, 4143-4160), which only its 1D branch assigns
   * (:222-475); in 2D they keep their zero initial value. */
  if (a >= 0.0 || (!(a <= 0.0 && 0.0 <= v) && !(v <= 0.0 && 0.0 <= b) && b < 0.0)) {
    for (int k = 0; k < 6; k++) out[k] = 0.0;
  } else if (a <= 0.0 && 0.0 <= v) {
    out[0] = F->q_f; out[1] = fdL; out[2] = ftL; out[3] = fsL[0]; out[4] = fsL[1]; out[5] = fsL[2];
  } else if (v <= 0.0 && 0.0 <= b) {
    out[0] = F->q_f; out[1] = fdR; out[2] = ftR; out[3] = fsR[0]; out[4] = fsR[1]; out[5] = fsR[2];
  } else {
    return 8; /* "Revise Vxstar velocity" stop */
  }
  return 0;
}

/* HLLE x and y faces of one (i,j).  The first Newton point of the x contact
 * solve in the flang -O2 reference build is the y face's (E x B)_y of the
 * HLL states (still in the register); that of the y solve is the x solve's
 * last iterate.  16_hlle_flow_mpx.f95:2263-4618. */
static int hlle_xy(double gam, const side_t *xl, const side_t *xr, const side_t *yl, const side_t *yr,
                   double fx[FX_N], double fy[FY_N]) {
  /* electromagnetic HLL states, s = -1, +1 (:2520-2585) */
  double xpsi_h = hll_state(xl->psi, xr->psi, xl->ex, xr->ex);
  double xphi_h = hll_state(xl->phi, xr->phi, xl->bx, xr->bx);
  double xex_h = hll_state(xl->ex, xr->ex, xl->psi, xr->psi);
  double xey_h = hll_state(xl->ey, xr->ey, xl->bz, xr->bz);
  double xez_h = hll_state(xl->ez, xr->ez, xl->bym, xr->bym);
  double xbx_h = hll_state(xl->bx, xr->bx, xl->phi, xr->phi);
  double xby_h = hll_state(xl->by, xr->by, xl->ezm, xr->ezm);
  double xbz_h = hll_state(xl->bz, xr->bz, xl->ey, xr->ey);
  double ypsi_h = hll_state(yl->psi, yr->psi, yl->ey, yr->ey);
  double yphi_h = hll_state(yl->phi, yr->phi, yl->by, yr->by);
  double yey_h = hll_state(yl->ey, yr->ey, yl->psi, yr->psi);
  double yez_h = hll_state(yl->ez, yr->ez, yl->bx, yr->bx);
  double yex_h = hll_state(yl->ex, yr->ex, yl->bzm, yr->bzm);
  double yby_h = hll_state(yl->by, yr->by, yl->phi, yr->phi);
  double ybz_h = hll_state(yl->bz, yr->bz, yl->exm, yr->exm);
  double ybx_h = hll_state(yl->bx, yr->bx, yl->ez, yr->ez);
  (void)xpsi_h; (void)xphi_h; (void)ypsi_h; (void)yphi_h;

  /* electromagnetic HLL fluxes (:2587-2708) are the face fluxes */
  fx[FX_EXPSI] = hll_flux(xl->ex, xr->ex, xl->psi, xr->psi);
  fx[FX_BXPHI] = hll_flux(xl->bx, xr->bx, xl->phi, xr->phi);
  fx[FX_PHIBX] = hll_flux(xl->phi, xr->phi, xl->bx, xr->bx);
  fx[FX_EZBY] = -hll_flux(xl->ezm, xr->ezm, xl->by, xr->by);
  fx[FX_EYBZ] = hll_flux(xl->ey, xr->ey, xl->bz, xr->bz);
  fx[FX_PSIEX] = hll_flux(xl->psi, xr->psi, xl->ex, xr->ex);
  fx[FX_BZEY] = hll_flux(xl->bz, xr->bz, xl->ey, xr->ey);
  fx[FX_BYEZ] = -hll_flux(xl->bym, xr->bym, xl->ez, xr->ez);
  fy[FY_EYPSI] = hll_flux(yl->ey, yr->ey, yl->psi, yr->psi);
  fy[FY_BYPHI] = hll_flux(yl->by, yr->by, yl->phi, yr->phi);
  fy[FY_PHIBY] = hll_flux(yl->phi, yr->phi, yl->by, yr->by);
  fy[FY_EXBZ] = -hll_flux(yl->exm, yr->exm, yl->bz, yr->bz);
  fy[FY_EZBX] = hll_flux(yl->ez, yr->ez, yl->bx, yr->bx);
  fy[FY_PSIEY] = hll_flux(yl->psi, yr->psi, yl->ey, yr->ey);
  fy[FY_BXEZ] = hll_flux(yl->bx, yr->bx, yl->ez, yr->ez);
  fy[FY_BZEX] = -hll_flux(yl->bzm, yr->bzm, yl->ex, yr->ex);

  /* x fluid part (:2715-2862) */
  double a, b;
  int st = hlle_speeds(gam, xl, xr, 0, &a, &b);
  if (st) return st;
  hlle_face_t X;
  X.q_f = hlle_flux(a, b, xl->j1, xr->j1, xl->q, xr->q);
  X.dd_h = hlle_state(a, b, xl->d, xr->d, xl->fd, xr->fd);
  X.tau_h = hlle_state(a, b, xl->tau, xr->tau, xl->ftau, xr->ftau);
  X.s_h[0] = hlle_state(a, b, xl->sx, xr->sx, xl->fs1, xr->fs1);
  X.s_h[1] = hlle_state(a, b, xl->sy, xr->sy, xl->fs2, xr->fs2);
  X.s_h[2] = hlle_state(a, b, xl->sz, xr->sz, xl->fs3, xr->fs3);
  X.dd_f = hlle_flux(a, b, xl->fd, xr->fd, xl->d, xr->d);
  X.tau_f = hlle_flux(a, b, xl->ftau, xr->ftau, xl->tau, xr->tau);
  X.s_f[0] = hlle_flux(a, b, xl->fs1, xr->fs1, xl->sx, xr->sx);
  X.s_f[1] = hlle_flux(a, b, xl->fs2, xr->fs2, xl->sy, xr->sy);
  X.s_f[2] = hlle_flux(a, b, xl->fs3, xr->fs3, xl->sz, xr->sz);
  X.e[0] = xex_h; X.e[1] = xey_h; X.e[2] = xez_h;
  X.b[0] = xbx_h; X.b[1] = xby_h; X.b[2] = xbz_h;
  X.er[0] = xbz_h * xey_h - xby_h * xez_h;
  X.er[1] = xbx_h * xez_h - xbz_h * xex_h;
  X.er[2] = xby_h * xex_h - xbx_h * xey_h;
  X.e2b2s = 0.5 * (((sq(xex_h) + sq(xey_h)) + sq(xez_h)) + ((sq(xbx_h) + sq(xby_h)) + sq(xbz_h)));

  /* y fluid part (:2893-3040) */
  double c, d;
  st = hlle_speeds(gam, yl, yr, 1, &c, &d);
  if (st) return st;
  hlle_face_t Y;
  Y.q_f = hlle_flux(c, d, yl->j1, yr->j1, yl->q, yr->q);
  Y.dd_h = hlle_state(c, d, yl->d, yr->d, yl->fd, yr->fd);
  Y.tau_h = hlle_state(c, d, yl->tau, yr->tau, yl->ftau, yr->ftau);
  Y.s_h[0] = hlle_state(c, d, yl->sx, yr->sx, yl->fs1, yr->fs1);
  Y.s_h[1] = hlle_state(c, d, yl->sy, yr->sy, yl->fs2, yr->fs2);
  Y.s_h[2] = hlle_state(c, d, yl->sz, yr->sz, yl->fs3, yr->fs3);
  Y.dd_f = hlle_flux(c, d, yl->fd, yr->fd, yl->d, yr->d);
  Y.tau_f = hlle_flux(c, d, yl->ftau, yr->ftau, yl->tau, yr->tau);
  Y.s_f[0] = hlle_flux(c, d, yl->fs1, yr->fs1, yl->sx, yr->sx);
  Y.s_f[1] = hlle_flux(c, d, yl->fs2, yr->fs2, yl->sy, yr->sy);
  Y.s_f[2] = hlle_flux(c, d, yl->fs3, yr->fs3, yl->sz, yr->sz);
  Y.e[0] = yex_h; Y.e[1] = yey_h; Y.e[2] = yez_h;
  Y.b[0] = ybx_h; Y.b[1] = yby_h; Y.b[2] = ybz_h;
  Y.er[0] = ybz_h * yey_h - yby_h * yez_h;
  Y.er[1] = ybx_h * yez_h - ybz_h * yex_h;
  Y.er[2] = yby_h * yex_h - ybx_h * yey_h;
  Y.e2b2s = 0.5 * (((sq(yex_h) + sq(yey_h)) + sq(yez_h)) + ((sq(ybx_h) + sq(yby_h)) + sq(ybz_h)));

  /* contact speeds (:3225-3334) */
  double ax = X.tau_f - X.er[0];
  double bx = ((((sq(xey_h) + sq(xez_h)) + (sq(xby_h) + sq(xbz_h))) - X.tau_h) - X.s_f[0]);
  double cx = X.s_h[0] - X.er[0];
  double vx = hllc_contact(ax, bx, cx, g_hllc_seed_mode ? Y.er[1] : NAN);
  if (g_hllc_fail) return 7;
  double ay = Y.tau_f - Y.er[1];
  double by = ((((sq(yex_h) + sq(yez_h)) + (sq(ybx_h) + sq(ybz_h))) - Y.tau_h) - Y.s_f[1]);
  double cy = Y.s_h[1] - Y.er[1];
  double vy = hllc_contact(ay, by, cy, g_hllc_seed_mode ? g_hllc_last_iterate : NAN);
  if (g_hllc_fail) return 7;

  double o[6];
  st = hlle_fluid(&X, 0, a, b, vx, o);
  if (st) return st;
  fx[FX_J] = o[0]; fx[FX_FD] = o[1]; fx[FX_FTAU] = o[2];
  fx[FX_FSX] = o[3]; fx[FX_FSY] = o[4]; fx[FX_FSZ] = o[5];
  st = hlle_fluid(&Y, 1, c, d, vy, o);
  if (st) return st;
  fy[FY_J] = o[0]; fy[FY_FD] = o[1]; fy[FY_FTAU] = o[2];
  fy[FY_FSX] = o[3]; fy[FY_FSY] = o[4]; fy[FY_FSZ] = o[5];
  return 0;
}

/* face loop: 16_hll_flow_mpx.f95:1358-1958, 16_hllc_flow_mpx.f95:2401-4612,
 * 16_hlle_flow_mpx.f95:2258-4618 */
static void mcl_face(const oracle_state *s, int i, int j, int dir, double *f);
static void lax_face(const oracle_state *s, int i, int j, int dir, double *f);
static int hllc_mcl_faces(const oracle_state *s, int i, int j, double *fx, double *fy);
static int hlle_mcl_faces(const oracle_state *s, int i, int j, double *fx, double *fy);

/* one printed face event (16_hll_flow_mpx.f95:1419,1453 "FIXED velocity",
 * 16_hllc_flow_mpx.f95:2665-2691 "Negative density"): kind 0 fallback
 * from the x faces, 1 from the y faces only, 2..5 density fix of y_L, y_R,
 * x_L, x_R; value: the negative density (0 for a fallback) */
static void face_event(oracle_state *s, int kind, int i, int j, double value) {
  if (s->nfe == s->capfe) {
    int cap = s->capfe ? 2 * s->capfe : 256;
    int32_t *k = realloc(s->fe_kind, sizeof(int32_t) * 4 * (size_t)cap);
    if (!k) return;
    s->fe_kind = k;
    double *v = realloc(s->fe_val, sizeof(double) * (size_t)cap);
    if (!v) return;
    s->fe_val = v;
    s->capfe = cap;
  }
  int n = s->nfe++;
  s->fe_kind[4 * n] = kind;
  s->fe_kind[4 * n + 1] = s->cur_l;
  s->fe_kind[4 * n + 2] = i;
  s->fe_kind[4 * n + 3] = j;
  s->fe_val[n] = value;
}

int oracle_get_face_events(const oracle_state *s, int32_t *kind_l_i_j, double *vals, int cap) {
  for (int e = 0; e < s->nfe && e < cap; e++) {
    for (int q = 0; q < 4; q++) kind_l_i_j[4 * e + q] = s->fe_kind[4 * e + q];
    vals[e] = s->fe_val[e];
  }
  return s->nfe;
}

/* where the reference stops: the face (i, j) of its loop */
static int face_stop(oracle_state *s, int i, int j, int st) {
  s->stop[2] = i;
  s->stop[3] = j;
  return st;
}

static int face_loop(oracle_state *s) {
  const int im = s->c.imax, jm = s->c.jmax, fs = s->c.flux_solver;
  /* hll_flow_mpx sets the shared module variable eps1 (16_hll_flow_mpx.f95:1375,
   * not in its PRIVATE list); the recovery's superluminal clamp reads it */
  if (fs == OR_HLL) s->eps1 = 1.0e-6;
  double **P = s->prim;
  for (int j = -1; j <= jm; j++)
    for (int i = -1; i <= im; i++) {
      size_t k = IX(s, i, j);
      if (fs == OR_HLL_MCL || fs == OR_LAX || fs == OR_HLLC_MCL || fs == OR_HLLE_MCL) {
        double fx[FX_N], fy[FY_N];
        if (fs == OR_HLLC_MCL || fs == OR_HLLE_MCL) {
          int st = fs == OR_HLLC_MCL ? hllc_mcl_faces(s, i, j, fx, fy) : hlle_mcl_faces(s, i, j, fx, fy);
          if (st) return face_stop(s, i, j, st);
        } else if (fs == OR_LAX) {
          lax_face(s, i, j, 0, fx);
          lax_face(s, i, j, 1, fy);
        } else {
          mcl_face(s, i, j, 0, fx);
          mcl_face(s, i, j, 1, fy);
        }
        for (int f = 0; f < FX_N; f++) s->fx[f][k] = fx[f];
        for (int f = 0; f < FY_N; f++) s->fy[f][k] = fy[f];
        continue;
      }
      side_t xl, xr, yl, yr;
      load_x(s, s->xL, k, &xl);
      load_x(s, s->xR, k, &xr);
      load_y(s, s->yL, k, &yl);
      load_y(s, s->yR, k, &yr);
      /* superluminal fallback (16_hll_flow_mpx.f95:1392-1461): any of the
       * four face speeds >= 1 replaces all of them by cell values, with the
       * reference's (i-1,j) and (i,j-1) right states.  HLLE only prints
       * (16_hlle_flow_mpx.f95:2270-2300). */
      double v2xl = (sq(xl.vx) + sq(xl.vy)) + sq(xl.vz);
      double v2xr = (sq(xr.vx) + sq(xr.vy)) + sq(xr.vz);
      double v2yl = (sq(yl.vx) + sq(yl.vy)) + sq(yl.vz);
      double v2yr = (sq(yr.vx) + sq(yr.vy)) + sq(yr.vz);
      if (fs != OR_HLLE && (v2xl >= 1.0 || v2xr >= 1.0 || v2yl >= 1.0 || v2yr >= 1.0)) {
        s->face_events[0]++;
        face_event(s, (v2xl >= 1.0 || v2xr >= 1.0) ? 0 : 1, i, j, 0.0);
        size_t kc = IX(s, i, j), km = IX(s, i - 1, j), kj = IX(s, i, j - 1);
        xl.vx = P[OR_VX][kc]; xl.vy = P[OR_VY][kc]; xl.vz = P[OR_VZ][kc];
        xr.vx = P[OR_VX][km]; xr.vy = P[OR_VY][km]; xr.vz = P[OR_VZ][km];
        yl.vx = P[OR_VX][kc]; yl.vy = P[OR_VY][kc]; yl.vz = P[OR_VZ][kc];
        yr.vx = P[OR_VX][kj]; yr.vy = P[OR_VY][kj]; yr.vz = P[OR_VZ][kj];
      }
      if (fs == OR_HLLC) { /* 16_hllc_flow_mpx.f95:2663-2693 */
        double rc = 0.1 * P[OR_RHO][k];
        if (yl.rho < 0.0 || yr.rho < 0.0 || xl.rho < 0.0 || xr.rho < 0.0) s->face_events[1]++;
        if (yl.rho < 0.0) { face_event(s, 2, i, j, yl.rho); yl.rho = rc; }
        else if (yr.rho < 0.0) { face_event(s, 3, i, j, yr.rho); yr.rho = rc; }
        else if (xl.rho < 0.0) { face_event(s, 4, i, j, xl.rho); xl.rho = rc; }
        else if (xr.rho < 0.0) { face_event(s, 5, i, j, xr.rho); xr.rho = rc; }
      }
      int he = fs == OR_HLLE;
      side_physics(s->c.gamma, 0, he, &xl);
      side_physics(s->c.gamma, 0, he, &xr);
      side_physics(s->c.gamma, 1, he, &yl);
      side_physics(s->c.gamma, 1, he, &yr);
      double fx[FX_N], fy[FY_N];
      if (fs == OR_HLL) {
        hll_x(&xl, &xr, fx);
        hll_y(&yl, &yr, fy);
      } else if (fs == OR_HLLC) {
        int st = hllc_x(&xl, &xr, fx);
        if (st) return face_stop(s, i, j, st);
        st = hllc_y(&yl, &yr, fy);
        if (st) return face_stop(s, i, j, st);
      } else {
        int st = hlle_xy(s->c.gamma, &xl, &xr, &yl, &yr, fx, fy);
        if (st) return face_stop(s, i, j, st);
      }
      for (int f = 0; f < FX_N; f++) s->fx[f][k] = fx[f];
      for (int f = 0; f < FY_N; f++) s->fy[f][k] = fy[f];
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* HLL with the monotonised-central slope limiter, flux_solver = 2:     */
/* hll_flow, 16_hll_flow_mcl.f95:554-1402 (DIM = 2, REC_PRIM = 0).      */
/* Face states of the stage fields and of the t^n velocity, pressure    */
/* and density; Sp = 1 (const_sp = 0, parameters.f95:274).               */
/* ------------------------------------------------------------------ */
/* mcl(fa, fb), funciones.f95:57-95, and limiter(ga, gb), :100-142: kind 1
 * MinMod, 2 monotonised central, 3 Superbee (slop / limit,
 * parameters.f95:17-18); min(a, b, c) evaluated as min(min(a, b), c) */
static int g_slop = 2, g_limit = 2;
static inline double limiter_of(int kind, double fa, double fb) {
  double s = 0.5 * (copysign(1.0, fa) + copysign(1.0, fb));
  if (kind == 1) return s * fmin2(fabs(fa), fabs(fb));
  if (kind == 3) return s * fmin2(fmin2(2.0 * fabs(fa), 2.0 * fabs(fb)), fmax2(fabs(fa), fabs(fb)));
  return s * fmin2(fmin2(2.0 * fabs(fa), 2.0 * fabs(fb)), 0.5 * fabs(fa + fb));
}
static inline double mcl(double fa, double fb) { return limiter_of(g_slop, fa, fb); }

typedef struct {
  double psi, phi, ex, ey, ez, bx, by, bz, q, d, tau, sx, sy, sz, vx, vy, vz, p, rho;
  double ezm, bym, exm, bzm; /* the "_m" states the fluxes use */
} mcl_side_t;

/* ------------------------------------------------------------------ */
/* recons_var_primitive, 05_reconstruct.f95:525-740 (DIM = 2), with     */
/* primitive_rec = 1 (parameters.f95:26) and limit = 2, whose limiter   */
/* (funciones.f95:100-142) is the MC slope above.  Along the diagonal,  */
/* L(i,j) = u + 0.5 lim(u(i+1,j+1) - u, u - u(i-1,j-1)) and R(i+1,j+1)  */
/* from the same differences, and every interior point becomes          */
/* 0.5 (R + L): u again up to rounding.  Ghosts are left to the caller. */
/* ------------------------------------------------------------------ */
static int recons_prims(oracle_state *s) {
  const int64_t dd = IX(s, 1, 1) - IX(s, 0, 0);
  double *tmp = malloc(s->n * sizeof(double));
  if (!tmp) return 90;
  for (int k = 0; k < OR_NPRIM; k++) {
    double *a = s->prim[k];
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) {
        size_t c = IX(s, i, j);
        double u = a[c];
        double lim = limiter_of(g_limit, a[c + dd] - u, u - a[c - dd]);
        double r = u - 0.5 * lim, l = u + 0.5 * lim;
        tmp[c] = 0.5 * (r + l);
      }
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) a[IX(s, i, j)] = tmp[IX(s, i, j)];
  }
  free(tmp);
  return 0;
}

/* u_L = u(i) + FAC mcl(u(i+1)-u(i), u(i)-u(i-1)), u_R = u(i+1) - FAC
 * mcl(u(i+2)-u(i+1), u(i+1)-u(i)), FAC = 0.5 (16_hll_flow_mcl.f95:572-575);
 * the "_m" states negate the cell value only (:634-637) */
static void mcl_rec(const double *a, size_t km, size_t k0, size_t k1, size_t k2, double *l, double *r,
                    double *lm, double *rm) {
  double u0 = a[k0], u1 = a[k1];
  double sl = 0.5 * mcl(u1 - u0, u0 - a[km]);
  double sr = 0.5 * mcl(a[k2] - u1, u1 - u0);
  *l = u0 + sl;
  *r = u1 - sr;
  if (lm) {
    *lm = (-u0) + sl;
    *rm = (-u1) - sr;
  }
}

/* J, FD, Ftau, FS of one side: :821 (E2B2), :900 (W), :995 (E.v),
 * :1033 (J), :1050 (FD), :1124 (eps, h), :1181 (Ftau), :1198-1241 (FS) */
static void mcl_side_fluxes(double gam, double sigma, int dir, const mcl_side_t *o, double out[6]) {
  double e2b2 = ((((sq(o->ex) + sq(o->ey)) + sq(o->ez)) + sq(o->bx)) + sq(o->by)) + sq(o->bz);
  double w = 1.0 / sqrt(((1.0 - sq(o->vx)) - sq(o->vy)) - sq(o->vz));
  double edv = (o->ex * o->vx + o->ey * o->vy) + o->ez * o->vz;
  double eps = o->p / ((gam - 1.0) * o->rho);
  double h = o->rho * (1.0 + eps) + o->p;
  double hw2 = h * sq(w);
  if (dir == 0) {
    double rx = o->bz * o->vy - o->by * o->vz;
    out[0] = (sigma * w) * ((o->ex + rx) - edv * o->vx) + o->q * o->vx;
    out[1] = o->d * o->vx;
    out[2] = (o->bz * o->ey - o->by * o->ez) + hw2 * o->vx;
    out[3] = ((((-sq(o->ex)) - sq(o->bx)) + hw2 * sq(o->vx)) + 0.5 * e2b2) + o->p;
    out[4] = ((-(o->ex * o->ey)) - o->bx * o->by) + (hw2 * o->vx) * o->vy;
    out[5] = ((-(o->ex * o->ez)) - o->bx * o->bz) + (hw2 * o->vx) * o->vz;
  } else {
    double ry = o->bx * o->vz - o->bz * o->vx;
    out[0] = (sigma * w) * ((o->ey + ry) - edv * o->vy) + o->q * o->vy;
    out[1] = o->d * o->vy;
    out[2] = (o->bx * o->ez - o->bz * o->ex) + hw2 * o->vy;
    out[3] = ((-(o->ex * o->ey)) - o->bx * o->by) + (hw2 * o->vx) * o->vy;
    out[4] = ((((-sq(o->ey)) - sq(o->by)) + hw2 * sq(o->vy)) + 0.5 * e2b2) + o->p;
    out[5] = ((-(o->ey * o->ez)) - o->by * o->bz) + (hw2 * o->vy) * o->vz;
  }
}

/* 0.5 (A_L + A_R - Sp (U_R - U_L)), Sp = 1 (:1269-1396) */
static inline double mcl_lf(double al, double ar, double ul, double ur) { return 0.5 * ((al + ar) - (ur - ul)); }

/* the face between cell (i,j) and its +x (dir 0) or +y (dir 1) neighbour */
static void mcl_face(const oracle_state *s, int i, int j, int dir, double *f) {
  double *const *S = s->st;
  const int di = dir == 0, dj = dir == 1;
  size_t km = IX(s, i - di, j - dj), k0 = IX(s, i, j), k1 = IX(s, i + di, j + dj),
         k2 = IX(s, i + 2 * di, j + 2 * dj);
  mcl_side_t L, R;
  mcl_rec(S[ST_PSI], km, k0, k1, k2, &L.psi, &R.psi, NULL, NULL);
  mcl_rec(S[ST_PHI], km, k0, k1, k2, &L.phi, &R.phi, NULL, NULL);
  mcl_rec(S[ST_EX], km, k0, k1, k2, &L.ex, &R.ex, &L.exm, &R.exm);
  mcl_rec(S[ST_EY], km, k0, k1, k2, &L.ey, &R.ey, NULL, NULL);
  mcl_rec(S[ST_EZ], km, k0, k1, k2, &L.ez, &R.ez, &L.ezm, &R.ezm);
  mcl_rec(S[ST_BX], km, k0, k1, k2, &L.bx, &R.bx, NULL, NULL);
  mcl_rec(S[ST_BY], km, k0, k1, k2, &L.by, &R.by, &L.bym, &R.bym);
  mcl_rec(S[ST_BZ], km, k0, k1, k2, &L.bz, &R.bz, &L.bzm, &R.bzm);
  mcl_rec(S[ST_Q], km, k0, k1, k2, &L.q, &R.q, NULL, NULL);
  mcl_rec(S[ST_D], km, k0, k1, k2, &L.d, &R.d, NULL, NULL);
  mcl_rec(S[ST_TAU], km, k0, k1, k2, &L.tau, &R.tau, NULL, NULL);
  mcl_rec(S[ST_SX], km, k0, k1, k2, &L.sx, &R.sx, NULL, NULL);
  mcl_rec(S[ST_SY], km, k0, k1, k2, &L.sy, &R.sy, NULL, NULL);
  mcl_rec(S[ST_SZ], km, k0, k1, k2, &L.sz, &R.sz, NULL, NULL);
  mcl_rec(s->kin[K_VX], km, k0, k1, k2, &L.vx, &R.vx, NULL, NULL);
  mcl_rec(s->kin[K_VY], km, k0, k1, k2, &L.vy, &R.vy, NULL, NULL);
  mcl_rec(s->kin[K_VZ], km, k0, k1, k2, &L.vz, &R.vz, NULL, NULL);
  mcl_rec(s->kin[K_P], km, k0, k1, k2, &L.p, &R.p, NULL, NULL);
  mcl_rec(s->kin[K_RHO], km, k0, k1, k2, &L.rho, &R.rho, NULL, NULL);
  double fl[6], fr[6];
  mcl_side_fluxes(s->c.gamma, s->c.sigma, dir, &L, fl);
  mcl_side_fluxes(s->c.gamma, s->c.sigma, dir, &R, fr);
  if (dir == 0) {
    f[FX_EXPSI] = mcl_lf(L.ex, R.ex, L.psi, R.psi);             /* Exlaxpsi */
    f[FX_BXPHI] = mcl_lf(L.bx, R.bx, L.phi, R.phi);             /* Bxlaxphi */
    f[FX_PHIBX] = mcl_lf(L.phi, R.phi, L.bx, R.bx);             /* philaxBx */
    f[FX_EZBY] = -mcl_lf(L.ezm, R.ezm, L.by, R.by);             /* EzlaxBy, :1292 */
    f[FX_EYBZ] = mcl_lf(L.ey, R.ey, L.bz, R.bz);                /* EylaxBz */
    f[FX_PSIEX] = mcl_lf(L.psi, R.psi, L.ex, R.ex);             /* psilaxEx */
    f[FX_BZEY] = mcl_lf(L.bz, R.bz, L.ey, R.ey);                /* BzlaxEy */
    f[FX_BYEZ] = -mcl_lf(L.bym, R.bym, L.ez, R.ez);             /* BylaxEz */
  } else {
    f[FY_EYPSI] = mcl_lf(L.ey, R.ey, L.psi, R.psi);             /* Eylaxpsi */
    f[FY_BYPHI] = mcl_lf(L.by, R.by, L.phi, R.phi);             /* Bylaxphi */
    f[FY_EZBX] = mcl_lf(L.ez, R.ez, L.bx, R.bx);                /* EzlaxBx */
    f[FY_PHIBY] = mcl_lf(L.phi, R.phi, L.by, R.by);             /* philaxBy */
    f[FY_EXBZ] = -mcl_lf(L.exm, R.exm, L.bz, R.bz);             /* ExlaxBz, :1299 */
    f[FY_BZEX] = -mcl_lf(L.bzm, R.bzm, L.ex, R.ex);             /* BzlaxEx, :1308 */
    f[FY_PSIEY] = mcl_lf(L.psi, R.psi, L.ey, R.ey);             /* psilaxEy */
    f[FY_BXEZ] = mcl_lf(L.bx, R.bx, L.ez, R.ez);                /* BxlaxEz */
  }
  /* J, FD, Ftau, FS sit at the same slots in both directions */
  _Static_assert((int)FX_J == (int)FY_J && (int)FX_FD == (int)FY_FD && (int)FX_FTAU == (int)FY_FTAU &&
                     (int)FX_FSX == (int)FY_FSX && (int)FX_FSZ == (int)FY_FSZ,
                 "slot layout");
  f[FX_J] = mcl_lf(fl[0], fr[0], L.q, R.q);
  f[FX_FD] = mcl_lf(fl[1], fr[1], L.d, R.d);
  f[FX_FTAU] = mcl_lf(fl[2], fr[2], L.tau, R.tau);
  f[FX_FSX] = mcl_lf(fl[3], fr[3], L.sx, R.sx);
  f[FX_FSY] = mcl_lf(fl[4], fr[4], L.sy, R.sy);
  f[FX_FSZ] = mcl_lf(fl[5], fr[5], L.sz, R.sz);
}

/* ------------------------------------------------------------------ */
/* HLLC with the MC limiter, flux_solver = 4: hllc_flow,                */
/* 16_hllc_flow_mcl.f95:1751-4140 (DIM = 2).  The face states and their */
/* physical fluxes are those of hll_flow (:1760-2479 repeats              */
/* 16_hll_flow_mcl.f95:563-1258 but for log lines); the Riemann solve is */
/* that of hllc_flow_mpx statement for statement (:2480-4140) with the   */
/* "_m" states of the MC reconstruction and |v*| <= 1e-6 -> 0 (:2596).    */
/* Its get_hllc_vel is an internal copy (:4072-4143) whose first Newton   */
/* evaluation reads v before setting it (:4113), like the module one: in  */
/* the flang -O2 build that is the register the caller last wrote, which */
/* for the x solve holds the y face's (E x B)_y of the HLL states and for */
/* the y solve the x solve's last iterate (read from the assembly of the  */
/* binary built here; hllc_seed = 0 starts both from clamp(rt2) instead). */
/* ------------------------------------------------------------------ */
static void mcl_sides(const oracle_state *s, int i, int j, int dir, side_t *L, side_t *R) {
  double *const *S = s->st;
  const int di = dir == 0, dj = dir == 1;
  size_t km = IX(s, i - di, j - dj), k0 = IX(s, i, j), k1 = IX(s, i + di, j + dj),
         k2 = IX(s, i + 2 * di, j + 2 * dj);
  memset(L, 0, sizeof(*L));
  memset(R, 0, sizeof(*R));
  mcl_rec(S[ST_PSI], km, k0, k1, k2, &L->psi, &R->psi, NULL, NULL);
  mcl_rec(S[ST_PHI], km, k0, k1, k2, &L->phi, &R->phi, NULL, NULL);
  mcl_rec(S[ST_EX], km, k0, k1, k2, &L->ex, &R->ex, &L->exm, &R->exm);
  mcl_rec(S[ST_EY], km, k0, k1, k2, &L->ey, &R->ey, &L->eym, &R->eym);
  mcl_rec(S[ST_EZ], km, k0, k1, k2, &L->ez, &R->ez, &L->ezm, &R->ezm);
  mcl_rec(S[ST_BX], km, k0, k1, k2, &L->bx, &R->bx, &L->bxm, &R->bxm);
  mcl_rec(S[ST_BY], km, k0, k1, k2, &L->by, &R->by, &L->bym, &R->bym);
  mcl_rec(S[ST_BZ], km, k0, k1, k2, &L->bz, &R->bz, &L->bzm, &R->bzm);
  mcl_rec(S[ST_Q], km, k0, k1, k2, &L->q, &R->q, NULL, NULL);
  mcl_rec(S[ST_D], km, k0, k1, k2, &L->d, &R->d, NULL, NULL);
  mcl_rec(S[ST_TAU], km, k0, k1, k2, &L->tau, &R->tau, NULL, NULL);
  mcl_rec(S[ST_SX], km, k0, k1, k2, &L->sx, &R->sx, NULL, NULL);
  mcl_rec(S[ST_SY], km, k0, k1, k2, &L->sy, &R->sy, NULL, NULL);
  mcl_rec(S[ST_SZ], km, k0, k1, k2, &L->sz, &R->sz, NULL, NULL);
  mcl_rec(s->kin[K_VX], km, k0, k1, k2, &L->vx, &R->vx, NULL, NULL);
  mcl_rec(s->kin[K_VY], km, k0, k1, k2, &L->vy, &R->vy, NULL, NULL);
  mcl_rec(s->kin[K_VZ], km, k0, k1, k2, &L->vz, &R->vz, NULL, NULL);
  mcl_rec(s->kin[K_P], km, k0, k1, k2, &L->p, &R->p, NULL, NULL);
  mcl_rec(s->kin[K_RHO], km, k0, k1, k2, &L->rho, &R->rho, NULL, NULL);
  side_t *sd[2] = {L, R};
  for (int t = 0; t < 2; t++) {
    mcl_side_t m;
    side_t *o = sd[t];
    m.psi = o->psi; m.phi = o->phi; m.ex = o->ex; m.ey = o->ey; m.ez = o->ez;
    m.bx = o->bx; m.by = o->by; m.bz = o->bz; m.q = o->q; m.d = o->d; m.tau = o->tau;
    m.sx = o->sx; m.sy = o->sy; m.sz = o->sz; m.vx = o->vx; m.vy = o->vy; m.vz = o->vz;
    m.p = o->p; m.rho = o->rho;
    double fl[6];
    mcl_side_fluxes(s->c.gamma, s->c.sigma, dir, &m, fl);
    o->j1 = fl[0]; o->fd = fl[1]; o->ftau = fl[2]; o->fs1 = fl[3]; o->fs2 = fl[4]; o->fs3 = fl[5];
  }
}

/* (E x B)_y of the y face's HLL states (16_hllc_flow_mcl.f95:2601-2609,
 * :2774), the x solve's first Newton point in the reference build */
static double mcl_y_ery(const side_t *L, const side_t *R) {
  double ez = hll_state(L->ez, R->ez, L->bx, R->bx);
  double ex = hll_state(L->ex, R->ex, L->bzm, R->bzm);
  double bz = hll_state(L->bz, R->bz, L->exm, R->exm);
  double bx = hll_state(L->bx, R->bx, L->ez, R->ez);
  return bx * ez - bz * ex;
}

static int hllc_mcl_faces(const oracle_state *s, int i, int j, double *fx, double *fy) {
  side_t xl, xr, yl, yr;
  mcl_sides(s, i, j, 0, &xl, &xr);
  mcl_sides(s, i, j, 1, &yl, &yr);
  const double zero = g_hllc_zero;
  const int seed = g_hllc_seed_mode;
  g_hllc_zero = 1.0e-6;
  if (seed) {
    g_mcl_seed_x = mcl_y_ery(&yl, &yr);
    g_mcl_seed_y = 1;
  }
  g_hllc_seed_mode = 0; /* otherwise clamp(rt2) */
  int st = hllc_x(&xl, &xr, fx);
  if (!st) st = hllc_y(&yl, &yr, fy);
  g_hllc_zero = zero;
  g_hllc_seed_mode = seed;
  g_mcl_seed_x = NAN;
  g_mcl_seed_y = 0;
  return st;
}

/* ------------------------------------------------------------------ */
/* HLLE with the MC limiter, flux_solver = 3: hlle_flow,                */
/* 16_hlle_flow_mcl.f95:1638-4348 (DIM = 2).  The face states and their */
/* physical fluxes are those of hll_flow (:1647-2373); the wave speeds,  */
/* fluid HLL states and fluxes and the contact solve follow              */
/* hlle_flow_mpx, but its star states and fluxes are its own:            */
/*  - S*_n carries the star pressure and tau* is written in terms of the */
/*    HLL tau state (:3187-3243);                                         */
/*  - fluid star fluxes are F_K + s_K (U*_K - U_K) (:3395-3450);          */
/*  - the EM star "fluxes" are star states (:3342-3364, :3539-3561), so   */
/*    a face with v* != 0 carries the EM HLL states as its EM fluxes,     */
/*    and v* = 0 the EM HLL fluxes;                                        */
/*  - the fluid selection uses s1 > 0, s1 <= 0 <= v*, v* < 0 <= s2,        */
/*    s2 < 0 with the sides' own physical fluxes outside the fan          */
/*    (:3719-3800), which this routine does assign in 2D.                 */
/* Its get_hllc_vel copy allows 200 iterations (:4385).                  */
/* ------------------------------------------------------------------ */
typedef struct {
  double psi_h, phi_h, e[3], b[3];   /* EM HLL states, s = -1, +1 */
  double f_em[8];                    /* EM HLL fluxes in FX or FY slot order 0..7 */
  double a, b_, q_f, dd_h, dd_f, tau_h, tau_f, s_h[3], s_f[3];
  double er[3], e2b2s;
} hlle_mcl_t;

/* the x (dir 0) or y face's HLL states and fluxes from its MC sides */
static int hlle_mcl_prep(double gam, int dir, const side_t *L, const side_t *R, hlle_mcl_t *F) {
  if (dir == 0) {
    F->psi_h = hll_state(L->psi, R->psi, L->ex, R->ex);
    F->phi_h = hll_state(L->phi, R->phi, L->bx, R->bx);
    F->e[0] = hll_state(L->ex, R->ex, L->psi, R->psi);
    F->e[1] = hll_state(L->ey, R->ey, L->bz, R->bz);
    F->e[2] = hll_state(L->ez, R->ez, L->bym, R->bym);
    F->b[0] = hll_state(L->bx, R->bx, L->phi, R->phi);
    F->b[1] = hll_state(L->by, R->by, L->ezm, R->ezm);
    F->b[2] = hll_state(L->bz, R->bz, L->ey, R->ey);
    F->f_em[FX_EXPSI] = hll_flux(L->ex, R->ex, L->psi, R->psi);
    F->f_em[FX_BXPHI] = hll_flux(L->bx, R->bx, L->phi, R->phi);
    F->f_em[FX_PHIBX] = hll_flux(L->phi, R->phi, L->bx, R->bx);
    F->f_em[FX_EZBY] = -hll_flux(L->ezm, R->ezm, L->by, R->by);
    F->f_em[FX_EYBZ] = hll_flux(L->ey, R->ey, L->bz, R->bz);
    F->f_em[FX_PSIEX] = hll_flux(L->psi, R->psi, L->ex, R->ex);
    F->f_em[FX_BZEY] = hll_flux(L->bz, R->bz, L->ey, R->ey);
    F->f_em[FX_BYEZ] = -hll_flux(L->bym, R->bym, L->ez, R->ez);
  } else {
    F->psi_h = hll_state(L->psi, R->psi, L->ey, R->ey);
    F->phi_h = hll_state(L->phi, R->phi, L->by, R->by);
    F->e[1] = hll_state(L->ey, R->ey, L->psi, R->psi);
    F->e[2] = hll_state(L->ez, R->ez, L->bx, R->bx);
    F->e[0] = hll_state(L->ex, R->ex, L->bzm, R->bzm);
    F->b[1] = hll_state(L->by, R->by, L->phi, R->phi);
    F->b[2] = hll_state(L->bz, R->bz, L->exm, R->exm);
    F->b[0] = hll_state(L->bx, R->bx, L->ez, R->ez);
    F->f_em[FY_EYPSI] = hll_flux(L->ey, R->ey, L->psi, R->psi);
    F->f_em[FY_BYPHI] = hll_flux(L->by, R->by, L->phi, R->phi);
    F->f_em[FY_EZBX] = hll_flux(L->ez, R->ez, L->bx, R->bx);
    F->f_em[FY_PHIBY] = hll_flux(L->phi, R->phi, L->by, R->by);
    F->f_em[FY_EXBZ] = -hll_flux(L->exm, R->exm, L->bz, R->bz);
    F->f_em[FY_BZEX] = -hll_flux(L->bzm, R->bzm, L->ex, R->ex);
    F->f_em[FY_PSIEY] = hll_flux(L->psi, R->psi, L->ey, R->ey);
    F->f_em[FY_BXEZ] = hll_flux(L->bx, R->bx, L->ez, R->ez);
  }
  double a, b;
  int st = hlle_speeds(gam, L, R, dir, &a, &b);
  if (st) return st;
  F->a = a;
  F->b_ = b;
  F->q_f = hlle_flux(a, b, L->j1, R->j1, L->q, R->q);
  F->dd_h = hlle_state(a, b, L->d, R->d, L->fd, R->fd);
  F->tau_h = hlle_state(a, b, L->tau, R->tau, L->ftau, R->ftau);
  F->s_h[0] = hlle_state(a, b, L->sx, R->sx, L->fs1, R->fs1);
  F->s_h[1] = hlle_state(a, b, L->sy, R->sy, L->fs2, R->fs2);
  F->s_h[2] = hlle_state(a, b, L->sz, R->sz, L->fs3, R->fs3);
  F->dd_f = hlle_flux(a, b, L->fd, R->fd, L->d, R->d);
  F->tau_f = hlle_flux(a, b, L->ftau, R->ftau, L->tau, R->tau);
  F->s_f[0] = hlle_flux(a, b, L->fs1, R->fs1, L->sx, R->sx);
  F->s_f[1] = hlle_flux(a, b, L->fs2, R->fs2, L->sy, R->sy);
  F->s_f[2] = hlle_flux(a, b, L->fs3, R->fs3, L->sz, R->sz);
  const double *e = F->e, *bb = F->b;
  F->er[0] = bb[2] * e[1] - bb[1] * e[2];
  F->er[1] = bb[0] * e[2] - bb[2] * e[0];
  F->er[2] = bb[1] * e[0] - bb[0] * e[1];
  F->e2b2s = 0.5 * (((sq(e[0]) + sq(e[1])) + sq(e[2])) + ((sq(bb[0]) + sq(bb[1])) + sq(bb[2])));
  return 0;
}

/* contact speed of one face (:2924-3056 x, :3064-3166 y) */
static double hlle_mcl_contact(int dir, const hlle_mcl_t *F, double v0) {
  const int n = dir, t1 = dir == 0 ? 1 : 0, t2 = 2;
  double a = F->tau_f - F->er[n];
  double b = (((sq(F->e[t1]) + sq(F->e[t2])) + (sq(F->b[t1]) + sq(F->b[t2]))) - F->tau_h) - F->s_f[n];
  double c = F->s_h[n] - F->er[n];
  return hllc_contact(a, b, c, v0);
}

/* the fluid and EM face fluxes of one face from its contact speed v */
static int hlle_mcl_fluxes(int dir, const side_t *L, const side_t *R, const hlle_mcl_t *F, double v, double *f) {
  const int n = dir, t1 = dir == 0 ? 1 : 2, t2 = dir == 0 ? 2 : 0; /* transverse momenta in reference order */
  const double a = F->a, b = F->b_;
  /* EM: the star "fluxes" of a face with v != 0 are its HLL states */
  if (v == 0.0) {
    for (int k = 0; k < 8; k++) f[k] = F->f_em[k];
  } else if (dir == 0) {
    f[FX_EXPSI] = F->e[0]; f[FX_BXPHI] = F->b[0]; f[FX_PHIBX] = F->phi_h;
    f[FX_EZBY] = -(-F->e[2]); f[FX_EYBZ] = F->e[1]; f[FX_PSIEX] = F->psi_h;
    f[FX_BZEY] = F->b[2]; f[FX_BYEZ] = -(-F->b[1]);
  } else {
    f[FY_EYPSI] = F->e[1]; f[FY_BYPHI] = F->b[1]; f[FY_EZBX] = F->e[2];
    f[FY_PHIBY] = F->phi_h; f[FY_EXBZ] = -(-F->e[0]); f[FY_BZEX] = -(-F->b[2]);
    f[FY_PSIEY] = F->psi_h; f[FY_BXEZ] = F->b[0];
  }
  /* fluid star states (:3174-3243) */
  double pstr = (F->s_f[n] + (sq(F->e[n]) + sq(F->b[n]))) - (F->tau_f - F->er[n]) * v;
  double sR[3], sL[3];
  sR[n] = (((((b * F->s_h[n] - F->s_f[n]) + pstr) - sq(F->b[n])) - sq(F->e[n])) - F->er[n] * v) / (b - v);
  sL[n] = (((((-(a * F->s_h[n]) + F->s_f[n]) - pstr) + sq(F->b[n])) + sq(F->e[n])) + F->er[n] * v) / (v - a);
  const int ts[2] = {t1, t2};
  for (int q = 0; q < 2; q++) {
    int t = ts[q];
    sR[t] = ((((b * F->s_h[t] - F->s_f[t]) - F->er[t] * v) - F->e[n] * F->e[t]) - F->b[n] * F->b[t]) / (b - v);
    sL[t] = ((((-(a * F->s_h[t]) + F->s_f[t]) + F->er[t] * v) + F->e[n] * F->e[t]) + F->b[n] * F->b[t]) / (v - a);
  }
  double tauR = (((b * F->tau_h - F->tau_f) + (pstr - F->e2b2s) * v) + F->er[n]) / (b - v);
  double tauL = ((((-(a * F->tau_h)) + F->tau_f) - (pstr - F->e2b2s) * v) - F->er[n]) / (v - a);
  double ddR = (b * F->dd_h - F->dd_f) / (b - v);
  double ddL = (-(a * F->dd_h) + F->dd_f) / (v - a);
  /* fluid star fluxes (:3287-3335, :3395-3450) */
  double fdR, fdL, ftR, ftL, fsR[3], fsL[3];
  double sideR[3] = {R->sx, R->sy, R->sz}, sideL[3] = {L->sx, L->sy, L->sz};
  double fsideR[3] = {R->fs1, R->fs2, R->fs3}, fsideL[3] = {L->fs1, L->fs2, L->fs3};
  if (v == 0.0) {
    fdR = fdL = F->dd_f; ftR = ftL = F->tau_f;
    for (int k = 0; k < 3; k++) fsR[k] = fsL[k] = F->s_f[k];
  } else {
    fdR = R->fd + b * (ddR - R->d);
    fdL = L->fd + a * (ddL - L->d);
    ftR = R->ftau + b * (tauR - R->tau);
    ftL = L->ftau + a * (tauL - L->tau);
    for (int k = 0; k < 3; k++) {
      fsR[k] = fsideR[k] + b * (sR[k] - sideR[k]);
      fsL[k] = fsideL[k] + a * (sL[k] - sideL[k]);
    }
  }
  /* selection (:3719-3800 x, :3905-3990 y) */
  double *o = f + FX_J; /* J, FD, Ftau, FS x, FS y, FS z: the same slots in both directions */
  if (a > 0.0) {
    o[0] = L->j1; o[1] = L->fd; o[2] = L->ftau; o[3] = L->fs1; o[4] = L->fs2; o[5] = L->fs3;
  } else if (a <= 0.0 && 0.0 <= v) {
    o[0] = F->q_f; o[1] = fdL; o[2] = ftL; o[3] = fsL[0]; o[4] = fsL[1]; o[5] = fsL[2];
  } else if (v < 0.0 && 0.0 <= b) {
    o[0] = F->q_f; o[1] = fdR; o[2] = ftR; o[3] = fsR[0]; o[4] = fsR[1]; o[5] = fsR[2];
  } else if (b < 0.0) {
    o[0] = R->j1; o[1] = R->fd; o[2] = R->ftau; o[3] = R->fs1; o[4] = R->fs2; o[5] = R->fs3;
  } else {
    return 8;
  }
  return 0;
}

/* seeds of the two contact solves in the reference build: see hllc_mcl_faces */
static int hlle_mcl_faces(const oracle_state *s, int i, int j, double *fx, double *fy) {
  side_t xl, xr, yl, yr;
  mcl_sides(s, i, j, 0, &xl, &xr);
  mcl_sides(s, i, j, 1, &yl, &yr);
  /* the sound speeds read the sides' enthalpy (:2243-2256) */
  side_t *sides[4] = {&xl, &xr, &yl, &yr};
  for (int k = 0; k < 4; k++) {
    double eps = sides[k]->p / ((s->c.gamma - 1.0) * sides[k]->rho);
    sides[k]->h = sides[k]->rho * (1.0 + eps) + sides[k]->p;
  }
  hlle_mcl_t X, Y;
  int st = hlle_mcl_prep(s->c.gamma, 0, &xl, &xr, &X);
  if (st) return st;
  st = hlle_mcl_prep(s->c.gamma, 1, &yl, &yr, &Y);
  if (st) return st;
  g_hllc_maxiter = 200;
  double vx = hlle_mcl_contact(0, &X, g_hllc_seed_mode ? Y.er[1] : NAN);
  int fail = g_hllc_fail;
  double vy = fail ? NAN : hlle_mcl_contact(1, &Y, g_hllc_seed_mode ? g_hllc_last_iterate : NAN);
  fail = fail || g_hllc_fail;
  g_hllc_maxiter = 100;
  if (fail) return 7;
  st = hlle_mcl_fluxes(0, &xl, &xr, &X, vx, fx);
  if (!st) st = hlle_mcl_fluxes(1, &yl, &yr, &Y, vy, fy);
  return st;
}

/* ------------------------------------------------------------------ */
/* Lax-Friedrichs flux splitting, flux_solver = 1: lax_flow,            */
/* 15_lax_flow.f95:733-1460 (DIM = 2).  At every cell the split fluxes  */
/* F+- = A +- Sp U (Sp = 1) of the stage fields and of the cell fluxes  */
/* flowinit forms (14_flowinit.f95:264-373); the face value is          */
/* 0.5 (L + R), L = F+(i) + FAC mcl(F+(i+1)-F+(i), F+(i)-F+(i-1)),        */
/* R = F-(i+1) - FAC mcl(F+(i+2)-F+(i+1), F+(i+1)-F+(i)): the slope of R */
/* is taken from F+ as well.  flowinit fills only the interior and      */
/* boundary_flux_conserved fills no ghosts for this solver, so the cell  */
/* fluxes are 0 on ghost cells (cleanvarint, 07_clean_var.f95:341-350).  */
/* ------------------------------------------------------------------ */
/* the six fluid cell fluxes along dir at cell k (14_flowinit.f95:291-373) */
static void lax_cell_fluxes(const oracle_state *s, int i, int j, int dir, double out[6]) {
  if (i < 0 || i > s->c.imax || j < 0 || j > s->c.jmax) {
    for (int t = 0; t < 6; t++) out[t] = 0.0;
    return;
  }
  double *const *S = s->st;
  size_t k = IX(s, i, j);
  /* v^n, p^n, rho^n, or with REC_PRIM = 1 the stage's primitives
   * (14_flowinit.f95:376-489: Wint and enthpyint as interconvar forms them) */
  double vx = s->kin[K_VX][k], vy = s->kin[K_VY][k], vz = s->kin[K_VZ][k], p = s->kin[K_P][k],
         rho = s->kin[K_RHO][k];
  double ex = S[ST_EX][k], ey = S[ST_EY][k], ez = S[ST_EZ][k];
  double bx = S[ST_BX][k], by = S[ST_BY][k], bz = S[ST_BZ][k];
  double w = s->c.rec_prim ? s->wint[s->cur_l - 1][k] : s->W[k]; /* 06_conserved_var.f95:92 */
  double eps = p / ((s->c.gamma - 1.0) * rho);                  /* :105 */
  double h = rho * (1.0 + eps) + p;                             /* :109 */
  double e2 = (sq(ex) + sq(ey)) + sq(ez), b2 = (sq(bx) + sq(by)) + sq(bz);
  double hw2 = h * sq(w);
  out[0] = S[dir == 0 ? ST_JX : ST_JY][k];
  if (dir == 0) {
    out[1] = S[ST_D][k] * vx;
    out[2] = (bz * ey - by * ez) + hw2 * vx;
    out[3] = ((((-sq(ex)) - sq(bx)) + hw2 * sq(vx)) + 0.5 * (e2 + b2)) + p;
    out[4] = ((-(ex * ey)) - bx * by) + (hw2 * vx) * vy;
    out[5] = ((-(ex * ez)) - bx * bz) + (hw2 * vx) * vz;
  } else {
    out[1] = S[ST_D][k] * vy;
    out[2] = (bx * ez - bz * ex) + hw2 * vy;
    out[3] = ((-(ex * ey)) - bx * by) + (hw2 * vx) * vy;
    out[4] = ((((-sq(ey)) - sq(by)) + hw2 * sq(vy)) + 0.5 * (e2 + b2)) + p;
    out[5] = ((-(ey * ez)) - by * bz) + (hw2 * vy) * vz;
  }
}

/* one face value from the split fluxes of cells i-1 .. i+2 (a: the flux,
 * u: the conserved field it is split with) */
static inline double lax_face_value(const double a[4], const double u[4]) {
  double p0 = a[0] + u[0], p1 = a[1] + u[1], p2 = a[2] + u[2], p3 = a[3] + u[3], m2 = a[2] - u[2];
  double l = p1 + 0.5 * mcl(p2 - p1, p1 - p0);
  double r = m2 - 0.5 * mcl(p3 - p2, p2 - p1);
  return 0.5 * (l + r);
}

static void lax_face(const oracle_state *s, int i, int j, int dir, double *f) {
  double *const *S = s->st;
  const int di = dir == 0, dj = dir == 1;
  size_t k[4];
  double fl[4][6];
  for (int c = 0; c < 4; c++) {
    k[c] = IX(s, i + (c - 1) * di, j + (c - 1) * dj);
    lax_cell_fluxes(s, i + (c - 1) * di, j + (c - 1) * dj, dir, fl[c]);
  }
#define LAX_FIELD(name, plane) double name[4]; for (int c = 0; c < 4; c++) name[c] = S[plane][k[c]];
  LAX_FIELD(psi, ST_PSI) LAX_FIELD(phi, ST_PHI) LAX_FIELD(ex, ST_EX) LAX_FIELD(ey, ST_EY) LAX_FIELD(ez, ST_EZ)
  LAX_FIELD(bx, ST_BX) LAX_FIELD(by, ST_BY) LAX_FIELD(bz, ST_BZ) LAX_FIELD(q, ST_Q) LAX_FIELD(d, ST_D)
  LAX_FIELD(tau, ST_TAU) LAX_FIELD(sx, ST_SX) LAX_FIELD(sy, ST_SY) LAX_FIELD(sz, ST_SZ)
#undef LAX_FIELD
  double nex[4], nez[4], nby[4], nbz[4], a[4];
  for (int c = 0; c < 4; c++) { nex[c] = -ex[c]; nez[c] = -ez[c]; nby[c] = -by[c]; nbz[c] = -bz[c]; }
  if (dir == 0) { /* 15_lax_flow.f95:930-1180 */
    f[FX_EXPSI] = lax_face_value(ex, psi);
    f[FX_BXPHI] = lax_face_value(bx, phi);
    f[FX_PHIBX] = lax_face_value(phi, bx);
    f[FX_EZBY] = -lax_face_value(nez, by);
    f[FX_EYBZ] = lax_face_value(ey, bz);
    f[FX_PSIEX] = lax_face_value(psi, ex);
    f[FX_BZEY] = lax_face_value(bz, ey);
    f[FX_BYEZ] = -lax_face_value(nby, ez);
  } else {
    f[FY_EYPSI] = lax_face_value(ey, psi);
    f[FY_BYPHI] = lax_face_value(by, phi);
    f[FY_EZBX] = lax_face_value(ez, bx);
    f[FY_PHIBY] = lax_face_value(phi, by);
    f[FY_EXBZ] = -lax_face_value(nex, bz);
    f[FY_BZEX] = -lax_face_value(nbz, ex);
    f[FY_PSIEY] = lax_face_value(psi, ey);
    f[FY_BXEZ] = lax_face_value(bx, ez);
  }
  const double *us[6] = {q, d, tau, sx, sy, sz};
  const int slot[6] = {FX_J, FX_FD, FX_FTAU, FX_FSX, FX_FSY, FX_FSZ};
  for (int t = 0; t < 6; t++) {
    for (int c = 0; c < 4; c++) a[c] = fl[c][t];
    f[slot[t]] = lax_face_value(a, us[t]);
  }
}

/* fluxandsource, 2D IMEX branch: 17_fluxandsource.f95:442-542 */
static void fluxandsource(oracle_state *s, int l) {
  const double dx = s->c.dx, dy = s->c.dy, sig = s->c.sigma;
  double **P = s->prim, **S = s->st, **X = s->fx, **Y = s->fy, **R = s->rhs[l];
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j), kw = IX(s, i - 1, j), ks = IX(s, i, j - 1);
#define DXF(a) ((X[a][k] - X[a][kw]) / dx)
#define DYF(a) ((Y[a][k] - Y[a][ks]) / dy)
      double vx = P[OR_VX][k], vy = P[OR_VY][k], vz = P[OR_VZ][k], w = s->W[k];
      if (s->c.rec_prim) { /* 17_fluxandsource.f95:460-474 */
        vx = s->vint[l][0][k]; vy = s->vint[l][1][k]; vz = s->vint[l][2][k];
        w = s->wint[l][k];
      }
      double ex = S[ST_EX][k], ey = S[ST_EY][k], ez = S[ST_EZ][k];
      double bx = S[ST_BX][k], by = S[ST_BY][k], bz = S[ST_BZ][k], q = S[ST_Q][k];
      double edv = (ex * vx + ey * vy) + ez * vz;
      double rx = bz * vy - by * vz, ry = bx * vz - bz * vx, rz = by * vx - bx * vy;
      R[0][k] = ((((-DXF(FX_EXPSI)) - DYF(FY_EYPSI)) - 0.0) + q) - s->c.kappa_psi * S[ST_PSI][k];
      R[1][k] = (((-DXF(FX_BXPHI)) - DYF(FY_BYPHI)) - 0.0) - s->c.kappa_phi * S[ST_PHI][k];
      R[2][k] = ((-DYF(FY_EZBX)) + 0.0) - DXF(FX_PHIBX);
      R[3][k] = ((-0.0) + DXF(FX_EZBY)) - DYF(FY_PHIBY);
      R[4][k] = ((-DXF(FX_EYBZ)) + DYF(FY_EXBZ)) - 0.0;
      R[5][k] = ((DYF(FY_BZEX) - 0.0) - DXF(FX_PSIEX)) - q * vx;
      R[6][k] = ((0.0 - DXF(FX_BZEY)) - DYF(FY_PSIEY)) - q * vy;
      R[7][k] = ((DXF(FX_BYEZ) - DYF(FY_BXEZ)) - 0.0) - q * vz;
      R[8][k] = ((-DXF(FX_J)) - DYF(FY_J)) - 0.0;
      R[9][k] = ((-DXF(FX_FD)) - DYF(FY_FD)) - 0.0;
      R[10][k] = ((-DXF(FX_FTAU)) - DYF(FY_FTAU)) - 0.0;
      R[11][k] = ((-DXF(FX_FSX)) - DYF(FY_FSX)) - 0.0;
      R[12][k] = ((-DXF(FX_FSY)) - DYF(FY_FSY)) - 0.0;
      R[13][k] = ((-DXF(FX_FSZ)) - DYF(FY_FSZ)) - 0.0;
      R[14][k] = (sig * w) * ((ex + rx) - edv * vx);
      R[15][k] = (sig * w) * ((ey + ry) - edv * vy);
      R[16][k] = (sig * w) * ((ez + rz) - edv * vz);
#undef DXF
#undef DYF
    }
}

static void sumfin(oracle_state *s, int l) {
  const double wb = s->c.wb[l], wbt = s->c.wbt[l];
  for (int a = 0; a < AC_N; a++) {
    double *acc = s->finsum[a];
    const double *f = s->rhs[l][a];
    const double *src = (a >= AC_EX && a <= AC_EZ) ? s->rhs[l][14 + (a - AC_EX)] : NULL;
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) {
        size_t k = IX(s, i, j);
        if (src)
          acc[k] = acc[k] + (wb * f[k] - wbt * src[k]);
        else
          acc[k] = acc[k] + wb * f[k];
      }
  }
}

/* varfin (19_varfin.f95:20-50), which also keeps B^n and E^n (Bxn..Ezn,
 * :24-30) for the MIRK electric field, and electricfield with MIRK = 0
 * (20_electricfield.f95:153-167): U^{n+1} = U^n + dt finsum.  The MIRK
 * electric field follows in electricfield_mirk. */
static void varfin(oracle_state *s) {
  const double dt = s->c.dt;
  for (int a = 0; a < AC_N; a++) {
    double *u = (double *)un_plane(s, a);
    const int e = a >= AC_EX && a <= AC_EZ;
    if (e && s->c.mirk) continue;
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) {
        size_t k = IX(s, i, j);
        u[k] = u[k] + dt * s->finsum[a][k];
      }
  }
}

/* electricfield, MIRK = 1 (20_electricfield.f95:13-48) and MIRK = 2
 * (:56-146), over the interior after varfin: bn, en are B^n and E^n;
 * vst, wst the last stage's velocity and its W, which the MIRK2 branch
 * recomputes from the velocity and stores (:66-71). */
static void electricfield_mirk(oracle_state *s, double *const bn[3], double *const en[3], double *const vst[3],
                               double *wst) {
  const double sig = s->c.sigma, dt = s->c.dt, c4 = cm4();
  double **P = s->prim, **S = s->st, **F = s->finsum;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      if (s->c.mirk == 1) {
        double w = s->W[k], vx = P[OR_VX][k], vy = P[OR_VY][k], vz = P[OR_VZ][k];
        double ex = P[OR_EX][k], ey = P[OR_EY][k], ez = P[OR_EZ][k];
        double sb = (dt * sig) * w;
        double edv = (ex * vx + ey * vy) + ez * vz;
        double nx = bn[2][k] * vy - bn[1][k] * vz, ny = bn[0][k] * vz - bn[2][k] * vx,
               nz = bn[1][k] * vx - bn[0][k] * vy;
        double bx = P[OR_BX][k], by = P[OR_BY][k], bz = P[OR_BZ][k];
        double rx = bz * vy - by * vz, ry = bx * vz - bz * vx, rz = by * vx - bx * vy;
        double inv = 1.0 / (1.0 + sb);
        P[OR_EX][k] = inv * (((ex + (sb * vx) * edv) - sb * (CM2 * nx + (1.0 - CM2) * rx)) + dt * F[AC_EX][k]);
        P[OR_EY][k] = inv * (((ey + (sb * vy) * edv) - sb * (CM2 * ny + (1.0 - CM2) * ry)) + dt * F[AC_EY][k]);
        P[OR_EZ][k] = inv * (((ez + (sb * vz) * edv) - sb * (CM2 * nz + (1.0 - CM2) * rz)) + dt * F[AC_EZ][k]);
        continue;
      }
      double vx = vst[0][k], vy = vst[1][k], vz = vst[2][k];
      double v2 = (sq(vx) + sq(vy)) + sq(vz);
      double w = 1.0 / sqrt(1.0 - v2);
      wst[k] = w;
      double sdb = (dt * sig) * w;
      double eix = S[ST_EX][k], eiy = S[ST_EY][k], eiz = S[ST_EZ][k];
      double edv = (eix * vx + eiy * vy) + eiz * vz;
      double bx = S[ST_BX][k], by = S[ST_BY][k], bz = S[ST_BZ][k];
      double rx = bz * vy - by * vz, ry = bx * vz - bz * vx, rz = by * vx - bx * vy;
      double a = -1.0 + sdb * (1.0 - CM1);
      double den = 1.0 + (0.5 * CM1 - c4) * sdb;
      P[OR_EX][k] = (eix + ((0.5 * a) / den) * (eix - en[0][k])) +
                    (0.5 / den) * ((((-(sdb * eix)) + (sdb * edv) * vx) + dt * F[AC_EX][k]) - sdb * rx);
      P[OR_EY][k] = (eiy + ((0.5 * a) / den) * (eiy - en[1][k])) +
                    (0.5 / den) * ((((-(sdb * eiy)) + (sdb * edv) * vy) + dt * F[AC_EY][k]) - sdb * ry);
      P[OR_EZ][k] = (eiz + ((0.5 * a) / den) * (eiz - en[2][k])) +
                    (0.5 / den) * ((((-(sdb * eiz)) + (sdb * edv) * vz) + dt * F[AC_EZ][k]) - sdb * rz);
    }
}

/* ------------------------------------------------------------------ */
/* varprimitivecardano, l_swap = 1: 12_varprimitive.f95:48-813          */
/* in:  psi phi Bx By Bz Vx Vy Vz Ex Ey Ez D tau S..., see below        */
/* ------------------------------------------------------------------ */

/* in[14] = {D, tau, Sx, Sy, Sz, Ex, Ey, Ez, Bx, By, Bz, Vx0, Vy0, Vz0} plus
 * the old rho and p through out[3], out[4] on entry.
 * out[5] = {Vx, Vy, Vz, rho, p}; flags = fix-ups {rho, p, drift}.
 * Returns 0 or a stop status. */
int oracle_c2p_point(double gamma, double eps1, const double in[14], double out[5], int flags[3], double rec[6]) {
  const double TOL = 1.0e-10;
  const int nrmax = 40;
  double D = in[0], tau = in[1], Sx = in[2], Sy = in[3], Sz = in[4];
  double Ex = in[5], Ey = in[6], Ez = in[7], Bx = in[8], By = in[9], Bz = in[10];
  double vx0 = in[11], vy0 = in[12], vz0 = in[13];
  double rho_old = out[3], p_old = out[4];
  flags[0] = flags[1] = flags[2] = 0;

  double E2B2 = 0.5 * (((((sq(Ex) + sq(Ey)) + sq(Ez)) + sq(Bx)) + sq(By)) + sq(Bz));
  double g1 = (gamma - 1.0) / gamma;
  double V2c = (sq(vx0) + sq(vy0)) + sq(vz0);
  /* :236-241: a superluminal previous velocity is clamped (it enters only the
   * pressure loop below; W is taken from the quartic) */
  if (V2c > 1.0) V2c = fmax2(-1.0 + eps1, fmin2(1.0 - eps1, V2c));
  for (int k = 0; k < 6; k++) rec[k] = 0.0;

  double C1 = (sq((Sx - Ey * Bz) + Ez * By) + sq((Sy - Ez * Bx) + Ex * Bz)) + sq((Sz - Ex * By) + Ey * Bx);
  double C2 = tau - E2B2;
  double a4 = C1 - sq(C2);
  double a3 = ((2.0 * g1) * C2) * D;
  double a2 = (sq(C2) - (2.0 * g1) * C1) - sq(g1) * sq(D);
  double a1 = -(((2.0 * g1) * C2) * D);
  double a0 = sq(g1) * (C1 + sq(D));
  double b3 = a3 / a4, b2 = a2 / a4, b1 = a1 / a4, b0 = a0 / a4;
  double d2 = b2 - (3.0 * sq(b3)) / 8.0;
  double d1 = (b1 - (b3 * b2) / 2.0) + pw3(b3) / 8.0;
  double d0 = ((b0 - (b3 * b1) / 4.0) + (sq(b3) * b2) / 16.0) - (3.0 * pw4(b3)) / 256.0;
  double e = (20.0 * d2) / 8.0;
  double f = -(d0 - 2.0 * sq(d2));
  double g = -((sq(d1) + (4.0 * d2) * d0) - 4.0 * pw3(d2)) / 8.0;
  double Q1 = (sq(e) - 3.0 * f) / 9.0;
  double R1 = ((2.0 * pw3(e) - (9.0 * e) * f) + 27.0 * g) / 54.0;
  double s, t, det, z1;
  const double pi = 3.14159265358979323846264338327950288;
  if (sq(R1) < pw3(Q1)) { /* :340-425 */
    double T1 = o_acos(R1 / fabs(sqrt(pw3(Q1))));
    double sQ = fabs(sqrt(Q1));
    double za = (-((2.0 * sQ) * o_cos(T1 / 3.0))) - e / 3.0;
    double sa = d2 + 2.0 * za, ta = ((sq(za) + (2.0 * d2) * za) + sq(d2)) - d0;
    double zb = (-((2.0 * sQ) * o_cos((T1 + 2.0 * pi) / 3.0))) - e / 3.0;
    double sb = d2 + 2.0 * zb, tb = ((sq(zb) + (2.0 * d2) * zb) + sq(d2)) - d0;
    double zc = (-((2.0 * sQ) * o_cos((T1 - 2.0 * pi) / 3.0))) - e / 3.0;
    double sc = d2 + 2.0 * zc, tc = ((sq(zc) + (2.0 * d2) * zc) + sq(d2)) - d0;
    if (sa >= 0.0 && ta >= 0.0) { s = fabs(sqrt(sa)); t = fabs(sqrt(ta)); z1 = za; }
    else if (sb >= 0.0 && tb >= 0.0) { s = fabs(sqrt(sb)); t = fabs(sqrt(tb)); z1 = zb; }
    else if (sc >= 0.0 && tc >= 0.0) { s = fabs(sqrt(sc)); t = fabs(sqrt(tc)); z1 = zc; }
    else return 21; /* "todo cardano IMAGINARIO" */
    det = sq(s) - 4.0 * ((d2 + z1) - t);
    if (det < 0.0) return 22;
  } else { /* :427-452 */
    double xx1 = fabs(R1) + fabs(sqrt(sq(R1) - pw3(Q1)));
    double AA = -(copysign(1.0, R1) * fsign(fabs(o_pow(xx1, 1.0 / 3.0)), xx1));
    double BB = AA != 0.0 ? Q1 / AA : 0.0;
    z1 = (AA + BB) - e / 3.0;
    s = fabs(sqrt(d2 + 2.0 * z1));
    t = fabs(sqrt(((sq(z1) + (2.0 * d2) * z1) + sq(d2)) - d0));
    det = sq(s) - 4.0 * ((d2 + z1) - t);
  }
  double Q2 = -(0.5 * ((-s) + copysign(1.0, -s) * fabs(sqrt(det))));
  double y1 = Q2 - b3 / 4.0;
  double y2 = (((d2 + z1) - t) / Q2) - b3 / 4.0;
  double Wn = NAN;
  if (y1 >= 1.0 && y2 >= 1.0) Wn = fmin2(y1, y2);
  if (y1 <= 1.0 && y2 <= 1.0) return 23;
  if (y1 >= 1.0 && y2 <= 1.0) Wn = y1;
  if (y1 <= 1.0 && y2 >= 1.0) Wn = y2;
  /* No root selected (NaN roots): the reference then keeps Wnro from the
   * previous cell's Newton loop (:463-500 assign nothing).  With a NaN or
   * infinite quartic coefficient every Newton update is NaN whatever the
   * start, so the cell goes on to NaN results as in the reference; with
   * finite coefficients the previous cell's W would matter, and the
   * restatement stops instead (status 24). */
  if (Wn != Wn && isfinite(b3) && isfinite(b2) && isfinite(b1) && isfinite(b0)) return 24;

  /* Newton on W, :510-526 */
  int nrw = 1;
  double cha1 = 1.0;
  while (cha1 > TOL && nrw < nrmax) {
    double Wr = Wn - ((((pw4(Wn) + b3 * pw3(Wn)) + b2 * sq(Wn)) + b1 * Wn) + b0) /
                         (((4.0 * pw3(Wn) + (3.0 * b3) * sq(Wn)) + (2.0 * b2) * Wn) + b1);
    cha1 = fabs(Wr - Wn) / (0.5 * (Wr + Wn));
    Wn = Wr;
    nrw = nrw + 1;
  }
  double W = Wn;
  double rho = D / W;
  if (rho < 0.0) { flags[0] = 1; rec[0] = rho; rec[1] = rho_old; rho = fmax2(rho_old * 1.0e-1, rho); }
  double h = ((tau - E2B2) - g1 * rho) / (sq(W) - g1);
  double p = g1 * (h - rho);

  /* pressure fixed point, :558-587 */
  double pn = p;
  int nr = 1;
  double epsn = (((tau - E2B2) - D * W) + pn * (1.0 - sq(W))) / (D * W);
  double cs2 = ((gamma * (gamma - 1.0)) * epsn) / (1.0 + gamma * epsn);
  double cha = 1.0;
  while (cha >= TOL && nr <= nrmax) {
    double pr = pn - (((gamma - 1.0) * rho) * epsn - pn) / (V2c * cs2 - 1.0);
    cha = fabs(pr - pn) / (0.5 * (pr + pn));
    pn = pr;
    nr = nr + 1;
  }
  if (pn < 0.0) { flags[1] = 1; rec[2] = pn; rec[3] = p_old; pn = fmax2(p_old * 1.0e-4, pn); }
  p = pn;
  tau = (E2B2 + h * sq(W)) - p;
  double den = (tau - E2B2) + p;
  double vx = ((Sx - Ey * Bz) + Ez * By) / den;
  double vy = ((Sy - Ez * Bx) + Ex * Bz) / den;
  double vz = ((Sz - Ex * By) + Ey * Bx) / den;
  V2c = (sq(vx) + sq(vy)) + sq(vz);
  if (V2c >= 1.0) { /* drift branch :627-696 */
    double Wsl = 1.0 / sqrt(1.0 - V2c);
    double erb = ((((Ey * Bz - Ez * By) + Ez * Bx) - Ex * Bz) + Ex * By) - Ey * Bx;
    double b2d = (sq(Bx) + sq(By)) + sq(Bz);
    if (!(b2d / rho >= 10.0)) return 25;
    double vd = erb / b2d;
    double Wd = 1.0 / sqrt(1.0 - sq(vd));
    rho = D / Wd;
    h = ((tau - E2B2) - g1 * rho) / (sq(Wd) - g1);
    p = g1 * (h - rho);
    tau = (E2B2 + h * sq(Wsl)) - p;
    den = (tau - E2B2) + p;
    vx = ((Sx - Ey * Bz) + Ez * By) / den;
    vy = ((Sy - Ez * Bx) + Ex * Bz) / den;
    vz = ((Sz - Ex * By) + Ey * Bx) / den;
    V2c = (sq(vx) + sq(vy)) + sq(vz);
    flags[2] = 1;
    rec[4] = V2c;
    rec[5] = sq(vd);
    if (V2c >= 1.0) return 26;
  }
  out[0] = vx; out[1] = vy; out[2] = vz; out[3] = rho; out[4] = p;
  return 0;
}

static int add_fixup(oracle_state *s, int kind, int i, int j, double a, double b) {
  if (s->nfix[kind] == s->capfix[kind]) {
    int cap = s->capfix[kind] ? 2 * s->capfix[kind] : 64;
    int32_t *ij = realloc(s->fix_ij[kind], sizeof(int32_t) * 2 * (size_t)cap);
    if (!ij) return 90;
    s->fix_ij[kind] = ij;
    double *v = realloc(s->fix_val[kind], sizeof(double) * 2 * (size_t)cap);
    if (!v) return 90;
    s->fix_val[kind] = v;
    s->capfix[kind] = cap;
  }
  int n = s->nfix[kind]++;
  s->fix_ij[kind][2 * n] = i;
  s->fix_ij[kind][2 * n + 1] = j;
  s->fix_val[kind][2 * n] = a;
  s->fix_val[kind][2 * n + 1] = b;
  return 0;
}

int oracle_get_fixups(const oracle_state *s, int kind, int32_t *ij, double *vals, int cap) {
  if (kind < 0 || kind > 2) return -1;
  int n = s->nfix[kind];
  for (int e = 0; e < n && e < cap; e++) {
    ij[2 * e] = s->fix_ij[kind][2 * e];
    ij[2 * e + 1] = s->fix_ij[kind][2 * e + 1];
    vals[2 * e] = s->fix_val[kind][2 * e];
    vals[2 * e + 1] = s->fix_val[kind][2 * e + 1];
  }
  return n;
}

/* varprimitivecardano with l_swap = 0 on stage l (12_varprimitive.f95:111-135,
 * 792-806; REC_PRIM = 1, main_cueva_code.f95:220-224), started from the
 * stage's primitives as they stand (the previous step's, zero at first),
 * then interconvar (13_intercon_var.f95:13-72) and its ghost fills */
static int c2p_stage(oracle_state *s, int l) {
  double **S = s->st;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      double in[14] = {S[ST_D][k], S[ST_TAU][k], S[ST_SX][k], S[ST_SY][k], S[ST_SZ][k],
                       S[ST_EX][k], S[ST_EY][k], S[ST_EZ][k], S[ST_BX][k], S[ST_BY][k], S[ST_BZ][k],
                       s->vint[l][0][k], s->vint[l][1][k], s->vint[l][2][k]};
      double out[5] = {0, 0, 0, s->rhoint[l][k], s->pint[l][k]};
      int fl[3];
      double rec[6];
      int st = oracle_c2p_point(s->c.gamma, s->eps1, in, out, fl, rec);
      if (st) {
        s->stop[2] = i;
        s->stop[3] = j;
        return st;
      }
      s->cnt[0] += fl[0]; s->cnt[1] += fl[1]; s->cnt[2] += fl[2];
      for (int kind = 0; kind < 3; kind++)
        if (fl[kind] && add_fixup(s, kind, i, j, rec[2 * kind], rec[2 * kind + 1])) return 90;
      s->vint[l][0][k] = out[0]; s->vint[l][1][k] = out[1]; s->vint[l][2][k] = out[2];
      s->rhoint[l][k] = out[3]; s->pint[l][k] = out[4];
    }
  const double gam = s->c.gamma;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      double vx = s->vint[l][0][k], vy = s->vint[l][1][k], vz = s->vint[l][2][k];
      double rho = s->rhoint[l][k], p = s->pint[l][k];
      double ex = S[ST_EX][k], ey = S[ST_EY][k], ez = S[ST_EZ][k];
      double bx = S[ST_BX][k], by = S[ST_BY][k], bz = S[ST_BZ][k];
      double v2 = (sq(vx) + sq(vy)) + sq(vz);
      double w = 1.0 / sqrt(1.0 - v2);
      double e2 = (sq(ex) + sq(ey)) + sq(ez);
      double b2 = (sq(bx) + sq(by)) + sq(bz);
      double eps = p / ((gam - 1.0) * rho);
      double h = rho * (1.0 + eps) + p;
      s->wint[l][k] = w;
      s->hint[l][k] = h;
      S[ST_D][k] = rho * w;
      S[ST_TAU][k] = (0.5 * (e2 + b2) + h * sq(w)) - p;
      S[ST_SX][k] = (bz * ey - by * ez) + (h * sq(w)) * vx;
      S[ST_SY][k] = (bx * ez - bz * ex) + (h * sq(w)) * vy;
      S[ST_SZ][k] = (by * ex - bx * ey) + (h * sq(w)) * vz;
    }
  static const int cons[11] = {ST_PSI, ST_PHI, ST_BX, ST_BY, ST_BZ, ST_Q, ST_D, ST_TAU, ST_SX, ST_SY, ST_SZ};
  for (int k = 0; k < 11; k++) fill_ghosts(s, S[cons[k]]);
  fill_ghosts(s, S[ST_EX]);
  fill_ghosts(s, S[ST_EY]);
  fill_ghosts(s, S[ST_EZ]);
  return 0;
}

/* varprimitivecardano after varfin: l_swap = 1 (the step's primitives as
 * the start), or with REC_PRIM = 1 l_swap = 2 (the last stage's velocity,
 * rho and p, 12_varprimitive.f95:59-83) */
static int c2p(oracle_state *s) {
  double **P = s->prim;
  const int L = s->c.lmax - 1, rp = s->c.rec_prim;
  if (!rp) s->nfix[0] = s->nfix[1] = s->nfix[2] = 0;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      double in[14] = {s->D[k], s->tau[k], s->Sx[k], s->Sy[k], s->Sz[k],
                       P[OR_EX][k], P[OR_EY][k], P[OR_EZ][k], P[OR_BX][k], P[OR_BY][k], P[OR_BZ][k],
                       rp ? s->vint[L][0][k] : P[OR_VX][k], rp ? s->vint[L][1][k] : P[OR_VY][k],
                       rp ? s->vint[L][2][k] : P[OR_VZ][k]};
      double out[5] = {0, 0, 0, rp ? s->rhoint[L][k] : P[OR_RHO][k], rp ? s->pint[L][k] : P[OR_P][k]};
      int fl[3];
      double rec[6];
      int st = oracle_c2p_point(s->c.gamma, s->eps1, in, out, fl, rec);
      if (st) {
        s->stop[2] = i;
        s->stop[3] = j;
        return st;
      }
      s->cnt[0] += fl[0]; s->cnt[1] += fl[1]; s->cnt[2] += fl[2];
      for (int kind = 0; kind < 3; kind++)
        if (fl[kind] && add_fixup(s, kind, i, j, rec[2 * kind], rec[2 * kind + 1])) return 90;
      P[OR_VX][k] = out[0]; P[OR_VY][k] = out[1]; P[OR_VZ][k] = out[2];
      P[OR_RHO][k] = out[3]; P[OR_P][k] = out[4];
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* one time step: main_cueva_code.f95:47-332                            */
/* ------------------------------------------------------------------ */

static void install_modes(const oracle_state *s);

int oracle_recover(oracle_state *s) {
  install_modes(s);
  s->cnt[0] = s->cnt[1] = s->cnt[2] = 0;
  s->nfix[0] = s->nfix[1] = s->nfix[2] = 0;
  s->stop[0] = 1;
  s->stop[1] = s->c.lmax;
  return c2p(s);
}

void oracle_face_events(const oracle_state *s, int64_t out[2]) {
  out[0] = s->face_events[0];
  out[1] = s->face_events[1];
}

void oracle_last_stop(const oracle_state *s, int32_t out[4]) {
  for (int k = 0; k < 4; k++) out[k] = s->stop[k];
}

static void install_modes(const oracle_state *s) {
  g_libm_mode = s->c.libm_mode;
  g_hllc_seed_mode = s->c.hllc_seed;
  g_slop = s->c.slop ? s->c.slop : 2;
  g_limit = s->c.limit ? s->c.limit : 2;
}

/* swap1, 08_swap1.f95:13-33: the primitives of the step start into the
 * stage-1 arrays (interior) */
static void swap1(oracle_state *s) {
  double **P = s->prim;
  for (int j = 0; j <= s->c.jmax; j++)
    for (int i = 0; i <= s->c.imax; i++) {
      size_t k = IX(s, i, j);
      double gam = s->c.gamma, rho = P[OR_RHO][k], p = P[OR_P][k];
      double eps = p / ((gam - 1.0) * rho);
      s->rhoint[0][k] = rho;
      s->pint[0][k] = p;
      s->vint[0][0][k] = P[OR_VX][k]; s->vint[0][1][k] = P[OR_VY][k]; s->vint[0][2][k] = P[OR_VZ][k];
      s->wint[0][k] = s->W[k];
      s->hint[0][k] = rho * (1.0 + eps) + p;
    }
}

/* B^n and E^n of the interior, as varfin keeps them (19_varfin.f95:24-30) */
static void keep_bn_en(oracle_state *s) {
  for (int d = 0; d < 3; d++)
    for (int j = 0; j <= s->c.jmax; j++)
      for (int i = 0; i <= s->c.imax; i++) {
        size_t k = IX(s, i, j);
        s->bn[d][k] = s->prim[OR_BX + d][k];
        s->en[d][k] = s->prim[OR_EX + d][k];
      }
}

int oracle_step(oracle_state *s, int nsteps) {
  install_modes(s);
  for (int h = 0; h < nsteps; h++) {
    for (int k = 0; k < OR_NPRIM; k++) fill_ghosts(s, s->prim[k]);
    if (s->c.primitive_rec) { /* main_cueva_code.f95:170-175 */
      if (recons_prims(s)) return 90;
      for (int k = 0; k < OR_NPRIM; k++) fill_ghosts(s, s->prim[k]);
    }
    startconvar(s);
    cleanvarint(s);
    s->cnt[0] = s->cnt[1] = s->cnt[2] = 0;
    s->nfix[0] = s->nfix[1] = s->nfix[2] = 0;
    if (s->vint[0][0]) swap1(s);
    for (int l = 0; l < s->c.lmax; l++) {
      for (int m = 0; m < l; m++) sumint(s, l, m);
      varint(s);
      electricvarint(s, l);
      if (s->c.rec_prim) { /* main_cueva_code.f95:220-228 */
        int st = c2p_stage(s, l);
        if (st) {
          s->stop[0] = h + 1;
          s->stop[1] = l + 1;
          return st;
        }
      }
      set_kin(s, l);
      flowinit(s, l);
      if (s->c.flux_solver == OR_HLL || s->c.flux_solver == OR_HLLE || s->c.flux_solver == OR_HLLC) reconstruct(s);
      s->cur_l = l + 1;
      int st = face_loop(s);
      if (st) {
        s->stop[0] = h + 1;
        s->stop[1] = l + 1;
        return st;
      }
      fluxandsource(s, l);
      sumfin(s, l);
    }
    if (s->c.mirk) keep_bn_en(s);
    varfin(s);
    if (s->c.mirk) {
      const int L = s->c.lmax - 1;
      electricfield_mirk(s, s->bn, s->en, s->vint[L], s->wint[L]);
    }
    int st = c2p(s);
    if (st) {
      s->stop[0] = h + 1;
      s->stop[1] = s->c.lmax;
      return st;
    }
    for (int k = 0; k < OR_NPRIM; k++) fill_ghosts(s, s->prim[k]);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* global diagnostics, 22_writer.f95:818-991: sequential sums with i    */
/* outer and j inner, each statement evaluated left to right.           */
/* ------------------------------------------------------------------ */
void oracle_globals(const oracle_state *s, double out[37]) {
  const double a = s->c.dx, b = s->c.dy, gm = s->c.gamma, B0 = 1.0e-1; /* parameters.f95:131 */
  double *const *P = s->prim;
  double sm[35] = {0};
  double mbx = -INFINITY, mphi = -INFINITY;
  for (int i = 0; i <= s->c.imax; i++)
    for (int j = 0; j <= s->c.jmax; j++) {
      size_t k = IX(s, i, j);
      double psi = P[OR_PSI][k], phi = P[OR_PHI][k], ex = P[OR_EX][k], ey = P[OR_EY][k], ez = P[OR_EZ][k];
      double bx = P[OR_BX][k], by = P[OR_BY][k], bz = P[OR_BZ][k], vx = P[OR_VX][k], vy = P[OR_VY][k];
      double vz = P[OR_VZ][k], q = P[OR_Q][k], p = P[OR_P][k], rho = P[OR_RHO][k];
      double d = s->D[k], tau = s->tau[k];
      sm[0] = sm[0] + psi * a * b;
      sm[1] = sm[1] + phi * a * b;
      sm[2] = sm[2] + sq(ex) * a * b;
      sm[3] = sm[3] + sq(ey) * a * b;
      sm[4] = sm[4] + sq(ez) * a * b;
      sm[5] = sm[5] + ex * ey * a * b;
      sm[6] = sm[6] + ex * ez * a * b;
      sm[7] = sm[7] + ey * ez * a * b;
      sm[8] = sm[8] + sq(bx) * a * b;
      sm[9] = sm[9] + sq(by) * a * b;
      sm[10] = sm[10] + sq(bz) * a * b;
      double m1 = bx / B0;
      mbx = m1 > mbx ? m1 : mbx;
      mphi = phi > mphi ? phi : mphi;
      sm[11] = sm[11] + bx * by * a * b;
      sm[12] = sm[12] + bx * bz * a * b;
      sm[13] = sm[13] + by * bz * a * b;
      sm[15] = sm[15] + (sq(bx) + sq(by) + sq(bz)) * a * b; /* sum_bxbybz */
      sm[14] = sm[14] + (sq(vx) + sq(vy) + sq(vz)) * a * b; /* sum_vxvyvz */
      sm[16] = sm[16] + sq(vx) * a * b;
      sm[17] = sm[17] + sq(vy) * a * b;
      sm[18] = sm[18] + sq(vz) * a * b;
      sm[19] = sm[19] + vx * vy * a * b;
      sm[20] = sm[20] + vx * vz * a * b;
      sm[21] = sm[21] + vy * vz * a * b;
      sm[22] = sm[22] + q * a * b;
      sm[23] = sm[23] + p * a * b;
      sm[24] = sm[24] + rho * a * b;
      sm[33] = sm[33] + d * a * b; /* sum_D */
      double eps = p / ((gm - 1.0) * rho);
      double h = rho * (1.0 + eps) + p;
      double w = 1.0 / sqrt(1.0 - (sq(vx) + sq(vy) + sq(vz)));
      double e2 = sq(ex) + sq(ey) + sq(ez);
      double b2 = sq(by) + sq(bz);
      sm[25] = sm[25] + 0.5 * rho * (sq(vx) + sq(vy) + sq(vz)) * a * b; /* ekin_c */
      sm[26] = sm[26] + d * (w - 1.0) * a * b;                       /* ekin_r */
      sm[30] = sm[30] + rho * eps * a * b;                           /* eint_c */
      sm[31] = sm[31] + sq(w) * (rho * eps + p) - p * a * b;         /* eint_r */
      sm[27] = sm[27] + tau * a * b;                                 /* etot */
      sm[28] = sm[28] + h * sq(w) - p * a * b;                       /* efluid_r */
      sm[29] = sm[29] + 0.5 * (e2 + b2) * a * b;                     /* emag */
      sm[32] = sm[32] + w * a * b;                                   /* lorentz */
    }
  int imed1 = (int)floor(0.25 * s->c.imax), imed2 = (int)floor(0.75 * s->c.imax);
  for (int i = imed1; i <= imed2; i++)
    for (int j = 0; j <= s->c.jmax; j++) sm[34] = sm[34] + s->D[IX(s, i, j)] * a * b;
  for (int k = 0; k < 35; k++) out[k] = sm[k];
  out[35] = mbx;
  out[36] = mphi;
}
