// Device-side data of the CUEVA step: plane sets, scheme constants and the
// status word shared by every kernel.
#pragma once

#include "cueva_device.hpp"

namespace cueva {

// ---- field planes --------------------------------------------------------
// primitive/EM planes, restart order (22_writer.f95:1116-1172)
enum { P_PSI, P_PHI, P_BX, P_BY, P_BZ, P_VX, P_VY, P_VZ, P_EX, P_EY, P_EZ, P_Q, P_RHO, P_P, P_N };
// startconvar outputs (06_conserved_var.f95)
enum { C_D, C_TAU, C_SX, C_SY, C_SZ, C_W, C_N };
// stage planes (the l-indexed arrays of fourvectors.f95, one stage at a time)
enum { S_PSI, S_PHI, S_BX, S_BY, S_BZ, S_EXA, S_EYA, S_EZA, S_Q, S_D, S_TAU, S_SX, S_SY, S_SZ,
       S_EX, S_EY, S_EZ, S_JX, S_JY, S_JZ, S_N };
// accumulated fields of sumint/sumfin (09_sumint.f95, 18_sumfin.f95)
enum { A_PSI, A_PHI, A_BX, A_BY, A_BZ, A_EX, A_EY, A_EZ, A_Q, A_D, A_TAU, A_SX, A_SY, A_SZ, A_N };
// stage right-hand side: 14 flux divergences (A_* order) + 3 stiff sources
enum { R_SEX = 14, R_SEY = 15, R_SEZ = 16, R_N = 17 };

constexpr int MAXL = 4;

// boundary_flux_conserved fills the ghosts of J only for flux_solver 5 and 7
// (24_boundary_flux_conserved.f95:13-14); otherwise they keep their zeros
__host__ __device__ inline bool j_ghosts_filled(int flux_solver) { return flux_solver == 5 || flux_solver == 7; }

struct Planes {
  double* p[P_N];
  double* c[C_N];
  double* s[S_N];
  double* sn[S_N];  // the next stage's state, written by the marching kernel while s is read
  double* fx[FX_N];
  double* fy[FY_N];
  float* fxf[FX_N];  // the same memory as fx/fy, as FP32 planes (mixed form)
  float* fyf[FY_N];
  double* rhs[MAXL][R_N];
  double* ints[A_N];
  double* fins[A_N];
  // where a primitive recovery writes Vx Vy Vz rho p: the primitive planes
  // themselves, or (inside the marching step, whose stencils still read the
  // t^n values) scratch planes that then take their place
  double* pn[5];
  // stage primitives of REC_PRIM = 1, per stage (fourvectors.f95:12-13:
  // Vxint..Vzint, Wint, rhoint, pint, enthpyint), kept from step to step;
  // null without REC_PRIM
  double* vint[MAXL][3];
  double* wint[MAXL];
  double* rhoint[MAXL];
  double* pint[MAXL];
  double* hint[MAXL];
  // where the face solves and the stage right-hand side read the velocity,
  // pressure and density they reconstruct (K_* order): null = the primitive
  // planes (REC_PRIM = 0); the stage primitives with REC_PRIM = 1, set only in
  // the launch's own copy of the plane set
  const double* kin[5];
};

enum { K_VX, K_VY, K_VZ, K_P, K_RHO, K_N };

// the plane a face solve or the stage right-hand side reads for kinematic
// quantity q (K_*); STAGE false: the primitive plane, for kernels that never
// run with REC_PRIM = 1 (the marching kernel), which then skip the test
template <bool STAGE = true>
__device__ inline const double* kin_plane(const Planes& P, int q) {
  const int pq = q == K_VX ? P_VX : q == K_VY ? P_VY : q == K_VZ ? P_VZ : q == K_P ? P_P : P_RHO;
  if constexpr (!STAGE) return P.p[pq];
  return P.kin[q] ? P.kin[q] : P.p[pq];
}

// the recovery writes in place
inline void recover_in_place(Planes& P) {
  P.pn[0] = P.p[P_VX];
  P.pn[1] = P.p[P_VY];
  P.pn[2] = P.p[P_VZ];
  P.pn[3] = P.p[P_RHO];
  P.pn[4] = P.p[P_P];
}

struct Scheme {
  double dx, dy, dt, sigma, kpsi, kphi, gamma;
  double inv_dx, inv_dy;
  int div_exact;  // dx, dy powers of two: x/d == x*(1/d)
  int lmax, flux_solver, order;
  double wb[MAXL], wbt[MAXL], Ab[MAXL][MAXL], Abt[MAXL][MAXL];
  int hllc_seed_ref;
  int crmath;
  int slop;   // slope limiter of the second-order reconstructions (parameters.f95:17)
  // the reference's module variable eps1 that clamps a superluminal previous
  // velocity in the recovery (12_varprimitive.f95:236-241): 1e-6 once
  // hll_flow_mpx has run in the process (16_hll_flow_mpx.f95:1375), else 0
  double eps1;
  // electric field integration (parameters.f95:11): 0 IMEX, 1 MIRK1, 2 MIRK2,
  // with the MIRK coefficients cm_1, cm_2, cm_4 (:233-235)
  int mirk;
  int rec_prim;  // stage primitive recovery (parameters.f95:23)
  double cm1, cm2, cm4;
};

// status words in device memory
struct DevStatus {
  // the first failure, kept until the host clears it
  unsigned long long seq;  // launch number of the launch that failed first (0: none)
  unsigned long long key;  // ~(first failing cell of that launch in the reference's loop order, see fail_key)
  // counters, cleared per step by the host (main_cueva_code.f95:189-193)
  unsigned long long cnt[3];   // contador_r, contador_p, contador_v
  unsigned long long fallback;
  unsigned int nrec[3];        // fix-up records of each kind (may exceed the capacity)
  unsigned int pad;
};
constexpr size_t STATUS_COUNTERS_OFFSET = 2 * sizeof(unsigned long long);

// one record of the reference's fix-up files 301-303 (12_varprimitive.f95:544,601,663)
struct FixRec {
  int h, i, j;
  int l;  // the recovery that wrote it: stage l (REC_PRIM = 1), or 0 for the one after varfin
  double a, b;  // rho* and rho^n / p* and p^n / v*^2 and v_drift^2
};

// what a launch that can fail carries: where to report, its launch number
// (launches are numbered in stream order) and the step it belongs to
struct StatusRef {
  DevStatus* st;
  FixRec* rec;  // [3][cap]
  unsigned long long seq;
  int h;
  int cap;
  int l;  // stage of a stage recovery (REC_PRIM = 1), else 0
};

// Cells are ordered as the reference's loops visit them (j outer, i inner),
// after the phase of the step they belong to (0: face solves, 1: primitive
// recovery, which the reference runs after every face of the stage), so the
// smallest key of a launch is the cell at which the reference stops.
__device__ inline unsigned long long fail_key(int code, int i, int j, int phase) {
  return ((unsigned long long)phase << 48) | ((unsigned long long)(j + 8) << 28) |
         ((unsigned long long)(i + 8) << 8) | (unsigned long long)code;
}

// Record a failure.  Only the launch that failed first reports; within it the
// first cell in the reference's order wins.
__device__ inline void report(const StatusRef& r, int code, int i, int j, int phase = 0) {
  const unsigned long long s0 = atomicCAS(&r.st->seq, 0ULL, r.seq);
  if (s0 != 0ULL && s0 != r.seq) return;
  atomicMax(&r.st->key, ~fail_key(code, i, j, phase));
}

// An earlier launch failed: the step stops there and later launches leave
// the state alone.  The answer is the same for every thread of a launch (only
// earlier launches can have set a smaller launch number), so kernels with
// barriers may return on it.
__device__ inline bool halted(const StatusRef& r) {
  const unsigned long long s = __atomic_load_n(&r.st->seq, __ATOMIC_RELAXED);
  return s != 0ULL && s < r.seq;
}

// a fix-up: counter and record (kind 0 rho, 1 p, 2 drift)
__device__ inline void add_fixup(const StatusRef& r, int kind, int i, int j, double a, double b) {
  atomicAdd(&r.st->cnt[kind], 1ULL);
  const unsigned n = atomicAdd(&r.st->nrec[kind], 1u);
  if ((int)n < r.cap) {
    FixRec& e = r.rec[(size_t)kind * r.cap + n];
    e.h = r.h;
    e.i = i;
    e.j = j;
    e.l = r.l;
    e.a = a;
    e.b = b;
  }
}

// XCD-aware tile order for a 1D launch of 8 * ceil(n/8) workgroups over n =
// tiles_x * tiles_y tiles: the hardware hands workgroup w to XCD w % 8, and
// w takes tile (w % 8) * ceil(n/8) + w / 8, so each XCD sweeps one
// contiguous band of tile rows and the stencil rows neighbouring tiles share
// stay in that XCD's L2.  Returns false for the padding workgroups.
__device__ inline bool xcd_tile(int tiles_x, int tiles_y, int& bx, int& by) {
  const int n = tiles_x * tiles_y;
  const int per = (n + 7) / 8;
  const int w = (int)blockIdx.x;
  const int t = (w % 8) * per + w / 8;
  if (t >= n) return false;
  bx = t % tiles_x;
  by = t / tiles_x;
  return true;
}

// U^n of accumulated field a (in A_* order) at offset k
__device__ inline double un_value(const Planes& P, int a, int64_t k) {
  switch (a) {
    case A_PSI: return P.p[P_PSI][k];
    case A_PHI: return P.p[P_PHI][k];
    case A_BX: return P.p[P_BX][k];
    case A_BY: return P.p[P_BY][k];
    case A_BZ: return P.p[P_BZ][k];
    case A_EX: return P.p[P_EX][k];
    case A_EY: return P.p[P_EY][k];
    case A_EZ: return P.p[P_EZ][k];
    case A_Q: return P.p[P_Q][k];
    case A_D: return P.c[C_D][k];
    case A_TAU: return P.c[C_TAU][k];
    case A_SX: return P.c[C_SX][k];
    case A_SY: return P.c[C_SY][k];
    default: return P.c[C_SZ][k];
  }
}


}  // namespace cueva
