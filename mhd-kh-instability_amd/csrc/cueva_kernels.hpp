// Kernels of the CUEVA step for gfx950.
//
// Two ways through one step share these kernels:
//  - one launch per reference subroutine (the drop-in entry points of
//    include/cueva.h), with the reference's accumulators materialised;
//  - cueva_step, which computes the stage state directly at every point the
//    stencils read (ghosts through the boundary index map), forms the IMEX
//    sums on the fly and fuses the last stage with varfin + primitive
//    recovery.  Both apply the same operations in the same order.
#pragma once

#include "cueva_common.hpp"
#include "cueva_faces.hpp"
#include "cueva_update.hpp"
#include "cueva_mlp.hpp"

namespace cueva {
// ---------------------------------------------------------------------------
// ghost fill: every ghost point copies the interior point of the index map
// ---------------------------------------------------------------------------
struct PlaneList {
  double* a[32];
  int n;
  int nzero;  // the last nzero planes get 0 at every ghost point instead of a copy
};

__global__ void k_fill_ghosts(Geom g, PlaneList pl, StatusRef sr) {
  if (halted(sr)) return;
  const int im = g.imax, jm = g.jmax;
  const int64_t nb = 6LL * (im + 13);  // one horizontal band
  const int64_t nv = 6LL * (jm + 1);   // one vertical band
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j;
  if (t < nb) { j = -6 + (int)(t / (im + 13)); i = -6 + (int)(t % (im + 13)); }
  else if (t < 2 * nb) { t -= nb; j = jm + 1 + (int)(t / (im + 13)); i = -6 + (int)(t % (im + 13)); }
  else if (t < 2 * nb + nv) { t -= 2 * nb; i = -6 + (int)(t / (jm + 1)); j = (int)(t % (jm + 1)); }
  else if (t < 2 * nb + 2 * nv) { t -= 2 * nb + nv; i = im + 1 + (int)(t / (jm + 1)); j = (int)(t % (jm + 1)); }
  else return;
  int64_t dst = g.at(i, j);
  int64_t src = g.at(src_index(i, im, g.bcx), src_index(j, jm, g.bcy));
  for (int k = 0; k < pl.n; k++) pl.a[k][dst] = k < pl.n - pl.nzero ? pl.a[k][src] : 0.0;
}

inline int64_t ghost_count(const Geom& g) { return 12LL * (g.imax + 13) + 12LL * (g.jmax + 1); }

// ---------------------------------------------------------------------------
// pointwise kernels over the interior [0,imax] x [0,jmax]
// ---------------------------------------------------------------------------
#define INTERIOR_IJ                                         \
  const int i = blockIdx.x * blockDim.x + threadIdx.x;      \
  const int j = blockIdx.y * blockDim.y + threadIdx.y;      \
  if (i > g.imax || j > g.jmax) return;                     \
  const int64_t k = g.at(i, j);

// startconvar, 06_conserved_var.f95:77-140
__global__ void k_startconvar(Geom g, Planes P, Scheme S) {
  INTERIOR_IJ
  Cons c = prim_to_cons(S.gamma, P.p[P_VX][k], P.p[P_VY][k], P.p[P_VZ][k], P.p[P_EX][k], P.p[P_EY][k],
                        P.p[P_EZ][k], P.p[P_BX][k], P.p[P_BY][k], P.p[P_BZ][k], P.p[P_RHO][k], P.p[P_P][k]);
  P.c[C_D][k] = c.d; P.c[C_TAU][k] = c.tau;
  P.c[C_SX][k] = c.sx; P.c[C_SY][k] = c.sy; P.c[C_SZ][k] = c.sz; P.c[C_W][k] = c.w;
}

// cleanvarint, 07_clean_var.f95:18-60
__global__ void k_cleanvarint(Geom g, Planes P) {
  INTERIOR_IJ
  for (int a = 0; a < A_N; a++) { P.ints[a][k] = 0.0; P.fins[a][k] = 0.0; }
}

// sumint, 09_sumint.f95:18-61 (stage l, substep m; 0-based here)
__global__ void k_sumint(Geom g, Planes P, Scheme S, int l, int m) {
  INTERIOR_IJ
  const double ab = S.Ab[l][m], abt = S.Abt[l][m];
  for (int a = 0; a < A_N; a++) {
    double f = P.rhs[m][a][k];
    if (a >= A_EX && a <= A_EZ)
      P.ints[a][k] = P.ints[a][k] + (ab * f - abt * P.rhs[m][R_SEX + (a - A_EX)][k]);
    else
      P.ints[a][k] = P.ints[a][k] + ab * f;
  }
}

// varint, 10_varint.f95:18-63
__global__ void k_varint(Geom g, Planes P, Scheme S) {
  INTERIOR_IJ
  static constexpr int dst[A_N] = {S_PSI, S_PHI, S_BX, S_BY, S_BZ, S_EXA, S_EYA, S_EZA,
                                   S_Q, S_D, S_TAU, S_SX, S_SY, S_SZ};
  for (int a = 0; a < A_N; a++) P.s[dst[a]][k] = un_value(P, a, k) + S.dt * P.ints[a][k];
}

// electricvarint, 11_elecvarint.f95:210-328 (DIM = 2): the IMEX implicit
// solve (:279-328), or the MIRK stage E (:210-277) from E^n, v^n, B^n, W^n
// and the stage sums of the E equation
__global__ void k_electricvarint(Geom g, Planes P, Scheme S, int l) {
  INTERIOR_IJ
  double vx = P.p[P_VX][k], vy = P.p[P_VY][k], vz = P.p[P_VZ][k];
  double w = 1.0 / sqrt(1.0 - ((sq(vx) + sq(vy)) + sq(vz)));
  if (S.rec_prim && !S.mirk) {
    // REC_PRIM = 1 (:331-378): this stage's primitives as they stand before
    // its own recovery -- from the previous step, zero in the first
    vx = P.vint[l][0][k]; vy = P.vint[l][1][k]; vz = P.vint[l][2][k];
    w = P.wint[l][k];
  }
  double ex, ey, ez;
  if (S.mirk) {
    const double ea[3] = {P.s[S_EXA][k], P.s[S_EYA][k], P.s[S_EZA][k]};
    const double acc[3] = {P.ints[A_EX][k], P.ints[A_EY][k], P.ints[A_EZ][k]};
    mirk_stage_e(S, l, w, vx, vy, vz, P.p[P_EX][k], P.p[P_EY][k], P.p[P_EZ][k], P.p[P_BX][k], P.p[P_BY][k],
                 P.p[P_BZ][k], ea, acc, ex, ey, ez);
  } else {
    e_solve(S.Abt[l][l], S.sigma, S.dt, w, vx, vy, vz, P.s[S_BX][k], P.s[S_BY][k], P.s[S_BZ][k],
            P.s[S_EXA][k], P.s[S_EYA][k], P.s[S_EZA][k], ex, ey, ez);
  }
  P.s[S_EX][k] = ex; P.s[S_EY][k] = ey; P.s[S_EZ][k] = ez;
}

// flowinit (current density), 14_flowinit.f95:258-301; with REC_PRIM = 1
// from the stage primitives (:376-420)
__global__ void k_flowinit(Geom g, Planes P, Scheme S, int l) {
  INTERIOR_IJ
  double vx = P.p[P_VX][k], vy = P.p[P_VY][k], vz = P.p[P_VZ][k];
  double w = 1.0 / sqrt(1.0 - ((sq(vx) + sq(vy)) + sq(vz)));
  if (S.rec_prim) {
    vx = P.vint[l][0][k]; vy = P.vint[l][1][k]; vz = P.vint[l][2][k];
    w = P.wint[l][k];
  }
  double sx, sy, sz;
  ohm_source(S.sigma, w, vx, vy, vz, P.s[S_EX][k], P.s[S_EY][k], P.s[S_EZ][k], P.s[S_BX][k], P.s[S_BY][k],
             P.s[S_BZ][k], sx, sy, sz);
  double q = P.s[S_Q][k];
  P.s[S_JX][k] = sx + q * vx;
  P.s[S_JY][k] = sy + q * vy;
  P.s[S_JZ][k] = sz + q * vz;
}

// sumfin, 18_sumfin.f95:20-60
__global__ void k_sumfin(Geom g, Planes P, Scheme S, int l) {
  INTERIOR_IJ
  const double wb = S.wb[l], wbt = S.wbt[l];
  for (int a = 0; a < A_N; a++) {
    double f = P.rhs[l][a][k];
    if (a >= A_EX && a <= A_EZ)
      P.fins[a][k] = P.fins[a][k] + (wb * f - wbt * P.rhs[l][R_SEX + (a - A_EX)][k]);
    else
      P.fins[a][k] = P.fins[a][k] + wb * f;
  }
}

// varfin (19_varfin.f95:20-50) and electricfield (20_electricfield.f95);
// which = 0: the 11 non-E fields, 1: E.  With MIRK, varfin keeps B^n and
// E^n (Bxn..Ezn, 19_varfin.f95:24-30) in the stage-sum planes of B and E,
// which no later part of the step reads, and electricfield applies the MIRK
// update (:13-146) instead of E^n + dt finsum.
__global__ void k_varfin(Geom g, Planes P, Scheme S, int which) {
  INTERIOR_IJ
  if (S.mirk && which == 0)
    for (int d = 0; d < 3; d++) {
      P.ints[A_BX + d][k] = P.p[P_BX + d][k];
      P.ints[A_EX + d][k] = P.p[P_EX + d][k];
    }
  if (S.mirk && which == 1) {
    const int L = S.lmax - 1;
    const double v[3] = {P.p[P_VX][k], P.p[P_VY][k], P.p[P_VZ][k]};
    const double vl[3] = {S.rec_prim ? P.vint[L][0][k] : 0.0, S.rec_prim ? P.vint[L][1][k] : 0.0,
                          S.rec_prim ? P.vint[L][2][k] : 0.0};
    const double en[3] = {P.ints[A_EX][k], P.ints[A_EY][k], P.ints[A_EZ][k]};
    const double bn[3] = {P.ints[A_BX][k], P.ints[A_BY][k], P.ints[A_BZ][k]};
    const double b[3] = {P.p[P_BX][k], P.p[P_BY][k], P.p[P_BZ][k]};
    const double acc[3] = {P.fins[A_EX][k], P.fins[A_EY][k], P.fins[A_EZ][k]};
    double e[3];
    mirk_final_e(S, P.c[C_W][k], v, vl, en, bn, b, P.s[S_EX][k], P.s[S_EY][k], P.s[S_EZ][k], P.s[S_BX][k],
                 P.s[S_BY][k], P.s[S_BZ][k], acc, e, S.rec_prim ? P.wint[L] + k : nullptr);
    P.p[P_EX][k] = e[0]; P.p[P_EY][k] = e[1]; P.p[P_EZ][k] = e[2];
    return;
  }
  for (int a = 0; a < A_N; a++) {
    bool e = (a >= A_EX && a <= A_EZ);
    if (e != (which == 1)) continue;
    double* u;
    switch (a) {
      case A_PSI: u = P.p[P_PSI]; break;
      case A_PHI: u = P.p[P_PHI]; break;
      case A_BX: u = P.p[P_BX]; break;
      case A_BY: u = P.p[P_BY]; break;
      case A_BZ: u = P.p[P_BZ]; break;
      case A_EX: u = P.p[P_EX]; break;
      case A_EY: u = P.p[P_EY]; break;
      case A_EZ: u = P.p[P_EZ]; break;
      case A_Q: u = P.p[P_Q]; break;
      case A_D: u = P.c[C_D]; break;
      case A_TAU: u = P.c[C_TAU]; break;
      case A_SX: u = P.c[C_SX]; break;
      case A_SY: u = P.c[C_SY]; break;
      default: u = P.c[C_SZ]; break;
    }
    u[k] = u[k] + S.dt * P.fins[a][k];
  }
}

// swap1, 08_swap1.f95:13-33: the step's primitives into the stage-1
// primitives (interior), with W and the enthalpy as startconvar forms them
__global__ void k_swap1(Geom g, Planes P, Scheme S) {
#pragma clang fp contract(off)
  INTERIOR_IJ
  const double rho = P.p[P_RHO][k], p = P.p[P_P][k];
  const double eps = p / ((S.gamma - 1.0) * rho);
  P.rhoint[0][k] = rho;
  P.pint[0][k] = p;
  P.vint[0][0][k] = P.p[P_VX][k]; P.vint[0][1][k] = P.p[P_VY][k]; P.vint[0][2][k] = P.p[P_VZ][k];
  P.wint[0][k] = P.c[C_W][k];
  P.hint[0][k] = rho * (1.0 + eps) + p;
}

// varprimitivecardano with l_swap = 0 (12_varprimitive.f95:111-135, 792-806):
// the recovery of stage l's state, started from that stage's primitives as
// they stand (the previous step's, zero in the first); E is kept
__global__ void k_c2p_stage(Geom g, Planes P, Scheme S, int l, StatusRef sr) {
  if (halted(sr)) return;
  INTERIOR_IJ
  C2PIn in;
  in.d = P.s[S_D][k]; in.tau = P.s[S_TAU][k];
  in.sx = P.s[S_SX][k]; in.sy = P.s[S_SY][k]; in.sz = P.s[S_SZ][k];
  in.ex = P.s[S_EX][k]; in.ey = P.s[S_EY][k]; in.ez = P.s[S_EZ][k];
  in.bx = P.s[S_BX][k]; in.by = P.s[S_BY][k]; in.bz = P.s[S_BZ][k];
  in.vx0 = P.vint[l][0][k]; in.vy0 = P.vint[l][1][k]; in.vz0 = P.vint[l][2][k];
  in.rho_old = P.rhoint[l][k]; in.p_old = P.pint[l][k];
  C2POut o;
  const int code = c2p_cardano_full(S.gamma, S.eps1, in, o, S.crmath != 0);
  if (code) { report(sr, code, i, j, 1); return; }
  c2p_count(sr, i, j, o);
  P.vint[l][0][k] = o.vx; P.vint[l][1][k] = o.vy; P.vint[l][2][k] = o.vz;
  P.rhoint[l][k] = o.rho;
  P.pint[l][k] = o.p;
  P.wint[l][k] = 1.0 / sqrt(1.0 - ((sq(o.vx) + sq(o.vy)) + sq(o.vz)));
}

// interconvar, 13_intercon_var.f95:13-72: stage l's W, enthalpy and
// conserved D tau S from its recovered primitives, E and B (interior)
__global__ void k_interconvar(Geom g, Planes P, Scheme S, int l) {
#pragma clang fp contract(off)
  INTERIOR_IJ
  const double vx = P.vint[l][0][k], vy = P.vint[l][1][k], vz = P.vint[l][2][k];
  const double rho = P.rhoint[l][k], p = P.pint[l][k];
  Cons cs = prim_to_cons(S.gamma, vx, vy, vz, P.s[S_EX][k], P.s[S_EY][k], P.s[S_EZ][k], P.s[S_BX][k],
                         P.s[S_BY][k], P.s[S_BZ][k], rho, p);
  const double eps = p / ((S.gamma - 1.0) * rho);
  P.wint[l][k] = cs.w;
  P.hint[l][k] = rho * (1.0 + eps) + p;
  P.s[S_D][k] = cs.d; P.s[S_TAU][k] = cs.tau;
  P.s[S_SX][k] = cs.sx; P.s[S_SY][k] = cs.sy; P.s[S_SZ][k] = cs.sz;
}

// primitive recovery over the interior; reads D tau S and E B of the new step
__device__ inline void c2p_store(const Geom& g, Planes& P, const Scheme& S, const StatusRef& sr, int i, int j,
                                 int64_t k) {
  C2POut o;
  c2p_apply(P, S, sr, i, j, k, c2p_load(P, k), o);
}

// STAGE0: also the next step's stage-0 state (fused step).  With REC_PRIM = 1
// the recovery starts from the last stage's primitives (l_swap = 2,
// 12_varprimitive.f95:59-83) instead of the step's.
template <bool STAGE0 = false>
__global__ void k_c2p(Geom g, Planes P, Scheme S, StatusRef sr) {
  if (halted(sr)) return;
  INTERIOR_IJ
  C2PIn in = c2p_load(P, k);
  if (S.rec_prim) {
    const int L = S.lmax - 1;
    in.vx0 = P.vint[L][0][k]; in.vy0 = P.vint[L][1][k]; in.vz0 = P.vint[L][2][k];
    in.rho_old = P.rhoint[L][k]; in.p_old = P.pint[L][k];
  }
  C2POut o;
  if (c2p_apply(P, S, sr, i, j, k, in, o) && STAGE0) stage0_after_c2p(P, S, k, in, o);
}

// ---------------------------------------------------------------------------
// flux differences and stiff sources from the face-flux planes,
// 17_fluxandsource.f95:442-542 (arithmetic in rhs_of)
// ---------------------------------------------------------------------------

// face-flux planes of the flux arithmetic FT (double, or float in the mixed form)
template <typename FT> __device__ inline const FT* flux_x(const Planes& P, int a);
template <typename FT> __device__ inline const FT* flux_y(const Planes& P, int a);
template <> __device__ inline const double* flux_x<double>(const Planes& P, int a) { return P.fx[a]; }
template <> __device__ inline const double* flux_y<double>(const Planes& P, int a) { return P.fy[a]; }
template <> __device__ inline const float* flux_x<float>(const Planes& P, int a) { return P.fxf[a]; }
template <> __device__ inline const float* flux_y<float>(const Planes& P, int a) { return P.fyf[a]; }

template <typename FT>
__device__ inline Rhs flux_source(const Geom& g, const Planes& P, const Scheme& S, int i, int j, int64_t k) {
  const int64_t kw = g.at(i - 1, j), ks = g.at(i, j - 1);
  FT dfx[FX_N], dfy[FY_N];
#pragma unroll
  for (int a = 0; a < FX_N; a++) dfx[a] = ddx<FT>(S, flux_x<FT>(P, a)[k] - flux_x<FT>(P, a)[kw]);
#pragma unroll
  for (int a = 0; a < FY_N; a++) dfy[a] = ddy<FT>(S, flux_y<FT>(P, a)[k] - flux_y<FT>(P, a)[ks]);
  return rhs_of<FT>(P, S, k, dfx, dfy);
}

// recons_var_primitive, 05_reconstruct.f95:525-740 (DIM = 2, limit = 2,
// whose limiter is the MC slope): each interior primitive u becomes
// 0.5 (R + L) with L = u + 0.5 lim and R = u - 0.5 lim from the diagonal
// differences u(i+1,j+1) - u and u - u(i-1,j-1).  The 14 results go to the
// x-face planes (dead between steps); k_recons_copy writes them back, so
// every read sees the old values.
__global__ void __launch_bounds__(256) k_recons_prim(Geom g, Planes P, StatusRef sr, int limit) {
  if (halted(sr)) return;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > g.imax || j > g.jmax) return;
  const int64_t c = g.at(i, j), dd = g.pitch + 1;
#pragma unroll 2
  for (int k = 0; k < P_N; k++) {
    const double* a = P.p[k];
    const double u = a[c];
    const double lim = mcl_slope(a[c + dd] - u, u - a[c - dd], limit);
    const double r = u - 0.5 * lim, l = u + 0.5 * lim;
    P.fx[k][c] = 0.5 * (r + l);
  }
}
__global__ void __launch_bounds__(256) k_recons_copy(Geom g, Planes P, StatusRef sr) {
  if (halted(sr)) return;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > g.imax || j > g.jmax) return;
  const int64_t c = g.at(i, j);
#pragma unroll
  for (int k = 0; k < P_N; k++) P.p[k][c] = P.fx[k][c];
}
static_assert(P_N <= FX_N, "the reconstruction stages its results in the x-face planes");

template <typename FT>
__global__ void k_fluxandsource(Geom g, Planes P, Scheme S, int l, StatusRef sr) {
  if (halted(sr)) return;
  INTERIOR_IJ
  Rhs o = flux_source<FT>(g, P, S, i, j, k);
#pragma unroll
  for (int t = 0; t < R_N; t++) P.rhs[l][t][k] = o.r[t];
}

// ---------------------------------------------------------------------------
// fused step pieces
// ---------------------------------------------------------------------------

// Stage state at every point the stencils read: the point's ghost map gives
// the interior point whose values the reference's ghost fills copy.
__global__ void k_stage_fused(Geom g, Planes P, Scheme S, int l, int halo, StatusRef sr) {
  if (halted(sr)) return;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - halo;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - halo;
  if (i > g.imax + halo || j > g.jmax + halo) return;
  const int si = src_index(i, g.imax, g.bcx), sj = src_index(j, g.jmax, g.bcy);
  const int64_t k = g.at(i, j), ks = g.at(si, sj);
  const bool ghost = (si != i) || (sj != j);
  double out[S_N];
  stage_point(P, S, l, ks, ghost, P.p[P_VX][ks], P.p[P_VY][ks], P.p[P_VZ][ks], P.p[P_EX][ks], P.p[P_EY][ks],
              P.p[P_EZ][ks], P.p[P_BX][ks], P.p[P_BY][ks], P.p[P_BZ][ks], P.p[P_RHO][ks], P.p[P_P][ks],
              P.p[P_PSI][ks], P.p[P_PHI][ks], P.p[P_Q][ks], out);
  store_stage(P.s, k, out);
}

// Last stage right-hand side, sumfin, varfin and electricfield at one
// interior point: U^{n+1} in A_* order, from registers.
template <typename FT>
__device__ inline void final_update(const Geom& g, const Planes& P, const Scheme& S, int i, int j, int64_t k,
                                    double* nu) {
  final_update_from(P, S, k, flux_source<FT>(g, P, S, i, j, k), nu);
}

// Last stage + sumfin + varfin + electricfield + varprimitivecardano
// (C2P false: without the primitive recovery, which then runs on its own).
// Bounded to its 256-thread blocks: under the default 1024-thread bound the
// FP64 form was held to 128 VGPRs and spilled 112 B per lane (1.70 ms per
// step at 2048x4096 against 1.66 ms at 150 VGPRs and 3 waves per SIMD).
template <typename FT, bool C2P = true>
__global__ void __launch_bounds__(256) k_final_fused(Geom g, Planes P, Scheme S, StatusRef sr) {
  if (halted(sr)) return;
  INTERIOR_IJ
  double nu[A_N];
  final_update<FT>(g, P, S, i, j, k, nu);
  store_update(P, k, nu);
  if (C2P) c2p_store(g, P, S, sr, i, j, k);
}

// The same with the learned primitive recovery.  Every lane of a wavefront
// (one row of 64 points, block 64 x 4) stays active for the network: points
// past the grid edge compute on the edge point and store nothing.  The new
// state is stored before the network runs and read back after it, so that
// it does not occupy registers across the network.
template <typename FT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) k_final_learned(Geom g, Planes P, Scheme S, const float* __restrict__ wp,
                                                       StatusRef sr) {
  if (halted(sr)) return;
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * blockDim.y + threadIdx.y;
  const bool valid = i0 <= g.imax && j0 <= g.jmax;
  const int i = i0 <= g.imax ? i0 : g.imax, j = j0 <= g.jmax ? j0 : g.jmax;
  const int64_t k = g.at(i, j);
  float f0, f1, f2;
  {
    double nu[A_N];
    final_update<FT>(g, P, S, i, j, k, nu);
    C2PIn in;
    in.d = nu[A_D]; in.tau = nu[A_TAU]; in.sx = nu[A_SX]; in.sy = nu[A_SY]; in.sz = nu[A_SZ];
    in.ex = nu[A_EX]; in.ey = nu[A_EY]; in.ez = nu[A_EZ];
    in.bx = nu[A_BX]; in.by = nu[A_BY]; in.bz = nu[A_BZ];
    in.vx0 = in.vy0 = in.vz0 = 0.0;  // not used by the features
    in.rho_old = in.p_old = 0.0;
    c2p_features(S.gamma, S.eps1, in, wp, f0, f1, f2);
    if (valid) store_update(P, k, nu);
  }
  const int lane = threadIdx.x & 63;
  float y0 = 0.0f, y1 = 0.0f;
  {
    const mlp::Operands o = mlp::load_operands(wp, lane);
    mlp::forward_wave(o, lane, f0, f1, f2, y0, y1);
  }
  if (!valid) return;
  C2PIn in;
  in.d = P.c[C_D][k]; in.tau = P.c[C_TAU][k];
  in.sx = P.c[C_SX][k]; in.sy = P.c[C_SY][k]; in.sz = P.c[C_SZ][k];
  in.ex = P.p[P_EX][k]; in.ey = P.p[P_EY][k]; in.ez = P.p[P_EZ][k];
  in.bx = P.p[P_BX][k]; in.by = P.p[P_BY][k]; in.bz = P.p[P_BZ][k];
  in.vx0 = P.p[P_VX][k]; in.vy0 = P.p[P_VY][k]; in.vz0 = P.p[P_VZ][k];
  in.rho_old = P.p[P_RHO][k]; in.p_old = P.p[P_P][k];
  C2POut out;
  bool fallback = false;
  const int code = c2p_learned_finish(S.gamma, S.eps1, in, wp, y0, y1, S.crmath != 0, out, fallback);
  if (fallback) atomicAdd(&sr.st->fallback, 1ULL);
  if (code) { report(sr, code, i, j, 1); return; }
  c2p_count(sr, i, j, out);
  c2p_store_out(P, k, out);
}

}  // namespace cueva
