// Pointwise pieces of the stage right-hand side and of the final update,
// shared by the plane-reading kernels (one launch per reference subroutine)
// and the marching face kernel, which forms the flux differences from
// registers and LDS.  Both evaluate the same expressions in the same order.
#pragma once

#include "cueva_common.hpp"

namespace cueva {

struct Rhs { double r[R_N]; };

template <typename FT> __device__ inline FT ddx(const Scheme& S, FT a) {
  return S.div_exact ? a * (FT)S.inv_dx : a / (FT)S.dx;
}
template <typename FT> __device__ inline FT ddy(const Scheme& S, FT a) {
  return S.div_exact ? a * (FT)S.inv_dy : a / (FT)S.dy;
}

// Right-hand side of one interior point, 17_fluxandsource.f95:442-542:
// dfx[a] = (Fx(i+1/2) - Fx(i-1/2))/dx and dfy likewise, formed in the flux
// arithmetic FT (held in FT until used: a float difference widens exactly),
// combined with the FP64 sources of the stage state.
template <typename FT, bool STAGE_KIN = true>
__device__ inline Rhs rhs_of(const Planes& P, const Scheme& S, int64_t k, const FT* dfx, const FT* dfy) {
#pragma clang fp contract(off)  // FP64: as the reference build, also inside the FP32 face unit
#define DX(a) ((double)dfx[a])
#define DY(a) ((double)dfy[a])
  // v^n, or with REC_PRIM = 1 the stage velocity (17_fluxandsource.f95:460-474);
  // W as 06_conserved_var.f95 / 13_intercon_var.f95:27 form it from that velocity
  double vx = kin_plane<STAGE_KIN>(P, K_VX)[k], vy = kin_plane<STAGE_KIN>(P, K_VY)[k],
         vz = kin_plane<STAGE_KIN>(P, K_VZ)[k];
  double w = 1.0 / sqrt(1.0 - ((sq(vx) + sq(vy)) + sq(vz)));
  double q = P.s[S_Q][k];
  Rhs o;
  o.r[0] = ((((-DX(FX_EXPSI)) - DY(FY_EYPSI)) - 0.0) + q) - S.kpsi * P.s[S_PSI][k];
  o.r[1] = (((-DX(FX_BXPHI)) - DY(FY_BYPHI)) - 0.0) - S.kphi * P.s[S_PHI][k];
  o.r[2] = ((-DY(FY_EZBX)) + 0.0) - DX(FX_PHIBX);
  o.r[3] = ((-0.0) + DX(FX_EZBY)) - DY(FY_PHIBY);
  o.r[4] = ((-DX(FX_EYBZ)) + DY(FY_EXBZ)) - 0.0;
  o.r[5] = ((DY(FY_BZEX) - 0.0) - DX(FX_PSIEX)) - q * vx;
  o.r[6] = ((0.0 - DX(FX_BZEY)) - DY(FY_PSIEY)) - q * vy;
  o.r[7] = ((DX(FX_BYEZ) - DY(FY_BXEZ)) - 0.0) - q * vz;
  o.r[8] = ((-DX(FX_J)) - DY(FY_J)) - 0.0;
  o.r[9] = ((-DX(FX_FD)) - DY(FY_FD)) - 0.0;
  o.r[10] = ((-DX(FX_FTAU)) - DY(FY_FTAU)) - 0.0;
  o.r[11] = ((-DX(FX_FSX)) - DY(FY_FSX)) - 0.0;
  o.r[12] = ((-DX(FX_FSY)) - DY(FY_FSY)) - 0.0;
  o.r[13] = ((-DX(FX_FSZ)) - DY(FY_FSZ)) - 0.0;
#undef DX
#undef DY
  ohm_source(S.sigma, w, vx, vy, vz, P.s[S_EX][k], P.s[S_EY][k], P.s[S_EZ][k], P.s[S_BX][k], P.s[S_BY][k],
             P.s[S_BZ][k], o.r[R_SEX], o.r[R_SEY], o.r[R_SEZ]);
  return o;
}

// The MIRK electric field of one interior point (20_electricfield.f95:13-146):
// w = W^n, v = v^n, en / bn = E^n / B^n, b = B^{n+1}, ei* / bi* = the last
// stage's E and B, vl = its velocity, acc = the three final sums; the new E
// goes to e.  MIRK2 also stores the last stage's W (:71) to wl when given.
__device__ inline void mirk_final_e(const Scheme& S, double w, const double* v, const double* vl, const double* en,
                                    const double* bn, const double* b, double eix, double eiy, double eiz, double bix,
                                    double biy, double biz, const double* acc, double* e, double* wl) {
#pragma clang fp contract(off)
  if (S.mirk == 1) {
    const double sb = (S.dt * S.sigma) * w;
    const double edv = (en[0] * v[0] + en[1] * v[1]) + en[2] * v[2];
    const double rn[3] = {bn[2] * v[1] - bn[1] * v[2], bn[0] * v[2] - bn[2] * v[0], bn[1] * v[0] - bn[0] * v[1]};
    const double r[3] = {b[2] * v[1] - b[1] * v[2], b[0] * v[2] - b[2] * v[0], b[1] * v[0] - b[0] * v[1]};
    for (int d = 0; d < 3; d++) e[d] = mirk1_final_e(sb, S.cm2, S.dt, en[d], v[d], edv, rn[d], r[d], acc[d]);
    return;
  }
  const double v2 = (sq(vl[0]) + sq(vl[1])) + sq(vl[2]);
  const double wi = 1.0 / sqrt(1.0 - v2);
  if (wl) *wl = wi;
  const double sdb = (S.dt * S.sigma) * wi;
  const double edv = (eix * vl[0] + eiy * vl[1]) + eiz * vl[2];
  const double r[3] = {biz * vl[1] - biy * vl[2], bix * vl[2] - biz * vl[0], biy * vl[0] - bix * vl[1]};
  const double ei[3] = {eix, eiy, eiz};
  for (int d = 0; d < 3; d++) e[d] = mirk2_final_e(sdb, S.cm1, S.cm4, S.dt, ei[d], en[d], vl[d], edv, r[d], acc[d]);
}

// sumfin + varfin + electricfield of one interior point (18_sumfin.f95:20-60,
// 19_varfin.f95:35-50, 20_electricfield.f95:157-163) with the last stage's
// right-hand side in registers: U^{n+1} in A_* order.
__device__ inline void final_update_from(const Planes& P, const Scheme& S, int64_t k, const Rhs& last,
                                         double* nu) {
#pragma clang fp contract(off)  // FP64: as the reference build, also inside the FP32 face unit
  const int L = S.lmax - 1;
  double vx = P.p[P_VX][k], vy = P.p[P_VY][k], vz = P.p[P_VZ][k];
  double ex0 = P.p[P_EX][k], ey0 = P.p[P_EY][k], ez0 = P.p[P_EZ][k];
  double bx0 = P.p[P_BX][k], by0 = P.p[P_BY][k], bz0 = P.p[P_BZ][k];
  Cons cs = prim_to_cons(S.gamma, vx, vy, vz, ex0, ey0, ez0, bx0, by0, bz0, P.p[P_RHO][k], P.p[P_P][k]);
  double un[A_N] = {P.p[P_PSI][k], P.p[P_PHI][k], bx0, by0, bz0, ex0, ey0, ez0, P.p[P_Q][k],
                    cs.d, cs.tau, cs.sx, cs.sy, cs.sz};
  double acc_e[3];
#pragma unroll
  for (int a = 0; a < A_N; a++) {
    double acc = 0.0;
    for (int l = 0; l <= L; l++) {
      double f = l == L ? last.r[a] : P.rhs[l][a][k];
      if (a >= A_EX && a <= A_EZ) {
        double s = l == L ? last.r[R_SEX + (a - A_EX)] : P.rhs[l][R_SEX + (a - A_EX)][k];
        acc = acc + (S.wb[l] * f - S.wbt[l] * s);
      } else {
        acc = acc + S.wb[l] * f;
      }
    }
    nu[a] = un[a] + S.dt * acc;
    if (a >= A_EX && a <= A_EZ) acc_e[a - A_EX] = acc;
  }
  if (S.mirk) {
    const double ve[3] = {vx, vy, vz};
    // the last stage's velocity: the stage primitives with REC_PRIM = 1, else
    // the never-written stage array (zero)
    const double vl[3] = {S.rec_prim ? P.vint[L][0][k] : 0.0, S.rec_prim ? P.vint[L][1][k] : 0.0,
                          S.rec_prim ? P.vint[L][2][k] : 0.0};
    mirk_final_e(S, cs.w, ve, vl, un + A_EX, un + A_BX, nu + A_BX, P.s[S_EX][k], P.s[S_EY][k], P.s[S_EZ][k],
                 P.s[S_BX][k], P.s[S_BY][k], P.s[S_BZ][k], acc_e, nu + A_EX, S.rec_prim ? P.wint[L] + k : nullptr);
  }
}

__device__ inline void store_update(Planes& P, int64_t k, const double* nu) {
  P.p[P_PSI][k] = nu[A_PSI]; P.p[P_PHI][k] = nu[A_PHI];
  P.p[P_BX][k] = nu[A_BX]; P.p[P_BY][k] = nu[A_BY]; P.p[P_BZ][k] = nu[A_BZ];
  P.p[P_EX][k] = nu[A_EX]; P.p[P_EY][k] = nu[A_EY]; P.p[P_EZ][k] = nu[A_EZ];
  P.p[P_Q][k] = nu[A_Q];
  P.c[C_D][k] = nu[A_D]; P.c[C_TAU][k] = nu[A_TAU];
  P.c[C_SX][k] = nu[A_SX]; P.c[C_SY][k] = nu[A_SY]; P.c[C_SZ][k] = nu[A_SZ];
}

// The MIRK stage E (11_elecvarint.f95:210-277, DIM = 2): MIRK1 E^n; MIRK2
// the explicit E* in the first stage (ea), later the linearly implicit
// update from E^n, v^n, B^n, W^n and the stage sums acc.
__device__ inline void mirk_stage_e(const Scheme& S, int l, double w, double vx, double vy, double vz, double ex0,
                                    double ey0, double ez0, double bx0, double by0, double bz0, const double* ea,
                                    const double* acc, double& ex, double& ey, double& ez) {
#pragma clang fp contract(off)
  if (S.mirk == 1) {
    ex = ex0; ey = ey0; ez = ez0;
    return;
  }
  if (l == 0) {
    ex = ea[0]; ey = ea[1]; ez = ea[2];
    return;
  }
  const double sb = (S.dt * S.sigma) * w;
  const double edv = (ex0 * vx + ey0 * vy) + ez0 * vz;
  const double rx = bz0 * vy - by0 * vz, ry = bx0 * vz - bz0 * vx, rz = by0 * vx - bx0 * vy;
  ex = mirk2_stage_e(sb, S.cm1, S.dt, ex0, vx, edv, rx, acc[0]);
  ey = mirk2_stage_e(sb, S.cm1, S.dt, ey0, vy, edv, ry, acc[1]);
  ez = mirk2_stage_e(sb, S.cm1, S.dt, ez0, vz, edv, rz, acc[2]);
}

// Stage-l state of one point from its U^n primitives (values) and the
// earlier stages' right-hand sides at ks: sumint (cumulative, as the
// reference never resets the sums within a step), varint, electricvarint and
// flowinit.  out: S_PSI..S_SZ (S_EXA..S_EZA unused), S_EX..S_JZ.
__device__ inline void stage_point(const Planes& P, const Scheme& S, int l, int64_t ks, bool ghost, double vx,
                                   double vy, double vz, double ex0, double ey0, double ez0, double bx0, double by0,
                                   double bz0, double rho, double p, double psi, double phi, double q,
                                   double* out) {
#pragma clang fp contract(off)  // FP64: as the reference build, also inside the FP32 face unit
  Cons cs = prim_to_cons(S.gamma, vx, vy, vz, ex0, ey0, ez0, bx0, by0, bz0, rho, p);
  double un[A_N] = {psi, phi, bx0, by0, bz0, ex0, ey0, ez0, q, cs.d, cs.tau, cs.sx, cs.sy, cs.sz};
  double stv[A_N], acc_e[3];
#pragma unroll
  for (int a = 0; a < A_N; a++) {
    double acc = 0.0;
    for (int lp = 1; lp <= l; lp++)
      for (int m = 0; m < lp; m++) {
        double f = P.rhs[m][a][ks];
        if (a >= A_EX && a <= A_EZ)
          acc = acc + (S.Ab[lp][m] * f - S.Abt[lp][m] * P.rhs[m][R_SEX + (a - A_EX)][ks]);
        else
          acc = acc + S.Ab[lp][m] * f;
      }
    stv[a] = un[a] + S.dt * acc;
    if (a >= A_EX && a <= A_EZ) acc_e[a - A_EX] = acc;
  }
  double ex, ey, ez;
  if (S.mirk == 0) {
    e_solve(S.Abt[l][l], S.sigma, S.dt, cs.w, vx, vy, vz, stv[A_BX], stv[A_BY], stv[A_BZ], stv[A_EX], stv[A_EY],
            stv[A_EZ], ex, ey, ez);
  } else {
    mirk_stage_e(S, l, cs.w, vx, vy, vz, ex0, ey0, ez0, bx0, by0, bz0, stv + A_EX, acc_e, ex, ey, ez);
  }
  double jx, jy, jz;
  ohm_source(S.sigma, cs.w, vx, vy, vz, ex, ey, ez, stv[A_BX], stv[A_BY], stv[A_BZ], jx, jy, jz);
  jx = jx + stv[A_Q] * vx;
  jy = jy + stv[A_Q] * vy;
  jz = jz + stv[A_Q] * vz;
  if (ghost && !j_ghosts_filled(S.flux_solver)) jx = jy = jz = 0.0;  // 24_boundary_flux_conserved.f95:13
  out[S_PSI] = stv[A_PSI]; out[S_PHI] = stv[A_PHI];
  out[S_BX] = stv[A_BX]; out[S_BY] = stv[A_BY]; out[S_BZ] = stv[A_BZ];
  out[S_EX] = ex; out[S_EY] = ey; out[S_EZ] = ez;
  out[S_Q] = stv[A_Q]; out[S_D] = stv[A_D]; out[S_TAU] = stv[A_TAU];
  out[S_SX] = stv[A_SX]; out[S_SY] = stv[A_SY]; out[S_SZ] = stv[A_SZ];
  out[S_JX] = jx; out[S_JY] = jy; out[S_JZ] = jz;
}

__device__ inline void store_stage(double* const* s, int64_t k, const double* v) {
#pragma unroll
  for (int t = 0; t < S_N; t++)
    if (t < S_EXA || t > S_EZA) s[t][k] = v[t];
}

// primitive recovery of one point (Cardano, 12_varprimitive.f95) and the
// store of its primitives and W; false on a failure (reported)
__device__ inline void c2p_count(const StatusRef& sr, int i, int j, const C2POut& o) {
  if (o.fix_rho) add_fixup(sr, 0, i, j, o.rec[0], o.rec[1]);
  if (o.fix_p) add_fixup(sr, 1, i, j, o.rec[2], o.rec[3]);
  if (o.fix_v) add_fixup(sr, 2, i, j, o.rec[4], o.rec[5]);
}

// the recovered primitives of one point and its W (12_varprimitive.f95:793-802)
__device__ inline void c2p_store_out(const Planes& P, int64_t k, const C2POut& o) {
#pragma clang fp contract(off)
  P.pn[0][k] = o.vx; P.pn[1][k] = o.vy; P.pn[2][k] = o.vz;
  P.pn[3][k] = o.rho; P.pn[4][k] = o.p;
  P.c[C_W][k] = 1.0 / sqrt(1.0 - ((sq(o.vx) + sq(o.vy)) + sq(o.vz)));
}

__device__ inline bool c2p_apply(const Planes& P, const Scheme& S, const StatusRef& sr, int i, int j, int64_t k,
                                 const C2PIn& in, C2POut& o) {
  int code = c2p_cardano_full(S.gamma, S.eps1, in, o, S.crmath != 0);
  if (code) { report(sr, code, i, j, 1); return false; }
  c2p_count(sr, i, j, o);
  c2p_store_out(P, k, o);
  return true;
}

__device__ inline C2PIn c2p_load(const Planes& P, int64_t k) {
  C2PIn in;
  in.d = P.c[C_D][k]; in.tau = P.c[C_TAU][k];
  in.sx = P.c[C_SX][k]; in.sy = P.c[C_SY][k]; in.sz = P.c[C_SZ][k];
  in.ex = P.p[P_EX][k]; in.ey = P.p[P_EY][k]; in.ez = P.p[P_EZ][k];
  in.bx = P.p[P_BX][k]; in.by = P.p[P_BY][k]; in.bz = P.p[P_BZ][k];
  in.vx0 = P.p[P_VX][k]; in.vy0 = P.p[P_VY][k]; in.vz0 = P.p[P_VZ][k];
  in.rho_old = P.p[P_RHO][k]; in.p_old = P.p[P_P][k];
  return in;
}

// Primitive recovery of one point from its U^{n+1} still in registers (the
// marching kernel's last stage, which thereby replaces the recovery pass):
// the t^n velocity, rho and p come from the primitive planes, which the
// launch's stencils still read, so the results go to P.pn; with stage0 the
// next step's stage-0 state goes to P.sn (P.s is the launch's input).
__device__ inline void recover_point(const Planes& P, const Scheme& S, const StatusRef& sr, int i, int j, int64_t k,
                                     const double* nu, int stage0) {
  C2PIn in;
  in.d = nu[A_D]; in.tau = nu[A_TAU];
  in.sx = nu[A_SX]; in.sy = nu[A_SY]; in.sz = nu[A_SZ];
  in.ex = nu[A_EX]; in.ey = nu[A_EY]; in.ez = nu[A_EZ];
  in.bx = nu[A_BX]; in.by = nu[A_BY]; in.bz = nu[A_BZ];
  in.vx0 = P.p[P_VX][k]; in.vy0 = P.p[P_VY][k]; in.vz0 = P.p[P_VZ][k];
  in.rho_old = P.p[P_RHO][k]; in.p_old = P.p[P_P][k];
  C2POut o;
  if (!c2p_apply(P, S, sr, i, j, k, in, o)) return;
  if (stage0) {
    double sv[S_N];
    stage_point(P, S, 0, k, false, o.vx, o.vy, o.vz, in.ex, in.ey, in.ez, in.bx, in.by, in.bz, o.rho, o.p,
                nu[A_PSI], nu[A_PHI], nu[A_Q], sv);
    store_stage(P.sn, k, sv);
  }
}

// The next step's stage-0 state of an interior point from its new primitives
// (the stage kernel's arithmetic; the ghost band is then a plain copy).
__device__ inline void stage0_after_c2p(Planes& P, const Scheme& S, int64_t k, const C2PIn& in, const C2POut& o) {
  double out[S_N];
  stage_point(P, S, 0, k, false, o.vx, o.vy, o.vz, in.ex, in.ey, in.ez, in.bx, in.by, in.bz, o.rho, o.p,
              P.p[P_PSI][k], P.p[P_PHI][k], P.p[P_Q][k], out);
  store_stage(P.s, k, out);
}

}  // namespace cueva
