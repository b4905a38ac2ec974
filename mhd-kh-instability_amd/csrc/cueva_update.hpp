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
// dxf[a] = (Fx(i+1/2) - Fx(i-1/2))/dx and dyf likewise (formed in the flux
// arithmetic, widened), combined with the FP64 sources of the stage state.
__device__ inline Rhs rhs_of(const Planes& P, const Scheme& S, int64_t k, const double* dxf, const double* dyf) {
#pragma clang fp contract(off)  // FP64: as the reference build, also inside the FP32 face unit
  double vx = P.p[P_VX][k], vy = P.p[P_VY][k], vz = P.p[P_VZ][k];
  double w = 1.0 / sqrt(1.0 - ((sq(vx) + sq(vy)) + sq(vz)));
  double q = P.s[S_Q][k];
  Rhs o;
  o.r[0] = ((((-dxf[FX_EXPSI]) - dyf[FY_EYPSI]) - 0.0) + q) - S.kpsi * P.s[S_PSI][k];
  o.r[1] = (((-dxf[FX_BXPHI]) - dyf[FY_BYPHI]) - 0.0) - S.kphi * P.s[S_PHI][k];
  o.r[2] = ((-dyf[FY_EZBX]) + 0.0) - dxf[FX_PHIBX];
  o.r[3] = ((-0.0) + dxf[FX_EZBY]) - dyf[FY_PHIBY];
  o.r[4] = ((-dxf[FX_EYBZ]) + dyf[FY_EXBZ]) - 0.0;
  o.r[5] = ((dyf[FY_BZEX] - 0.0) - dxf[FX_PSIEX]) - q * vx;
  o.r[6] = ((0.0 - dxf[FX_BZEY]) - dyf[FY_PSIEY]) - q * vy;
  o.r[7] = ((dxf[FX_BYEZ] - dyf[FY_BXEZ]) - 0.0) - q * vz;
  o.r[8] = ((-dxf[FX_J]) - dyf[FY_J]) - 0.0;
  o.r[9] = ((-dxf[FX_FD]) - dyf[FY_FD]) - 0.0;
  o.r[10] = ((-dxf[FX_FTAU]) - dyf[FY_FTAU]) - 0.0;
  o.r[11] = ((-dxf[FX_FSX]) - dyf[FY_FSX]) - 0.0;
  o.r[12] = ((-dxf[FX_FSY]) - dyf[FY_FSY]) - 0.0;
  o.r[13] = ((-dxf[FX_FSZ]) - dyf[FY_FSZ]) - 0.0;
  ohm_source(S.sigma, w, vx, vy, vz, P.s[S_EX][k], P.s[S_EY][k], P.s[S_EZ][k], P.s[S_BX][k], P.s[S_BY][k],
             P.s[S_BZ][k], o.r[R_SEX], o.r[R_SEY], o.r[R_SEZ]);
  return o;
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

}  // namespace cueva
