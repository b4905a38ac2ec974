// Face kernel of the CUEVA step: MPx reconstruction of every quantity at the
// x and y faces of a face pair and the HLL / HLLE / HLLC solve, written to
// the face-flux planes.  T is the arithmetic of reconstruction and Riemann
// solve: double (bitwise with the reference) or float (the mixed form, built
// in its own translation unit).
#pragma once

#include "cueva_common.hpp"

#ifndef CUEVA_XCD_FACES
#define CUEVA_XCD_FACES 1
#endif
#ifndef CUEVA_FACES_WAVES_PER_EU
#define FACES_WAVES
#else
#define FACES_WAVES __attribute__((amdgpu_waves_per_eu(CUEVA_FACES_WAVES_PER_EU, CUEVA_FACES_WAVES_PER_EU)))
#endif

namespace cueva {

// launch grid of k_faces over the face pairs [-1,imax] x [-1,jmax]
inline dim3 faces_grid(const Geom& g, dim3 blk) {
  const unsigned tx = (unsigned)((g.imax + 2 + blk.x - 1) / blk.x), ty = (unsigned)((g.jmax + 2 + blk.y - 1) / blk.y);
#if CUEVA_XCD_FACES
  return dim3(8 * ((tx * ty + 7) / 8));
#else
  return dim3(tx, ty);
#endif
}

// ---------------------------------------------------------------------------
// faces: reconstruction + Riemann solve for the face pair (i,j), i.e. the x
// face between (i,j),(i+1,j) and the y face between (i,j),(i,j+1), for
// i in [-1,imax], j in [-1,jmax] (16_*_flow_mpx.f95 face loops)
// ---------------------------------------------------------------------------
template <int ORD, typename T>
__device__ inline void face_rec(const double* __restrict__ a, int64_t c, int64_t st, T& L, T& R) {
  constexpr int K = MPc<ORD>::K;
  T v[2 * K + 2];
#pragma unroll
  for (int t = 0; t < 2 * K + 2; t++) v[t] = (T)a[c + (int64_t)(t - K) * st];
  mp_face<ORD>(v, L, R);
}
template <int ORD, typename T>
__device__ inline void face_rec_pos(const double* __restrict__ a, int64_t c, int64_t st, T& L, T& R) {
  constexpr int K = MPc<ORD>::K;
  T v[2 * K + 2];
#pragma unroll
  for (int t = 0; t < 2 * K + 2; t++) v[t] = (T)a[c + (int64_t)(t - K) * st];
  mp_face_pos<ORD>(v, L, R);
}

// reconstruct every non-kinematic quantity of one face (slots 1-20, 26/27/63)
template <int ORD, int DIR, typename T>
__device__ inline void face_fields(const Planes& P, int64_t c, int64_t st, Side<T>& L, Side<T>& R) {
  face_rec<ORD>(P.s[S_PSI], c, st, L.psi, R.psi);
  face_rec<ORD>(P.s[S_PHI], c, st, L.phi, R.phi);
  face_rec<ORD>(P.s[S_EX], c, st, L.ex, R.ex);
  face_rec<ORD>(P.s[S_EY], c, st, L.ey, R.ey);
  face_rec<ORD>(P.s[S_EZ], c, st, L.ez, R.ez);
  face_rec<ORD>(P.s[S_BX], c, st, L.bx, R.bx);
  face_rec<ORD>(P.s[S_BY], c, st, L.by, R.by);
  face_rec<ORD>(P.s[S_BZ], c, st, L.bz, R.bz);
  face_rec<ORD>(P.s[S_Q], c, st, L.q, R.q);
  face_rec<ORD>(P.s[S_D], c, st, L.d, R.d);
  face_rec<ORD>(P.s[S_TAU], c, st, L.tau, R.tau);
  face_rec<ORD>(P.s[S_SX], c, st, L.sx, R.sx);
  face_rec<ORD>(P.s[S_SY], c, st, L.sy, R.sy);
  face_rec<ORD>(P.s[S_SZ], c, st, L.sz, R.sz);
  face_rec<ORD>(P.s[DIR == 0 ? S_JX : S_JY], c, st, L.jn, R.jn);
}

template <int ORD, int SOLVER, typename T = double>
__global__ void __launch_bounds__(256) FACES_WAVES k_faces(Geom g, Planes P, Scheme S, DevStatus* st) {
#if CUEVA_XCD_FACES
  int bx, by;
  if (!xcd_tile((g.imax + 2 + (int)blockDim.x - 1) / (int)blockDim.x,
                (g.jmax + 2 + (int)blockDim.y - 1) / (int)blockDim.y, bx, by))
    return;
#else
  const int bx = blockIdx.x, by = blockIdx.y;
#endif
  const int i = bx * blockDim.x + threadIdx.x - 1;
  const int j = by * blockDim.y + threadIdx.y - 1;
  if (i > g.imax || j > g.jmax) return;
  const int64_t c = g.at(i, j);
  const int64_t sx = 1, sy = g.pitch;
  Side<T> xl, xr, yl, yr;
  // kinematics of both faces first: the fallbacks couple them
  face_rec<ORD>(P.p[P_VX], c, sx, xl.vx, xr.vx);
  face_rec<ORD>(P.p[P_VY], c, sx, xl.vy, xr.vy);
  face_rec<ORD>(P.p[P_VZ], c, sx, xl.vz, xr.vz);
  face_rec<ORD>(P.p[P_VX], c, sy, yl.vx, yr.vx);
  face_rec<ORD>(P.p[P_VY], c, sy, yl.vy, yr.vy);
  face_rec<ORD>(P.p[P_VZ], c, sy, yl.vz, yr.vz);
  face_rec_pos<ORD>(P.p[P_P], c, sx, xl.p, xr.p);
  face_rec_pos<ORD>(P.p[P_RHO], c, sx, xl.rho, xr.rho);
  face_rec_pos<ORD>(P.p[P_P], c, sy, yl.p, yr.p);
  face_rec_pos<ORD>(P.p[P_RHO], c, sy, yl.rho, yr.rho);
  // superluminal fallback, 16_hll_flow_mpx.f95:1392-1461 (same in hllc;
  // hlle only prints, 16_hlle_flow_mpx.f95:2270-2300)
  if (SOLVER != 6) {
    T v2xl = (sq(xl.vx) + sq(xl.vy)) + sq(xl.vz);
    T v2xr = (sq(xr.vx) + sq(xr.vy)) + sq(xr.vz);
    T v2yl = (sq(yl.vx) + sq(yl.vy)) + sq(yl.vz);
    T v2yr = (sq(yr.vx) + sq(yr.vy)) + sq(yr.vz);
    if (v2xl >= T(1) || v2xr >= T(1) || v2yl >= T(1) || v2yr >= T(1)) {
      const int64_t cm = g.at(i - 1, j), cj = g.at(i, j - 1);
      xl.vx = (T)P.p[P_VX][c]; xl.vy = (T)P.p[P_VY][c]; xl.vz = (T)P.p[P_VZ][c];
      xr.vx = (T)P.p[P_VX][cm]; xr.vy = (T)P.p[P_VY][cm]; xr.vz = (T)P.p[P_VZ][cm];
      yl.vx = (T)P.p[P_VX][c]; yl.vy = (T)P.p[P_VY][c]; yl.vz = (T)P.p[P_VZ][c];
      yr.vx = (T)P.p[P_VX][cj]; yr.vy = (T)P.p[P_VY][cj]; yr.vz = (T)P.p[P_VZ][cj];
    }
  }
  if (SOLVER == 7) {  // density fix-up, 16_hllc_flow_mpx.f95:2663-2693
    T rc = (T)(0.1 * P.p[P_RHO][c]);
    if (yl.rho < 0.0) yl.rho = rc;
    else if (yr.rho < 0.0) yr.rho = rc;
    else if (xl.rho < 0.0) xl.rho = rc;
    else if (xr.rho < 0.0) xr.rho = rc;
  }
  constexpr bool HW = SOLVER == 6;
  // HLLE: first Newton point of the x contact solve in the reference build
  // is the y face's (E x B)_y; that of the y solve is the x solve's last iterate
  T seed_x = T(0), seed_y = T(0);
  if (SOLVER == 6 && S.hllc_seed_ref) {
    face_rec<ORD>(P.s[S_EX], c, sy, yl.ex, yr.ex);
    face_rec<ORD>(P.s[S_EZ], c, sy, yl.ez, yr.ez);
    face_rec<ORD>(P.s[S_BX], c, sy, yl.bx, yr.bx);
    face_rec<ORD>(P.s[S_BZ], c, sy, yl.bz, yr.bz);
    seed_x = hlle_y_ery(yl, yr);
  }
  // x face
  {
    face_fields<ORD, 0>(P, c, sx, xl, xr);
    side_physics<0, HW>((T)S.gamma, xl);
    side_physics<0, HW>((T)S.gamma, xr);
    T f[FX_N];
    if (SOLVER == 7) {
      int code = hllc_x(xl, xr, f, S.hllc_seed_ref != 0);
      if (code) report(st, code, i, j);
    } else if (SOLVER == 6) {
      int code = hlle_face<0>((T)S.gamma, xl, xr, f, S.hllc_seed_ref != 0, seed_x, seed_y);
      if (code) report(st, code, i, j);
    } else {
      hll_x(xl, xr, f);
    }
    if constexpr (sizeof(T) == 8) {
#pragma unroll
      for (int t = 0; t < FX_N; t++) P.fx[t][c] = f[t];
    } else {
#pragma unroll
      for (int t = 0; t < FX_N; t++) P.fxf[t][c] = f[t];
    }
  }
  // y face
  {
    face_fields<ORD, 1>(P, c, sy, yl, yr);
    side_physics<1, HW>((T)S.gamma, yl);
    side_physics<1, HW>((T)S.gamma, yr);
    T f[FY_N];
    if (SOLVER == 7) {
      int code = hllc_y(yl, yr, f, S.hllc_seed_ref != 0);
      if (code) report(st, code, i, j);
    } else if (SOLVER == 6) {
      T unused;
      int code = hlle_face<1>((T)S.gamma, yl, yr, f, S.hllc_seed_ref != 0, seed_y, unused);
      if (code) report(st, code, i, j);
    } else {
      hll_y(yl, yr, f);
    }
    if constexpr (sizeof(T) == 8) {
#pragma unroll
      for (int t = 0; t < FY_N; t++) P.fy[t][c] = f[t];
    } else {
#pragma unroll
      for (int t = 0; t < FY_N; t++) P.fyf[t][c] = f[t];
    }
  }
}

}  // namespace cueva
