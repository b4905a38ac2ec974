// Face kernel of the CUEVA step: MPx reconstruction of every quantity at the
// x and y faces of a face pair and the HLL / HLLE / HLLC solve, written to
// the face-flux planes.  T is the arithmetic of reconstruction and Riemann
// solve: double (bitwise with the reference) or float (the mixed form, built
// in its own translation unit).
#pragma once

#include <stddef.h>

#include <type_traits>

#include "cueva_common.hpp"
#include "cueva_update.hpp"

#ifndef CUEVA_XCD_FACES
#define CUEVA_XCD_FACES 1
#endif
// minimum waves per SIMD the marching kernel is compiled for (0: the
// default of march_min_waves)
#ifndef CUEVA_MARCH_F32_WAVES
#define CUEVA_MARCH_F32_WAVES 0
#endif
#ifndef CUEVA_MARCH_F64_WAVES
#define CUEVA_MARCH_F64_WAVES 0
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

// ---------------------------------------------------------------------------
// HLL with the monotonised-central slope limiter (flux_solver = 2,
// 16_hll_flow_mcl.f95:554-1402, REC_PRIM = 0): second-order face states of
// the stage fields and of the t^n velocity, pressure and density, and the
// Lax-Friedrichs-form HLL flux with signal speed Sp = 1 (const_sp = 0,
// parameters.f95:274).  The x face uses only x-reconstructed states and the
// y face only y-reconstructed ones.
// ---------------------------------------------------------------------------
// mcl(fa, fb) of funciones.f95:57-95 (kind = slop, parameters.f95:17) and
// limiter(ga, gb) of funciones.f95:100-142 (kind = limit, :18): 1 MinMod,
// 2 monotonised central, 3 Superbee; the three-argument min as flang
// evaluates it, min(min(a, b), c)
template <typename T> __device__ inline T mcl_slope(T fa, T fb, int kind = 2) {
  const T s = T(0.5) * (csign(fa) + csign(fb));
  if (kind == 1) return s * mn(fab(fa), fab(fb));
  if (kind == 3) return s * mn(mn(T(2) * fab(fa), T(2) * fab(fb)), mx(fab(fa), fab(fb)));
  return s * mn(mn(T(2) * fab(fa), T(2) * fab(fb)), T(0.5) * fab(fa + fb));
}

// left and right face values between the cells at c and c+st
// (u_L = u_i + FAC mcl(.), u_R = u_{i+1} - FAC mcl(.), FAC = 0.5); the
// half-slopes are returned too, for the "_m" states
template <typename T>
__device__ inline void mcl_rec(int slop, const double* __restrict__ a, int64_t c, int64_t st, T& l, T& r, T* u01 = nullptr,
                               T* hs = nullptr) {
  const T um = (T)a[c - st], u0 = (T)a[c], u1 = (T)a[c + st], u2 = (T)a[c + 2 * st];
  const T sl = T(0.5) * mcl_slope(u1 - u0, u0 - um, slop);
  const T sr = T(0.5) * mcl_slope(u2 - u1, u1 - u0, slop);
  l = u0 + sl;
  r = u1 - sr;
  if (u01) { u01[0] = u0; u01[1] = u1; hs[0] = sl; hs[1] = sr; }
}

template <typename T> struct McSide { T psi, phi, ex, ey, ez, bx, by, bz, q, d, tau, sx, sy, sz, vx, vy, vz, p, rho; };

// J, FD, Ftau and the three FS of one side: E^2+B^2 (16_hll_flow_mcl.f95:821),
// W (:900), E.v (:995), J (:1033), FD (:1050), eps and h (:1124), Ftau
// (:1181), FS (:1198-1241)
template <int DIR, typename T>
__device__ inline void mcl_side_fluxes(T gam, T sigma, const McSide<T>& o, T& jn, T& fd, T& ftau, T& fs1, T& fs2,
                                       T& fs3) {
  const T e2b2 = ((((sq(o.ex) + sq(o.ey)) + sq(o.ez)) + sq(o.bx)) + sq(o.by)) + sq(o.bz);
  const T w = T(1) / sqrt(((T(1) - sq(o.vx)) - sq(o.vy)) - sq(o.vz));
  const T edv = (o.ex * o.vx + o.ey * o.vy) + o.ez * o.vz;
  const T eps = o.p / ((gam - T(1)) * o.rho);
  const T h = o.rho * (T(1) + eps) + o.p;
  const T hw2 = h * sq(w);
  if constexpr (DIR == 0) {
    jn = (sigma * w) * ((o.ex + (o.bz * o.vy - o.by * o.vz)) - edv * o.vx) + o.q * o.vx;
    fd = o.d * o.vx;
    ftau = (o.bz * o.ey - o.by * o.ez) + hw2 * o.vx;
    fs1 = ((((-sq(o.ex)) - sq(o.bx)) + hw2 * sq(o.vx)) + T(0.5) * e2b2) + o.p;
    fs2 = ((-(o.ex * o.ey)) - o.bx * o.by) + (hw2 * o.vx) * o.vy;
    fs3 = ((-(o.ex * o.ez)) - o.bx * o.bz) + (hw2 * o.vx) * o.vz;
  } else {
    jn = (sigma * w) * ((o.ey + (o.bx * o.vz - o.bz * o.vx)) - edv * o.vy) + o.q * o.vy;
    fd = o.d * o.vy;
    ftau = (o.bx * o.ez - o.bz * o.ex) + hw2 * o.vy;
    fs1 = ((-(o.ex * o.ey)) - o.bx * o.by) + (hw2 * o.vx) * o.vy;
    fs2 = ((((-sq(o.ey)) - sq(o.by)) + hw2 * sq(o.vy)) + T(0.5) * e2b2) + o.p;
    fs3 = ((-(o.ey * o.ez)) - o.by * o.bz) + (hw2 * o.vy) * o.vz;
  }
}

// 0.5 (A_L + A_R - Sp (U_R - U_L)) with Sp = 1 (16_hll_flow_mcl.f95:1259-1396)
template <typename T> __device__ inline T mcl_lf(T al, T ar, T ul, T ur) { return T(0.5) * ((al + ar) - (ur - ul)); }

// the fluxes of the face between the cells at c and c+st, DIR 0 = x (FX_*
// slots), 1 = y (FY_* slots)
template <int DIR, typename T>
__device__ inline void mcl_face(const Planes& P, const Scheme& S, int64_t c, int64_t st, T* f) {
  McSide<T> L, R;
  T e0[2], es[2], b0[2], bs[2];  // cell values and half-slopes of the "_m" pair
  mcl_rec(S.slop, P.s[S_PSI], c, st, L.psi, R.psi);
  mcl_rec(S.slop, P.s[S_PHI], c, st, L.phi, R.phi);
  mcl_rec(S.slop, P.s[S_EX], c, st, L.ex, R.ex, DIR == 1 ? e0 : (T*)nullptr, DIR == 1 ? es : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_EY], c, st, L.ey, R.ey);
  mcl_rec(S.slop, P.s[S_EZ], c, st, L.ez, R.ez, DIR == 0 ? e0 : (T*)nullptr, DIR == 0 ? es : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_BX], c, st, L.bx, R.bx);
  mcl_rec(S.slop, P.s[S_BY], c, st, L.by, R.by, DIR == 0 ? b0 : (T*)nullptr, DIR == 0 ? bs : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_BZ], c, st, L.bz, R.bz, DIR == 1 ? b0 : (T*)nullptr, DIR == 1 ? bs : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_Q], c, st, L.q, R.q);
  mcl_rec(S.slop, P.s[S_D], c, st, L.d, R.d);
  mcl_rec(S.slop, P.s[S_TAU], c, st, L.tau, R.tau);
  mcl_rec(S.slop, P.s[S_SX], c, st, L.sx, R.sx);
  mcl_rec(S.slop, P.s[S_SY], c, st, L.sy, R.sy);
  mcl_rec(S.slop, P.s[S_SZ], c, st, L.sz, R.sz);
  mcl_rec(S.slop, kin_plane(P, K_VX), c, st, L.vx, R.vx);
  mcl_rec(S.slop, kin_plane(P, K_VY), c, st, L.vy, R.vy);
  mcl_rec(S.slop, kin_plane(P, K_VZ), c, st, L.vz, R.vz);
  mcl_rec(S.slop, kin_plane(P, K_P), c, st, L.p, R.p);
  mcl_rec(S.slop, kin_plane(P, K_RHO), c, st, L.rho, R.rho);
  // "_m" states: the cell value negated, the slope kept
  // (16_hll_flow_mcl.f95:634-655, 708-737)
  const T em_l = (-e0[0]) + es[0], em_r = (-e0[1]) - es[1];
  const T bm_l = (-b0[0]) + bs[0], bm_r = (-b0[1]) - bs[1];
  T jl, fdl, ftl, f1l, f2l, f3l, jr, fdr, ftr, f1r, f2r, f3r;
  mcl_side_fluxes<DIR>((T)S.gamma, (T)S.sigma, L, jl, fdl, ftl, f1l, f2l, f3l);
  mcl_side_fluxes<DIR>((T)S.gamma, (T)S.sigma, R, jr, fdr, ftr, f1r, f2r, f3r);
  if constexpr (DIR == 0) {
    f[FX_EXPSI] = mcl_lf(L.ex, R.ex, L.psi, R.psi);
    f[FX_BXPHI] = mcl_lf(L.bx, R.bx, L.phi, R.phi);
    f[FX_PHIBX] = mcl_lf(L.phi, R.phi, L.bx, R.bx);
    f[FX_EZBY] = -mcl_lf(em_l, em_r, L.by, R.by);
    f[FX_EYBZ] = mcl_lf(L.ey, R.ey, L.bz, R.bz);
    f[FX_PSIEX] = mcl_lf(L.psi, R.psi, L.ex, R.ex);
    f[FX_BZEY] = mcl_lf(L.bz, R.bz, L.ey, R.ey);
    f[FX_BYEZ] = -mcl_lf(bm_l, bm_r, L.ez, R.ez);
    f[FX_J] = mcl_lf(jl, jr, L.q, R.q);
    f[FX_FD] = mcl_lf(fdl, fdr, L.d, R.d);
    f[FX_FTAU] = mcl_lf(ftl, ftr, L.tau, R.tau);
    f[FX_FSX] = mcl_lf(f1l, f1r, L.sx, R.sx);
    f[FX_FSY] = mcl_lf(f2l, f2r, L.sy, R.sy);
    f[FX_FSZ] = mcl_lf(f3l, f3r, L.sz, R.sz);
  } else {
    f[FY_EYPSI] = mcl_lf(L.ey, R.ey, L.psi, R.psi);
    f[FY_BYPHI] = mcl_lf(L.by, R.by, L.phi, R.phi);
    f[FY_EZBX] = mcl_lf(L.ez, R.ez, L.bx, R.bx);
    f[FY_PHIBY] = mcl_lf(L.phi, R.phi, L.by, R.by);
    f[FY_EXBZ] = -mcl_lf(em_l, em_r, L.bz, R.bz);
    f[FY_BZEX] = -mcl_lf(bm_l, bm_r, L.ex, R.ex);
    f[FY_PSIEY] = mcl_lf(L.psi, R.psi, L.ey, R.ey);
    f[FY_BXEZ] = mcl_lf(L.bx, R.bx, L.ez, R.ez);
    f[FY_J] = mcl_lf(jl, jr, L.q, R.q);
    f[FY_FD] = mcl_lf(fdl, fdr, L.d, R.d);
    f[FY_FTAU] = mcl_lf(ftl, ftr, L.tau, R.tau);
    f[FY_FSX] = mcl_lf(f1l, f1r, L.sx, R.sx);
    f[FY_FSY] = mcl_lf(f2l, f2r, L.sy, R.sy);
    f[FY_FSZ] = mcl_lf(f3l, f3r, L.sz, R.sz);
  }
}

// The face states of HLLC with the MC limiter (flux_solver = 4,
// 16_hllc_flow_mcl.f95:1760-2479, the states and fluxes of hll_flow), as
// the HLLC solve reads them
template <int DIR, typename T>
__device__ inline void mcl_sides(const Planes& P, const Scheme& S, int64_t c, int64_t st, Side<T>& L, Side<T>& R) {
  McSide<T> l, r;
  T e0[2], es[2], b0[2], bs[2];
  mcl_rec(S.slop, P.s[S_PSI], c, st, l.psi, r.psi);
  mcl_rec(S.slop, P.s[S_PHI], c, st, l.phi, r.phi);
  mcl_rec(S.slop, P.s[S_EX], c, st, l.ex, r.ex, DIR == 1 ? e0 : (T*)nullptr, DIR == 1 ? es : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_EY], c, st, l.ey, r.ey);
  mcl_rec(S.slop, P.s[S_EZ], c, st, l.ez, r.ez, DIR == 0 ? e0 : (T*)nullptr, DIR == 0 ? es : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_BX], c, st, l.bx, r.bx);
  mcl_rec(S.slop, P.s[S_BY], c, st, l.by, r.by, DIR == 0 ? b0 : (T*)nullptr, DIR == 0 ? bs : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_BZ], c, st, l.bz, r.bz, DIR == 1 ? b0 : (T*)nullptr, DIR == 1 ? bs : (T*)nullptr);
  mcl_rec(S.slop, P.s[S_Q], c, st, l.q, r.q);
  mcl_rec(S.slop, P.s[S_D], c, st, l.d, r.d);
  mcl_rec(S.slop, P.s[S_TAU], c, st, l.tau, r.tau);
  mcl_rec(S.slop, P.s[S_SX], c, st, l.sx, r.sx);
  mcl_rec(S.slop, P.s[S_SY], c, st, l.sy, r.sy);
  mcl_rec(S.slop, P.s[S_SZ], c, st, l.sz, r.sz);
  mcl_rec(S.slop, kin_plane(P, K_VX), c, st, l.vx, r.vx);
  mcl_rec(S.slop, kin_plane(P, K_VY), c, st, l.vy, r.vy);
  mcl_rec(S.slop, kin_plane(P, K_VZ), c, st, l.vz, r.vz);
  mcl_rec(S.slop, kin_plane(P, K_P), c, st, l.p, r.p);
  mcl_rec(S.slop, kin_plane(P, K_RHO), c, st, l.rho, r.rho);
  Side<T>* sd[2] = {&L, &R};
  const McSide<T>* ms[2] = {&l, &r};
#pragma unroll
  for (int t = 0; t < 2; t++) {
    Side<T>& o = *sd[t];
    const McSide<T>& m = *ms[t];
    o.psi = m.psi; o.phi = m.phi; o.ex = m.ex; o.ey = m.ey; o.ez = m.ez;
    o.bx = m.bx; o.by = m.by; o.bz = m.bz; o.q = m.q; o.d = m.d; o.tau = m.tau;
    o.sx = m.sx; o.sy = m.sy; o.sz = m.sz; o.vx = m.vx; o.vy = m.vy; o.vz = m.vz; o.p = m.p; o.rho = m.rho;
    mcl_side_fluxes<DIR>((T)S.gamma, (T)S.sigma, m, o.jn, o.fd, o.ftau, o.fs1, o.fs2, o.fs3);
  }
  // "_m" states: the cell value negated, the slope kept
  L.em = (-e0[0]) + es[0];
  R.em = (-e0[1]) - es[1];
  L.bm = (-b0[0]) + bs[0];
  R.bm = (-b0[1]) - bs[1];
}

// ---------------------------------------------------------------------------
// Lax-Friedrichs flux splitting (flux_solver = 1, 15_lax_flow.f95:733-1460):
// at every cell the split fluxes F+- = A +- Sp U (Sp = 1) of the stage fields
// and of the cell fluxes flowinit forms (14_flowinit.f95:264-373); the face
// value is 0.5 (L + R), L = F+(i) + FAC mcl(F+(i+1)-F+(i), F+(i)-F+(i-1)),
// R = F-(i+1) - FAC mcl(F+(i+2)-F+(i+1), F+(i+1)-F+(i)) (the slope of R is
// also taken from F+).  flowinit fills only the interior and
// boundary_flux_conserved no ghosts for this solver, so the cell fluxes are
// 0 on ghost cells (cleanvarint, 07_clean_var.f95:341-350).
// ---------------------------------------------------------------------------
// J, FD, Ftau, FS along DIR at cell (i,j) = k, with the t^n W and enthalpy
// of startconvar (06_conserved_var.f95:92-109)
template <int DIR, typename T>
__device__ inline void lax_cell_fluxes(const Geom& g, const Planes& P, const Scheme& S, int i, int j, int64_t k,
                                       T* out) {
  if (i < 0 || i > g.imax || j < 0 || j > g.jmax) {
#pragma unroll
    for (int t = 0; t < 6; t++) out[t] = T(0);
    return;
  }
  // flowinit's cell fluxes: of v^n, p^n, rho^n, or with REC_PRIM = 1 of the
  // stage's primitives (14_flowinit.f95:376-489)
  const T vx = (T)kin_plane(P, K_VX)[k], vy = (T)kin_plane(P, K_VY)[k], vz = (T)kin_plane(P, K_VZ)[k];
  const T p = (T)kin_plane(P, K_P)[k], rho = (T)kin_plane(P, K_RHO)[k];
  const T ex = (T)P.s[S_EX][k], ey = (T)P.s[S_EY][k], ez = (T)P.s[S_EZ][k];
  const T bx = (T)P.s[S_BX][k], by = (T)P.s[S_BY][k], bz = (T)P.s[S_BZ][k];
  const T w = T(1) / sqrt(T(1) - ((sq(vx) + sq(vy)) + sq(vz)));
  const T eps = p / (((T)S.gamma - T(1)) * rho);
  const T h = rho * (T(1) + eps) + p;
  const T e2 = (sq(ex) + sq(ey)) + sq(ez), b2 = (sq(bx) + sq(by)) + sq(bz);
  const T hw2 = h * sq(w);
  out[0] = (T)P.s[DIR == 0 ? S_JX : S_JY][k];
  if constexpr (DIR == 0) {
    out[1] = (T)P.s[S_D][k] * vx;
    out[2] = (bz * ey - by * ez) + hw2 * vx;
    out[3] = ((((-sq(ex)) - sq(bx)) + hw2 * sq(vx)) + T(0.5) * (e2 + b2)) + p;
    out[4] = ((-(ex * ey)) - bx * by) + (hw2 * vx) * vy;
    out[5] = ((-(ex * ez)) - bx * bz) + (hw2 * vx) * vz;
  } else {
    out[1] = (T)P.s[S_D][k] * vy;
    out[2] = (bx * ez - bz * ex) + hw2 * vy;
    out[3] = ((-(ex * ey)) - bx * by) + (hw2 * vx) * vy;
    out[4] = ((((-sq(ey)) - sq(by)) + hw2 * sq(vy)) + T(0.5) * (e2 + b2)) + p;
    out[5] = ((-(ey * ez)) - by * bz) + (hw2 * vy) * vz;
  }
}

// the face value from the split flux of cells i-1 .. i+2 (a: the flux, u:
// the field it is split with, sa = -1 for the fluxes the reference splits
// as -X +- U and negates afterwards)
template <typename T> __device__ inline T lax_face_value(int slop, const T* a, const T* u, T sa = T(1)) {
  const T p0 = sa * a[0] + u[0], p1 = sa * a[1] + u[1], p2 = sa * a[2] + u[2], p3 = sa * a[3] + u[3];
  const T m2 = sa * a[2] - u[2];
  const T l = p1 + T(0.5) * mcl_slope(p2 - p1, p1 - p0, slop);
  const T r = m2 - T(0.5) * mcl_slope(p3 - p2, p2 - p1, slop);
  return T(0.5) * (l + r);
}

template <typename T> __device__ inline void lax_ld(const double* __restrict__ a, const int64_t* k, T* v) {
#pragma unroll
  for (int c = 0; c < 4; c++) v[c] = (T)a[k[c]];
}

// the fluxes of the face between cell (i,j) and its +x (DIR 0) or +y
// (DIR 1) neighbour
template <int DIR, typename T>
__device__ inline void lax_face(const Geom& g, const Planes& P, const Scheme& S, int i, int j, T* f) {
  int64_t k[4];
  T fl[4][6];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int ii = DIR == 0 ? i + c - 1 : i, jj = DIR == 1 ? j + c - 1 : j;
    k[c] = g.at(ii, jj);
    lax_cell_fluxes<DIR>(g, P, S, ii, jj, k[c], fl[c]);
  }
  T psi[4], phi[4], ex[4], ey[4], ez[4], bx[4], by[4], bz[4];
  lax_ld(P.s[S_PSI], k, psi); lax_ld(P.s[S_PHI], k, phi);
  lax_ld(P.s[S_EX], k, ex); lax_ld(P.s[S_EY], k, ey); lax_ld(P.s[S_EZ], k, ez);
  lax_ld(P.s[S_BX], k, bx); lax_ld(P.s[S_BY], k, by); lax_ld(P.s[S_BZ], k, bz);
  if constexpr (DIR == 0) {  // 15_lax_flow.f95:930-1180
    f[FX_EXPSI] = lax_face_value(S.slop, ex, psi);
    f[FX_BXPHI] = lax_face_value(S.slop, bx, phi);
    f[FX_PHIBX] = lax_face_value(S.slop, phi, bx);
    f[FX_EZBY] = -lax_face_value(S.slop, ez, by, T(-1));
    f[FX_EYBZ] = lax_face_value(S.slop, ey, bz);
    f[FX_PSIEX] = lax_face_value(S.slop, psi, ex);
    f[FX_BZEY] = lax_face_value(S.slop, bz, ey);
    f[FX_BYEZ] = -lax_face_value(S.slop, by, ez, T(-1));
  } else {
    f[FY_EYPSI] = lax_face_value(S.slop, ey, psi);
    f[FY_BYPHI] = lax_face_value(S.slop, by, phi);
    f[FY_EZBX] = lax_face_value(S.slop, ez, bx);
    f[FY_PHIBY] = lax_face_value(S.slop, phi, by);
    f[FY_EXBZ] = -lax_face_value(S.slop, ex, bz, T(-1));
    f[FY_BZEX] = -lax_face_value(S.slop, bz, ex, T(-1));
    f[FY_PSIEY] = lax_face_value(S.slop, psi, ey);
    f[FY_BXEZ] = lax_face_value(S.slop, bx, ez);
  }
  static_assert((int)FX_J == (int)FY_J && (int)FX_FSZ == (int)FY_FSZ, "fluid slots at the same place");
  const int us[6] = {S_Q, S_D, S_TAU, S_SX, S_SY, S_SZ};
#pragma unroll
  for (int t = 0; t < 6; t++) {
    T a[4], u[4];
#pragma unroll
    for (int c = 0; c < 4; c++) a[c] = fl[c][t];
    lax_ld(P.s[us[t]], k, u);
    f[FX_J + t] = lax_face_value(S.slop, a, u);
  }
}

// both faces of the face pair (i,j): put_x(Fx(i+1/2,j)) as soon as the x
// face is solved, put_y(Fy(i,j+1/2)) as soon as the y face is; each flux
// array is dead after its put, which keeps the other solve's registers free
// of it.  YFIRST solves the y face first (the two solves are independent
// except under HLLE, whose y solve starts from the x solve's last iterate).
template <int ORD, int SOLVER, typename T, bool YFIRST = false, bool STAGE_KIN = true, typename PutX, typename PutY>
__device__ __forceinline__ void face_pair(const Geom& g, const Planes& P, const Scheme& S, const StatusRef& st, int i,
                                          int j, PutX put_x, PutY put_y) {
  static_assert(!(YFIRST && SOLVER == 6), "the HLLE y solve needs the x solve first");
  const int64_t c = g.at(i, j);
  const int64_t sx = 1, sy = g.pitch;
  if constexpr (SOLVER == 3 || SOLVER == 4) {
    // HLLE / HLLC + MC limiter: x solve first (the y solve starts from its last iterate)
    T f[FX_N > FY_N ? FX_N : FY_N];
    const bool seed = S.hllc_seed_ref != 0;
    T xseed = T(0), vlast = T(0);
    Side<T> sl, sr;
    if (seed) {
      mcl_sides<1>(P, S, c, sy, sl, sr);
      xseed = hlle_y_ery<T, true>(sl, sr);
    }
    mcl_sides<0>(P, S, c, sx, sl, sr);
    int code = SOLVER == 3 ? hlle_mcl_face<0>((T)S.gamma, sl, sr, f, seed, xseed, vlast)
                           : hllc_x<T, true>(sl, sr, f, seed, xseed, &vlast);
    if (code) report(st, code, i, j);
    put_x(f);
    mcl_sides<1>(P, S, c, sy, sl, sr);
    T unused;
    code = SOLVER == 3 ? hlle_mcl_face<1>((T)S.gamma, sl, sr, f, seed, vlast, unused)
                       : hllc_y<T, true>(sl, sr, f, seed, vlast);
    if (code) report(st, code, i, j);
    put_y(f);
    return;
  }
  if constexpr (SOLVER == 1 || SOLVER == 2) {  // second-order solvers: the faces are independent
    T f[FX_N > FY_N ? FX_N : FY_N];
    auto face = [&](auto dir) {
      constexpr int D = decltype(dir)::value;
      if constexpr (SOLVER == 1) lax_face<D>(g, P, S, i, j, f);
      else mcl_face<D>(P, S, c, D == 0 ? sx : sy, f);
      if constexpr (D == 0) put_x(f);
      else put_y(f);
    };
    if constexpr (YFIRST) {
      face(std::integral_constant<int, 1>());
      face(std::integral_constant<int, 0>());
    } else {
      face(std::integral_constant<int, 0>());
      face(std::integral_constant<int, 1>());
    }
    return;
  }
  Side<T> xl, xr, yl, yr;
  // kinematics of both faces first: the fallbacks couple them
  face_rec<ORD>(kin_plane<STAGE_KIN>(P, K_VX), c, sx, xl.vx, xr.vx);
  face_rec<ORD>(kin_plane<STAGE_KIN>(P, K_VY), c, sx, xl.vy, xr.vy);
  face_rec<ORD>(kin_plane<STAGE_KIN>(P, K_VZ), c, sx, xl.vz, xr.vz);
  face_rec<ORD>(kin_plane<STAGE_KIN>(P, K_VX), c, sy, yl.vx, yr.vx);
  face_rec<ORD>(kin_plane<STAGE_KIN>(P, K_VY), c, sy, yl.vy, yr.vy);
  face_rec<ORD>(kin_plane<STAGE_KIN>(P, K_VZ), c, sy, yl.vz, yr.vz);
  face_rec_pos<ORD>(kin_plane<STAGE_KIN>(P, K_P), c, sx, xl.p, xr.p);
  face_rec_pos<ORD>(kin_plane<STAGE_KIN>(P, K_RHO), c, sx, xl.rho, xr.rho);
  face_rec_pos<ORD>(kin_plane<STAGE_KIN>(P, K_P), c, sy, yl.p, yr.p);
  face_rec_pos<ORD>(kin_plane<STAGE_KIN>(P, K_RHO), c, sy, yl.rho, yr.rho);
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
  auto xface = [&]() {
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
    put_x(f);
  };
  auto yface = [&]() {
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
    put_y(f);
  };
  if constexpr (YFIRST) {
    yface();
    xface();
  } else {
    xface();
    yface();
  }
}

// Waves per SIMD k_faces is built for (a register bound).  Unbounded, the
// MCL solvers' face pairs need 177-181 VGPRs in FP32 (2 waves per SIMD) and
// past 256 in FP64 (1 wave): bounded to 3 (FP32, 14-16 VGPRs spilled) the
// face pass measured 18-19 % faster at 2048x4096, and to 2 (FP64) HLL-MCL
// 8 % and HLLC-MCL 5 % faster, HLLE-MCL (176 spilled) 3 % slower, so it
// stays unbounded.  The other solvers fit their waves unbounded.
template <typename T, int SOLVER>
constexpr int faces_min_waves() {
  if constexpr (sizeof(T) == 8) return SOLVER == 2 || SOLVER == 4 ? 2 : 1;
  else return SOLVER >= 2 && SOLVER <= 4 ? 3 : 1;
}

template <int ORD, int SOLVER, typename T = double>
__global__ void __launch_bounds__(256, (faces_min_waves<T, SOLVER>())) FACES_WAVES
k_faces(Geom g, Planes P, Scheme S, StatusRef st) {
  if (halted(st)) return;
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
  face_pair<ORD, SOLVER, T>(
      g, P, S, st, i, j,
      [&](const T* f) {
#pragma unroll
        for (int t = 0; t < FX_N; t++) {
          if constexpr (sizeof(T) == 8) P.fx[t][c] = f[t];
          else P.fxf[t][c] = f[t];
        }
      },
      [&](const T* f) {
#pragma unroll
        for (int t = 0; t < FY_N; t++) {
          if constexpr (sizeof(T) == 8) P.fy[t][c] = f[t];
          else P.fyf[t][c] = f[t];
        }
      });
}

// ---------------------------------------------------------------------------
// Marching face kernel: reconstruction + Riemann solve + flux difference in
// one pass.  A block of MW wavefronts owns a strip of MARCH_COLS cell
// columns and a segment of `seg` cell rows (seg = MW m - 1).  Lane L of every
// wavefront computes the face pair of column i0 - 1 + L; per tile step the
// MW wavefronts take MW consecutive face rows.  Fx(i-1/2) is the left lane's x flux (a
// cross-lane shuffle of the registers the x solve leaves; HLLE, which must
// solve x first, goes through LDS); Fy(j-1/2) is the LDS row of the
// wavefront below, or for the lowest wavefront the previous tile step's top
// row.  Lane 0 only supplies Fx(i0-1/2), the segment's first face row only
// Fy(j0-1/2): each face is solved once per strip and segment, plus one
// column and one row per block.
//
// One barrier per tile step: the y rows alternate between two LDS buffers,
// and wavefront 0 copies the top row it needs next into its own row after
// the barrier; a buffer is rewritten two tile steps later, after every
// reader has passed the barrier in between.  HLLE keeps one buffer and
// takes a second barrier at the end of each tile step instead.
//   FINAL false: the stage right-hand side -> rhs[l] (17_fluxandsource.f95),
//                and the next stage's state (sumint, varint, electricvarint,
//                flowinit) -> sn, interior points;
//   FINAL true:  sumfin + varfin + electricfield of the last stage -> U^{n+1}
//                into rhs[l][0..13]; with C2P also the primitive recovery
//                of every point (-> pn, W) and, with args.stage0, the next
//                step's stage-0 state (-> sn); else the recovery pass follows.
// ---------------------------------------------------------------------------
constexpr int MARCH_COLS = 63;

inline int march_strips(const Geom& g) { return (g.imax + 1 + MARCH_COLS - 1) / MARCH_COLS; }

// the marching kernel's one argument (read through the kernel-argument
// pointer inside the loop, so it must be the first and only argument)
struct MarchArgs {
  Planes P;
  Geom g;
  Scheme S;
  int l, seg;
  StatusRef st;
  int stage0;  // C2P: also the next step's stage-0 state
  int strips, segs;  // the tiles of the launch
  int xcd;           // 1D grid in XCD-aware tile order (xcd_tile), else x = strip, y = segment
};
static_assert(offsetof(MarchArgs, P) == 0, "argument block layout");

// Waves per SIMD the marching kernel is built for, a register bound (the
// compiler spills past it).  Measured per step at 2048x4096:
//   FP64: 2 for every solver.  HLLC fits 237 VGPRs; the others need a few
//     more than 256 and ran 1 wave per SIMD unbounded: bounded, HLL spills
//     2-12 VGPRs and runs 36 % faster, HLLE (one y-row buffer, below) 36 %,
//     Lax-Friedrichs 32 %, the MCL solvers (63-114 spilled) 17-20 %;
//   FP32: 4 for HLLC (2 VGPRs spill) and HLLE (8-22 spill; 3 %), 3 for the
//     others: the MCL solvers' 178-192 VGPRs ran 2 waves per SIMD unbounded
//     and 9-19 % slower; HLL and Lax-Friedrichs fit 3 unbounded, and at 4
//     HLL spills 25-56 VGPRs and ran 19 % slower, Lax-Friedrichs 29 and 6 %.
template <typename T, int SOLVER>
constexpr int march_min_waves() {
  if constexpr (sizeof(T) == 8) return CUEVA_MARCH_F64_WAVES > 0 ? CUEVA_MARCH_F64_WAVES : 2;
  else return CUEVA_MARCH_F32_WAVES > 0 ? CUEVA_MARCH_F32_WAVES : (SOLVER == 7 || SOLVER == 6 ? 4 : 3);
}

// y-face row buffers: 2 (one barrier per tile step), or 1 for HLLE (a second
// barrier at the end of the tile step, before the rows are rewritten), whose
// x rows also pass through LDS: 63 KB per FP64 block instead of 91 KB, so
// that two blocks fit a CU's 160 KB
template <int SOLVER>
constexpr int march_y_buffers() {
  return SOLVER == 6 ? 1 : 2;
}

// Wavefronts per block = face rows per tile step.  4 everywhere by default;
// FP32 HLL/HLLC may take 8 (a tile step then re-reads 5 halo rows per 8 face
// rows instead of per 4; two 61 KB blocks per CU still give 4 waves per SIMD)
// but measured the same step time at 2048x4096, so the smaller block stays.
#ifndef CUEVA_MARCH_F32_BLOCK_WAVES
#define CUEVA_MARCH_F32_BLOCK_WAVES 4
#endif
#ifndef CUEVA_MARCH_F64_BLOCK_WAVES
#define CUEVA_MARCH_F64_BLOCK_WAVES 4
#endif
__host__ __device__ constexpr int march_block_waves(bool fp32, int solver) {
  return solver == 6 ? 4 : (fp32 ? CUEVA_MARCH_F32_BLOCK_WAVES : CUEVA_MARCH_F64_BLOCK_WAVES);
}

template <int ORD, int SOLVER, typename T, bool FINAL, bool C2P = false>
__global__ void __launch_bounds__(64 * march_block_waves(sizeof(T) == 4, SOLVER), (march_min_waves<T, SOLVER>()))
k_faces_march(MarchArgs args) {
  constexpr int MW = march_block_waves(sizeof(T) == 4, SOLVER);
  constexpr bool XLDS = SOLVER == 6;            // x solved first, its fluxes through LDS
  constexpr int YB = march_y_buffers<SOLVER>();
  __shared__ T fyb[YB][MW][FY_N][64];           // y-face rows of this (and the previous) tile step
  __shared__ T fyp[FY_N][64];                   // wavefront 0's row below: previous step's top row
  __shared__ T fxb[XLDS ? MW : 1][FX_N][64];    // x-face rows (HLLE)
  if (halted(args.st)) return;  // uniform over the launch: before any barrier
  int strip = (int)blockIdx.x, segment = (int)blockIdx.y;
  if (args.xcd && !xcd_tile(args.strips, args.segs, strip, segment)) return;  // uniform over the block
  const int lane = threadIdx.x, w = threadIdx.y;
  const int i = strip * MARCH_COLS - 1 + lane;
  const int jc0 = segment * args.seg;
  const int jend = min(jc0 + args.seg - 1, args.g.jmax);
  int b = 0;
  for (int jt = jc0 - 1; jt <= jend; jt += MW, b ^= YB - 1) {  // uniform over the block: the barriers below
    // The arguments (plane pointers, geometry, scheme constants) are read
    // again from the kernel's argument block every tile step rather than
    // held in scalar registers across the loop, where they do not fit and
    // spill into vector-register lanes (FP64 HLLC: 156 SGPRs spilled, none
    // this way; FP32 HLLC: 147 VGPRs against 139).  FP64 HLL keeps them:
    // re-reading measured 24 % slower there.
    constexpr bool REREAD = sizeof(T) == 4 || SOLVER != 5;
    const __attribute__((address_space(4))) MarchArgs* ap =
        (const __attribute__((address_space(4))) MarchArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    if constexpr (REREAD) asm volatile("" : "+s"(ap));
    const MarchArgs& A = REREAD ? *(const MarchArgs*)ap : args;
    const Planes& P = A.P;
    const Geom& g = A.g;
    const Scheme& S = A.S;
    const int l = A.l;
    const StatusRef& st = A.st;
    const int j = jt + w;
    const bool face_ok = i <= g.imax && j <= jend;
    T fx[FX_N];
    auto put_y = [&](const T* f) {
#pragma unroll
      for (int a = 0; a < FY_N; a++) fyb[b][w][a][lane] = f[a];
    };
    auto put_x = [&](const T* f) {
#pragma unroll
      for (int a = 0; a < FX_N; a++) {
        if constexpr (XLDS) fxb[w][a][lane] = f[a];
        else fx[a] = f[a];
      }
    };
    if (face_ok) {
      face_pair<ORD, SOLVER, T, !XLDS, false>(g, P, S, st, i, j, put_x, put_y);
    } else {
      T z[FX_N > FY_N ? FX_N : FY_N];
#pragma unroll
      for (int a = 0; a < (FX_N > FY_N ? FX_N : FY_N); a++) z[a] = T(0);
      put_x(z);
      put_y(z);
    }
    T dfx[FX_N];
    if constexpr (!XLDS) {
#pragma unroll
      for (int a = 0; a < FX_N; a++) {
        const T left = __shfl_up(fx[a], 1, 64);
        dfx[a] = ddx<T>(S, fx[a] - left);
      }
    }
    __syncthreads();
    const bool cell_ok = face_ok && lane >= 1 && j >= jc0;
    if (cell_ok) {
      if constexpr (XLDS) {
#pragma unroll
        for (int a = 0; a < FX_N; a++) dfx[a] = ddx<T>(S, fxb[w][a][lane] - fxb[w][a][lane - 1]);
      }
      T dfy[FY_N];
#pragma unroll
      for (int a = 0; a < FY_N; a++) {
        const T below = w == 0 ? fyp[a][lane] : fyb[b][w - 1][a][lane];
        dfy[a] = ddy<T>(S, fyb[b][w][a][lane] - below);
      }
      const int64_t k = g.at(i, j);
      const Rhs r = rhs_of<T, false>(P, S, k, dfx, dfy);  // never with REC_PRIM = 1
      if constexpr (!FINAL) {
#pragma unroll
        for (int t = 0; t < R_N; t++) P.rhs[l][t][k] = r.r[t];
        // the next stage's state of this point (pointwise in U^n and the
        // stage right-hand sides, the last of which was just stored)
        double sv[S_N];
        stage_point(P, S, l + 1, k, false, P.p[P_VX][k], P.p[P_VY][k], P.p[P_VZ][k], P.p[P_EX][k], P.p[P_EY][k],
                    P.p[P_EZ][k], P.p[P_BX][k], P.p[P_BY][k], P.p[P_BZ][k], P.p[P_RHO][k], P.p[P_P][k],
                    P.p[P_PSI][k], P.p[P_PHI][k], P.p[P_Q][k], sv);
        store_stage(P.sn, k, sv);
      } else {
        double nu[A_N];
        final_update_from(P, S, k, r, nu);
#pragma unroll
        for (int a = 0; a < A_N; a++) P.rhs[l][a][k] = nu[a];
        if constexpr (C2P) recover_point(P, S, st, i, j, k, nu, A.stage0);
      }
    }
    if (w == 0) {
#pragma unroll
      for (int a = 0; a < FY_N; a++) fyp[a][lane] = fyb[b][MW - 1][a][lane];
    }
    if constexpr (YB == 1) __syncthreads();  // every read of this step's rows before the next step's writes
    else if constexpr (XLDS) __builtin_amdgcn_wave_barrier();  // this step's fxb reads before the next step's writes
  }
}

// kernel of (rec_order, solver, stage kind), for a launch by pointer
template <typename T>
inline const void* march_kernel(int rec_order, int solver, bool final_stage, bool c2p) {
#define CUEVA_MK(O, SV)                                                                          \
  (final_stage ? (c2p ? reinterpret_cast<const void*>(&k_faces_march<O, SV, T, true, true>)      \
                      : reinterpret_cast<const void*>(&k_faces_march<O, SV, T, true, false>))    \
               : reinterpret_cast<const void*>(&k_faces_march<O, SV, T, false, false>))
  if (solver == 1) return CUEVA_MK(5, 1);  // second order: the MP order does not enter
  if (solver == 2) return CUEVA_MK(5, 2);
  if (solver == 3) return CUEVA_MK(5, 3);
  if (solver == 4) return CUEVA_MK(5, 4);
  switch (rec_order * 10 + solver) {
    case 55: return CUEVA_MK(5, 5);
    case 56: return CUEVA_MK(5, 6);
    case 57: return CUEVA_MK(5, 7);
    case 75: return CUEVA_MK(7, 5);
    case 76: return CUEVA_MK(7, 6);
    case 77: return CUEVA_MK(7, 7);
    case 95: return CUEVA_MK(9, 5);
    case 96: return CUEVA_MK(9, 6);
    case 97: return CUEVA_MK(9, 7);
    default: return nullptr;
  }
#undef CUEVA_MK
}

}  // namespace cueva
