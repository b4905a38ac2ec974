// Device functions of the CUEVA step for gfx950.
//
// Every function keeps the floating-point operation order of the reference
// routine it cites (Fortran evaluates left to right; x**n is x*(x*(...x))),
// and the library is compiled with -ffp-contract=off, so the FP64 path
// reproduces the reference's roundings.  File:line citations are relative to
// the reference's cueva_2D_base/ directory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cueva_crmath.hpp"

namespace cueva {

constexpr int GH = 6;  // ghost layers (scalar.f95:9)

// ---------------------------------------------------------------------------
// grid geometry: plane of doubles with reference indices (-6:imax+6,-6:jmax+6)
// ---------------------------------------------------------------------------
struct Geom {
  int imax, jmax;
  int pitch;         // doubles per row (>= imax+15, multiple of 8)
  int rows;          // jmax + 13
  int64_t plane;     // doubles per plane (multiple of 64)
  int bcx, bcy;      // 0 copy, 1 periodic
  // element (i,j) lives at offset (j+6)*pitch + (i+8)
  __host__ __device__ inline int64_t at(int i, int j) const {
    return (int64_t)(j + GH) * pitch + (i + 8);
  }
};

// Ghost index map of the reference boundary routines (04_boundary.f95,
// 24_boundary_conserved.f95, 25_boundary_electric.f95,
// 24_boundary_flux_conserved.f95): copy clamps to [0,n]; periodic wraps with
// the duplicated edge, ghost -k <- n-k and ghost n+k <- k.
__host__ __device__ inline int src_index(int i, int n, int periodic) {
  if (i >= 0 && i <= n) return i;
  if (periodic) return i < 0 ? n + i : i - n;
  return i < 0 ? 0 : n;
}

template <typename T> __device__ inline T sq(T x) { return x * x; }
template <typename T> __device__ inline T pw3(T x) { return x * (x * x); }
template <typename T> __device__ inline T pw4(T x) { return x * (x * (x * x)); }
// min/max as the reference build evaluates them (FP64, bitwise path); the
// FP32 form uses the IEEE-754 2019 minimum/maximum (one v_minimum3_f32 /
// v_maximum3_f32 instead of a compare, a hazard wait and a select; they
// differ only in the sign of a zero result and propagate every NaN)
// Fortran min / max as the reference's compiler (flang) evaluates them: the
// first argument when it compares strictly less (greater), else the second,
// so min(NaN, x) = x, min(x, NaN) = NaN, min(-0, +0) = +0; three arguments
// fold from the left
template <typename T> __device__ inline T mn(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ inline T mx(T a, T b) { return a > b ? a : b; }
template <> __device__ inline float mn<float>(float a, float b) { return __builtin_elementwise_minimum(a, b); }
template <> __device__ inline float mx<float>(float a, float b) { return __builtin_elementwise_maximum(a, b); }
// max(lo, min(hi, x)) as the reference build evaluates it (a NaN passes)
template <typename T> __device__ inline T clampf(T x, T lo, T hi) {
  T t = x > hi ? hi : x;
  return t < lo ? lo : t;
}
__device__ inline double csign(double x) { return copysign(1.0, x); }
__device__ inline float csign(float x) { return copysignf(1.0f, x); }
__device__ inline double fab(double x) { return fabs(x); }
__device__ inline float fab(float x) { return fabsf(x); }

// Quotients of the face solves: the IEEE quotient in FP64 (the reference's
// arithmetic, bit for bit); in FP32 (mixed form) the product with the
// hardware reciprocal (v_rcp_f32, within 1 ulp), taken once for a
// denominator several quotients share.  inv_sqrt: 1/sqrt(x) likewise
// (v_rsq_f32 in FP32).
template <typename T> struct Quot {
  T d;
  __device__ explicit Quot(T den) : d(den) {}
  __device__ T operator()(T num) const { return num / d; }
};
template <> struct Quot<float> {
  float r;
  __device__ explicit Quot(float den) : r(__builtin_amdgcn_rcpf(den)) {}
  __device__ float operator()(float num) const { return num * r; }
};
template <typename T> __device__ inline T quot(T num, T den) { return Quot<T>(den)(num); }
__device__ inline double inv_sqrt(double x) { return 1.0 / sqrt(x); }
__device__ inline float inv_sqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// square root of the face solves: IEEE in FP64, v_sqrt_f32 (within 1 ulp) in FP32
__device__ inline double face_sqrt(double x) { return sqrt(x); }
__device__ inline float face_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// ---------------------------------------------------------------------------
// MP5 / MP7 / MP9 interface values (mp5.f95:40-264, mp7.f95, mp9.f95)
// ---------------------------------------------------------------------------
// minmod, funciones.f95:148-162
template <typename T> __device__ inline T dmm(T x, T y) {
  return (T(0.5) * (csign(x) + csign(y))) * mn(fab(x), fab(y));
}
// 4-argument minmod, funciones.f95:165-185
template <typename T> __device__ inline T dm4(T w, T x, T y, T z) {
  T sw = csign(w);
  T m = mn(mn(mn(fab(w), fab(x)), fab(y)), fab(z));
  return ((T(0.125) * (sw + csign(x))) * fab((sw + csign(y)) * (sw + csign(z)))) * m;
}
// FP32 (mixed form): both minmods are medians with zero -- min(x, y) where
// both are positive, max(x, y) where both are negative, 0 where the signs
// differ, the sign factors being exactly 1 -- so one v_med3_f32 per pair
// gives the same values (up to the sign of a zero result)
template <> __device__ inline float dmm<float>(float x, float y) { return __builtin_amdgcn_fmed3f(x, y, 0.0f); }
template <> __device__ inline float dm4<float>(float w, float x, float y, float z) {
  return __builtin_amdgcn_fmed3f(__builtin_amdgcn_fmed3f(w, x, 0.0f), __builtin_amdgcn_fmed3f(y, z, 0.0f), 0.0f);
}

template <int ORD> struct MPc;
template <> struct MPc<5> { static constexpr int K = 2; };
template <> struct MPc<7> { static constexpr int K = 3; };
template <> struct MPc<9> { static constexpr int K = 4; };

// unlimited interface values at cell c of the register array v
template <int ORD, typename T> __device__ inline T vor_plus(const T* v, int c) {
  if constexpr (ORD == 5)  // mp5.f95:79-80
    return T(1.0 / 60.0) * ((((T(2) * v[c - 2] - T(13) * v[c - 1]) + T(47) * v[c]) + T(27) * v[c + 1]) -
                            T(3) * v[c + 2]);
  else if constexpr (ORD == 7)  // mp7.f95:111-113
    return T(1.0 / 420.0) * ((((((T(-3) * v[c - 3] + T(25) * v[c - 2]) - T(101) * v[c - 1]) + T(319) * v[c]) +
                               T(214) * v[c + 1]) - T(38) * v[c + 2]) + T(4) * v[c + 3]);
  else  // mp9.f95:116-118
    return T(1.0 / 2520.0) *
           ((((((((T(4) * v[c - 4] - T(41) * v[c - 3]) + T(199) * v[c - 2]) - T(641) * v[c - 1]) + T(1879) * v[c]) +
               T(1375) * v[c + 1]) - T(305) * v[c + 2]) + T(55) * v[c + 3]) - T(5) * v[c + 4]);
}
template <int ORD, typename T> __device__ inline T vor_minus(const T* v, int c) {
  if constexpr (ORD == 5)  // mp5.f95:82-83
    return T(1.0 / 60.0) * ((((T(-3) * v[c - 2] + T(27) * v[c - 1]) + T(47) * v[c]) - T(13) * v[c + 1]) +
                            T(2) * v[c + 2]);
  else if constexpr (ORD == 7)  // mp7.f95:115-117
    return T(1.0 / 420.0) * ((((((T(-3) * v[c + 3] + T(25) * v[c + 2]) - T(101) * v[c + 1]) + T(319) * v[c]) +
                               T(214) * v[c - 1]) - T(38) * v[c - 2]) + T(4) * v[c - 3]);
  else  // mp9.f95:121-123
    return T(1.0 / 2520.0) *
           ((((((((T(4) * v[c + 4] - T(41) * v[c + 3]) + T(199) * v[c + 2]) - T(641) * v[c + 1]) + T(1879) * v[c]) +
               T(1375) * v[c - 1]) - T(305) * v[c - 2]) + T(55) * v[c - 3]) - T(5) * v[c - 4]);
}

// U^L_{i+1/2} at cell c (mp5.f95:104-136).  The limiter is evaluated only
// when some lane of the wavefront needs it; lanes that keep the unlimited
// value select it, so every lane gets the reference's value.
template <int ORD, typename T> __device__ inline T mp_up(const T* v, int c) {
  const T alpha = T(2), epsm = T(1.0e-10), b2 = T(4.0 / 3.0);
  T u0 = v[c];
  T du0 = v[c] - v[c - 1], du1 = v[c + 1] - v[c];
  T vorp = vor_plus<ORD>(v, c);
  T vmpp = u0 + dmm(du1, alpha * du0);
  const bool keep = (vorp - u0) * (vorp - vmpp) <= epsm;
  if (__all(keep)) return vorp;  // no lane of the wavefront needs the limiter
  T d2m = (v[c - 2] - T(2) * v[c - 1]) + v[c];
  T d20 = (v[c - 1] - T(2) * v[c]) + v[c + 1];
  T d2p = (v[c] - T(2) * v[c + 1]) + v[c + 2];
  T dm4jph = dm4(T(4) * d20 - d2p, T(4) * d2p - d20, d20, d2p);
  T dm4jmh = dm4(T(4) * d20 - d2m, T(4) * d2m - d20, d20, d2m);
  T vul = u0 + alpha * du0;
  T vav = T(0.5) * (u0 + v[c + 1]);
  T vmd = vav - T(0.5) * dm4jph;
  T vlc = (u0 + T(0.5) * du0) + b2 * dm4jmh;
  T vmin = mx(mn(mn(u0, v[c + 1]), vmd), mn(mn(u0, vul), vlc));
  T vmax = mn(mx(mx(u0, v[c + 1]), vmd), mx(mx(u0, vul), vlc));
  return keep ? vorp : vorp + dmm(vmin - vorp, vmax - vorp);
}

// U^R_{i-1/2} at cell c (mp5.f95:142-167)
template <int ORD, typename T> __device__ inline T mp_um(const T* v, int c) {
  const T alpha = T(2), epsm = T(1.0e-10), b2 = T(4.0 / 3.0);
  T u0 = v[c];
  T du0 = v[c] - v[c - 1], du1 = v[c + 1] - v[c];
  T vorm = vor_minus<ORD>(v, c);
  T vmpm = u0 + dmm(-du0, -(alpha * du1));
  const bool keep = (vorm - u0) * (vorm - vmpm) <= epsm;
  if (__all(keep)) return vorm;
  T d2m = (v[c - 2] - T(2) * v[c - 1]) + v[c];
  T d20 = (v[c - 1] - T(2) * v[c]) + v[c + 1];
  T d2p = (v[c] - T(2) * v[c + 1]) + v[c + 2];
  T dm4jph = dm4(T(4) * d20 - d2p, T(4) * d2p - d20, d20, d2p);
  T dm4jmh = dm4(T(4) * d20 - d2m, T(4) * d2m - d20, d20, d2m);
  T vul = u0 - alpha * du1;
  T vav = T(0.5) * (u0 + v[c - 1]);
  T vmd = vav - T(0.5) * dm4jmh;
  T vlc = (u0 - T(0.5) * du1) + b2 * dm4jph;
  T vmin = mx(mn(mn(u0, v[c - 1]), vmd), mn(mn(u0, vul), vlc));
  T vmax = mn(mx(mx(u0, v[c - 1]), vmd), mx(mx(u0, vul), vlc));
  return keep ? vorm : vorm + dmm(vmin - vorm, vmax - vorm);
}

// Face value pair at the face between cells i and i+1: L = up(i), R = um(i+1).
// v holds u[i-K .. i+1+K] (2K+2 values), K = MPc<ORD>::K.
template <int ORD, typename T> __device__ inline void mp_face(const T* v, T& L, T& R) {
  constexpr int K = MPc<ORD>::K;
  L = mp_up<ORD>(v, K);
  R = mp_um<ORD>(v, K + 1);
}

// Same with the positivity fallback of mp5.f95:177-216 (p and rho slots): at
// each cell, first "up < 0" then "um < 0" rewrites both interface values.
template <int ORD, typename T> __device__ inline void mp_face_pos(const T* v, T& L, T& R) {
  constexpr int K = MPc<ORD>::K;
  // cell i
  T up0 = mp_up<ORD>(v, K), um0 = mp_um<ORD>(v, K);
  T du0 = v[K] - v[K - 1], du1 = v[K + 1] - v[K];
  if (up0 < T(0)) { up0 = v[K] + T(0.5) * dmm(du0, du1); um0 = v[K]; }
  if (um0 < T(0)) { um0 = v[K] - T(0.5) * dmm(du0, du1); up0 = v[K]; }
  // cell i+1
  T up1 = mp_up<ORD>(v, K + 1), um1 = mp_um<ORD>(v, K + 1);
  T eu0 = v[K + 1] - v[K], eu1 = v[K + 2] - v[K + 1];
  if (up1 < T(0)) { up1 = v[K + 1] + T(0.5) * dmm(eu0, eu1); um1 = v[K + 1]; }
  if (um1 < T(0)) { um1 = v[K + 1] - T(0.5) * dmm(eu0, eu1); up1 = v[K + 1]; }
  (void)um0;
  (void)up1;
  L = up0;
  R = um1;
}

// ---------------------------------------------------------------------------
// one reconstructed face side and its physical fluxes
// ---------------------------------------------------------------------------
template <typename T> struct Side {
  T psi, phi, ex, ey, ez, bx, by, bz, q, d, tau, sx, sy, sz;
  T vx, vy, vz, p, rho, jn;  // jn: normal current (Jx on x faces, Jy on y faces)
  T w, e2b2, h;
  T fd, ftau, fs1, fs2, fs3;  // normal physical fluxes
  // the reference's "_m" states of the tangential E and B pair (x faces: Ez,
  // By; y faces: Ex, Bz), read only by the HLLC solve of the MC-limited
  // solver; with MPx they are the negated states
  T em, bm;
};

// W, E^2+B^2, enthalpy and normal fluxes (16_hll_flow_mpx.f95:1463-1466,
// 1522-1532, 1603-1617, 1627-1715).  DIR 0 = x face, 1 = y face.  HLLE_W:
// W as 16_hlle_flow_mpx.f95:2302-2305 forms it (squares subtracted one by one).
template <int DIR, bool HLLE_W = false, typename T> __device__ inline void side_physics(T gam, Side<T>& o) {
  if constexpr (HLLE_W) {
    o.w = inv_sqrt(((T(1) - sq(o.vx)) - sq(o.vy)) - sq(o.vz));
  } else {
    T v2 = (sq(o.vx) + sq(o.vy)) + sq(o.vz);
    o.w = inv_sqrt(T(1) - v2);
  }
  o.e2b2 = ((((sq(o.ex) + sq(o.ey)) + sq(o.ez)) + sq(o.bx)) + sq(o.by)) + sq(o.bz);
  T eps = quot(o.p, (gam - T(1)) * o.rho);
  o.h = o.rho * (T(1) + eps) + o.p;
  T hw2 = o.h * sq(o.w);
  if constexpr (DIR == 0) {
    o.fd = o.d * o.vx;
    o.ftau = (o.bz * o.ey - o.by * o.ez) + hw2 * o.vx;
    o.fs1 = (((-sq(o.ex)) - sq(o.bx)) + hw2 * sq(o.vx)) + T(0.5) * o.e2b2 + o.p;
    o.fs2 = ((-(o.ex * o.ey)) - o.bx * o.by) + (hw2 * o.vx) * o.vy;
    o.fs3 = ((-(o.ex * o.ez)) - o.bx * o.bz) + (hw2 * o.vx) * o.vz;
  } else {
    o.fd = o.d * o.vy;
    o.ftau = (o.bx * o.ez - o.bz * o.ex) + hw2 * o.vy;
    o.fs1 = ((-(o.ex * o.ey)) - o.bx * o.by) + (hw2 * o.vx) * o.vy;
    o.fs2 = (((-sq(o.ey)) - sq(o.by)) + hw2 * sq(o.vy)) + T(0.5) * o.e2b2 + o.p;
    o.fs3 = ((-(o.ey * o.ez)) - o.by * o.bz) + (hw2 * o.vy) * o.vz;
  }
}

// face flux indices (order of the divergence in 17_fluxandsource.f95:433-637)
enum FxIdx { FX_EXPSI, FX_BXPHI, FX_PHIBX, FX_EZBY, FX_EYBZ, FX_PSIEX, FX_BZEY, FX_BYEZ, FX_J, FX_FD,
             FX_FTAU, FX_FSX, FX_FSY, FX_FSZ, FX_N };
enum FyIdx { FY_EYPSI, FY_BYPHI, FY_EZBX, FY_PHIBY, FY_EXBZ, FY_BZEX, FY_PSIEY, FY_BXEZ, FY_J, FY_FD,
             FY_FTAU, FY_FSX, FY_FSY, FY_FSZ, FY_N };

// (s2*FL - s1*FR + s1*s2*(UR-UL))/(s2-s1), s1 = -1, s2 = 1 (parameters.f95:283-284)
// The reference's "_m" slot of a field: the MPx reconstruction of -u, which
// equals -MPx(u) except that a zero comes out as +0 (16_hll_flow_mpx.f95:
// 784-917 packs -E and -B into their own slots); 0 - x is exactly that.
template <typename T> __device__ inline T msl(T x) { return T(0) - x; }

template <typename T> __device__ inline T hllf(T fl, T fr, T ul, T ur) {
  const T S1 = T(-1), S2 = T(1);
  return ((S2 * fl - S1 * fr) + (S1 * S2) * (ur - ul)) / (S2 - S1);
}
template <typename T> __device__ inline T hll_state(T ul, T ur, T fl, T fr) {
  const T S1 = T(-1), S2 = T(1);
  return (((S2 * ur - S1 * ul) + fl) - fr) / (S2 - S1);
}
template <typename T> __device__ inline T hll_flux(T fl, T fr, T ul, T ur) {
  return T(0.5) * ((fl + fr) - (ur - ul));
}

// HLL face fluxes, 16_hll_flow_mpx.f95:1738-1958.  The "_m" slots of the
// reference hold MPx(-u) = -MPx(u); they enter as negated values.
template <typename T> __device__ inline void hll_x(const Side<T>& L, const Side<T>& R, T* f) {
  f[FX_EXPSI] = hllf(L.ex, R.ex, L.psi, R.psi);
  f[FX_BXPHI] = hllf(L.bx, R.bx, L.phi, R.phi);
  f[FX_PHIBX] = hllf(L.phi, R.phi, L.bx, R.bx);
  f[FX_EZBY] = -hllf(msl(L.ez), msl(R.ez), L.by, R.by);
  f[FX_EYBZ] = hllf(L.ey, R.ey, L.bz, R.bz);
  f[FX_PSIEX] = hllf(L.psi, R.psi, L.ex, R.ex);
  f[FX_BZEY] = hllf(L.bz, R.bz, L.ey, R.ey);
  f[FX_BYEZ] = -hllf(msl(L.by), msl(R.by), L.ez, R.ez);
  f[FX_J] = hllf(L.jn, R.jn, L.q, R.q);
  f[FX_FD] = hllf(L.fd, R.fd, L.d, R.d);
  f[FX_FTAU] = hllf(L.ftau, R.ftau, L.tau, R.tau);
  f[FX_FSX] = hllf(L.fs1, R.fs1, L.sx, R.sx);
  f[FX_FSY] = hllf(L.fs2, R.fs2, L.sy, R.sy);
  f[FX_FSZ] = hllf(L.fs3, R.fs3, L.sz, R.sz);
}
template <typename T> __device__ inline void hll_y(const Side<T>& L, const Side<T>& R, T* f) {
  f[FY_EYPSI] = hllf(L.ey, R.ey, L.psi, R.psi);
  f[FY_BYPHI] = hllf(L.by, R.by, L.phi, R.phi);
  f[FY_EZBX] = hllf(L.ez, R.ez, L.bx, R.bx);
  f[FY_PHIBY] = hllf(L.phi, R.phi, L.by, R.by);
  f[FY_EXBZ] = -hllf(msl(L.ex), msl(R.ex), L.bz, R.bz);
  f[FY_BZEX] = -hllf(msl(L.bz), msl(R.bz), L.ex, R.ex);
  f[FY_PSIEY] = hllf(L.psi, R.psi, L.ey, R.ey);
  f[FY_BXEZ] = hllf(L.bx, R.bx, L.ez, R.ez);
  f[FY_J] = hllf(L.jn, R.jn, L.q, R.q);
  f[FY_FD] = hllf(L.fd, R.fd, L.d, R.d);
  f[FY_FTAU] = hllf(L.ftau, R.ftau, L.tau, R.tau);
  f[FY_FSX] = hllf(L.fs1, R.fs1, L.sx, R.sx);
  f[FY_FSY] = hllf(L.fs2, R.fs2, L.sy, R.sy);
  f[FY_FSZ] = hllf(L.fs3, R.fs3, L.sz, R.sz);
}

// get_hllc_vel (funciones.f95:188-259): Newton on a v^2 - b v + c = 0 in
// [-1,1].  The reference evaluates its first step at the iterate before
// assigning it; v0 is that point (clamp(rt) in the seed mode).  Tolerances
// are the reference's for double and scaled to the precision for float.
// vlast receives the last iterate before the final clamp.  maxiter: 100
// (funciones.f95), 200 in the copy inside 16_hlle_flow_mcl.f95:4385.
template <typename T>
__device__ inline T hllc_vel(T a, T b, T c, T rt, T v0, int& fail, T& vlast, int maxiter = 100) {
  const T minerr = sizeof(T) == 8 ? T(1.0e-12) : T(2.0e-6);
  T oldv = clampf(rt, T(-1), T(1));
  T v = v0;
  int iter = 0;
  for (;;) {
    T fv = (a * v * v - b * v) + c;
    T dfdv = (T(2) * a) * v - b;
    v = clampf(oldv - quot(fv, dfdv), T(-1), T(1));
    T err = fab(oldv - v);
    oldv = v;
    iter = iter + 1;
    if (iter >= maxiter) { fail = 1; vlast = v; return v; }
    if (!(err > minerr)) break;
  }
  vlast = v;
  return clampf(v, T(-1) + T(1.0e-3), T(1) - T(1.0e-3));
}

// contact speed, 16_hllc_flow_mpx.f95:3185-3334
template <typename T>
__device__ inline T hllc_contact(T a, T b, T c, bool seed_ref, T vref, int& fail, T zero = T(1.0e-9),
                                 T* vlast_out = nullptr) {
  T vlast;
  T det = sq(b) - (T(4) * a) * c;
  T qr = -(T(0.5) * (b + csign(b) * fab(face_sqrt(det))));
  T rt2 = quot(c, qr);
  if (det < T(0)) rt2 = quot(c, b);
  T v0 = seed_ref ? vref : clampf(rt2, T(-1), T(1));
  T v = hllc_vel(a, -b, c, rt2, v0, fail, vlast);
  if (vlast_out) *vlast_out = vlast;
  if (fab(v) <= zero)
    v = T(0);
  else if (fab(v) > T(1))
    v = clampf(v, T(-1) + T(1.0e-6), T(1) - T(1.0e-6));
  return v;
}

// The star total energy's (S*_n - (E x B)*_n) / v* (16_hllc_flow_mpx.f95:
// 3536-3574): FP64 divides by v* as the reference does.  FP32 (mixed form)
// uses (S*_n - (E x B)*_n) / v* = (S2 c/v* - a)/(S2 - v*) (right state) or
// (-S1 c/v* + a)/(v* - S1) (left), with c/v* = -(a v* + b) at the root of
// the contact quadratic a v^2 + b v + c = 0 (c = S_n - (E x B)_n): where the
// normal flow vanishes c and v* are both rounding noise in FP32 and their
// quotient is not.  Only for |v*| < 1e-3: there the division amplifies
// the FP32 rounding of c by 1/|v*| > 1000, while the root form carries the
// FP32 convergence error of Newton's v* times a, which is the larger error
// at moderate |v*|.
template <typename T> __device__ inline bool star_tau_from_root(T v) {
  return sizeof(T) == 4 && fab(v) < T(1.0e-3);
}

// HLLC x face, 16_hllc_flow_mpx.f95:2811-4295.  Returns 0 or a status.
// seed_ref: first Newton point as the flang -O2 reference build (R.bz).
// MC: the MC-limited solver (16_hllc_flow_mcl.f95:2480-4140): its own "_m"
// states, |v*| <= 1e-6 -> 0, and, when seed_ref, Newton's first point mcref
// (the reference build's: the y face's (E x B)_y of the HLL states for the
// x solve, the x solve's last iterate for the y solve); *vlast receives the
// solve's last iterate.
template <typename T, bool MC = false>
__device__ inline int hllc_x(const Side<T>& L, const Side<T>& R, T* f, bool seed_ref, T mcref = T(0),
                             T* vlast = nullptr) {
  const T S1 = T(-1), S2 = T(1);
  const T ezmL = MC ? L.em : msl(L.ez), ezmR = MC ? R.em : msl(R.ez);
  const T bymL = MC ? L.bm : msl(L.by), bymR = MC ? R.bm : msl(R.by);
  T psi_h = hll_state(L.psi, R.psi, L.ex, R.ex);
  T phi_h = hll_state(L.phi, R.phi, L.bx, R.bx);
  T ex_h = hll_state(L.ex, R.ex, L.psi, R.psi);
  T ey_h = hll_state(L.ey, R.ey, L.bz, R.bz);
  T ez_h = hll_state(L.ez, R.ez, bymL, bymR);
  T bx_h = hll_state(L.bx, R.bx, L.phi, R.phi);
  T by_h = hll_state(L.by, R.by, ezmL, ezmR);
  T bz_h = hll_state(L.bz, R.bz, L.ey, R.ey);
  T dd_h = hll_state(L.d, R.d, L.fd, R.fd);
  T tau_h = hll_state(L.tau, R.tau, L.ftau, R.ftau);
  T sx_h = hll_state(L.sx, R.sx, L.fs1, R.fs1);
  T sy_h = hll_state(L.sy, R.sy, L.fs2, R.fs2);
  T sz_h = hll_state(L.sz, R.sz, L.fs3, R.fs3);
  T psi_f = hll_flux(L.ex, R.ex, L.psi, R.psi);
  T phi_f = hll_flux(L.bx, R.bx, L.phi, R.phi);
  T bx_f = hll_flux(L.phi, R.phi, L.bx, R.bx);
  T by_f = hll_flux(ezmL, ezmR, L.by, R.by);
  T bz_f = hll_flux(L.ey, R.ey, L.bz, R.bz);
  T ex_f = hll_flux(L.psi, R.psi, L.ex, R.ex);
  T ey_f = hll_flux(L.bz, R.bz, L.ey, R.ey);
  T ez_f = hll_flux(bymL, bymR, L.ez, R.ez);
  T q_f = hll_flux(L.jn, R.jn, L.q, R.q);
  T dd_f = hll_flux(L.fd, R.fd, L.d, R.d);
  T tau_f = hll_flux(L.ftau, R.ftau, L.tau, R.tau);
  T sx_f = hll_flux(L.fs1, R.fs1, L.sx, R.sx);
  T sy_f = hll_flux(L.fs2, R.fs2, L.sy, R.sy);
  T sz_f = hll_flux(L.fs3, R.fs3, L.sz, R.sz);
  T erx = bz_h * ey_h - by_h * ez_h;
  T ery = bx_h * ez_h - bz_h * ex_h;
  T erz = by_h * ex_h - bx_h * ey_h;
  T e2p = sq(ey_h) + sq(ez_h);
  T b2p = sq(by_h) + sq(bz_h);
  T e2b2s = T(0.5) * (((sq(ex_h) + sq(ey_h)) + sq(ez_h)) + ((sq(bx_h) + sq(by_h)) + sq(bz_h)));
  T a = tau_f - erx;
  T b = ((e2p + b2p) - tau_h) - sx_f;
  T c = sx_h - erx;
  int fail = 0;
  T v = MC ? hllc_contact(a, b, c, seed_ref, mcref, fail, T(1.0e-6), vlast) : hllc_contact(a, b, c, seed_ref, R.bz, fail);
  if (fail) return 7;  // a NaN v* goes on to the branches on v*, which stop (8)
  const Quot<T> dR(S2 - v), dL(v - S1), dV(v);
  T pstr = (sx_f + (sq(ex_h) + sq(bx_h))) - (tau_f - erx) * v;
  T psiR = dR(S2 * psi_h - v * ex_f);
  T psiL = dL(-(S1 * psi_h) + v * ex_f);
  T phiR = dR(S2 * phi_h - v * bx_f);
  T phiL = dL(-(S1 * phi_h) + v * bx_f);
  T sxR = dR(S2 * sx_h - tau_f * v);
  T sxL = dL(-(S1 * sx_h) + tau_f * v);
  T syR = dR((((S2 * sy_h - sy_f) - ery * v) - ex_h * ey_h) - bx_h * by_h);
  T syL = dL((((-(S1 * sy_h) + sy_f) + ery * v) + ex_h * ey_h) + bx_h * by_h);
  T szR = dR((((S2 * sz_h - sz_f) - erz * v) - ex_h * ez_h) - bx_h * bz_h);
  T szL = dL((((-(S1 * sz_h) + sz_f) + erz * v) + ex_h * ez_h) + bx_h * bz_h);
  T tauR = (dV(sxR - erx) - pstr) + e2b2s;
  T tauL = (dV(sxL - erx) - pstr) + e2b2s;
  if (star_tau_from_root(v)) {
    const T cv = -(a * v + b);
    tauR = (dR(S2 * cv - a) - pstr) + e2b2s;
    tauL = (dL(-(S1 * cv) + a) - pstr) + e2b2s;
  }
  T ddR = dR(S2 * dd_h - dd_f);
  T ddL = dL(-(S1 * dd_h) + dd_f);
  T psifR, psifL, phifR, phifL, fdR, fdL, ftR, ftL, fsxR, fsxL, fsyR, fsyL, fszR, fszL;
  if (v == T(0)) {
    psifR = psifL = psi_f; phifR = phifL = phi_f;
    fdR = fdL = dd_f; ftR = ftL = tau_f;
    fsxR = fsxL = sx_f; fsyR = fsyL = sy_f; fszR = fszL = sz_f;
  } else {
    T cr = S2 * (v - S1), cl = S1 * (S2 - v), den = S2 - S1;
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
  f[FX_J] = q_f;
  f[FX_PHIBX] = bx_f; f[FX_EZBY] = -by_f; f[FX_EYBZ] = bz_f;
  f[FX_PSIEX] = ex_f; f[FX_BZEY] = ey_f; f[FX_BYEZ] = -ez_f;
  if (S1 < T(0) && T(0) < v) {
    f[FX_EXPSI] = psifL; f[FX_BXPHI] = phifL;
    f[FX_FD] = fdL; f[FX_FTAU] = ftL; f[FX_FSX] = fsxL; f[FX_FSY] = fsyL; f[FX_FSZ] = fszL;
  } else if (v < T(0) && T(0) < S2) {
    f[FX_EXPSI] = psifR; f[FX_BXPHI] = phifR;
    f[FX_FD] = fdR; f[FX_FTAU] = ftR; f[FX_FSX] = fsxR; f[FX_FSY] = fsyR; f[FX_FSZ] = fszR;
  } else if (v == T(0)) {
    // 0.5*(x+x) == x and -(0.5*(x+x)) == -x exactly
    f[FX_EXPSI] = T(0.5) * (psifR + psifL); f[FX_BXPHI] = T(0.5) * (phifR + phifL);
    f[FX_FD] = T(0.5) * (fdR + fdL); f[FX_FTAU] = T(0.5) * (ftR + ftL);
    f[FX_FSX] = T(0.5) * (fsxR + fsxL); f[FX_FSY] = T(0.5) * (fsyR + fsyL); f[FX_FSZ] = T(0.5) * (fszR + fszL);
  } else {
    return 8;
  }
  return 0;
}

// HLLC y face, 16_hllc_flow_mpx.f95:2950-3095, 3338-3489, 3536-3574,
// 3742-3901, 4305-4606.  seed_ref: first Newton point = -b (the b argument).
template <typename T, bool MC = false>
__device__ inline int hllc_y(const Side<T>& L, const Side<T>& R, T* f, bool seed_ref, T mcref = T(0)) {
  const T S1 = T(-1), S2 = T(1);
  const T exmL = MC ? L.em : msl(L.ex), exmR = MC ? R.em : msl(R.ex);
  const T bzmL = MC ? L.bm : msl(L.bz), bzmR = MC ? R.bm : msl(R.bz);
  T psi_h = hll_state(L.psi, R.psi, L.ey, R.ey);
  T phi_h = hll_state(L.phi, R.phi, L.by, R.by);
  T ey_h = hll_state(L.ey, R.ey, L.psi, R.psi);
  T ez_h = hll_state(L.ez, R.ez, L.bx, R.bx);
  T ex_h = hll_state(L.ex, R.ex, bzmL, bzmR);
  T by_h = hll_state(L.by, R.by, L.phi, R.phi);
  T bz_h = hll_state(L.bz, R.bz, exmL, exmR);
  T bx_h = hll_state(L.bx, R.bx, L.ez, R.ez);
  T dd_h = hll_state(L.d, R.d, L.fd, R.fd);
  T tau_h = hll_state(L.tau, R.tau, L.ftau, R.ftau);
  T sy_h = hll_state(L.sy, R.sy, L.fs2, R.fs2);
  T sz_h = hll_state(L.sz, R.sz, L.fs3, R.fs3);
  T sx_h = hll_state(L.sx, R.sx, L.fs1, R.fs1);
  T psi_f = hll_flux(L.ey, R.ey, L.psi, R.psi);
  T phi_f = hll_flux(L.by, R.by, L.phi, R.phi);
  T by_f = hll_flux(L.phi, R.phi, L.by, R.by);
  T bz_f = hll_flux(exmL, exmR, L.bz, R.bz);
  T bx_f = hll_flux(L.ez, R.ez, L.bx, R.bx);
  T ey_f = hll_flux(L.psi, R.psi, L.ey, R.ey);
  T ez_f = hll_flux(L.bx, R.bx, L.ez, R.ez);
  T ex_f = hll_flux(bzmL, bzmR, L.ex, R.ex);
  T q_f = hll_flux(L.jn, R.jn, L.q, R.q);
  T dd_f = hll_flux(L.fd, R.fd, L.d, R.d);
  T tau_f = hll_flux(L.ftau, R.ftau, L.tau, R.tau);
  T sy_f = hll_flux(L.fs2, R.fs2, L.sy, R.sy);
  T sz_f = hll_flux(L.fs3, R.fs3, L.sz, R.sz);
  T sx_f = hll_flux(L.fs1, R.fs1, L.sx, R.sx);
  T erx = bz_h * ey_h - by_h * ez_h;
  T ery = bx_h * ez_h - bz_h * ex_h;
  T erz = by_h * ex_h - bx_h * ey_h;
  T e2p = sq(ex_h) + sq(ez_h);
  T b2p = sq(bx_h) + sq(bz_h);
  T e2b2s = T(0.5) * (((sq(ex_h) + sq(ey_h)) + sq(ez_h)) + ((sq(bx_h) + sq(by_h)) + sq(bz_h)));
  T a = tau_f - ery;
  T b = ((e2p + b2p) - tau_h) - sy_f;
  T c = sy_h - ery;
  int fail = 0;
  T v = MC ? hllc_contact(a, b, c, seed_ref, mcref, fail, T(1.0e-6)) : hllc_contact(a, b, c, seed_ref, -b, fail);
  if (fail) return 7;  // a NaN v* goes on to the branches on v*, which stop (8)
  const Quot<T> dR(S2 - v), dL(v - S1), dV(v);
  T pstr = (sy_f + (sq(ey_h) + sq(by_h))) - (tau_f - ery) * v;
  T psiR = dR(S2 * psi_h - v * ey_f);
  T psiL = dL(-(S1 * psi_h) + v * ey_f);
  T phiR = dR(S2 * phi_h - v * by_f);
  T phiL = dL(-(S1 * phi_h) + v * by_f);
  T syR = dR(S2 * sy_h - tau_f * v);
  T syL = dL(-(S1 * sy_h) + tau_f * v);
  T szR = dR((((S2 * sz_h - sz_f) - erz * v) - ey_h * ez_h) - by_h * bz_h);
  T szL = dL((((-(S1 * sz_h) + sz_f) + erz * v) + ey_h * ez_h) + by_h * bz_h);
  T sxR = dR((((S2 * sx_h - sx_f) - erx * v) - ey_h * ex_h) - by_h * bx_h);
  T sxL = dL((((-(S1 * sx_h) + sx_f) + erx * v) + ey_h * ex_h) + by_h * bx_h);
  T tauR = (dV(syR - ery) - pstr) + e2b2s;
  T tauL = (dV(syL - ery) - pstr) + e2b2s;
  if (star_tau_from_root(v)) {
    const T cv = -(a * v + b);
    tauR = (dR(S2 * cv - a) - pstr) + e2b2s;
    tauL = (dL(-(S1 * cv) + a) - pstr) + e2b2s;
  }
  T ddR = dR(S2 * dd_h - dd_f);
  T ddL = dL(-(S1 * dd_h) + dd_f);
  T psifR, psifL, phifR, phifL, fdR, fdL, ftR, ftL, fsxR, fsxL, fsyR, fsyL, fszR, fszL;
  if (v == T(0)) {
    psifR = psifL = psi_f; phifR = phifL = phi_f;
    fdR = fdL = dd_f; ftR = ftL = tau_f;
    fsxR = fsxL = sx_f; fsyR = fsyL = sy_f; fszR = fszL = sz_f;
  } else {
    T cr = S2 * (v - S1), cl = S1 * (S2 - v), den = S2 - S1;
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
  f[FY_J] = q_f;
  f[FY_EZBX] = bx_f; f[FY_PHIBY] = by_f; f[FY_EXBZ] = -bz_f;
  f[FY_BZEX] = -ex_f; f[FY_PSIEY] = ey_f; f[FY_BXEZ] = ez_f;
  if (S1 < T(0) && T(0) < v) {
    f[FY_EYPSI] = psifL; f[FY_BYPHI] = phifL;
    f[FY_FD] = fdL; f[FY_FTAU] = ftL; f[FY_FSX] = fsxL; f[FY_FSY] = fsyL; f[FY_FSZ] = fszL;
  } else if (v < T(0) && T(0) < S2) {
    f[FY_EYPSI] = psifR; f[FY_BYPHI] = phifR;
    f[FY_FD] = fdR; f[FY_FTAU] = ftR; f[FY_FSX] = fsxR; f[FY_FSY] = fsyR; f[FY_FSZ] = fszR;
  } else if (v == T(0)) {
    f[FY_EYPSI] = T(0.5) * (psifR + psifL); f[FY_BYPHI] = T(0.5) * (phifR + phifL);
    f[FY_FD] = T(0.5) * (fdR + fdL); f[FY_FTAU] = T(0.5) * (ftR + ftL);
    f[FY_FSX] = T(0.5) * (fsxR + fsxL); f[FY_FSY] = T(0.5) * (fsyR + fsyL); f[FY_FSZ] = T(0.5) * (fszR + fszL);
  } else {
    return 8;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// HLLE, 16_hlle_flow_mpx.f95:2258-4618: the electromagnetic subsystem is HLL
// with s = -1, +1; q, D, tau, S use relativistic sound-speed wave estimates
// (sa, sb) and an HLLC-structured contact for D, tau, S.
// ---------------------------------------------------------------------------

// wave estimates of one side, 16_hlle_flow_mpx.f95:2715-2740: the root
// holds -vn*(1-cs^2), the first power of the normal speed
template <typename T> __device__ inline void hlle_lambda(T gam, const Side<T>& o, T vn, T& lm, T& lp) {
  T cs = face_sqrt(quot(gam * o.p, o.h));
  T cs2 = sq(cs);
  T v2 = (sq(o.vx) + sq(o.vy)) + sq(o.vz);
  T root = face_sqrt((T(1) - v2) * ((T(1) - v2 * cs2) - vn * (T(1) - cs2)));
  const Quot<T> den(T(1) - cs2 * v2);
  lp = den(vn * (T(1) - cs2) + cs * root);
  lm = den(vn * (T(1) - cs2) - cs * root);
}

// min/max over the sides and the clamp to |s| < 1 (:2789-2818); 9 = NaN stop
template <int DIR, typename T>
__device__ inline int hlle_speeds(T gam, const Side<T>& L, const Side<T>& R, T& sa, T& sb) {
  T lmL, lpL, lmR, lpR;
  hlle_lambda(gam, L, DIR == 0 ? L.vx : L.vy, lmL, lpL);
  hlle_lambda(gam, R, DIR == 0 ? R.vx : R.vy, lmR, lpR);
  T a = lmL < lmR ? lmL : lmR;
  T b = lpL > lpR ? lpL : lpR;
  if (a != a || b != b) return 9;
  const T e2 = T(1.0e-9);
  if (fab(a) > T(1)) a = clampf(a, T(-1) + e2, T(1) - e2);
  if (fab(b) > T(1)) b = clampf(b, T(-1) + e2, T(1) - e2);
  sa = a;
  sb = b;
  return 0;
}

template <typename T> __device__ inline T hlle_state(T a, T b, T ul, T ur, T fl, T fr) {
  return quot(((b * ur - a * ul) + fl) - fr, b - a);
}
template <typename T> __device__ inline T hlle_flux(T a, T b, T fl, T fr, T ul, T ur) {
  return quot((b * fl - a * fr) + (a * b) * (ur - ul), b - a);
}

// (E x B)_y of the y face's HLL states: the first Newton point of the x
// contact solve in the flang -O2 reference build (still in the register)
template <typename T, bool MC = false> __device__ inline T hlle_y_ery(const Side<T>& L, const Side<T>& R) {
  T ez_h = hll_state(L.ez, R.ez, L.bx, R.bx);
  T ex_h = hll_state(L.ex, R.ex, MC ? L.bm : msl(L.bz), MC ? R.bm : msl(R.bz));
  T bz_h = hll_state(L.bz, R.bz, MC ? L.em : msl(L.ex), MC ? R.em : msl(R.ex));
  T bx_h = hll_state(L.bx, R.bx, L.ez, R.ez);
  return bx_h * ez_h - bz_h * ex_h;
}

// One HLLE face.  DIR 0: x face (16_hlle_flow_mpx.f95:2520-2862, 3225-3290,
// 3336-3412, 3459-3540, 4021-4175); DIR 1: y face (:2560-2708, 2893-3040,
// 3292-3334, 3413-3458, 3573-3654, 4413-4500).  vref: first Newton point
// when seed_ref; vlast: the solve's last iterate (the y solve's first point).
// A face whose waves all leave one side takes the reference's never-assigned
// (zero) one-sided scalars (:4021-4033, 4143-4160, which only its 1D branch
// sets at :222-475).
template <int DIR, typename T>
__device__ inline int hlle_face(T gam, const Side<T>& L, const Side<T>& R, T* f, bool seed_ref, T vref, T& vlast) {
  T e[3], b[3];  // electromagnetic HLL states (x, y, z)
  if constexpr (DIR == 0) {
    e[0] = hll_state(L.ex, R.ex, L.psi, R.psi);
    e[1] = hll_state(L.ey, R.ey, L.bz, R.bz);
    e[2] = hll_state(L.ez, R.ez, msl(L.by), msl(R.by));
    b[0] = hll_state(L.bx, R.bx, L.phi, R.phi);
    b[1] = hll_state(L.by, R.by, msl(L.ez), msl(R.ez));
    b[2] = hll_state(L.bz, R.bz, L.ey, R.ey);
    f[FX_EXPSI] = hll_flux(L.ex, R.ex, L.psi, R.psi);
    f[FX_BXPHI] = hll_flux(L.bx, R.bx, L.phi, R.phi);
    f[FX_PHIBX] = hll_flux(L.phi, R.phi, L.bx, R.bx);
    f[FX_EZBY] = -hll_flux(msl(L.ez), msl(R.ez), L.by, R.by);
    f[FX_EYBZ] = hll_flux(L.ey, R.ey, L.bz, R.bz);
    f[FX_PSIEX] = hll_flux(L.psi, R.psi, L.ex, R.ex);
    f[FX_BZEY] = hll_flux(L.bz, R.bz, L.ey, R.ey);
    f[FX_BYEZ] = -hll_flux(msl(L.by), msl(R.by), L.ez, R.ez);
  } else {
    e[1] = hll_state(L.ey, R.ey, L.psi, R.psi);
    e[2] = hll_state(L.ez, R.ez, L.bx, R.bx);
    e[0] = hll_state(L.ex, R.ex, msl(L.bz), msl(R.bz));
    b[1] = hll_state(L.by, R.by, L.phi, R.phi);
    b[2] = hll_state(L.bz, R.bz, msl(L.ex), msl(R.ex));
    b[0] = hll_state(L.bx, R.bx, L.ez, R.ez);
    f[FY_EYPSI] = hll_flux(L.ey, R.ey, L.psi, R.psi);
    f[FY_BYPHI] = hll_flux(L.by, R.by, L.phi, R.phi);
    f[FY_PHIBY] = hll_flux(L.phi, R.phi, L.by, R.by);
    f[FY_EXBZ] = -hll_flux(msl(L.ex), msl(R.ex), L.bz, R.bz);
    f[FY_EZBX] = hll_flux(L.ez, R.ez, L.bx, R.bx);
    f[FY_PSIEY] = hll_flux(L.psi, R.psi, L.ey, R.ey);
    f[FY_BXEZ] = hll_flux(L.bx, R.bx, L.ez, R.ez);
    f[FY_BZEX] = -hll_flux(msl(L.bz), msl(R.bz), L.ex, R.ex);
  }
  T sa, sb;
  int code = hlle_speeds<DIR>(gam, L, R, sa, sb);
  if (code) return code;
  constexpr int n = DIR, t1 = DIR == 0 ? 1 : 2, t2 = DIR == 0 ? 2 : 0;
  T q_f = hlle_flux(sa, sb, L.jn, R.jn, L.q, R.q);
  T dd_h = hlle_state(sa, sb, L.d, R.d, L.fd, R.fd);
  T tau_h = hlle_state(sa, sb, L.tau, R.tau, L.ftau, R.ftau);
  T s_h[3], s_f[3];
  s_h[0] = hlle_state(sa, sb, L.sx, R.sx, L.fs1, R.fs1);
  s_h[1] = hlle_state(sa, sb, L.sy, R.sy, L.fs2, R.fs2);
  s_h[2] = hlle_state(sa, sb, L.sz, R.sz, L.fs3, R.fs3);
  T dd_f = hlle_flux(sa, sb, L.fd, R.fd, L.d, R.d);
  T tau_f = hlle_flux(sa, sb, L.ftau, R.ftau, L.tau, R.tau);
  s_f[0] = hlle_flux(sa, sb, L.fs1, R.fs1, L.sx, R.sx);
  s_f[1] = hlle_flux(sa, sb, L.fs2, R.fs2, L.sy, R.sy);
  s_f[2] = hlle_flux(sa, sb, L.fs3, R.fs3, L.sz, R.sz);
  T er[3];
  er[0] = b[2] * e[1] - b[1] * e[2];
  er[1] = b[0] * e[2] - b[2] * e[0];
  er[2] = b[1] * e[0] - b[0] * e[1];
  T e2b2s = T(0.5) * (((sq(e[0]) + sq(e[1])) + sq(e[2])) + ((sq(b[0]) + sq(b[1])) + sq(b[2])));
  constexpr int p1 = DIR == 0 ? 1 : 0;  // the transverse pair is (y,z) or (x,z)
  T ca = tau_f - er[n];
  T cb = (((sq(e[p1]) + sq(e[2])) + (sq(b[p1]) + sq(b[2]))) - tau_h) - s_f[n];
  T cc = s_h[n] - er[n];
  int fail = 0;
  T v;
  {
    T det = sq(cb) - (T(4) * ca) * cc;
    T qr = -(T(0.5) * (cb + csign(cb) * fab(face_sqrt(det))));
    T rt2 = quot(cc, qr);
    if (det < T(0)) rt2 = quot(cc, cb);
    T v0 = seed_ref ? vref : clampf(rt2, T(-1), T(1));
    v = hllc_vel(ca, -cb, cc, rt2, v0, fail, vlast);
    if (fab(v) <= T(1.0e-9))
      v = T(0);
    else if (fab(v) > T(1))
      v = clampf(v, T(-1) + T(1.0e-6), T(1) - T(1.0e-6));
  }
  if (fail) return 7;  // a NaN v* goes on to the branches on v*, which stop (8)
  const Quot<T> qR(sb - v), qL(v - sa), qV(v);
  T pstr = (s_f[n] + (sq(e[n]) + sq(b[n]))) - (tau_f - er[n]) * v;
  T sR[3], sL[3];
  sR[n] = qR(sb * s_h[n] - tau_f * v);
  sL[n] = qL(-(sa * s_h[n]) + tau_f * v);
  sR[t1] = qR((((sb * s_h[t1] - s_f[t1]) - er[t1] * v) - e[n] * e[t1]) - b[n] * b[t1]);
  sL[t1] = qL((((-(sa * s_h[t1]) + s_f[t1]) + er[t1] * v) + e[n] * e[t1]) + b[n] * b[t1]);
  sR[t2] = qR((((sb * s_h[t2] - s_f[t2]) - er[t2] * v) - e[n] * e[t2]) - b[n] * b[t2]);
  sL[t2] = qL((((-(sa * s_h[t2]) + s_f[t2]) + er[t2] * v) + e[n] * e[t2]) + b[n] * b[t2]);
  T tauR = (qV(sR[n] - er[n]) - pstr) + e2b2s;
  T tauL = (qV(sL[n] - er[n]) - pstr) + e2b2s;
  if (star_tau_from_root(v)) {  // the same with the wave speeds sa, sb
    const T cv = -(ca * v + cb);
    tauR = (qR(sb * cv - ca) - pstr) + e2b2s;
    tauL = (qL(-(sa * cv) + ca) - pstr) + e2b2s;
  }
  T ddR = qR(sb * dd_h - dd_f);
  T ddL = qL(-(sa * dd_h) + dd_f);
  // flux slots J, FD, Ftau, FSx, FSy, FSz are contiguous in both enums
  T* o = f + (DIR == 0 ? FX_J : FY_J);
  if (sa >= T(0)) {
    o[0] = o[1] = o[2] = o[3] = o[4] = o[5] = T(0);
  } else if (v == T(0)) {  // L branch with the HLL fluxes
    o[0] = q_f; o[1] = dd_f; o[2] = tau_f; o[3] = s_f[0]; o[4] = s_f[1]; o[5] = s_f[2];
  } else {
    const Quot<T> den(sb - sa);
    if (sa <= T(0) && T(0) <= v) {
      T cl = sa * (sb - v);
      o[0] = q_f;
      o[1] = dd_f - den(cl * (ddR - ddL));
      o[2] = tau_f - den(cl * (tauR - tauL));
#pragma unroll
      for (int k = 0; k < 3; k++) o[3 + k] = s_f[k] - den(cl * (sR[k] - sL[k]));
    } else if (v <= T(0) && T(0) <= sb) {
      T cr = sb * (v - sa);
      o[0] = q_f;
      o[1] = dd_f + den(cr * (ddR - ddL));
      o[2] = tau_f + den(cr * (tauR - tauL));
#pragma unroll
      for (int k = 0; k < 3; k++) o[3 + k] = s_f[k] + den(cr * (sR[k] - sL[k]));
    } else if (sb < T(0)) {
      o[0] = o[1] = o[2] = o[3] = o[4] = o[5] = T(0);
    } else {
      return 8;
    }
  }
  return 0;
}

// One face of HLLE with the MC limiter (flux_solver 3, hlle_flow of
// 16_hlle_flow_mcl.f95, DIM = 2), from MC-limited sides with their physical
// fluxes (mcl_sides).  DIR 0: x face (:2374-2388 EM HLL states, :2417-2455
// EM HLL fluxes, :2510-2665 wave speeds and fluid HLL states and fluxes,
// :2924-3056 contact, :3174-3212 star states, :3256-3474 star fluxes,
// :3654-4034 selection); DIR 1: y face (:2395-2409, 2463-2501, 2673-2831,
// 3064-3166, 3178, 3219-3243, 3478-3643, 4044-4340).  Unlike hlle_face:
//  - the normal star momentum carries the star pressure and the star tau
//    is written from the HLL tau state;
//  - fluid star fluxes are F_K + s_K (U*_K - U_K) with the sides' own
//    physical fluxes F_K, which also leave the fan when s1 > 0 or s2 < 0;
//  - the EM star "fluxes" are the EM HLL states (the HLL fluxes when
//    v* = 0);
//  - Newton allows 200 iterations.
// vref: first Newton point when seed_ref; vlast: the solve's last iterate.
template <int DIR, typename T>
__device__ inline int hlle_mcl_face(T gam, Side<T>& L, Side<T>& R, T* f, bool seed_ref, T vref, T& vlast) {
  T e[3], b[3], psi_h, phi_h;
  if constexpr (DIR == 0) {
    psi_h = hll_state(L.psi, R.psi, L.ex, R.ex);
    phi_h = hll_state(L.phi, R.phi, L.bx, R.bx);
    e[0] = hll_state(L.ex, R.ex, L.psi, R.psi);
    e[1] = hll_state(L.ey, R.ey, L.bz, R.bz);
    e[2] = hll_state(L.ez, R.ez, L.bm, R.bm);
    b[0] = hll_state(L.bx, R.bx, L.phi, R.phi);
    b[1] = hll_state(L.by, R.by, L.em, R.em);
    b[2] = hll_state(L.bz, R.bz, L.ey, R.ey);
  } else {
    psi_h = hll_state(L.psi, R.psi, L.ey, R.ey);
    phi_h = hll_state(L.phi, R.phi, L.by, R.by);
    e[1] = hll_state(L.ey, R.ey, L.psi, R.psi);
    e[2] = hll_state(L.ez, R.ez, L.bx, R.bx);
    e[0] = hll_state(L.ex, R.ex, L.bm, R.bm);
    b[1] = hll_state(L.by, R.by, L.phi, R.phi);
    b[2] = hll_state(L.bz, R.bz, L.em, R.em);
    b[0] = hll_state(L.bx, R.bx, L.ez, R.ez);
  }
  // the sound speeds read the sides' enthalpy (:2243-2256)
  L.h = L.rho * (T(1) + L.p / ((gam - T(1)) * L.rho)) + L.p;
  R.h = R.rho * (T(1) + R.p / ((gam - T(1)) * R.rho)) + R.p;
  T sa, sb;
  int code = hlle_speeds<DIR>(gam, L, R, sa, sb);
  if (code) return code;
  constexpr int n = DIR, t1 = DIR == 0 ? 1 : 2, t2 = DIR == 0 ? 2 : 0;
  T q_f = hlle_flux(sa, sb, L.jn, R.jn, L.q, R.q);
  T dd_h = hlle_state(sa, sb, L.d, R.d, L.fd, R.fd);
  T tau_h = hlle_state(sa, sb, L.tau, R.tau, L.ftau, R.ftau);
  T s_h[3], s_f[3];
  s_h[0] = hlle_state(sa, sb, L.sx, R.sx, L.fs1, R.fs1);
  s_h[1] = hlle_state(sa, sb, L.sy, R.sy, L.fs2, R.fs2);
  s_h[2] = hlle_state(sa, sb, L.sz, R.sz, L.fs3, R.fs3);
  T dd_f = hlle_flux(sa, sb, L.fd, R.fd, L.d, R.d);
  T tau_f = hlle_flux(sa, sb, L.ftau, R.ftau, L.tau, R.tau);
  s_f[0] = hlle_flux(sa, sb, L.fs1, R.fs1, L.sx, R.sx);
  s_f[1] = hlle_flux(sa, sb, L.fs2, R.fs2, L.sy, R.sy);
  s_f[2] = hlle_flux(sa, sb, L.fs3, R.fs3, L.sz, R.sz);
  T er[3];
  er[0] = b[2] * e[1] - b[1] * e[2];
  er[1] = b[0] * e[2] - b[2] * e[0];
  er[2] = b[1] * e[0] - b[0] * e[1];
  T e2b2s = T(0.5) * (((sq(e[0]) + sq(e[1])) + sq(e[2])) + ((sq(b[0]) + sq(b[1])) + sq(b[2])));
  constexpr int p1 = DIR == 0 ? 1 : 0;  // the transverse pair is (y,z) or (x,z)
  T ca = tau_f - er[n];
  T cb = (((sq(e[p1]) + sq(e[2])) + (sq(b[p1]) + sq(b[2]))) - tau_h) - s_f[n];
  T cc = s_h[n] - er[n];
  int fail = 0;
  T v;
  {
    T det = sq(cb) - (T(4) * ca) * cc;
    T qr = -(T(0.5) * (cb + csign(cb) * fab(face_sqrt(det))));
    T rt2 = quot(cc, qr);
    if (det < T(0)) rt2 = quot(cc, cb);
    T v0 = seed_ref ? vref : clampf(rt2, T(-1), T(1));
    v = hllc_vel(ca, -cb, cc, rt2, v0, fail, vlast, 200);
    if (fab(v) <= T(1.0e-9))
      v = T(0);
    else if (fab(v) > T(1))
      v = clampf(v, T(-1) + T(1.0e-6), T(1) - T(1.0e-6));
  }
  if (fail) return 7;  // a NaN v* goes on to the branches on v*, which stop (8)
  // EM face fluxes: the HLL states (their HLL fluxes when v* = 0)
  if (v == T(0)) {
    if constexpr (DIR == 0) {
      f[FX_EXPSI] = hll_flux(L.ex, R.ex, L.psi, R.psi);
      f[FX_BXPHI] = hll_flux(L.bx, R.bx, L.phi, R.phi);
      f[FX_PHIBX] = hll_flux(L.phi, R.phi, L.bx, R.bx);
      f[FX_EZBY] = -hll_flux(L.em, R.em, L.by, R.by);
      f[FX_EYBZ] = hll_flux(L.ey, R.ey, L.bz, R.bz);
      f[FX_PSIEX] = hll_flux(L.psi, R.psi, L.ex, R.ex);
      f[FX_BZEY] = hll_flux(L.bz, R.bz, L.ey, R.ey);
      f[FX_BYEZ] = -hll_flux(L.bm, R.bm, L.ez, R.ez);
    } else {
      f[FY_EYPSI] = hll_flux(L.ey, R.ey, L.psi, R.psi);
      f[FY_BYPHI] = hll_flux(L.by, R.by, L.phi, R.phi);
      f[FY_EZBX] = hll_flux(L.ez, R.ez, L.bx, R.bx);
      f[FY_PHIBY] = hll_flux(L.phi, R.phi, L.by, R.by);
      f[FY_EXBZ] = -hll_flux(L.em, R.em, L.bz, R.bz);
      f[FY_BZEX] = -hll_flux(L.bm, R.bm, L.ex, R.ex);
      f[FY_PSIEY] = hll_flux(L.psi, R.psi, L.ey, R.ey);
      f[FY_BXEZ] = hll_flux(L.bx, R.bx, L.ez, R.ez);
    }
  } else if constexpr (DIR == 0) {
    f[FX_EXPSI] = e[0]; f[FX_BXPHI] = b[0]; f[FX_PHIBX] = phi_h; f[FX_EZBY] = e[2];
    f[FX_EYBZ] = e[1]; f[FX_PSIEX] = psi_h; f[FX_BZEY] = b[2]; f[FX_BYEZ] = b[1];
  } else {
    f[FY_EYPSI] = e[1]; f[FY_BYPHI] = b[1]; f[FY_EZBX] = e[2]; f[FY_PHIBY] = phi_h;
    f[FY_EXBZ] = e[0]; f[FY_BZEX] = b[2]; f[FY_PSIEY] = psi_h; f[FY_BXEZ] = b[0];
  }
  // fluid: flux slots J, FD, Ftau, FSx, FSy, FSz are contiguous in both enums
  T* o = f + (DIR == 0 ? FX_J : FY_J);
  bool left;
  if (sa > T(0)) {
    o[0] = L.jn; o[1] = L.fd; o[2] = L.ftau; o[3] = L.fs1; o[4] = L.fs2; o[5] = L.fs3;
    return 0;
  } else if (sa <= T(0) && T(0) <= v) {
    left = true;
  } else if (v < T(0) && T(0) <= sb) {
    left = false;
  } else if (sb < T(0)) {
    o[0] = R.jn; o[1] = R.fd; o[2] = R.ftau; o[3] = R.fs1; o[4] = R.fs2; o[5] = R.fs3;
    return 0;
  } else {
    return 8;
  }
  o[0] = q_f;
  if (v == T(0)) {
    o[1] = dd_f; o[2] = tau_f; o[3] = s_f[0]; o[4] = s_f[1]; o[5] = s_f[2];
    return 0;
  }
  // the star state of the chosen side and its flux F_K + s_K (U*_K - U_K)
  const Side<T>& K = left ? L : R;
  const T sk = left ? sa : sb;
  const T den = left ? v - sa : sb - v;
  T pstr = (s_f[n] + (sq(e[n]) + sq(b[n]))) - (tau_f - er[n]) * v;
  T ss[3];
  if (left) {
    ss[n] = (((((-(sa * s_h[n]) + s_f[n]) - pstr) + sq(b[n])) + sq(e[n])) + er[n] * v) / den;
    ss[t1] = ((((-(sa * s_h[t1]) + s_f[t1]) + er[t1] * v) + e[n] * e[t1]) + b[n] * b[t1]) / den;
    ss[t2] = ((((-(sa * s_h[t2]) + s_f[t2]) + er[t2] * v) + e[n] * e[t2]) + b[n] * b[t2]) / den;
  } else {
    ss[n] = (((((sb * s_h[n] - s_f[n]) + pstr) - sq(b[n])) - sq(e[n])) - er[n] * v) / den;
    ss[t1] = ((((sb * s_h[t1] - s_f[t1]) - er[t1] * v) - e[n] * e[t1]) - b[n] * b[t1]) / den;
    ss[t2] = ((((sb * s_h[t2] - s_f[t2]) - er[t2] * v) - e[n] * e[t2]) - b[n] * b[t2]) / den;
  }
  T taus = left ? ((((-(sa * tau_h)) + tau_f) - (pstr - e2b2s) * v) - er[n]) / den
                : (((sb * tau_h - tau_f) + (pstr - e2b2s) * v) + er[n]) / den;
  T dds = left ? (-(sa * dd_h) + dd_f) / den : (sb * dd_h - dd_f) / den;
  o[1] = K.fd + sk * (dds - K.d);
  o[2] = K.ftau + sk * (taus - K.tau);
  o[3] = K.fs1 + sk * (ss[0] - K.sx);
  o[4] = K.fs2 + sk * (ss[1] - K.sy);
  o[5] = K.fs3 + sk * (ss[2] - K.sz);
  return 0;
}

// ---------------------------------------------------------------------------
// pointwise pieces of the stage (FP64)
// ---------------------------------------------------------------------------

// startconvar, 06_conserved_var.f95:88-133: W, D, tau, S from primitives
struct Cons { double w, d, tau, sx, sy, sz; };
__device__ inline Cons prim_to_cons(double gam, double vx, double vy, double vz, double ex, double ey,
                                    double ez, double bx, double by, double bz, double rho, double p) {
#pragma clang fp contract(off)  // FP64: products and sums rounded separately in every unit
  Cons c;
  double v2 = (sq(vx) + sq(vy)) + sq(vz);
  c.w = 1.0 / sqrt(1.0 - v2);
  double e2 = (sq(ex) + sq(ey)) + sq(ez);
  double b2 = (sq(bx) + sq(by)) + sq(bz);
  double eps = p / ((gam - 1.0) * rho);
  double h = rho * (1.0 + eps) + p;
  c.d = rho * c.w;
  c.tau = (0.5 * (e2 + b2) + h * sq(c.w)) - p;
  c.sx = (bz * ey - by * ez) + (h * sq(c.w)) * vx;
  c.sy = (bx * ez - bz * ex) + (h * sq(c.w)) * vy;
  c.sz = (by * ex - bx * ey) + (h * sq(c.w)) * vz;
  return c;
}

// implicit Ohm's-law step, 11_elecvarint.f95:289-323
__device__ inline void e_solve(double abt, double sig, double dt, double w, double vx, double vy, double vz,
                               double bx, double by, double bz, double eax, double eay, double eaz, double& ex,
                               double& ey, double& ez) {
#pragma clang fp contract(off)  // FP64: products and sums rounded separately in every unit
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
  ex = inv * (((txx * eax + ((aw2 * vx) * vy) * eay) + ((aw2 * vx) * vz) * eaz) -
              kk * ((txx * rx + ((aw2 * vx) * vy) * ry) + ((aw2 * vx) * vz) * rz));
  ey = inv * (((((aw2 * vx) * vy) * eax + tyy * eay) + ((aw2 * vy) * vz) * eaz) -
              kk * ((((aw2 * vx) * vy) * rx + tyy * ry) + ((aw2 * vy) * vz) * rz));
  ez = inv * (((((aw2 * vx) * vz) * eax + ((aw2 * vy) * vz) * eay) + tzz * eaz) -
              kk * ((((aw2 * vx) * vz) * rx + ((aw2 * vy) * vz) * ry) + tzz * rz));
}

// MIRK2 stage E of stages l >= 2 (11_elecvarint.f95:249-265): the linearly
// implicit update from E^n, v^n, B^n, W^n and the stage sum `acc` of the E
// equation; cm1 = cm_1 (parameters.f95:233)
__device__ inline double mirk2_stage_e(double sb, double cm1, double dt, double e, double v, double edv, double r,
                                       double acc) {
#pragma clang fp contract(off)
  return (1.0 / (1.0 + sb * (1.0 - cm1))) * ((((e * (1.0 - sb * cm1)) + (sb * v) * edv) - sb * r) + dt * acc);
}

// MIRK1 E^{n+1} (20_electricfield.f95:36-46): e = E^n, v = v^n, rn / r = the
// components of v^n x B^n and v^n x B^{n+1}, acc = the final sum
__device__ inline double mirk1_final_e(double sb, double cm2, double dt, double e, double v, double edv, double rn,
                                       double r, double acc) {
#pragma clang fp contract(off)
  return (1.0 / (1.0 + sb)) * (((e + (sb * v) * edv) - sb * (cm2 * rn + (1.0 - cm2) * r)) + dt * acc);
}

// MIRK2 E^{n+1} (20_electricfield.f95:117-141): ei = the last stage's E, en =
// E^n, v = the last stage's velocity, r = its v x B, acc = the final sum
__device__ inline double mirk2_final_e(double sdb, double cm1, double cm4, double dt, double ei, double en, double v,
                                       double edv, double r, double acc) {
#pragma clang fp contract(off)
  const double a = -1.0 + sdb * (1.0 - cm1);
  const double den = 1.0 + (0.5 * cm1 - cm4) * sdb;
  return (ei + ((0.5 * a) / den) * (ei - en)) + (0.5 / den) * ((((-(sdb * ei)) + (sdb * edv) * v) + dt * acc) - sdb * r);
}

// sigma W (E + v x B - (E.v) v) (14_flowinit.f95:281-301, 17_fluxandsource.f95:527-537)
__device__ inline void ohm_source(double sig, double w, double vx, double vy, double vz, double ex, double ey,
                                  double ez, double bx, double by, double bz, double& sx, double& sy, double& sz) {
#pragma clang fp contract(off)  // FP64: products and sums rounded separately in every unit
  double edv = (ex * vx + ey * vy) + ez * vz;
  double rx = bz * vy - by * vz, ry = bx * vz - bz * vx, rz = by * vx - bx * vy;
  sx = (sig * w) * ((ex + rx) - edv * vx);
  sy = (sig * w) * ((ey + ry) - edv * vy);
  sz = (sig * w) * ((ez + rz) - edv * vz);
}

// ---------------------------------------------------------------------------
// varprimitivecardano, l_swap = 1 (12_varprimitive.f95:48-813)
// ---------------------------------------------------------------------------
struct C2PIn {
  double d, tau, sx, sy, sz, ex, ey, ez, bx, by, bz, vx0, vy0, vz0, rho_old, p_old;
};
struct C2POut {
  double vx, vy, vz, rho, p;
  int fix_rho, fix_p, fix_v;
  double rec[6];  // the values each fix-up record prints: rho* rho^n, p* p^n, v*^2 v_drift^2
};

// Quartic coefficients in W (12_varprimitive.f95:245-296)
struct Quartic { double e2b2, g1, v2old, b3, b2, b1, b0, d2, d1, d0, c1, c2; };

// eps1: the clamp of a superluminal previous velocity (12_varprimitive.f95:236-241),
// whose square enters only the pressure loop
__device__ inline Quartic c2p_quartic(double gamma, double eps1, const C2PIn& in) {
  #pragma clang fp contract(off)  // FP64 as the reference build, also where the FP32 face unit inlines it
  Quartic qd;
  qd.e2b2 = 0.5 * (((((sq(in.ex) + sq(in.ey)) + sq(in.ez)) + sq(in.bx)) + sq(in.by)) + sq(in.bz));
  qd.g1 = (gamma - 1.0) / gamma;
  qd.v2old = (sq(in.vx0) + sq(in.vy0)) + sq(in.vz0);
  if (qd.v2old > 1.0) qd.v2old = mx(-1.0 + eps1, mn(1.0 - eps1, qd.v2old));
  double C1 = (sq((in.sx - in.ey * in.bz) + in.ez * in.by) + sq((in.sy - in.ez * in.bx) + in.ex * in.bz)) +
              sq((in.sz - in.ex * in.by) + in.ey * in.bx);
  double C2 = in.tau - qd.e2b2;
  qd.c1 = C1;
  qd.c2 = C2;
  double g1 = qd.g1, D = in.d;
  double a4 = C1 - sq(C2);
  double a3 = ((2.0 * g1) * C2) * D;
  double a2 = (sq(C2) - (2.0 * g1) * C1) - sq(g1) * sq(D);
  double a1 = -(((2.0 * g1) * C2) * D);
  double a0 = sq(g1) * (C1 + sq(D));
  qd.b3 = a3 / a4; qd.b2 = a2 / a4; qd.b1 = a1 / a4; qd.b0 = a0 / a4;
  qd.d2 = qd.b2 - (3.0 * sq(qd.b3)) / 8.0;
  qd.d1 = (qd.b1 - (qd.b3 * qd.b2) / 2.0) + pw3(qd.b3) / 8.0;
  qd.d0 = ((qd.b0 - (qd.b3 * qd.b1) / 4.0) + (sq(qd.b3) * qd.b2) / 16.0) - (3.0 * pw4(qd.b3)) / 256.0;
  return qd;
}

// Ferrari-Cardano root (12_varprimitive.f95:297-500); returns 0 or a status.
__device__ inline int c2p_cardano(const Quartic& qd, double& Wn, bool crm) {
  #pragma clang fp contract(off)  // FP64 as the reference build, also where the FP32 face unit inlines it
  const double pi = 3.14159265358979323846264338327950288;
  double d2 = qd.d2, d1 = qd.d1, d0 = qd.d0;
  double e = (20.0 * d2) / 8.0;
  double f = -(d0 - 2.0 * sq(d2));
  double g = -((sq(d1) + (4.0 * d2) * d0) - 4.0 * pw3(d2)) / 8.0;
  double Q1 = (sq(e) - 3.0 * f) / 9.0;
  double R1 = ((2.0 * pw3(e) - (9.0 * e) * f) + 27.0 * g) / 54.0;
  double s, t, det, z1;
  if (sq(R1) < pw3(Q1)) {
    double ac = R1 / fabs(sqrt(pw3(Q1)));
    double T1 = crm ? cr::acos_cr(ac) : acos(ac);
    double sQ = fabs(sqrt(Q1));
    double ta0 = T1 / 3.0, tb0 = (T1 + 2.0 * pi) / 3.0, tc0 = (T1 - 2.0 * pi) / 3.0;
    double ca = crm ? cr::cos_cr(ta0) : cos(ta0);
    double cb = crm ? cr::cos_cr(tb0) : cos(tb0);
    double cc = crm ? cr::cos_cr(tc0) : cos(tc0);
    double za = (-((2.0 * sQ) * ca)) - e / 3.0;
    double sa = d2 + 2.0 * za, ta = ((sq(za) + (2.0 * d2) * za) + sq(d2)) - d0;
    double zb = (-((2.0 * sQ) * cb)) - e / 3.0;
    double sb = d2 + 2.0 * zb, tb = ((sq(zb) + (2.0 * d2) * zb) + sq(d2)) - d0;
    double zc = (-((2.0 * sQ) * cc)) - e / 3.0;
    double sc = d2 + 2.0 * zc, tc = ((sq(zc) + (2.0 * d2) * zc) + sq(d2)) - d0;
    if (sa >= 0.0 && ta >= 0.0) { s = fabs(sqrt(sa)); t = fabs(sqrt(ta)); z1 = za; }
    else if (sb >= 0.0 && tb >= 0.0) { s = fabs(sqrt(sb)); t = fabs(sqrt(tb)); z1 = zb; }
    else if (sc >= 0.0 && tc >= 0.0) { s = fabs(sqrt(sc)); t = fabs(sqrt(tc)); z1 = zc; }
    else return 21;
    det = sq(s) - 4.0 * ((d2 + z1) - t);
    if (det < 0.0) return 22;
  } else {
    double xx1 = fabs(R1) + fabs(sqrt(sq(R1) - pw3(Q1)));
    double rt3 = crm ? cr::pow_third_cr(xx1) : pow(xx1, 1.0 / 3.0);
    double AA = -(copysign(1.0, R1) * copysign(fabs(rt3), xx1));
    double BB = AA != 0.0 ? Q1 / AA : 0.0;
    z1 = (AA + BB) - e / 3.0;
    s = fabs(sqrt(d2 + 2.0 * z1));
    t = fabs(sqrt(((sq(z1) + (2.0 * d2) * z1) + sq(d2)) - d0));
    det = sq(s) - 4.0 * ((d2 + z1) - t);
  }
  double Q2 = -(0.5 * ((-s) + copysign(1.0, -s) * fabs(sqrt(det))));
  double y1 = Q2 - qd.b3 / 4.0;
  double y2 = (((d2 + z1) - t) / Q2) - qd.b3 / 4.0;
  double W = __builtin_nan("");
  if (y1 >= 1.0 && y2 >= 1.0) W = mn(y1, y2);
  if (y1 <= 1.0 && y2 <= 1.0) return 23;
  if (y1 >= 1.0 && y2 <= 1.0) W = y1;
  if (y1 <= 1.0 && y2 >= 1.0) W = y2;
  // No root selected (NaN roots): the reference keeps the previous cell's W
  // (12_varprimitive.f95:463-500 assign nothing).  With a NaN or infinite
  // quartic coefficient every Newton update is NaN whatever the start, so the
  // point goes on to NaN results as there; with finite coefficients the
  // previous cell's W would matter, and the point stops instead (status 24).
  if (W != W && isfinite(qd.b3) && isfinite(qd.b2) && isfinite(qd.b1) && isfinite(qd.b0)) return 24;
  Wn = W;
  return 0;
}

// one Newton update of W on the quartic (12_varprimitive.f95:513-514)
__device__ inline double c2p_newton_w(const Quartic& qd, double Wn) {
  #pragma clang fp contract(off)  // FP64 as the reference build, also where the FP32 face unit inlines it
  return Wn - ((((pw4(Wn) + qd.b3 * pw3(Wn)) + qd.b2 * sq(Wn)) + qd.b1 * Wn) + qd.b0) /
                  (((4.0 * pw3(Wn) + (3.0 * qd.b3) * sq(Wn)) + (2.0 * qd.b2) * Wn) + qd.b1);
}

// quartic residual relative to its leading term, used by the learned guess
__device__ inline double c2p_residual(const Quartic& qd, double W) {
  #pragma clang fp contract(off)  // FP64 as the reference build, also where the FP32 face unit inlines it
  double f = (((pw4(W) + qd.b3 * pw3(W)) + qd.b2 * sq(W)) + qd.b1 * W) + qd.b0;
  return fabs(f) / pw4(W);
}

// Everything after W (12_varprimitive.f95:510-696).  newton: run the
// reference's Newton loop on W first.  Returns 0 or a status.
__device__ inline int c2p_finish(double gamma, const C2PIn& in, const Quartic& qd, double Wn, bool newton,
                                 C2POut& o) {
  #pragma clang fp contract(off)  // FP64 as the reference build, also where the FP32 face unit inlines it
  const double TOL = 1.0e-10;
  const int nrmax = 40;
  if (newton) {
    int nrw = 1;
    double cha1 = 1.0;
    while (cha1 > TOL && nrw < nrmax) {
      double Wr = c2p_newton_w(qd, Wn);
      cha1 = fabs(Wr - Wn) / (0.5 * (Wr + Wn));
      Wn = Wr;
      nrw = nrw + 1;
    }
  }
  double E2B2 = qd.e2b2, g1 = qd.g1, V2c = qd.v2old;
  double D = in.d, tau = in.tau;
  double W = Wn;
  o.fix_rho = o.fix_p = o.fix_v = 0;
  double rho = D / W;
  if (rho < 0.0) {
    o.fix_rho = 1;
    o.rec[0] = rho;
    o.rec[1] = in.rho_old;
    rho = mx(in.rho_old * 1.0e-1, rho);
  }
  double h = ((tau - E2B2) - g1 * rho) / (sq(W) - g1);
  double p = g1 * (h - rho);
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
  if (pn < 0.0) {
    o.fix_p = 1;
    o.rec[2] = pn;
    o.rec[3] = in.p_old;
    pn = mx(in.p_old * 1.0e-4, pn);
  }
  p = pn;
  tau = (E2B2 + h * sq(W)) - p;
  double den = (tau - E2B2) + p;
  double vx = ((in.sx - in.ey * in.bz) + in.ez * in.by) / den;
  double vy = ((in.sy - in.ez * in.bx) + in.ex * in.bz) / den;
  double vz = ((in.sz - in.ex * in.by) + in.ey * in.bx) / den;
  V2c = (sq(vx) + sq(vy)) + sq(vz);
  if (V2c >= 1.0) {
    double Wsl = 1.0 / sqrt(1.0 - V2c);
    double erb = ((((in.ey * in.bz - in.ez * in.by) + in.ez * in.bx) - in.ex * in.bz) + in.ex * in.by) -
                 in.ey * in.bx;
    double b2d = (sq(in.bx) + sq(in.by)) + sq(in.bz);
    if (!(b2d / rho >= 10.0)) return 25;
    double vd = erb / b2d;
    double Wd = 1.0 / sqrt(1.0 - sq(vd));
    rho = D / Wd;
    h = ((tau - E2B2) - g1 * rho) / (sq(Wd) - g1);
    p = g1 * (h - rho);
    tau = (E2B2 + h * sq(Wsl)) - p;
    den = (tau - E2B2) + p;
    vx = ((in.sx - in.ey * in.bz) + in.ez * in.by) / den;
    vy = ((in.sy - in.ez * in.bx) + in.ex * in.bz) / den;
    vz = ((in.sz - in.ex * in.by) + in.ey * in.bx) / den;
    V2c = (sq(vx) + sq(vy)) + sq(vz);
    o.fix_v = 1;
    o.rec[4] = V2c;
    o.rec[5] = sq(vd);
    if (V2c >= 1.0) return 26;
  }
  o.vx = vx; o.vy = vy; o.vz = vz; o.rho = rho; o.p = p;
  return 0;
}

__device__ inline int c2p_cardano_full(double gamma, double eps1, const C2PIn& in, C2POut& o, bool crm) {
  #pragma clang fp contract(off)  // FP64 as the reference build, also where the FP32 face unit inlines it
  Quartic qd = c2p_quartic(gamma, eps1, in);
  double Wn;
  int st = c2p_cardano(qd, Wn, crm);
  if (st) return st;
  return c2p_finish(gamma, in, qd, Wn, true, o);
}

}  // namespace cueva
