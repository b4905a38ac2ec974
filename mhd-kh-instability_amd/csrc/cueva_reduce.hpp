// Global diagnostics and the signal-speed maximum.
//
// cueva_reduce_globals: the 37 volume integrals and maxima that the
// reference writes every hglb steps (22_writer.f95:818-991), in the order of
// its output units 80..217.  Each cell term is rounded as the reference forms
// it (x * Delx * Dely = (x*Delx)*Dely, including the two terms where only p
// carries the cell area).  The sums are taken per thread over a fixed
// grid-stride set of cells, then over wavefronts with shuffles and over the
// block through LDS, and the per-block partials are added by one block in a
// fixed order: the result is deterministic run to run, and differs from the
// reference's sequential i-outer/j-inner sum only by summation order.
//
// cueva_max_speed: the largest relativistic acoustic signal speed |lambda|
// over the interior cells and both directions (Del Zanna-Bucciantini form
// with v_n^2 under the root).  The augmented resistive system also carries
// light-speed waves, so the CFL step of parameters.f95:312 (light speed 1)
// bounds it; the value lets a caller check that bound or pick an adaptive dt.
#pragma once

#include "cueva_kernels.hpp"

namespace cueva {

enum GlobalIdx {
  G_PSI, G_PHI, G_EX2, G_EY2, G_EZ2, G_EXEY, G_EXEZ, G_EYEZ, G_BX2, G_BY2, G_BZ2, G_BXBY, G_BXBZ, G_BYBZ,
  G_V2, G_B2, G_VX2, G_VY2, G_VZ2, G_VXVY, G_VXVZ, G_VYVZ, G_Q, G_P, G_RHO, G_EKIN_C, G_EKIN_R, G_ETOT,
  G_EFLUID_R, G_EMAG, G_EINT_C, G_EINT_R, G_LORENTZ, G_D, G_D_MED, G_NSUM,
  G_MAX_BX = G_NSUM, G_MAX_PHI, G_N
};

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
constexpr double REF_B0 = 1.0e-1;  // parameters.f95:131

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// partial[b * G_N + k]: block b's sum (or max) of quantity k
__global__ void __launch_bounds__(RED_THREADS) k_globals_partial(Geom g, Planes P, Scheme S, double* partial) {
  double acc[G_N];
#pragma unroll
  for (int k = 0; k < G_NSUM; k++) acc[k] = 0.0;
  acc[G_MAX_BX] = -INFINITY;
  acc[G_MAX_PHI] = -INFINITY;
  const int ni = g.imax + 1;
  const int64_t ncell = (int64_t)ni * (g.jmax + 1);
  const int imed1 = (int)floor(0.25 * g.imax), imed2 = (int)floor(0.75 * g.imax);
  const double a = S.dx, b = S.dy, gm1 = S.gamma - 1.0;
  for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < ncell; n += (int64_t)RED_BLOCKS * RED_THREADS) {
    const int j = (int)(n / ni), i = (int)(n - (int64_t)j * ni);
    const int64_t k = g.at(i, j);
    const double psi = P.p[P_PSI][k], phi = P.p[P_PHI][k];
    const double ex = P.p[P_EX][k], ey = P.p[P_EY][k], ez = P.p[P_EZ][k];
    const double bx = P.p[P_BX][k], by = P.p[P_BY][k], bz = P.p[P_BZ][k];
    const double vx = P.p[P_VX][k], vy = P.p[P_VY][k], vz = P.p[P_VZ][k];
    const double q = P.p[P_Q][k], p = P.p[P_P][k], rho = P.p[P_RHO][k];
    const double d = P.c[C_D][k], tau = P.c[C_TAU][k];
    const double v2 = (sq(vx) + sq(vy)) + sq(vz);
    const double eps = p / (gm1 * rho);
    const double h = rho * (1.0 + eps) + p;
    const double w = 1.0 / sqrt(1.0 - v2);
    const double e2 = (sq(ex) + sq(ey)) + sq(ez);
    const double b2 = sq(by) + sq(bz);  // B2 of 22_writer.f95:890 leaves out Bx
    acc[G_PSI] += (psi * a) * b;
    acc[G_PHI] += (phi * a) * b;
    acc[G_EX2] += (sq(ex) * a) * b;
    acc[G_EY2] += (sq(ey) * a) * b;
    acc[G_EZ2] += (sq(ez) * a) * b;
    acc[G_EXEY] += ((ex * ey) * a) * b;
    acc[G_EXEZ] += ((ex * ez) * a) * b;
    acc[G_EYEZ] += ((ey * ez) * a) * b;
    acc[G_BX2] += (sq(bx) * a) * b;
    acc[G_BY2] += (sq(by) * a) * b;
    acc[G_BZ2] += (sq(bz) * a) * b;
    acc[G_BXBY] += ((bx * by) * a) * b;
    acc[G_BXBZ] += ((bx * bz) * a) * b;
    acc[G_BYBZ] += ((by * bz) * a) * b;
    acc[G_V2] += (v2 * a) * b;
    acc[G_B2] += (((sq(bx) + sq(by)) + sq(bz)) * a) * b;
    acc[G_VX2] += (sq(vx) * a) * b;
    acc[G_VY2] += (sq(vy) * a) * b;
    acc[G_VZ2] += (sq(vz) * a) * b;
    acc[G_VXVY] += ((vx * vy) * a) * b;
    acc[G_VXVZ] += ((vx * vz) * a) * b;
    acc[G_VYVZ] += ((vy * vz) * a) * b;
    acc[G_Q] += (q * a) * b;
    acc[G_P] += (p * a) * b;
    acc[G_RHO] += (rho * a) * b;
    acc[G_EKIN_C] += (((0.5 * rho) * v2) * a) * b;
    acc[G_EKIN_R] += ((d * (w - 1.0)) * a) * b;
    acc[G_ETOT] += (tau * a) * b;
    acc[G_EFLUID_R] += h * sq(w) - (p * a) * b;  // as written: only p carries the area
    acc[G_EMAG] += ((0.5 * (e2 + b2)) * a) * b;
    acc[G_EINT_C] += ((rho * eps) * a) * b;
    acc[G_EINT_R] += sq(w) * (rho * eps + p) - (p * a) * b;  // as written
    acc[G_LORENTZ] += (w * a) * b;
    acc[G_D] += (d * a) * b;
    if (i >= imed1 && i <= imed2) acc[G_D_MED] += (d * a) * b;
    const double mbx = bx / REF_B0;
    acc[G_MAX_BX] = mbx > acc[G_MAX_BX] ? mbx : acc[G_MAX_BX];
    acc[G_MAX_PHI] = phi > acc[G_MAX_PHI] ? phi : acc[G_MAX_PHI];
  }
  __shared__ double red[RED_THREADS / 64][G_N];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < G_N; k++) {
    double v = k < G_NSUM ? wave_sum(acc[k]) : wave_max(acc[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < G_N) {
    const int k = threadIdx.x;
    double v = red[0][k];
    for (int t = 1; t < RED_THREADS / 64; t++) v = k < G_NSUM ? v + red[t][k] : (red[t][k] > v ? red[t][k] : v);
    partial[(int64_t)blockIdx.x * G_N + k] = v;
  }
}

// one block: out[k] = fixed-order combination of the RED_BLOCKS partials
__global__ void __launch_bounds__(RED_THREADS) k_globals_final(const double* partial, double* out) {
  __shared__ double red[RED_THREADS / 64][G_N];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < G_N; k++) {
    const bool is_sum = k < G_NSUM;
    double v = is_sum ? 0.0 : -INFINITY;
    for (int bI = threadIdx.x; bI < RED_BLOCKS; bI += RED_THREADS) {
      double x = partial[(int64_t)bI * G_N + k];
      v = is_sum ? v + x : (x > v ? x : v);
    }
    v = is_sum ? wave_sum(v) : wave_max(v);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < G_N) {
    const int k = threadIdx.x;
    double v = red[0][k];
    for (int t = 1; t < RED_THREADS / 64; t++) v = k < G_NSUM ? v + red[t][k] : (red[t][k] > v ? red[t][k] : v);
    out[k] = v;
  }
}

// relativistic acoustic wave speeds of one cell along a direction with
// normal speed vn (Del Zanna & Bucciantini 2002, eq. 35 for the fluid part)
__device__ inline double acoustic_speed(double gam, double rho, double p, double v2, double vn) {
  const double h = rho + (gam / (gam - 1.0)) * p;  // enthalpy density
  const double cs2 = (gam * p) / h;
  const double root = sqrt(cs2 * (1.0 - v2) * ((1.0 - v2 * cs2) - sq(vn) * (1.0 - cs2)));
  const double den = 1.0 - v2 * cs2;
  const double lp = (vn * (1.0 - cs2) + root) / den;
  const double lm = (vn * (1.0 - cs2) - root) / den;
  const double ap = fabs(lp), am = fabs(lm);
  return ap > am ? ap : am;
}

__global__ void __launch_bounds__(RED_THREADS) k_speed_partial(Geom g, Planes P, Scheme S, double* partial) {
  double m = 0.0;
  const int ni = g.imax + 1;
  const int64_t ncell = (int64_t)ni * (g.jmax + 1);
  for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < ncell; n += (int64_t)RED_BLOCKS * RED_THREADS) {
    const int j = (int)(n / ni), i = (int)(n - (int64_t)j * ni);
    const int64_t k = g.at(i, j);
    const double vx = P.p[P_VX][k], vy = P.p[P_VY][k], vz = P.p[P_VZ][k];
    const double rho = P.p[P_RHO][k], p = P.p[P_P][k];
    const double v2 = (sq(vx) + sq(vy)) + sq(vz);
    const double sx = acoustic_speed(S.gamma, rho, p, v2, vx);
    const double sy = acoustic_speed(S.gamma, rho, p, v2, vy);
    const double s = sx > sy ? sx : sy;
    m = s > m ? s : m;
  }
  __shared__ double red[RED_THREADS / 64];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = red[0];
    for (int t = 1; t < RED_THREADS / 64; t++) v = red[t] > v ? red[t] : v;
    partial[blockIdx.x] = v;
  }
}

__global__ void __launch_bounds__(RED_THREADS) k_speed_final(const double* partial, double* out) {
  double m = 0.0;
  for (int b = threadIdx.x; b < RED_BLOCKS; b += RED_THREADS) {
    double x = partial[b];
    m = x > m ? x : m;
  }
  __shared__ double red[RED_THREADS / 64];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = red[0];
    for (int t = 1; t < RED_THREADS / 64; t++) v = red[t] > v ? red[t] : v;
    out[0] = v;
  }
}

// minimum and maximum of one plane over the interior: partial[2b], [2b+1]
__global__ void __launch_bounds__(RED_THREADS) k_range_partial(Geom g, const double* __restrict__ a,
                                                              double* partial) {
  double lo = INFINITY, hi = -INFINITY;
  const int ni = g.imax + 1;
  const int64_t ncell = (int64_t)ni * (g.jmax + 1);
  for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < ncell; n += (int64_t)RED_BLOCKS * RED_THREADS) {
    const int j = (int)(n / ni), i = (int)(n - (int64_t)j * ni);
    const double x = a[g.at(i, j)];
    lo = x < lo ? x : lo;
    hi = x > hi ? x : hi;
  }
  __shared__ double rlo[RED_THREADS / 64], rhi[RED_THREADS / 64];
  lo = -wave_max(-lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0) { rlo[threadIdx.x >> 6] = lo; rhi[threadIdx.x >> 6] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < RED_THREADS / 64; t++) {
      lo = rlo[t] < lo ? rlo[t] : lo;
      hi = rhi[t] > hi ? rhi[t] : hi;
    }
    lo = rlo[0] < lo ? rlo[0] : lo;
    hi = rhi[0] > hi ? rhi[0] : hi;
    partial[2 * blockIdx.x] = lo;
    partial[2 * blockIdx.x + 1] = hi;
  }
}

__global__ void __launch_bounds__(RED_THREADS) k_range_final(const double* partial, double* out) {
  double lo = INFINITY, hi = -INFINITY;
  for (int b = threadIdx.x; b < RED_BLOCKS; b += RED_THREADS) {
    lo = partial[2 * b] < lo ? partial[2 * b] : lo;
    hi = partial[2 * b + 1] > hi ? partial[2 * b + 1] : hi;
  }
  __shared__ double rlo[RED_THREADS / 64], rhi[RED_THREADS / 64];
  lo = -wave_max(-lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0) { rlo[threadIdx.x >> 6] = lo; rhi[threadIdx.x >> 6] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    lo = rlo[0];
    hi = rhi[0];
    for (int t = 1; t < RED_THREADS / 64; t++) {
      lo = rlo[t] < lo ? rlo[t] : lo;
      hi = rhi[t] > hi ? rhi[t] : hi;
    }
    out[0] = lo;
    out[1] = hi;
  }
}

}  // namespace cueva
