// Seeded Kelvin-Helmholtz initial data (host).  The reference has no such
// setup (its KH tests 34/35 use another profile, 03_init_var_primitive.f95:
// 2297-2387); this one is the relativistic shear layer of the benchmark
// configurations.  Node convention x_i = -Lx/2 + i*dx as in
// 03_init_var_primitive.f95:2304-2305.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "cueva.h"

namespace {

// SplitMix64 finaliser (Steele, Lea & Flood 2014) of a counter built from
// (seed, i, j): the noise of a cell does not depend on the grid traversal.
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline double noise(uint64_t seed, int i, int j) {
  uint64_t key = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)(uint32_t)i * 0xD1B54A32D192ED03ULL +
                 (uint64_t)(uint32_t)j * 0x8CB92BA72F3D8DD7ULL;
  double u = (double)(mix64(key) >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
  return (2.0 * u - 1.0) * 1.0e-4;
}

inline int src_index(int i, int n, int periodic) {
  if (i >= 0 && i <= n) return i;
  if (periodic) return i < 0 ? n + i : i - n;
  return i < 0 ? 0 : n;
}

}  // namespace

extern "C" int cueva_kh_init(int imax, int jmax, double lx, double ly, int rotated, double v0, double bx0,
                             double bz0, uint64_t seed, double* prims14) {
  if (imax < 8 || jmax < 8 || !prims14 || !(lx > 0.0) || !(ly > 0.0)) return CUEVA_E_ARG;
  const int nxp = imax + 13, nyp = jmax + 13;
  const size_t plane = (size_t)nxp * (size_t)nyp;
  memset(prims14, 0, 14 * plane * sizeof(double));
  // native (unrotated) problem: periodic x with ni cells, outflow y with nj
  const int ni = rotated ? jmax : imax, nj = rotated ? imax : jmax;
  const double Lx = rotated ? ly : lx, Ly = rotated ? lx : ly;
  const double dx = Lx / (double)(float)ni, dy = Ly / (double)(float)nj;
  const double pi = 3.14159265358979323846264338327950288;
  for (int jn = 0; jn <= nj; jn++)
    for (int in = 0; in <= ni; in++) {
      const int ic = in == ni ? 0 : in;  // periodic duplicate column equals column 0
      const double x = -0.5 * Lx + ic * dx, y = -0.5 * Ly + jn * dy;
      double vx = v0 * tanh(y / 0.01) + noise(seed, ic, jn);
      double vy = 0.01 * sin(2.0 * pi * x / Lx) * exp(-(y * y) / 0.01);
      double vz = 0.0;
      double bx = bx0, by = 0.0, bz = bz0;
      int io = in, jo = jn;
      if (rotated) {  // x' = y, y' = -x: (i', j') = (jn, ni - in); v' = (vy, -vx, vz)
        io = jn;
        jo = ni - in;
        double t = vx; vx = vy; vy = -t;
        t = bx; bx = by; by = -t;
      }
      // E = -v x B
      double ex = -(vy * bz - vz * by), ey = -(vz * bx - vx * bz), ez = -(vx * by - vy * bx);
      const double val[14] = {0.0, 0.0, bx, by, bz, vx, vy, vz, ex, ey, ez, 0.0, 1.0, 1.0};
      const size_t k = (size_t)(jo + 6) * nxp + (size_t)(io + 6);
      for (int f = 0; f < 14; f++) prims14[f * plane + k] = val[f];
    }
  // ghosts by the boundary index maps (periodic along the shear, copy across)
  const int px = rotated ? 0 : 1, py = rotated ? 1 : 0;
  for (int f = 0; f < 14; f++) {
    double* a = prims14 + f * plane;
    for (int j = -6; j <= jmax + 6; j++)
      for (int i = -6; i <= imax + 6; i++) {
        if (i >= 0 && i <= imax && j >= 0 && j <= jmax) continue;
        int si = src_index(i, imax, px), sj = src_index(j, jmax, py);
        a[(size_t)(j + 6) * nxp + (i + 6)] = a[(size_t)(sj + 6) * nxp + (si + 6)];
      }
  }
  return CUEVA_OK;
}

// Circularly polarised Alfven wave along x, uniform in y, periodic box
// (03_init_var_primitive.f95:129-160, DIM = 1, TEST = 2, extended to 2D):
// B = B0 (1, eta cos kx, eta sin kx), v = -eta VA (0, cos kx, sin kx),
// E = -v x B = VA eta B0 (0, sin kx, -cos kx), k = 2 pi / Lx, with the
// reference's exact Alfven speed VA; rho, p uniform.
extern "C" int cueva_cpaw_init(int imax, int jmax, double lx, double gamma, double rho0, double p0, double b0,
                               double eta, double* prims14, double* va) {
  if (imax < 8 || jmax < 8 || !prims14 || !(lx > 0.0) || !(gamma > 1.0)) return CUEVA_E_ARG;
  const int nxp = imax + 13, nyp = jmax + 13;
  const size_t plane = (size_t)nxp * (size_t)nyp;
  const double pi = 3.14159265358979323846264338327950288;
  const double eps = p0 / ((gamma - 1.0) * rho0);
  const double h = rho0 * (1.0 + eps) + p0;
  const double a = h + b0 * b0 * (1.0 + eta * eta);
  const double r = 2.0 * eta * b0 * b0 / a;
  const double VA = sqrt(2.0 * b0 * b0 / (a * (1.0 + sqrt(1.0 - r * r))));
  const double dx = lx / (double)imax, kx = 2.0 * pi / lx;
  for (int j = -6; j <= jmax + 6; j++)
    for (int i = -6; i <= imax + 6; i++) {
      // periodic in x: ghost i takes the value of i mod imax (cell imax == cell 0)
      int ic = ((i % imax) + imax) % imax;
      const double x = -0.5 * lx + ic * dx;
      const double c = cos(kx * x), s = sin(kx * x);
      const double val[14] = {0.0, 0.0, b0, eta * b0 * c, eta * b0 * s, 0.0, -eta * VA * c, -eta * VA * s,
                              0.0, VA * eta * b0 * s, -VA * eta * b0 * c, 0.0, rho0, p0};
      const size_t k = (size_t)(j + 6) * nxp + (size_t)(i + 6);
      for (int f = 0; f < 14; f++) prims14[f * plane + k] = val[f];
    }
  if (va) *va = VA;
  return CUEVA_OK;
}
