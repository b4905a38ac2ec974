// Host build of the correctly rounded functions used by the device
// primitive recovery, for tests/test_crmath.py.
#include "cueva_crmath.hpp"

extern "C" {
void cr_cos_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = cueva::cr::cos_cr(x[i]); }
void cr_acos_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = cueva::cr::acos_cr(x[i]); }
void cr_pow3_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = cueva::cr::pow_third_cr(x[i]); }
void lib_cos_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = cos(x[i]); }
void lib_acos_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = acos(x[i]); }
void lib_pow3_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = pow(x[i], 0x1.5555555555555p-2); }
}

#include <quadmath.h>
extern "C" {
void q_cos_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = (double)cosq((__float128)x[i]); }
void q_acos_many(const double* x, double* y, long n) { for (long i = 0; i < n; i++) y[i] = (double)acosq((__float128)x[i]); }
void q_pow3_many(const double* x, double* y, long n) {
  for (long i = 0; i < n; i++) y[i] = (double)powq((__float128)x[i], (__float128)0x1.5555555555555p-2);
}
}
