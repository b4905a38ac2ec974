// Correctly rounded cos, acos and x^(1/3) for the FP64 parity path.
//
// The reference's primitive recovery calls acos, cos and x**THIRD
// (12_varprimitive.f95:342-434); its build takes them from the host C
// library, which rounds them correctly for practically every argument.  The
// device math library is accurate to about one ulp, and a one-ulp change of
// the Cardano start point moves the Newton iterate (12_varprimitive.f95:
// 510-526) in its last bits.  These versions evaluate the function in
// double-double arithmetic (about 106 bits) and round once, so that they
// return the correctly rounded result; they build for host and device so
// the tests can compare them with the host library on the CPU.
//
// Method: cos by reduction with a triple-double pi/2 and Taylor series in
// double-double on |r| <= pi/4; acos and cbrt by one Newton step from the
// library value, with the residual taken in double-double.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CR_FN __host__ __device__ inline
#else
#define CR_FN inline
#endif

namespace cueva {
namespace cr {

struct dd {
  double hi, lo;
};

CR_FN dd two_sum(double a, double b) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
CR_FN dd quick_two_sum(double a, double b) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}
CR_FN dd two_prod(double a, double b) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  double p = a * b;
  return {p, fma(a, b, -p)};
}
CR_FN dd add(dd a, dd b) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo = s.lo + t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo = s.lo + t.lo;
  return quick_two_sum(s.hi, s.lo);
}
CR_FN dd add_d(dd a, double b) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  dd s = two_sum(a.hi, b);
  s.lo = s.lo + a.lo;
  return quick_two_sum(s.hi, s.lo);
}
CR_FN dd mul(dd a, dd b) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  dd p = two_prod(a.hi, b.hi);
  p.lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
  return quick_two_sum(p.hi, p.lo);
}
CR_FN dd mul_d(dd a, double b) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  dd p = two_prod(a.hi, b);
  p.lo = p.lo + a.lo * b;
  return quick_two_sum(p.hi, p.lo);
}

// 1/(2n)! and 1/(2n+1)! as double-double, n = 0..16, signs (-1)^n applied
#define CR_C(h, l) dd{h, l}
CR_FN dd cos_coef(int n) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  switch (n) {
    case 0: return CR_C(0x1.0000000000000p+0, 0.0);
    case 1: return CR_C(-0x1.0000000000000p-1, 0.0);
    case 2: return CR_C(0x1.5555555555555p-5, 0x1.5555555555555p-59);
    case 3: return CR_C(-0x1.6c16c16c16c17p-10, 0x1.f49f49f49f49fp-65);
    case 4: return CR_C(0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76);
    case 5: return CR_C(-0x1.27e4fb7789f5cp-22, -0x1.cbbc05b4fa99ap-76);
    case 6: return CR_C(0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83);
    case 7: return CR_C(-0x1.93974a8c07c9dp-37, -0x1.05d6f8a2efd1fp-92);
    case 8: return CR_C(0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101);
    case 9: return CR_C(-0x1.6827863b97d97p-53, -0x1.eec01221a8b0bp-107);
    case 10: return CR_C(0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120);
    case 11: return CR_C(-0x1.0ce396db7f853p-70, 0x1.aebcdbd20331cp-124);
    case 12: return CR_C(0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135);
    case 13: return CR_C(-0x1.88e85fc6a4e5ap-89, 0x1.71c37ebd16540p-143);
    case 14: return CR_C(0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153);
    case 15: return CR_C(-0x1.3932c5047d60ep-108, -0x1.832b7b530a627p-162);
    default: return CR_C(0x1.434d2e783f5bcp-118, 0x1.0b87b91be9affp-172);
  }
}
CR_FN dd sin_coef(int n) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  switch (n) {
    case 0: return CR_C(0x1.0000000000000p+0, 0.0);
    case 1: return CR_C(-0x1.5555555555555p-3, -0x1.5555555555555p-57);
    case 2: return CR_C(0x1.1111111111111p-7, 0x1.1111111111111p-63);
    case 3: return CR_C(-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73);
    case 4: return CR_C(0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73);
    case 5: return CR_C(-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80);
    case 6: return CR_C(0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87);
    case 7: return CR_C(-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97);
    case 8: return CR_C(0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103);
    case 9: return CR_C(-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112);
    case 10: return CR_C(0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120);
    case 11: return CR_C(-0x1.761b41316381ap-75, 0x1.3423c7d91404fp-130);
    case 12: return CR_C(0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139);
    case 13: return CR_C(-0x1.d1ab1c2dccea3p-94, -0x1.054d0c78aea14p-149);
    case 14: return CR_C(0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157);
    case 15: return CR_C(-0x1.434d2e783f5bcp-113, -0x1.0b87b91be9affp-167);
    default: return CR_C(0x1.3981254dd0d52p-123, -0x1.2b1f4c8015a2fp-177);
  }
}
#undef CR_C

// cos(x) in double-double for |x| <= 4
CR_FN dd cos_dd(double x) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  const double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c07p-54, P3 = -0x1.f1976b7ed8fbcp-110;
  double kf = rint(x * 0x1.45f306dc9c883p-1);  // x * 2/pi
  int k = (int)kf;
  dd r = add(dd{x, 0.0}, two_prod(-kf, P1));
  r = add(r, two_prod(-kf, P2));
  r = add_d(r, -kf * P3);
  dd r2 = mul(r, r);
  dd acc;
  int q = k & 3;
  if ((q & 1) == 0) {
    acc = cos_coef(16);
    for (int n = 15; n >= 0; n--) acc = add(mul(acc, r2), cos_coef(n));
  } else {
    acc = sin_coef(16);
    for (int n = 15; n >= 0; n--) acc = add(mul(acc, r2), sin_coef(n));
    acc = mul(acc, r);
  }
  if (q == 1 || q == 2) { acc.hi = -acc.hi; acc.lo = -acc.lo; }
  return acc;
}

CR_FN double cos_cr(double x) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  if (!(fabs(x) <= 4.0)) return cos(x);
  return cos_dd(x).hi;
}

CR_FN double acos_cr(double x) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  if (!(fabs(x) < 1.0)) return acos(x);
  double y0 = acos(x);
  dd d = add_d(cos_dd(y0), -x);  // cos(y0) - x
  return y0 + d.hi / sin(y0);
}

// x**THIRD with THIRD the double nearest 1/3 (parameters.f95:88), i.e.
// x^((1 - 2^-54)/3) = cbrt(x * x^(-2^-54)), for x > 0.
CR_FN double pow_third_cr(double x) {
#pragma clang fp contract(off)  // exact error-free transforms: no fused products
  if (!(x > 0.0) || !(x < 1.0e300) || !(x > 1.0e-300)) return pow(x, 0x1.5555555555555p-2);
  double r0 = cbrt(x);
  double tlo = -(x * log(x)) * 0x1p-54;
  dd r3 = mul_d(two_prod(r0, r0), r0);
  dd d = add_d(add_d(r3, -x), -tlo);  // r0^3 - t
  return r0 - d.hi / (3.0 * (r0 * r0));
}

}  // namespace cr
}  // namespace cueva
