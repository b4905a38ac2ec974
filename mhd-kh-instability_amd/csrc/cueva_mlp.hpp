// Learned start point of the primitive recovery: a 3 -> 32 -> 32 -> 2 MLP
// (tanh hidden layers) evaluated on the matrix cores, BF16 operands and FP32
// accumulation, one wavefront = 64 cells.
//
// The layers are computed transposed, H^T = W X^T, with
// v_mfma_f32_16x16x16_bf16: the weights are the A operand (16 output units by
// 16 inputs), the activations of 16 cells the B operand.  In that form the
// accumulator of one layer (lane l holds rows 4*(l/16)..+3 of cell l%16) is
// exactly the B operand of the next layer's K block, so activations never
// leave registers.  A wavefront handles its 64 cells as four tiles of 16;
// lane l's cell features reach the tile's B operand by one shuffle, and the
// outputs go back by another.
//
// Packed parameters (FP32, CUEVA_MLP_NPACK floats):
//   W1[32][3] b1[32] W2[32][32] b2[32] W3[2][32] b3[2]   (trainable, 1250)
//   in_mean[3] in_std[3] out_mean[2] out_std[2]           (normalisation)
// Features of a conserved state (D, S - E x B, tau - (E^2+B^2)/2 = C2):
//   x = (log(D/C2), |S - E x B|/C2, log C2); targets y = (log W, log(p/C2)).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cueva_device.hpp"

namespace cueva {
namespace mlp {

constexpr int NIN = 3, NH = 32, NOUT = 2;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NOUT * NH, NPARAM = OFF_B3 + NOUT;
constexpr int OFF_IN_MEAN = NPARAM, OFF_IN_STD = OFF_IN_MEAN + NIN, OFF_OUT_MEAN = OFF_IN_STD + NIN,
              OFF_OUT_STD = OFF_OUT_MEAN + NOUT, NPACK = OFF_OUT_STD + NOUT;
static_assert(NPARAM == 1250 && NPACK == 1260, "packed layout");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (NaN kept quiet)
__host__ __device__ inline uint16_t bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_round(float f) {
  return __builtin_bit_cast(float, (uint32_t)bf16_bits(f) << 16);
}
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// four floats -> one BF16 operand, nearest even: two v_cvt_pk_bf16_f32
__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
  const f32x2 lo = {a, b}, hi = {c, d};
  const bf16x2 l = __builtin_convertvector(lo, bf16x2), h = __builtin_convertvector(hi, bf16x2);
  const uint32_t lu = __builtin_bit_cast(uint32_t, l), hu = __builtin_bit_cast(uint32_t, h);
  bf16x4 r;
  r[0] = (short)(lu & 0xffffu);
  r[1] = (short)(lu >> 16);
  r[2] = (short)(hu & 0xffffu);
  r[3] = (short)(hu >> 16);
  return r;
}

__device__ inline float tanh_f(float x) {
  // tanh(x) = 1 - 2/(exp(2x)+1) with the hardware exp2 and reciprocal (about
  // 1 ulp each, far below the BF16 rounding the activation goes through);
  // saturates cleanly for large |x|
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);  // 2/ln 2
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Weight operands of one wavefront, loaded once.
struct Operands {
  bf16x4 a1[2];     // W1 rows 16m.., inputs 0..3 (3 real)
  bf16x4 a2[2][2];  // W2 rows 16m.., inputs 16s..
  bf16x4 a3[2];     // W3 (rows 0,1 real), inputs 16s..
  float b1[2][4], b2[2][4], b3[2];
};

__device__ inline Operands load_operands(const float* __restrict__ p, int lane) {
  const int r = lane & 15, kb = lane >> 4;
  Operands o;
#pragma unroll
  for (int m = 0; m < 2; m++) {
    const int row = 16 * m + r;
    if (kb == 0)
      o.a1[m] = pack4(p[OFF_W1 + row * NIN + 0], p[OFF_W1 + row * NIN + 1], p[OFF_W1 + row * NIN + 2], 0.0f);
    else
      o.a1[m] = pack4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const float* w = p + OFF_W2 + row * NH + 16 * s + 4 * kb;
      o.a2[m][s] = pack4(w[0], w[1], w[2], w[3]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      o.b1[m][i] = p[OFF_B1 + 16 * m + 4 * kb + i];
      o.b2[m][i] = p[OFF_B2 + 16 * m + 4 * kb + i];
    }
  }
#pragma unroll
  for (int s = 0; s < 2; s++) {
    if (r < NOUT) {
      const float* w = p + OFF_W3 + r * NH + 16 * s + 4 * kb;
      o.a3[s] = pack4(w[0], w[1], w[2], w[3]);
    } else {
      o.a3[s] = pack4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  o.b3[0] = p[OFF_B3 + 0];
  o.b3[1] = p[OFF_B3 + 1];
  return o;
}

// Activations of one tile as the layers leave them (lane: rows 4*(l/16)+i of
// cell l%16 of the tile), FP32 before the BF16 rounding of the next layer.
struct TileAct {
  float h1[2][4], h2[2][4];
  float y[2];  // lanes 0..15 only
};

// forward of tile t (cells of lanes 16t..16t+15); x* are every lane's own
// normalised features
__device__ inline TileAct forward_tile(const Operands& o, int lane, int t, float x0, float x1, float x2) {
  const int r = lane & 15, kb = lane >> 4;
  const int src = 16 * t + r;
  float s0 = __shfl(x0, src, 64), s1 = __shfl(x1, src, 64), s2 = __shfl(x2, src, 64);
  bf16x4 bx = kb == 0 ? pack4(s0, s1, s2, 0.0f) : pack4(0.0f, 0.0f, 0.0f, 0.0f);
  TileAct a;
  bf16x4 bh1[2];
#pragma unroll
  for (int m = 0; m < 2; m++) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma16(o.a1[m], bx, acc);
#pragma unroll
    for (int i = 0; i < 4; i++) a.h1[m][i] = tanh_f(acc[i] + o.b1[m][i]);
    bh1[m] = pack4(a.h1[m][0], a.h1[m][1], a.h1[m][2], a.h1[m][3]);
  }
  bf16x4 bh2[2];
#pragma unroll
  for (int m = 0; m < 2; m++) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma16(o.a2[m][0], bh1[0], acc);
    acc = mfma16(o.a2[m][1], bh1[1], acc);
#pragma unroll
    for (int i = 0; i < 4; i++) a.h2[m][i] = tanh_f(acc[i] + o.b2[m][i]);
    bh2[m] = pack4(a.h2[m][0], a.h2[m][1], a.h2[m][2], a.h2[m][3]);
  }
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = mfma16(o.a3[0], bh2[0], acc);
  acc = mfma16(o.a3[1], bh2[1], acc);
  a.y[0] = acc[0] + o.b3[0];
  a.y[1] = acc[1] + o.b3[1];
  return a;
}

// normalised outputs of every lane's own cell (all 64 lanes must call)
__device__ inline void forward_wave(const Operands& o, int lane, float x0, float x1, float x2, float& y0,
                                    float& y1) {
#pragma unroll
  for (int t = 0; t < 4; t++) {
    TileAct a = forward_tile(o, lane, t, x0, x1, x2);
    // lane r (< 16) holds the outputs of cell 16t + r
    float v0 = __shfl(a.y[0], lane & 15, 64), v1 = __shfl(a.y[1], lane & 15, 64);
    if ((lane >> 4) == t) {
      y0 = v0;
      y1 = v1;
    }
  }
}

}  // namespace mlp

// Normalised network features of a conserved state: (log(D/C2),
// |S - E x B|/C2, log C2), C2 = tau - (E^2+B^2)/2, formed as the quartic of
// 12_varprimitive.f95:245-296 forms C1 and C2.
__device__ inline void c2p_features(double gamma, double eps1, const C2PIn& in, const float* __restrict__ wp, float& f0,
                                    float& f1, float& f2) {
  using namespace mlp;
  const Quartic qd = c2p_quartic(gamma, eps1, in);
  const double C2 = qd.c2, sf = sqrt(qd.c1);
  // the features only need single precision (the network reads them in BF16)
  f0 = (__logf((float)(in.d / C2)) - wp[OFF_IN_MEAN + 0]) / wp[OFF_IN_STD + 0];
  f1 = ((float)(sf / C2) - wp[OFF_IN_MEAN + 1]) / wp[OFF_IN_STD + 1];
  f2 = (__logf((float)C2) - wp[OFF_IN_MEAN + 2]) / wp[OFF_IN_STD + 2];
  if (!(f0 == f0) || !(f1 == f1) || !(f2 == f2)) f0 = f1 = f2 = 0.0f;  // unphysical: the start is rejected later
}

// Primitive recovery of one cell from the network outputs (y0, y1): start W
// -> two FP64 Newton steps on the quartic of 12_varprimitive.f95:245-296 ->
// accepted when the relative quartic residual is <= 1e-10, W >= 1, p > 0,
// |v| < 1 and p is within a factor 2 of the network's own p; otherwise the
// reference's Cardano root (fallback = true).  Then the reference's pressure
// loop, fix-ups and velocities (12_varprimitive.f95:510-696).  Returns 0 or a
// status.
__device__ inline int c2p_learned_finish(double gamma, double eps1, const C2PIn& in, const float* __restrict__ wp,
                                         float y0, float y1, bool crm, C2POut& out, bool& fallback) {
  using namespace mlp;
  const Quartic qd = c2p_quartic(gamma, eps1, in);
  fallback = false;
  const double C2 = qd.c2, sf = sqrt(qd.c1);
  // Start: W from the network's pressure through |S - E x B| = (C2 + p) v,
  // W = 1/sqrt(1 - v^2); its error is v^2 W^2 p/(C2+p) times that of p,
  // an order of magnitude below the error of the network's own W, which is
  // the start only where that v would not be subluminal.
  const float lp_net = y1 * wp[OFF_OUT_STD + 1] + wp[OFF_OUT_MEAN + 1];
  const double pn = (double)__expf(lp_net) * C2;
  const double ve = sf / (C2 + pn);
  double W = ve < 1.0 ? 1.0 / sqrt(1.0 - ve * ve) : (double)__expf(y0 * wp[OFF_OUT_STD + 0] + wp[OFF_OUT_MEAN + 0]);
  W = c2p_newton_w(qd, W);
  W = c2p_newton_w(qd, W);
  bool ok = W >= 1.0 && c2p_residual(qd, W) <= 1.0e-10;
  if (ok) {
    // physical and the root the network meant (h and p as 12_varprimitive.f95:550-553 form them)
    const double rho = in.d / W;
    const double h = (qd.c2 - qd.g1 * rho) / (sq(W) - qd.g1);
    const double pw = qd.g1 * (h - rho);
    ok = pw > 0.0 && qd.c1 < sq(qd.c2 + pw) && fabsf(__logf((float)(pw / qd.c2)) - lp_net) <= 0.6931472f;
  }
  if (ok) return c2p_finish(gamma, in, qd, W, false, out);
  fallback = true;
  double Wc;
  int code = c2p_cardano(qd, Wc, crm);
  if (!code) code = c2p_finish(gamma, in, qd, Wc, true, out);
  return code;
}

// Both together (every lane of the wavefront must call: the network runs
// wavefront-wide).
__device__ inline int c2p_learned_cell(double gamma, double eps1, const C2PIn& in, const float* __restrict__ wp,
                                       const mlp::Operands& o, int lane, bool crm, C2POut& out, bool& fallback) {
  float f0, f1, f2, y0 = 0.0f, y1 = 0.0f;
  c2p_features(gamma, eps1, in, wp, f0, f1, f2);
  mlp::forward_wave(o, lane, f0, f1, f2, y0, y1);
  return c2p_learned_finish(gamma, eps1, in, wp, y0, y1, crm, out, fallback);
}

}  // namespace cueva
