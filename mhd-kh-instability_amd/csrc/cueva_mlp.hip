// Learned primitive recovery: training and inference kernels of the
// 3 -> 32 -> 32 -> 2 start-point MLP (cueva_mlp.hpp) and the primitive
// recovery that uses it.
//
//  - k_mlp_grad: forward on the matrix cores, backward in FP32 through LDS;
//    the gradient of the mean squared error over a sample range, w.r.t. the
//    FP32 master weights (the forward uses their BF16 rounding).
//  - k_mlp_adam: one Adam update of the master weights; clears the gradient.
//  - k_mlp_forward: normalised outputs for normalised inputs (tests).
//  - k_c2p_learned: per cell, the network's start for W, at most two FP64 Newton
//    steps on the quartic of 12_varprimitive.f95:245-296, acceptance when the
//    relative quartic residual is <= 1e-10, W >= 1 and the state it implies is
//    physical and agrees with the network's pressure, otherwise the Cardano
//    path of the reference; then the reference's pressure loop, fix-ups and
//    velocities (12_varprimitive.f95:510-696).  Cardano fallbacks are counted.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cueva.h"
#include "cueva_common.hpp"
#include "cueva_mlp.hpp"
#include "cueva_update.hpp"

namespace cueva {
namespace mlp {

constexpr int GRAD_CELLS = 128;  // samples per block of k_mlp_grad (two wavefronts)
constexpr int LDP = NH + 1;      // padded LDS row

// Gradient of sum over [off, off+n) of |y - t|^2 / n_total.
__global__ void __launch_bounds__(GRAD_CELLS) k_mlp_grad(const float* __restrict__ p, const float* __restrict__ X,
                                                         const float* __restrict__ Y, int64_t off, int64_t n,
                                                         float inv_total, float* __restrict__ grad,
                                                         double* __restrict__ loss) {
  __shared__ float h1s[GRAD_CELLS][LDP];
  __shared__ float h2s[GRAD_CELLS][LDP];  // later dz1
  __shared__ float dz2s[GRAD_CELLS][LDP];
  __shared__ float xs[GRAD_CELLS][4];
  __shared__ float dys[GRAD_CELLS][2];
  __shared__ float w2s[NH][NH];
  __shared__ float w3s[NOUT][NH];
  __shared__ double lsum[GRAD_CELLS / 64];

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * GRAD_CELLS;
  const int64_t gi = off + c0 + tid;
  const bool valid = c0 + tid < n;
  for (int e = tid; e < NH * NH; e += GRAD_CELLS) w2s[e / NH][e % NH] = bf16_round(p[OFF_W2 + e]);
  for (int e = tid; e < NOUT * NH; e += GRAD_CELLS) w3s[e / NH][e % NH] = bf16_round(p[OFF_W3 + e]);

  float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f, t0 = 0.0f, t1 = 0.0f;
  if (valid) {
    x0 = X[gi * 3 + 0]; x1 = X[gi * 3 + 1]; x2 = X[gi * 3 + 2];
    t0 = Y[gi * 2 + 0]; t1 = Y[gi * 2 + 1];
  }
  xs[tid][0] = bf16_round(x0); xs[tid][1] = bf16_round(x1); xs[tid][2] = bf16_round(x2);
  // forward, wavefront-wide; scatter the activations to LDS rows of their cells
  Operands o = load_operands(p, lane);
  const int r = lane & 15, kb = lane >> 4;
  float y0 = 0.0f, y1 = 0.0f;
  for (int t = 0; t < 4; t++) {
    TileAct a = forward_tile(o, lane, t, x0, x1, x2);
    const int cell = wv * 64 + 16 * t + r;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        h1s[cell][16 * m + 4 * kb + i] = a.h1[m][i];
        h2s[cell][16 * m + 4 * kb + i] = a.h2[m][i];
      }
    float v0 = __shfl(a.y[0], r, 64), v1 = __shfl(a.y[1], r, 64);
    if (kb == t) { y0 = v0; y1 = v1; }
  }
  // output error
  double l2 = 0.0;
  if (valid) {
    float e0 = y0 - t0, e1 = y1 - t1;
    dys[tid][0] = 2.0f * e0 * inv_total;
    dys[tid][1] = 2.0f * e1 * inv_total;
    l2 = ((double)e0 * e0 + (double)e1 * e1) * inv_total;
  } else {
    dys[tid][0] = 0.0f;
    dys[tid][1] = 0.0f;
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) l2 += __shfl_down(l2, s, 64);
  if (lane == 0) lsum[wv] = l2;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < GRAD_CELLS / 64; w++) s += lsum[w];
    atomicAdd(loss, s);
  }
  // dz2 = (dy W3) * (1 - h2^2)
  for (int h = 0; h < NH; h++) {
    float g = dys[tid][0] * w3s[0][h] + dys[tid][1] * w3s[1][h];
    float hv = h2s[tid][h];
    dz2s[tid][h] = g * (1.0f - hv * hv);
  }
  __syncthreads();
  // dW3, db3 from dy and bf16(h2)
  if (tid < NOUT * NH) {
    const int oo = tid / NH, h = tid % NH;
    float s = 0.0f;
    for (int c = 0; c < GRAD_CELLS; c++) s += dys[c][oo] * bf16_round(h2s[c][h]);
    atomicAdd(&grad[OFF_W3 + oo * NH + h], s);
  }
  if (tid < NOUT) {
    float s = 0.0f;
    for (int c = 0; c < GRAD_CELLS; c++) s += dys[c][tid];
    atomicAdd(&grad[OFF_B3 + tid], s);
  }
  __syncthreads();  // h2s free from here: it receives dz1
  // dz1 = (dz2 W2) * (1 - h1^2)
  for (int h = 0; h < NH; h++) {
    float g = 0.0f;
    for (int k = 0; k < NH; k++) g += dz2s[tid][k] * w2s[k][h];
    float hv = h1s[tid][h];
    h2s[tid][h] = g * (1.0f - hv * hv);
  }
  __syncthreads();
  // dW2 (8 entries per thread), db2, dW1, db1
  for (int e = tid; e < NH * NH; e += GRAD_CELLS) {
    const int ho = e / NH, hi = e % NH;
    float s = 0.0f;
    for (int c = 0; c < GRAD_CELLS; c++) s += dz2s[c][ho] * bf16_round(h1s[c][hi]);
    atomicAdd(&grad[OFF_W2 + e], s);
  }
  if (tid < NH) {
    float s2 = 0.0f, s1 = 0.0f;
    for (int c = 0; c < GRAD_CELLS; c++) {
      s2 += dz2s[c][tid];
      s1 += h2s[c][tid];
    }
    atomicAdd(&grad[OFF_B2 + tid], s2);
    atomicAdd(&grad[OFF_B1 + tid], s1);
  }
  for (int e = tid; e < NH * NIN; e += GRAD_CELLS) {
    const int h = e / NIN, k = e % NIN;
    float s = 0.0f;
    for (int c = 0; c < GRAD_CELLS; c++) s += h2s[c][h] * xs[c][k];
    atomicAdd(&grad[OFF_W1 + e], s);
  }
}

__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ grad, float* __restrict__ m1,
                           float* __restrict__ m2, float lr, float b1t, float b2t) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1.0e-8f;
  for (int e = threadIdx.x; e < NPARAM; e += blockDim.x) {
    float g = grad[e];
    float a = b1 * m1[e] + (1.0f - b1) * g;
    float v = b2 * m2[e] + (1.0f - b2) * g * g;
    m1[e] = a;
    m2[e] = v;
    float ah = a / (1.0f - b1t), vh = v / (1.0f - b2t);
    p[e] -= lr * ah / (sqrtf(vh) + eps);
    grad[e] = 0.0f;
  }
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ p, const float* __restrict__ X,
                                                     int64_t n, float* __restrict__ Yo) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = c < n;
  float x0 = valid ? X[c * 3 + 0] : 0.0f, x1 = valid ? X[c * 3 + 1] : 0.0f, x2 = valid ? X[c * 3 + 2] : 0.0f;
  Operands o = load_operands(p, lane);
  float y0 = 0.0f, y1 = 0.0f;
  forward_wave(o, lane, x0, x1, x2, y0, y1);
  if (valid) {
    Yo[c * 2 + 0] = y0;
    Yo[c * 2 + 1] = y1;
  }
}

}  // namespace mlp

// ---------------------------------------------------------------------------
// learned primitive recovery over the interior, one thread per cell
// (STAGE0: also the next step's stage-0 state, as k_c2p<true>)
// ---------------------------------------------------------------------------
template <bool STAGE0>
__global__ void __launch_bounds__(256) k_c2p_learned(Geom g, Planes P, Scheme S, const float* __restrict__ wp,
                                                     StatusRef sr) {
  using namespace mlp;
  if (halted(sr)) return;
  const int lane = threadIdx.x & 63;
  const int ni = g.imax + 1;
  const int64_t ncell = (int64_t)ni * (g.jmax + 1);
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = n < ncell;
  const int j = valid ? (int)(n / ni) : 0, i = valid ? (int)(n - (int64_t)j * ni) : 0;
  const int64_t k = g.at(i, j);
  C2PIn in;
  in.d = P.c[C_D][k]; in.tau = P.c[C_TAU][k];
  in.sx = P.c[C_SX][k]; in.sy = P.c[C_SY][k]; in.sz = P.c[C_SZ][k];
  in.ex = P.p[P_EX][k]; in.ey = P.p[P_EY][k]; in.ez = P.p[P_EZ][k];
  in.bx = P.p[P_BX][k]; in.by = P.p[P_BY][k]; in.bz = P.p[P_BZ][k];
  in.vx0 = P.p[P_VX][k]; in.vy0 = P.p[P_VY][k]; in.vz0 = P.p[P_VZ][k];
  in.rho_old = P.p[P_RHO][k]; in.p_old = P.p[P_P][k];
  Operands o = load_operands(wp, lane);
  C2POut out;
  bool fallback = false;
  const int code = c2p_learned_cell(S.gamma, S.eps1, in, wp, o, lane, S.crmath != 0, out, fallback);
  if (!valid) return;
  if (fallback) atomicAdd(&sr.st->fallback, 1ULL);
  if (code) { report(sr, code, i, j, 1); return; }
  c2p_count(sr, i, j, out);
  c2p_store_out(P, k, out);
  if (STAGE0) stage0_after_c2p(P, S, k, in, out);
}

hipError_t launch_c2p_learned(hipStream_t s, const Geom& g, const Planes& P, const Scheme& S, const float* wp,
                              const StatusRef& sr, bool stage0) {
  const int64_t ncell = (int64_t)(g.imax + 1) * (g.jmax + 1);
  const dim3 grid((unsigned)((ncell + 255) / 256));
  if (stage0)
    hipLaunchKernelGGL(k_c2p_learned<true>, grid, dim3(256), 0, s, g, P, S, wp, sr);
  else
    hipLaunchKernelGGL(k_c2p_learned<false>, grid, dim3(256), 0, s, g, P, S, wp, sr);
  return hipGetLastError();
}

}  // namespace cueva

// ---------------------------------------------------------------------------
// C ABI of the model (declared in include/cueva.h)
// ---------------------------------------------------------------------------
using namespace cueva;

struct cueva_mlp {
  float* p;      // packed parameters (NPACK)
  float* grad;   // NPARAM
  float* m1;     // Adam moments
  float* m2;
  double* loss;
  float* X;      // training inputs (3 per sample), normalised
  float* Y;      // training targets (2 per sample), normalised
  int64_t n, cap;
  int64_t t;     // Adam step count
  hipStream_t stream;
};

#define MCHK(x)                                   \
  do {                                            \
    if ((x) != hipSuccess) return CUEVA_E_HIP;    \
  } while (0)

extern "C" int cueva_mlp_create(cueva_mlp** out) {
  if (!out) return CUEVA_E_ARG;
  cueva_mlp* m = new cueva_mlp();
  size_t bytes = sizeof(float) * (mlp::NPACK + 3 * mlp::NPARAM) + sizeof(double);
  char* base;
  if (hipMalloc(&base, bytes) != hipSuccess) { delete m; return CUEVA_E_NOMEM; }
  m->p = (float*)base;
  m->grad = m->p + mlp::NPACK;
  m->m1 = m->grad + mlp::NPARAM;
  m->m2 = m->m1 + mlp::NPARAM;
  m->loss = (double*)(m->m2 + mlp::NPARAM);  // NPACK + 3 NPARAM floats = 20040 bytes: 8-byte aligned
  static_assert(((mlp::NPACK + 3 * mlp::NPARAM) * sizeof(float)) % 8 == 0, "loss alignment");
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    hipFree(base);
    delete m;
    return CUEVA_E_HIP;
  }
  if (hipMemsetAsync(base, 0, bytes, m->stream) != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess) {
    hipStreamDestroy(m->stream);
    hipFree(base);
    delete m;
    return CUEVA_E_HIP;
  }
  *out = m;
  return CUEVA_OK;
}

extern "C" int cueva_mlp_destroy(cueva_mlp* m) {
  if (!m) return CUEVA_E_ARG;
  hipStreamSynchronize(m->stream);
  hipFree(m->p);
  if (m->X) hipFree(m->X);
  hipStreamDestroy(m->stream);
  delete m;
  return CUEVA_OK;
}

extern "C" int cueva_mlp_set_params(cueva_mlp* m, const float* packed) {
  if (!m || !packed) return CUEVA_E_ARG;
  MCHK(hipMemcpyAsync(m->p, packed, sizeof(float) * mlp::NPACK, hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemsetAsync(m->m1, 0, sizeof(float) * 2 * mlp::NPARAM, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  m->t = 0;
  return CUEVA_OK;
}

extern "C" int cueva_mlp_get_params(cueva_mlp* m, float* packed) {
  if (!m || !packed) return CUEVA_E_ARG;
  MCHK(hipMemcpyAsync(packed, m->p, sizeof(float) * mlp::NPACK, hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return CUEVA_OK;
}

extern "C" int cueva_mlp_set_data(cueva_mlp* m, const float* x, const float* y, int64_t n) {
  if (!m || !x || !y || n <= 0) return CUEVA_E_ARG;
  if (n > m->cap) {
    if (m->X) hipFree(m->X);
    m->X = nullptr;
    if (hipMalloc(&m->X, sizeof(float) * 5 * (size_t)n) != hipSuccess) { m->cap = 0; return CUEVA_E_NOMEM; }
    m->cap = n;
  }
  m->Y = m->X + 3 * n;
  m->n = n;
  MCHK(hipMemcpyAsync(m->X, x, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemcpyAsync(m->Y, y, sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return CUEVA_OK;
}

static void launch_grad(cueva_mlp* m, int64_t off, int64_t cnt) {
  const unsigned nb = (unsigned)((cnt + mlp::GRAD_CELLS - 1) / mlp::GRAD_CELLS);
  hipLaunchKernelGGL(mlp::k_mlp_grad, dim3(nb), dim3(mlp::GRAD_CELLS), 0, m->stream, (const float*)m->p,
                     (const float*)m->X, (const float*)m->Y, off, cnt, 1.0f / (float)cnt, m->grad, m->loss);
}

extern "C" int cueva_mlp_grad(cueva_mlp* m, int64_t off, int64_t cnt, float* grad_out, double* loss_out) {
  if (!m || off < 0 || cnt <= 0 || off + cnt > m->n) return CUEVA_E_ARG;
  MCHK(hipMemsetAsync(m->grad, 0, sizeof(float) * mlp::NPARAM, m->stream));
  MCHK(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
  launch_grad(m, off, cnt);
  MCHK(hipGetLastError());
  if (grad_out) MCHK(hipMemcpyAsync(grad_out, m->grad, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost, m->stream));
  if (loss_out) MCHK(hipMemcpyAsync(loss_out, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MCHK(hipMemsetAsync(m->grad, 0, sizeof(float) * mlp::NPARAM, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return CUEVA_OK;
}

extern "C" int cueva_mlp_train(cueva_mlp* m, int64_t batch, int steps, float lr, double* loss_out) {
  if (!m || batch <= 0 || steps <= 0 || m->n < batch) return CUEVA_E_ARG;
  const int64_t nb = m->n / batch;
  MCHK(hipMemsetAsync(m->grad, 0, sizeof(float) * mlp::NPARAM, m->stream));
  MCHK(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
  for (int s = 0; s < steps; s++) {
    const int64_t off = ((m->t) % nb) * batch;
    launch_grad(m, off, batch);
    m->t++;
    const float b1t = powf(0.9f, (float)m->t), b2t = powf(0.999f, (float)m->t);
    hipLaunchKernelGGL(mlp::k_mlp_adam, dim3(1), dim3(256), 0, m->stream, m->p, m->grad, m->m1, m->m2, lr, b1t,
                       b2t);
  }
  MCHK(hipGetLastError());
  double l = 0.0;
  MCHK(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  if (loss_out) *loss_out = l / steps;
  return CUEVA_OK;
}

extern "C" int cueva_mlp_forward(cueva_mlp* m, const float* x, int64_t n, float* y) {
  if (!m || !x || !y || n <= 0) return CUEVA_E_ARG;
  float* buf;
  if (hipMalloc(&buf, sizeof(float) * 5 * (size_t)n) != hipSuccess) return CUEVA_E_NOMEM;
  int rc = CUEVA_OK;
  if (hipMemcpyAsync(buf, x, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    rc = CUEVA_E_HIP;
  if (!rc) {
    hipLaunchKernelGGL(mlp::k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream,
                       (const float*)m->p, (const float*)buf, n, buf + 3 * n);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(y, buf + 3 * n, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      rc = CUEVA_E_HIP;
  }
  hipStreamSynchronize(m->stream);
  hipFree(buf);
  return rc;
}
