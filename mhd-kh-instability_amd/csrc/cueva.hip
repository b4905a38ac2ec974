// C ABI of the gfx950 CUEVA step (declarations and citations: include/cueva.h).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "cueva.h"
#include "cueva_kernels.hpp"
#include "cueva_reduce.hpp"

namespace cueva {
// FP32 face kernels (cueva_mixed.hip)
hipError_t launch_faces_fp32(int rec_order, dim3 grid, dim3 blk, hipStream_t s, const Geom& g, const Planes& P,
                             const Scheme& S, const StatusRef& st);
// FP32 marching face kernels (cueva_mixed.hip)
const void* march_kernel_fp32(int rec_order, int solver, bool final_stage, bool c2p);
// learned primitive recovery (cueva_mlp.hip)
hipError_t launch_c2p_learned(hipStream_t s, const Geom& g, const Planes& P, const Scheme& S, const float* wp,
                              const StatusRef& st, bool stage0);
}

using namespace cueva;

// A launch that can fail, remembered until the status is next read: its
// launch number, step and stage, and (fused step) the plane roles it ran
// with, which become the state's again if it was the first to fail.
struct LaunchRecord {
  uint64_t seq;
  int64_t h;
  int l;
  bool fused;
  Planes P;
};
constexpr size_t MAX_LAUNCH_RECORDS = 16384;
constexpr int FIX_CAP = 1 << 16;  // fix-up records kept per kind between reads

struct cueva_ctx {
  cueva_params prm;
  Geom g;
  Scheme S;
  Planes P;
  hipStream_t stream;
  double* base;
  int64_t bytes;
  DevStatus* dst;
  double* red;  // reduction partials (RED_BLOCKS * G_N) and results (G_N), after dst
  float* mlpw;  // packed learned-recovery parameters (CUEVA_MLP_NPACK), after red
  int has_mlp;
  int rec_order;
  int march;         // fused step uses the marching face kernel (default; CUEVA_MARCH=0: face planes)
  int march_rounds;  // resident-block rounds a marching launch is cut into (CUEVA_MARCH_ROUNDS)
  int march_xcd;     // marching blocks in XCD-aware tile order (CUEVA_MARCH_XCD, default on)
  int march_rows;    // rows per segment when set (CUEVA_MARCH_SEG; rounded up to 4m - 1)
  int c2p_in_march;  // Cardano recovery inside the last marching launch (CUEVA_C2P_IN_MARCH)
  int march_seg[3];  // rows per segment of the stage / last-stage / last-stage + recovery kernel (0: not sized yet)
  // failures and fix-up records
  FixRec* rec;          // device: [3][FIX_CAP], after mlpw
  uint64_t seq;         // launches numbered so far
  int64_t hstep;        // step number the next step carries (1 after create)
  std::vector<LaunchRecord>* launches;
  int err;              // first failure (sticky until the next upload of primitives)
  int32_t err_where[4]; // its step, stage, i, j
  // kernel timing (cueva_set_profiling)
  int prof;
  int nev, capev;
  hipEvent_t* ev;  // pairs (start, stop)
  int* evcls;
};

enum { K_STAGE = 0, K_FACES = 1, K_FLUXSRC = 2, K_FINAL = 3, K_GHOSTS = 4, K_OTHER = 5, K_NCLS = 6 };

static void prof_begin(cueva_ctx* c, int cls) {
  if (!c->prof) return;
  if (c->nev == c->capev) {
    int cap = c->capev ? 2 * c->capev : 256;
    hipEvent_t* ev = (hipEvent_t*)realloc(c->ev, sizeof(hipEvent_t) * 2 * cap);
    int* cl = (int*)realloc(c->evcls, sizeof(int) * cap);
    if (!ev || !cl) { c->prof = 0; return; }
    for (int k = 2 * c->capev; k < 2 * cap; k++) hipEventCreate(&ev[k]);
    c->ev = ev;
    c->evcls = cl;
    c->capev = cap;
  }
  c->evcls[c->nev] = cls;
  hipEventRecord(c->ev[2 * c->nev], c->stream);
}

static void prof_end(cueva_ctx* c) {
  if (!c->prof) return;
  hipEventRecord(c->ev[2 * c->nev + 1], c->stream);
  c->nev++;
}

#define LAUNCH(cls, kern, grid, blk, ...)                                   \
  do {                                                                    \
    prof_begin(c, cls);                                                   \
    hipLaunchKernelGGL(kern, grid, blk, 0, c->stream, __VA_ARGS__);       \
    prof_end(c);                                                          \
  } while (0)

#define HIPCHK(x)                                   \
  do {                                              \
    hipError_t e_ = (x);                            \
    if (e_ != hipSuccess) return CUEVA_E_HIP;       \
  } while (0)

static inline int is_pow2(double d) {
  int e;
  double m = frexp(d, &e);
  return m == 0.5;
}

extern "C" int cueva_abi_version(void) { return CUEVA_ABI_VERSION; }

extern "C" int cueva_params_default(cueva_params* p) {
  if (!p) return CUEVA_E_ARG;
  memset(p, 0, sizeof(*p));
  p->imax = 128;
  p->jmax = 256;
  p->order = 2;
  p->flux_solver = CUEVA_HLL;
  p->rec_mp = 0;
  p->bc_x = CUEVA_BC_PERIODIC;
  p->bc_y = CUEVA_BC_COPY;
  p->precision = CUEVA_FP64;
  p->c2p_mode = CUEVA_C2P_CARDANO;
  p->hllc_seed = 1;
  p->crmath = 1;
  p->slop = 2;
  p->limit = 2;
  p->longx = 1.0;
  p->longy = 2.0;
  p->cfl = 0.4;
  p->sigma = 1.0e3;
  p->gamma = 4.0 / 3.0;
  p->dt = 0.0;
  return CUEVA_OK;
}

// parameters.f95:306-334: Delx = Longx/real(imax) (real() is exact here),
// Delt = CFL*min(Delx,Dely,Delz) with Delz = 1, kappa = 1/cp^2,
// cp = sqrt(min(Delx,Dely)/alpha), alpha = 1.
extern "C" int cueva_derived(const cueva_params* p, double out[5]) {
  if (!p || !out) return CUEVA_E_ARG;
  if (p->imax < 8 || p->jmax < 8 || p->imax > (1 << 24) || p->jmax > (1 << 24)) return CUEVA_E_ARG;
  double dx = p->longx / (double)(float)p->imax;
  double dy = p->longy / (double)(float)p->jmax;
  double m = dy < dx ? dy : dx;
  double dt = p->dt > 0.0 ? p->dt : p->cfl * (1.0 < m ? 1.0 : m);
  double cp = sqrt(m / 1.0);
  out[0] = dx;
  out[1] = dy;
  out[2] = dt;
  out[3] = 1.0 / (cp * cp);
  out[4] = 1.0 / (cp * cp);
  return CUEVA_OK;
}

// Butcher tables of imex_scheme.f95:13-20 (first order), :24-36 (MIRK2),
// :40-49 (SSP2(2,2,2)), :77-86 (SSP2(3,3,2)-LUM) and :280-291 (SSP3(4,3,3))
static void imex_tables(int order, int mirk, Scheme& S) {
  memset(S.wb, 0, sizeof(S.wb));
  memset(S.wbt, 0, sizeof(S.wbt));
  memset(S.Ab, 0, sizeof(S.Ab));
  memset(S.Abt, 0, sizeof(S.Abt));
  if (order == 1) {  // imex_scheme.f95:13-20, the first-order scheme of MIRK1
    S.lmax = 1;
    S.wb[0] = 1.0;
    S.Ab[0][0] = 1.0;
  } else if (order == 2 && mirk == 2) {  // MIRK2, imex_scheme.f95:24-36
    S.lmax = 2;
    S.wb[0] = 0.5; S.wb[1] = 0.5;
    S.Ab[1][0] = 1.0;
  } else if (order == 2) {
    const double g = 1.0 - 1.0 / sqrt(2.0);  // parameters.f95:240
    S.lmax = 2;
    S.wb[0] = 0.5; S.wb[1] = 0.5;
    S.wbt[0] = 0.5; S.wbt[1] = 0.5;
    S.Ab[1][0] = 1.0;
    S.Abt[0][0] = g;
    S.Abt[1][0] = 1.0 - 2.0 * g;
    S.Abt[1][1] = g;
  } else if (order == 3) {
    S.lmax = 3;
    for (int l = 0; l < 3; l++) S.wb[l] = S.wbt[l] = 1.0 / 3.0;
    S.Ab[1][0] = 0.5;
    S.Ab[2][0] = 0.5; S.Ab[2][1] = 0.5;
    S.Abt[0][0] = 1.0 / 5.0;
    S.Abt[1][0] = 1.0 / 10.0; S.Abt[1][1] = 1.0 / 5.0;
    S.Abt[2][0] = 1.0 / 3.0; S.Abt[2][1] = 1.0 / 3.0; S.Abt[2][2] = 1.0 / 3.0;
  } else {
    const double a1 = 0.24169426078821, b1 = 0.06042356519705, e1 = 0.12915286960590;  // :242-244
    S.lmax = 4;
    S.wb[1] = 1.0 / 6.0; S.wb[2] = 1.0 / 6.0; S.wb[3] = 2.0 / 3.0;
    S.wbt[1] = 1.0 / 6.0; S.wbt[2] = 1.0 / 6.0; S.wbt[3] = 2.0 / 3.0;
    S.Ab[2][1] = 1.0;
    S.Ab[3][1] = 0.25; S.Ab[3][2] = 0.25;
    S.Abt[0][0] = a1;
    S.Abt[1][0] = -1.0 * a1; S.Abt[1][1] = a1;
    S.Abt[2][1] = 1.0 - a1; S.Abt[2][2] = a1;
    S.Abt[3][0] = b1; S.Abt[3][1] = e1; S.Abt[3][2] = ((0.5 - b1) - e1) - a1; S.Abt[3][3] = a1;
  }
}

static int check_params(const cueva_params* p) {
  if (!p) return CUEVA_E_ARG;
  if (p->imax < 8 || p->jmax < 8 || p->imax > 65536 || p->jmax > 65536) return CUEVA_E_ARG;
  if (p->order < 1 || p->order > 4) return CUEVA_E_ARG;
  if (p->mirk < 0 || p->mirk > 2 || (p->order == 2 && p->mirk == 1)) return CUEVA_E_ARG;
  if (p->rec_prim != 0 && p->rec_prim != 1) return CUEVA_E_ARG;
  // REC_PRIM = 1 with the second-order HLL / HLLE / HLLC: the reference's
  // REC_PRIM branch of those solvers (16_hll_flow_mcl.f95:911-958,
  // 16_hllc_flow_mcl.f95:2106-2153, 16_hlle_flow_mcl.f95:1993-2040) stores
  // the z-velocity's right states in Vy_x_R / Vy_y_R and never sets Vz_x_R /
  // Vz_y_R, OpenMP-private scalars it then reads: its results are undefined
  if (p->rec_prim && (p->flux_solver == CUEVA_HLL_MCL || p->flux_solver == CUEVA_HLLE_MCL ||
                      p->flux_solver == CUEVA_HLLC_MCL))
    return CUEVA_E_ARG;
  if (p->flux_solver != CUEVA_HLL && p->flux_solver != CUEVA_HLLC && p->flux_solver != CUEVA_HLLE &&
      p->flux_solver != CUEVA_HLL_MCL && p->flux_solver != CUEVA_LAX && p->flux_solver != CUEVA_HLLC_MCL &&
      p->flux_solver != CUEVA_HLLE_MCL)
    return CUEVA_E_ARG;
  if (p->rec_mp < 0 || p->rec_mp > 2) return CUEVA_E_ARG;
  if (p->primitive_rec != 0 && p->primitive_rec != 1) return CUEVA_E_ARG;
  if (p->slop < 0 || p->slop > 3 || p->limit < 0 || p->limit > 3) return CUEVA_E_ARG;
  if ((p->bc_x != 0 && p->bc_x != 1) || (p->bc_y != 0 && p->bc_y != 1)) return CUEVA_E_ARG;
  if (p->precision != CUEVA_FP64 && p->precision != CUEVA_MIXED) return CUEVA_E_ARG;
  if (p->c2p_mode != CUEVA_C2P_CARDANO && p->c2p_mode != CUEVA_C2P_LEARNED) return CUEVA_E_ARG;
  if (!(p->longx > 0.0) || !(p->longy > 0.0) || !(p->cfl > 0.0) || !(p->sigma >= 0.0)) return CUEVA_E_ARG;
  return CUEVA_OK;
}

extern "C" int cueva_create(const cueva_params* p, cueva_ctx** out) {
  if (!out) return CUEVA_E_ARG;
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  cueva_ctx* c = new (std::nothrow) cueva_ctx();
  if (!c) return CUEVA_E_NOMEM;
  memset(c, 0, sizeof(*c));
  c->prm = *p;
  Geom& g = c->g;
  g.imax = p->imax;
  g.jmax = p->jmax;
  g.pitch = ((p->imax + 16) + 7) / 8 * 8;
  g.rows = p->jmax + 13;
  g.plane = ((int64_t)g.pitch * g.rows + 63) / 64 * 64;
  g.bcx = p->bc_x;
  g.bcy = p->bc_y;
  double d[5];
  cueva_derived(p, d);
  Scheme& S = c->S;
  S.dx = d[0]; S.dy = d[1]; S.dt = d[2]; S.kpsi = d[3]; S.kphi = d[4];
  S.inv_dx = 1.0 / S.dx; S.inv_dy = 1.0 / S.dy;
  S.div_exact = is_pow2(S.dx) && is_pow2(S.dy);
  S.sigma = p->sigma;
  S.gamma = p->gamma;
  S.flux_solver = p->flux_solver;
  S.order = p->order;
  S.hllc_seed_ref = p->hllc_seed;
  S.crmath = p->crmath;
  S.slop = p->slop ? p->slop : 2;
  S.mirk = p->mirk;
  S.rec_prim = p->rec_prim;
  S.cm1 = -0.05;  // parameters.f95:233-235
  S.cm2 = 1.0;
  S.cm4 = (0.5 * ((1.0 - S.cm1) * (1.0 - S.cm1))) / S.cm1;
  imex_tables(p->order, p->mirk, S);
  c->rec_order = 5 + 2 * p->rec_mp;
  {
    const char* e = getenv("CUEVA_MARCH");
    c->march = e ? atoi(e) != 0 : 1;
    // resident-block rounds a marching launch is cut into: measured best at
    // 4 for the FP64 kernels (2 blocks per CU; 1.5-3.6 % faster than 2 at
    // 2048x4096 and 4096x8192) and 2 for the FP32 ones (4 blocks per CU):
    // about 2000 blocks of ~67-row segments at 2048x4096 either way
    e = getenv("CUEVA_MARCH_ROUNDS");
    c->march_rounds = e && atoi(e) > 0 ? atoi(e) : (p->precision == CUEVA_MIXED ? 2 : 4);
    // the hardware deals workgroups to the 8 XCDs round robin; in the XCD-aware
    // order each XCD marches a contiguous band of tiles, whose neighbours
    // share halo columns and rows in its L2: 19-25 % fewer bytes fetched and
    // FP32 steps 0.4 % faster, the other FP64 steps within 1 % of the plain
    // order's, but FP64 HLLE (MP) steps 5 % slower, which keep the plain order
    e = getenv("CUEVA_MARCH_XCD");
    c->march_xcd = e ? atoi(e) != 0 : !(p->precision == CUEVA_FP64 && p->flux_solver == CUEVA_HLLE);
    e = getenv("CUEVA_MARCH_SEG");
    c->march_rows = e && atoi(e) > 0 ? atoi(e) : 0;
    // default off: inside the last marching launch the recovery's registers
    // cost the whole launch occupancy or spill (FP32 faces at 2 waves per SIMD
    // instead of 4; FP64 faces 256 B of scratch per lane): steps measured
    // 0.83 ms (FP32 faces) and 0.12 ms (FP64 faces) slower at 2048x4096 than
    // with the recovery in its own 0.55 ms pass
    e = getenv("CUEVA_C2P_IN_MARCH");
    c->c2p_in_march = e ? atoi(e) != 0 : 0;
  }

  const int nstage = p->rec_prim ? S.lmax * 7 : 0;  // stage primitives of REC_PRIM = 1
  const int nplanes = P_N + C_N + 2 * S_N + FX_N + FY_N + MAXL * R_N + 2 * A_N + nstage;
  c->bytes = (int64_t)nplanes * g.plane * (int64_t)sizeof(double);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return CUEVA_E_HIP; }
  if (hipMalloc(&c->base, (size_t)c->bytes) != hipSuccess) {
    hipStreamDestroy(c->stream);
    delete c;
    return CUEVA_E_NOMEM;
  }
  const size_t status_bytes = (sizeof(DevStatus) + 255) / 256 * 256;
  const size_t red_bytes = (size_t)(RED_BLOCKS + 1) * G_N * sizeof(double);
  const size_t mlp_bytes = ((size_t)CUEVA_MLP_NPACK * sizeof(float) + 255) / 256 * 256;
  const size_t rec_bytes = (size_t)3 * FIX_CAP * sizeof(FixRec);
  if (hipMalloc(&c->dst, status_bytes + red_bytes + mlp_bytes + rec_bytes) != hipSuccess) {
    hipFree(c->base);
    hipStreamDestroy(c->stream);
    delete c;
    return CUEVA_E_NOMEM;
  }
  c->red = (double*)((char*)c->dst + status_bytes);
  c->mlpw = (float*)((char*)c->red + red_bytes);
  c->rec = (FixRec*)((char*)c->mlpw + mlp_bytes);
  c->hstep = 1;
  c->launches = new (std::nothrow) std::vector<LaunchRecord>();
  if (!c->launches) {
    hipFree(c->dst);
    hipFree(c->base);
    hipStreamDestroy(c->stream);
    delete c;
    return CUEVA_E_NOMEM;
  }
  double* q = c->base;
  auto take = [&]() { double* r = q; q += g.plane; return r; };
  for (int k = 0; k < P_N; k++) c->P.p[k] = take();
  for (int k = 0; k < C_N; k++) c->P.c[k] = take();
  for (int k = 0; k < S_N; k++) c->P.s[k] = take();
  for (int k = 0; k < S_N; k++) c->P.sn[k] = take();
  for (int k = 0; k < FX_N; k++) c->P.fx[k] = take();
  for (int k = 0; k < FY_N; k++) c->P.fy[k] = take();
  // the FP32 face-flux planes of the mixed form reuse the FP64 ones
  for (int k = 0; k < FX_N; k++) c->P.fxf[k] = (float*)c->P.fx[k];
  for (int k = 0; k < FY_N; k++) c->P.fyf[k] = (float*)c->P.fy[k];
  for (int l = 0; l < MAXL; l++)
    for (int k = 0; k < R_N; k++) c->P.rhs[l][k] = take();
  for (int k = 0; k < A_N; k++) c->P.ints[k] = take();
  for (int k = 0; k < A_N; k++) c->P.fins[k] = take();
  if (p->rec_prim)
    for (int l = 0; l < S.lmax; l++) {
      for (int d = 0; d < 3; d++) c->P.vint[l][d] = take();
      c->P.wint[l] = take();
      c->P.rhoint[l] = take();
      c->P.pint[l] = take();
      c->P.hint[l] = take();
    }
  recover_in_place(c->P);
  // every plane pointer the kernels receive must lie inside the allocation
  {
    const char* lo = (const char*)c->base;
    const char* hi = lo + c->bytes;
    void* const* ptrs = (void* const*)&c->P;
    const int nptr = (int)(sizeof(Planes) / sizeof(void*));
    for (int k = 0; k < nptr; k++) {
      const char* a = (const char*)ptrs[k];
      if (!a) continue;  // stage primitives without REC_PRIM
      if (a < lo || a + g.plane * (int64_t)sizeof(double) > hi) {
        delete c->launches;
        hipFree(c->dst);
        hipFree(c->base);
        hipStreamDestroy(c->stream);
        delete c;
        return CUEVA_E_HIP;
      }
    }
  }
  if (hipMemsetAsync(c->base, 0, (size_t)c->bytes, c->stream) != hipSuccess ||
      hipMemsetAsync(c->dst, 0, sizeof(DevStatus), c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    delete c->launches;
    hipFree(c->dst);
    hipFree(c->base);
    hipStreamDestroy(c->stream);
    delete c;
    return CUEVA_E_HIP;
  }
  *out = c;
  return CUEVA_OK;
}

extern "C" int cueva_destroy(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  hipStreamSynchronize(c->stream);
  delete c->launches;
  for (int k = 0; k < 2 * c->capev; k++) hipEventDestroy(c->ev[k]);
  free(c->ev);
  free(c->evcls);
  hipFree(c->dst);
  hipFree(c->base);
  hipStreamDestroy(c->stream);
  delete c;
  return CUEVA_OK;
}

extern "C" int cueva_sync(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  return CUEVA_OK;
}

extern "C" int cueva_device_bytes(const cueva_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return CUEVA_E_ARG;
  *bytes = c->bytes;
  return CUEVA_OK;
}

// ---------------------------------------------------------------------------
// host <-> device
// ---------------------------------------------------------------------------
static int copy_planes_h2d(cueva_ctx* c, double* const* dst, int n, const double* host) {
  const Geom& g = c->g;
  const size_t hw = (size_t)(g.imax + 13), hh = (size_t)(g.jmax + 13);
  for (int k = 0; k < n; k++) {
    HIPCHK(hipMemcpy2DAsync(dst[k] + 2, (size_t)g.pitch * sizeof(double), host + (size_t)k * hw * hh,
                            hw * sizeof(double), hw * sizeof(double), hh, hipMemcpyHostToDevice, c->stream));
  }
  return CUEVA_OK;
}

static int copy_planes_d2h(cueva_ctx* c, double* const* src, int n, double* host) {
  const Geom& g = c->g;
  const size_t hw = (size_t)(g.imax + 13), hh = (size_t)(g.jmax + 13);
  for (int k = 0; k < n; k++) {
    HIPCHK(hipMemcpy2DAsync(host + (size_t)k * hw * hh, hw * sizeof(double), src[k] + 2,
                            (size_t)g.pitch * sizeof(double), hw * sizeof(double), hh, hipMemcpyDeviceToHost,
                            c->stream));
  }
  return CUEVA_OK;
}

static dim3 interior_grid(const Geom& g, dim3 blk) {
  return dim3((unsigned)((g.imax + 1 + blk.x - 1) / blk.x), (unsigned)((g.jmax + 1 + blk.y - 1) / blk.y));
}

static int launch_ok() { return hipGetLastError() == hipSuccess ? CUEVA_OK : CUEVA_E_HIP; }

// Number the next launch.  record: a launch that can fail (face solves,
// primitive recovery), remembered with its step and stage l (1-based) and,
// for the fused step, the plane roles it runs with.
static StatusRef sref(cueva_ctx* c, int l, bool record = false, bool fused = false, const Planes* snap = nullptr) {
  StatusRef r;
  r.st = c->dst;
  r.rec = c->rec;
  r.seq = ++c->seq;
  r.h = (int)c->hstep;
  r.cap = FIX_CAP;
  r.l = 0;
  if (record && c->launches->size() < MAX_LAUNCH_RECORDS) {
    LaunchRecord lr;
    lr.seq = r.seq;
    lr.h = c->hstep;
    lr.l = l;
    lr.fused = fused;
    lr.P = snap ? *snap : c->P;
    c->launches->push_back(lr);
  }
  return r;
}

// copy every ghost point of the planes from its interior source point; the
// last nzero planes (at most 32) are set to 0 there instead
static int fill_ghosts(cueva_ctx* c, double* const* planes, int n, int nzero = 0) {
  const StatusRef sr = sref(c, 0);
  PlaneList pl;
  memset(&pl, 0, sizeof(pl));
  for (int start = 0; start < n; start += 32) {
    pl.n = n - start < 32 ? n - start : 32;
    pl.nzero = start + pl.n == n ? nzero : 0;
    for (int k = 0; k < pl.n; k++) pl.a[k] = planes[start + k];
    int64_t cnt = ghost_count(c->g);
    LAUNCH(K_GHOSTS, k_fill_ghosts, dim3((unsigned)((cnt + 255) / 256)), dim3(256), c->g, pl, sr);
    if (launch_ok()) return CUEVA_E_HIP;
  }
  return CUEVA_OK;
}

extern "C" int cueva_upload_prims(cueva_ctx* c, const double* prims14) {
  if (!c || !prims14) return CUEVA_E_ARG;
  // a new state: the first failure, counters and records start again
  HIPCHK(hipMemsetAsync(c->dst, 0, sizeof(DevStatus), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->err = 0;
  c->launches->clear();
  int rc = copy_planes_h2d(c, c->P.p, P_N, prims14);
  if (rc) return rc;
  return fill_ghosts(c, c->P.p, P_N);
}

extern "C" int cueva_download(cueva_ctx* c, double* prims14, double* cons5) {
  if (!c) return CUEVA_E_ARG;
  int rc;
  if (prims14 && (rc = copy_planes_d2h(c, c->P.p, P_N, prims14))) return rc;
  if (cons5 && (rc = copy_planes_d2h(c, c->P.c, 5, cons5))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return CUEVA_OK;
}

// ---------------------------------------------------------------------------
// one entry point per reference subroutine
// ---------------------------------------------------------------------------
static const dim3 BLK(64, 4);

extern "C" int cueva_boundary(cueva_ctx* c, int kind, int l) {
  if (!c) return CUEVA_E_ARG;
  double* list[20];
  int n = 0;
  switch (kind) {
    case CUEVA_BND_PRIM:
      for (int k = 0; k < P_N; k++) list[n++] = c->P.p[k];
      break;
    case CUEVA_BND_CONS: {
      static const int cons[11] = {S_PSI, S_PHI, S_BX, S_BY, S_BZ, S_Q, S_D, S_TAU, S_SX, S_SY, S_SZ};
      for (int k = 0; k < 11; k++) list[n++] = c->P.s[cons[k]];
      break;
    }
    case CUEVA_BND_ELEC:
      list[n++] = c->P.s[S_EX]; list[n++] = c->P.s[S_EY]; list[n++] = c->P.s[S_EZ];
      break;
    case CUEVA_BND_FLUX:
      // guarded by flux_solver == 5 .or. 7 (24_boundary_flux_conserved.f95:13)
      if (!j_ghosts_filled(c->prm.flux_solver)) return CUEVA_OK;
      list[n++] = c->P.s[S_JX]; list[n++] = c->P.s[S_JY]; list[n++] = c->P.s[S_JZ];
      // and the stage velocity (:1426-1560 and the like); the stage rho, p and
      // W keep zero ghosts: no routine of the reference fills them
      if (c->prm.rec_prim) {
        if (l < 1 || l > c->S.lmax) return CUEVA_E_ARG;
        for (int d = 0; d < 3; d++) list[n++] = c->P.vint[l - 1][d];
      }
      break;
    default:
      return CUEVA_E_ARG;
  }
  return fill_ghosts(c, list, n);
}

static int recons_prims(cueva_ctx* c) {
  LAUNCH(K_OTHER, k_recons_prim, interior_grid(c->g, BLK), BLK, c->g, c->P, sref(c, 0),
         c->prm.limit ? c->prm.limit : 2);
  if (launch_ok()) return CUEVA_E_HIP;
  LAUNCH(K_OTHER, k_recons_copy, interior_grid(c->g, BLK), BLK, c->g, c->P, sref(c, 0));
  return launch_ok();
}

extern "C" int cueva_recons_var_primitive(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  return recons_prims(c);
}

extern "C" int cueva_startconvar(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_startconvar, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S);
  return launch_ok();
}

// counters and fix-up records (not the first failure)
static int clear_counters(cueva_ctx* c) {
  HIPCHK(hipMemsetAsync((char*)c->dst + STATUS_COUNTERS_OFFSET, 0, sizeof(DevStatus) - STATUS_COUNTERS_OFFSET,
                        c->stream));
  return CUEVA_OK;
}

extern "C" int cueva_cleanvarint(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_cleanvarint, interior_grid(c->g, BLK), BLK, c->g, c->P);
  if (launch_ok()) return CUEVA_E_HIP;
  return clear_counters(c);  // contador_* = 0 (main_cueva_code.f95:189-193)
}

// swap1 copies the primitives into the stage-1 primitives (08_swap1.f95),
// which only REC_PRIM = 1 reads; without it there is nothing to do.
extern "C" int cueva_swap1(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  if (!c->prm.rec_prim) return CUEVA_OK;
  LAUNCH(K_OTHER, k_swap1, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S);
  return launch_ok();
}

static int stage_ok(const cueva_ctx* c, int l) { return c && l >= 1 && l <= c->S.lmax; }

extern "C" int cueva_sumint(cueva_ctx* c, int l, int m) {
  if (!stage_ok(c, l) || m < 1 || m >= l) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_sumint, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, l - 1, m - 1);
  return launch_ok();
}

extern "C" int cueva_varint(cueva_ctx* c, int l) {
  if (!stage_ok(c, l)) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_varint, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S);
  if (launch_ok()) return CUEVA_E_HIP;
  return cueva_boundary(c, CUEVA_BND_CONS, l);  // 10_varint.f95:72
}

extern "C" int cueva_electricvarint(cueva_ctx* c, int l) {
  if (!stage_ok(c, l)) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_electricvarint, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, l - 1);
  if (launch_ok()) return CUEVA_E_HIP;
  return cueva_boundary(c, CUEVA_BND_ELEC, l);  // 11_elecvarint.f95:402
}

extern "C" int cueva_flowinit(cueva_ctx* c, int l) {
  if (!stage_ok(c, l)) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_flowinit, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, l - 1);
  return launch_ok();
}

template <int ORD>
static int launch_faces_ord(cueva_ctx* c, const StatusRef& sr, const Planes& P) {
  const dim3 grid = faces_grid(c->g, BLK);
  if (c->prm.flux_solver == CUEVA_LAX)
    LAUNCH(K_FACES, (k_faces<ORD, 1>), grid, BLK, c->g, P, c->S, sr);
  else if (c->prm.flux_solver == CUEVA_HLLC_MCL)
    LAUNCH(K_FACES, (k_faces<ORD, 4>), grid, BLK, c->g, P, c->S, sr);
  else if (c->prm.flux_solver == CUEVA_HLLE_MCL)
    LAUNCH(K_FACES, (k_faces<ORD, 3>), grid, BLK, c->g, P, c->S, sr);
  else if (c->prm.flux_solver == CUEVA_HLL_MCL)
    LAUNCH(K_FACES, (k_faces<ORD, 2>), grid, BLK, c->g, P, c->S, sr);
  else if (c->prm.flux_solver == CUEVA_HLLC)
    LAUNCH(K_FACES, (k_faces<ORD, 7>), grid, BLK, c->g, P, c->S, sr);
  else if (c->prm.flux_solver == CUEVA_HLLE)
    LAUNCH(K_FACES, (k_faces<ORD, 6>), grid, BLK, c->g, P, c->S, sr);
  else
    LAUNCH(K_FACES, (k_faces<ORD, 5>), grid, BLK, c->g, P, c->S, sr);
  return launch_ok();
}

// The plane set of stage l's face solves and right-hand side: with REC_PRIM
// = 1 they read the stage's velocity, pressure and density (16_*_flow_*.f95
// REC_PRIM branches; Lax-Friedrichs through flowinit's cell fluxes,
// 14_flowinit.f95:376-489; 17_fluxandsource.f95:460-474).  l is 1-based.
static Planes stage_planes(const cueva_ctx* c, int l) {
  Planes P = c->P;
  for (int q = 0; q < K_N; q++) P.kin[q] = nullptr;
  if (c->prm.rec_prim && l >= 1 && l <= c->S.lmax) {
    const int s = l - 1;
    P.kin[K_VX] = c->P.vint[s][0];
    P.kin[K_VY] = c->P.vint[s][1];
    P.kin[K_VZ] = c->P.vint[s][2];
    P.kin[K_P] = c->P.pint[s];
    P.kin[K_RHO] = c->P.rhoint[s];
  }
  return P;
}

// hll_flow_mpx sets the module variable eps1 that the recovery's clamp of a
// superluminal previous velocity reads (16_hll_flow_mpx.f95:1375)
static void note_face_solver(cueva_ctx* c) {
  if (c->prm.flux_solver == CUEVA_HLL) c->S.eps1 = 1.0e-6;
}

// the face solves of stage l (1-based); fused: inside cueva_step
static int launch_faces(cueva_ctx* c, int l, bool fused) {
  const StatusRef sr = sref(c, l, true, fused);
  note_face_solver(c);
  const Planes P = stage_planes(c, l);
  if (c->prm.precision == CUEVA_MIXED) {
    const dim3 grid = faces_grid(c->g, BLK);
    prof_begin(c, K_FACES);
    hipError_t e = launch_faces_fp32(c->rec_order, grid, BLK, c->stream, c->g, P, c->S, sr);
    prof_end(c);
    return e == hipSuccess ? CUEVA_OK : CUEVA_E_HIP;
  }
  if (c->prm.flux_solver == CUEVA_HLL_MCL || c->prm.flux_solver == CUEVA_LAX ||
      c->prm.flux_solver == CUEVA_HLLC_MCL || c->prm.flux_solver == CUEVA_HLLE_MCL)
    return launch_faces_ord<5>(c, sr, P);  // the MP order does not enter
  switch (c->rec_order) {
    case 5: return launch_faces_ord<5>(c, sr, P);
    case 7: return launch_faces_ord<7>(c, sr, P);
    default: return launch_faces_ord<9>(c, sr, P);
  }
}

extern "C" int cueva_flux(cueva_ctx* c, int l) {
  if (!stage_ok(c, l)) return CUEVA_E_ARG;
  // the MPx solvers refresh the stage ghosts first (16_hll_flow_mpx.f95:739-741)
  int rc;
  if ((rc = cueva_boundary(c, CUEVA_BND_CONS, l))) return rc;
  if ((rc = cueva_boundary(c, CUEVA_BND_ELEC, l))) return rc;
  if ((rc = cueva_boundary(c, CUEVA_BND_FLUX, l))) return rc;
  return launch_faces(c, l, false);
}

extern "C" int cueva_fluxandsource(cueva_ctx* c, int l) {
  if (!stage_ok(c, l)) return CUEVA_E_ARG;
  const Planes P = stage_planes(c, l);
  if (c->prm.precision == CUEVA_MIXED)
    LAUNCH(K_FLUXSRC, k_fluxandsource<float>, interior_grid(c->g, BLK), BLK, c->g, P, c->S, l - 1, sref(c, l));
  else
    LAUNCH(K_FLUXSRC, k_fluxandsource<double>, interior_grid(c->g, BLK), BLK, c->g, P, c->S, l - 1, sref(c, l));
  return launch_ok();
}

// varprimitivecardano with l_swap = 0 on stage l (main_cueva_code.f95:220-224):
// REC_PRIM = 1 only; its fix-ups count towards the step's counters
extern "C" int cueva_c2p_stage(cueva_ctx* c, int l) {
  if (!stage_ok(c, l) || !c->prm.rec_prim) return CUEVA_E_ARG;
  if (c->err) return c->err;
  StatusRef sr = sref(c, l, true, false);
  sr.l = l;
  LAUNCH(K_OTHER, k_c2p_stage, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, l - 1, sr);
  return launch_ok();
}

// interconvar, 13_intercon_var.f95: stage l's conserved variables from its
// recovered primitives, then the ghosts of the stage's conserved and E
// planes (:74-75)
extern "C" int cueva_interconvar(cueva_ctx* c, int l) {
  if (!stage_ok(c, l) || !c->prm.rec_prim) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_interconvar, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, l - 1);
  if (launch_ok()) return CUEVA_E_HIP;
  int rc = cueva_boundary(c, CUEVA_BND_CONS, l);
  return rc ? rc : cueva_boundary(c, CUEVA_BND_ELEC, l);
}

extern "C" int cueva_sumfin(cueva_ctx* c, int l) {
  if (!stage_ok(c, l)) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_sumfin, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, l - 1);
  return launch_ok();
}

extern "C" int cueva_varfin(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_varfin, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, 0);
  return launch_ok();
}

extern "C" int cueva_electricfield(cueva_ctx* c) {
  if (!c) return CUEVA_E_ARG;
  LAUNCH(K_OTHER, k_varfin, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, 1);
  return launch_ok();
}

// Wait for the device and read the status words.  A first failure found here
// becomes the context's (code and where: step, stage, i, j); if it arose in
// the fused step, the plane roles go back to those the failing launch ran
// with (the launches after it changed no data, they only rotated roles).
static int read_status(cueva_ctx* c, DevStatus* h) {
  HIPCHK(hipMemcpyAsync(h, c->dst, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (h->seq != 0 && !c->err) {
    const unsigned long long key = ~h->key;
    c->err = (int)(key & 0xFF);
    c->err_where[0] = -1;
    c->err_where[1] = -1;
    c->err_where[2] = (int32_t)((key >> 8) & 0xFFFFF) - 8;
    c->err_where[3] = (int32_t)((key >> 28) & 0xFFFFF) - 8;
    for (const LaunchRecord& lr : *c->launches)
      if (lr.seq == h->seq) {
        c->err_where[0] = (int32_t)lr.h;
        c->err_where[1] = lr.l;
        if (lr.fused) {
          c->P = lr.P;
          recover_in_place(c->P);
        }
        break;
      }
  }
  c->launches->clear();
  return CUEVA_OK;
}

// Primitive recovery after the last stage (l_swap = 1).  fused: inside
// cueva_step; stage0: also the next step's stage-0 state at the interior.
// snap: the plane roles a failure restores (default: those of this launch)
static int launch_c2p(cueva_ctx* c, bool fused, int cls = K_OTHER, bool stage0 = false,
                      const Planes* snap = nullptr) {
  const StatusRef sr = sref(c, c->S.lmax, true, fused, snap);
  if (c->prm.c2p_mode == CUEVA_C2P_LEARNED) {
    if (!c->has_mlp) return CUEVA_E_ARG;
    prof_begin(c, cls);
    hipError_t e = launch_c2p_learned(c->stream, c->g, c->P, c->S, c->mlpw, sr, stage0);
    prof_end(c);
    return e == hipSuccess ? CUEVA_OK : CUEVA_E_HIP;
  }
  if (stage0)
    LAUNCH(cls, k_c2p<true>, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, sr);
  else
    LAUNCH(cls, k_c2p<false>, interior_grid(c->g, BLK), BLK, c->g, c->P, c->S, sr);
  return launch_ok() ? CUEVA_E_HIP : CUEVA_OK;
}

extern "C" int cueva_set_c2p_params(cueva_ctx* c, const float* packed) {
  if (!c || !packed) return CUEVA_E_ARG;
  HIPCHK(hipMemcpyAsync(c->mlpw, packed, (size_t)CUEVA_MLP_NPACK * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->has_mlp = 1;
  return CUEVA_OK;
}

extern "C" int cueva_c2p(cueva_ctx* c, int l_swap, int32_t counters[3], int32_t* fallback) {
  // l_swap = 2: from the last stage's primitives, as main_cueva_code.f95:301-303
  // selects with REC_PRIM = 1
  if (!c || (l_swap != 1 && l_swap != 2) || (l_swap == 2) != (c->prm.rec_prim != 0)) return CUEVA_E_ARG;
  if (c->err) return c->err;
  // the step's counters and records start here (the reference zeroes the
  // counters after cleanvarint and nothing else counts before the recovery),
  // except with REC_PRIM = 1, whose stage recoveries counted since cleanvarint
  int rc = c->prm.rec_prim ? CUEVA_OK : clear_counters(c);
  if (rc) return rc;
  if ((rc = launch_c2p(c, false))) return rc;
  DevStatus h;
  if ((rc = read_status(c, &h))) return rc;
  c->hstep++;  // a step of the per-subroutine sequence ends with this call
  if (counters) for (int k = 0; k < 3; k++) counters[k] = (int32_t)h.cnt[k];
  if (fallback) *fallback = (int32_t)h.fallback;
  return c->err;
}

// ---------------------------------------------------------------------------
// fused step
// ---------------------------------------------------------------------------
static int march_waves(const cueva_ctx* c) {
  return march_block_waves(c->prm.precision == CUEVA_MIXED, c->prm.flux_solver);
}

// ghost band of a stage state formed at the interior points only: a copy of
// the source point's (the state is pointwise in the point's own data), except
// J under HLLE and HLL+MCL, which the reference keeps 0 there
// (24_boundary_flux_conserved.f95:13);
// with_prims: the primitive planes' ghosts in the same launch
static int fill_stage_ghosts(cueva_ctx* c, bool with_prims = false) {
  double* list[P_N + 17];
  int n = 0;
  if (with_prims)
    for (int k = 0; k < P_N; k++) list[n++] = c->P.p[k];
  const int st[17] = {S_PSI, S_PHI, S_BX, S_BY, S_BZ, S_Q, S_D, S_TAU, S_SX, S_SY, S_SZ,
                      S_EX, S_EY, S_EZ, S_JX, S_JY, S_JZ};
  for (int k = 0; k < 17; k++) list[n++] = c->P.s[st[k]];
  return fill_ghosts(c, list, n, j_ghosts_filled(c->prm.flux_solver) ? 0 : 3);
}

// U^{n+1}, left by the last marching launch in rhs[L][0..13], becomes the
// state by exchanging plane pointers: psi phi B E q into the primitive
// planes, D tau S into the conserved ones; the planes they held become the
// scratch of rhs[L].  The primitive recovery then runs as after varfin.
static void adopt_update(cueva_ctx* c, int L) {
  static const int to_p[A_Q + 1] = {P_PSI, P_PHI, P_BX, P_BY, P_BZ, P_EX, P_EY, P_EZ, P_Q};
  static const int to_c[A_SZ - A_D + 1] = {C_D, C_TAU, C_SX, C_SY, C_SZ};
  for (int a = A_PSI; a <= A_Q; a++) {
    double* t = c->P.p[to_p[a]];
    c->P.p[to_p[a]] = c->P.rhs[L][a];
    c->P.rhs[L][a] = t;
  }
  for (int a = A_D; a <= A_SZ; a++) {
    double* t = c->P.c[to_c[a - A_D]];
    c->P.c[to_c[a - A_D]] = c->P.rhs[L][a];
    c->P.rhs[L][a] = t;
  }
}

static const void* march_kernel_of(cueva_ctx* c, bool fin, bool c2p) {
  if (c->prm.precision == CUEVA_MIXED) return march_kernel_fp32(c->rec_order, c->prm.flux_solver, fin, c2p);
  return march_kernel<double>(c->rec_order, c->prm.flux_solver, fin, c2p);
}

// The recovered Vx Vy Vz rho p, written to the x-face planes (unused by the
// marching kernel), take the place of the primitive planes they replace.
static void take_recovered(cueva_ctx* c) {
  static const int idx[5] = {P_VX, P_VY, P_VZ, P_RHO, P_P};
  for (int q = 0; q < 5; q++) {
    double* t = c->P.p[idx[q]];
    c->P.p[idx[q]] = c->P.fx[q];
    c->P.fx[q] = t;
    c->P.fxf[q] = (float*)t;
  }
  recover_in_place(c->P);
}

static void recover_to_scratch(cueva_ctx* c) {
  for (int q = 0; q < 5; q++) c->P.pn[q] = c->P.fx[q];
}

// Rows per segment of a marching launch: enough segments that the blocks
// fill the resident slots of every CU march_rounds times, seg = MW m - 1 (the
// segment's face rows, seg + 1, fill whole tile steps of MW rows).
static int march_seg(cueva_ctx* c, bool fin, bool c2p) {
  int& seg = c->march_seg[fin ? (c2p ? 2 : 1) : 0];
  if (seg) return seg;
  const int mw = march_waves(c);
  if (c->march_rows > 0) {
    seg = (int)(mw * ((c->march_rows + 1 + mw - 1) / mw) - 1);
    return seg;
  }
  int per_cu = 0, cus = 0, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, march_kernel_of(c, fin, c2p), 64 * mw, 0) != hipSuccess)
    return 0;
  const int64_t slots = (int64_t)(per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
  const int strips = march_strips(c->g), rows = c->g.jmax + 1;
  int64_t want = slots * c->march_rounds / strips;
  if (want < 1) want = 1;
  const int64_t per = (rows + want - 1) / want;
  const int64_t m = (per + 1 + mw - 1) / mw;
  seg = (int)(mw * m - 1);
  return seg;
}

// c2p (last stage only): with the primitive recovery (to P.pn) and, with
// stage0, the next step's stage-0 state (to P.sn)
static int launch_march(cueva_ctx* c, int l, bool fin, bool c2p = false, bool stage0 = false) {
  // its face solves and right-hand side read the t^n primitives (no stage
  // kinematics: REC_PRIM = 1 runs the per-subroutine sequence)
  if (c->prm.rec_prim) return CUEVA_E_ARG;
  const void* fn = march_kernel_of(c, fin, c2p);
  int seg = march_seg(c, fin, c2p);
  if (!fn || seg <= 0) return CUEVA_E_HIP;
  const int strips = march_strips(c->g), segs = (c->g.jmax + 1 + seg - 1) / seg;
  const int xcd = c->march_xcd;
  const dim3 grid = xcd ? dim3((unsigned)(8 * ((strips * segs + 7) / 8))) : dim3((unsigned)strips, (unsigned)segs);
  MarchArgs ma;
  ma.strips = strips;
  ma.segs = segs;
  ma.xcd = xcd;
  ma.P = c->P;
  ma.g = c->g;
  ma.S = c->S;
  ma.l = l;
  ma.seg = seg;
  ma.st = sref(c, l + 1, true, true);
  ma.stage0 = stage0 ? 1 : 0;
  note_face_solver(c);
  void* args[] = {(void*)&ma};
  prof_begin(c, K_FACES);
  hipError_t e = hipLaunchKernel(fn, grid, dim3(64, (unsigned)march_waves(c)), args, 0, c->stream);
  prof_end(c);
  return e == hipSuccess ? CUEVA_OK : CUEVA_E_HIP;
}

static int fused_steps(cueva_ctx* c, int nsteps, int halo, bool learned);
static int routine_steps(cueva_ctx* c, int nsteps);

extern "C" int cueva_step(cueva_ctx* c, int nsteps) {
  if (!c || nsteps < 0) return CUEVA_E_ARG;
  const int halo = (c->rec_order - 1) / 2 + 1;  // stencil radius K plus the face offset
  const bool learned = c->prm.c2p_mode == CUEVA_C2P_LEARNED;
  if (learned && !c->has_mlp) return CUEVA_E_ARG;
  if (c->err) return c->err;
  int rc0 = c->prm.rec_prim ? routine_steps(c, nsteps) : fused_steps(c, nsteps, halo, learned);
  // one wait per call: a failure stops the call's later launches on the
  // device, and its status, step, stage and cell are known here
  DevStatus h;
  int rc = read_status(c, &h);
  if (rc0) return rc0;
  if (rc) return rc;
  return c->err;
}

static int fused_steps(cueva_ctx* c, int nsteps, int halo, bool learned) {
  const Geom& g = c->g;
  bool s0_ready = false;  // this step's stage-0 state was formed by the previous step's recovery pass
  for (int h = 0; h < nsteps; h++, c->hstep++) {
    if (c->prm.primitive_rec) {  // main_cueva_code.f95:170-175
      int rc = recons_prims(c);
      if (!rc) rc = fill_ghosts(c, c->P.p, P_N);
      if (rc) return rc;
    }
    for (int l = 0; l < c->S.lmax; l++) {
      const bool last = l == c->S.lmax - 1;
      // the marching path forms stage l >= 1 in the previous stage's face pass
      const bool have_stage = c->march && (l > 0 || s0_ready);
      if (!have_stage) {
        dim3 gs((unsigned)((g.imax + 1 + 2 * halo + BLK.x - 1) / BLK.x),
                (unsigned)((g.jmax + 1 + 2 * halo + BLK.y - 1) / BLK.y));
        LAUNCH(K_STAGE, k_stage_fused, gs, BLK, c->g, c->P, c->S, l, halo, sref(c, l + 1));
        if (launch_ok()) return CUEVA_E_HIP;
      }
      if (c->march) {
        // reconstruction + Riemann + flux difference in one pass; then the
        // next stage's state (interior) or U^{n+1}; the Cardano recovery
        // runs inside the last launch (the learned one in its own pass)
        const bool in_kernel = last && !learned && c->c2p_in_march;
        const Planes pre = c->P;
        if (last) recover_to_scratch(c);
        int rc = launch_march(c, l, last, in_kernel, in_kernel && h + 1 < nsteps && !c->prm.primitive_rec);
        if (rc) return rc;
        if (!last) {
          for (int t = 0; t < S_N; t++) {
            double* x = c->P.s[t];
            c->P.s[t] = c->P.sn[t];
            c->P.sn[t] = x;
          }
          if ((rc = fill_stage_ghosts(c))) return rc;
          continue;
        }
        adopt_update(c, l);
        // the primitive recovery also forms the next step's stage-0 state
        // (not when the next step first re-forms the primitives)
        const bool fuse = h + 1 < nsteps && !c->prm.primitive_rec;
        if (!in_kernel) {
          if ((rc = launch_c2p(c, true, K_FINAL, fuse, &pre))) return rc;
        } else if (fuse) {
          for (int t = 0; t < S_N; t++) {
            double* x = c->P.s[t];
            c->P.s[t] = c->P.sn[t];
            c->P.sn[t] = x;
          }
        }
        take_recovered(c);
        // ghosts of the new primitives (and of the next stage-1 state)
        if ((rc = fuse ? fill_stage_ghosts(c, true) : fill_ghosts(c, c->P.p, P_N))) return rc;
        s0_ready = fuse;
        continue;
      }
      int rc = launch_faces(c, l + 1, true);
      if (rc) return rc;
      const bool mixed = c->prm.precision == CUEVA_MIXED;
      if (!last) {
        const StatusRef sr = sref(c, l + 1);
        if (mixed)
          LAUNCH(K_FLUXSRC, k_fluxandsource<float>, interior_grid(g, BLK), BLK, c->g, c->P, c->S, l, sr);
        else
          LAUNCH(K_FLUXSRC, k_fluxandsource<double>, interior_grid(g, BLK), BLK, c->g, c->P, c->S, l, sr);
      } else {
        const StatusRef sr = sref(c, l + 1, true, true);
        if (learned) {
          if (mixed)
            LAUNCH(K_FINAL, k_final_learned<float>, interior_grid(g, BLK), BLK, c->g, c->P, c->S,
                   (const float*)c->mlpw, sr);
          else
            LAUNCH(K_FINAL, k_final_learned<double>, interior_grid(g, BLK), BLK, c->g, c->P, c->S,
                   (const float*)c->mlpw, sr);
        } else {
          if (mixed)
            LAUNCH(K_FINAL, k_final_fused<float>, interior_grid(g, BLK), BLK, c->g, c->P, c->S, sr);
          else
            LAUNCH(K_FINAL, k_final_fused<double>, interior_grid(g, BLK), BLK, c->g, c->P, c->S, sr);
        }
      }
      if (launch_ok()) return CUEVA_E_HIP;
    }
    if (!c->march) {
      int rc = fill_ghosts(c, c->P.p, P_N);
      if (rc) return rc;
    }
  }
  return CUEVA_OK;
}

// Steps with REC_PRIM = 1 as the reference sequences them (main_cueva_code.f95:
// 53-328), one launch per subroutine: every stage recovers its primitives,
// which feed that stage's fluxes, so no stage can be formed ahead of its
// recovery.  The device is waited on after every launch that can fail, and
// a failure ends the call there (the state is then that of the failing
// launch, as where the reference stops).
static int routine_steps(cueva_ctx* c, int nsteps) {
  const int lmax = c->S.lmax;
  DevStatus st;
  int rc;
#define RS(x) do { if ((rc = (x))) return rc; } while (0)
#define RS_WAIT() do { RS(read_status(c, &st)); if (c->err) return c->err; } while (0)
  for (int h = 0; h < nsteps; h++) {
    RS(cueva_boundary(c, CUEVA_BND_PRIM, 1));
    if (c->prm.primitive_rec) {
      RS(recons_prims(c));
      RS(cueva_boundary(c, CUEVA_BND_PRIM, 1));
    }
    RS(cueva_startconvar(c));
    RS(cueva_cleanvarint(c));
    RS(cueva_swap1(c));
    for (int l = 1; l <= lmax; l++) {
      for (int m = 1; m < l; m++) RS(cueva_sumint(c, l, m));
      RS(cueva_varint(c, l));
      RS(cueva_electricvarint(c, l));
      RS(cueva_c2p_stage(c, l));
      RS_WAIT();
      RS(cueva_interconvar(c, l));
      RS(cueva_boundary(c, CUEVA_BND_CONS, l));
      RS(cueva_boundary(c, CUEVA_BND_ELEC, l));
      RS(cueva_flowinit(c, l));
      RS(cueva_boundary(c, CUEVA_BND_FLUX, l));
      RS(cueva_flux(c, l));
      RS_WAIT();
      RS(cueva_fluxandsource(c, l));
      RS(cueva_sumfin(c, l));
    }
    RS(cueva_varfin(c));
    RS(cueva_electricfield(c));
    RS(launch_c2p(c, false));
    RS_WAIT();
    c->hstep++;
    RS(cueva_boundary(c, CUEVA_BND_PRIM, 1));
  }
#undef RS_WAIT
#undef RS
  return CUEVA_OK;
}

extern "C" int cueva_step_counters(cueva_ctx* c, int64_t counters[3], int64_t* fallback) {
  if (!c) return CUEVA_E_ARG;
  DevStatus h;
  int rc = read_status(c, &h);
  if (rc) return rc;
  if (counters) for (int k = 0; k < 3; k++) counters[k] = (int64_t)h.cnt[k];
  if (fallback) *fallback = (int64_t)h.fallback;
  if ((rc = clear_counters(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return c->err;
}

extern "C" int cueva_last_error(cueva_ctx* c, int32_t where[4]) {
  if (!c) return CUEVA_E_ARG;
  DevStatus h;
  int rc = read_status(c, &h);
  if (rc) return rc;
  if (where)
    for (int k = 0; k < 4; k++) where[k] = c->err ? c->err_where[k] : 0;
  return c->err;
}

extern "C" int cueva_set_step_index(cueva_ctx* c, int64_t h) {
  if (!c || h < 1 || h > INT32_MAX) return CUEVA_E_ARG;
  c->hstep = h;
  return CUEVA_OK;
}

extern "C" int cueva_fixup_records(cueva_ctx* c, int kind, int32_t* hij, double* vals, int64_t cap, int64_t* n) {
  if (!c || kind < 0 || kind > 2 || !n || cap < 0 || (cap > 0 && (!hij || !vals))) return CUEVA_E_ARG;
  DevStatus h;
  int rc = read_status(c, &h);
  if (rc) return rc;
  const int64_t total = (int64_t)h.nrec[kind];
  *n = total;
  if (total > FIX_CAP) return CUEVA_E_NOMEM;  // more than the device kept
  std::vector<FixRec> r((size_t)total);
  if (total) {
    HIPCHK(hipMemcpyAsync(r.data(), c->rec + (size_t)kind * FIX_CAP, (size_t)total * sizeof(FixRec),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  // the reference's order: steps in turn; within a step the stage recoveries
  // (REC_PRIM = 1) in stage order, then the one after varfin; within a
  // recovery its loop order (j outer, i inner)
  const int after = c->S.lmax + 1;
  std::sort(r.begin(), r.end(), [after](const FixRec& a, const FixRec& b) {
    if (a.h != b.h) return a.h < b.h;
    const int la = a.l ? a.l : after, lb = b.l ? b.l : after;
    if (la != lb) return la < lb;
    if (a.j != b.j) return a.j < b.j;
    return a.i < b.i;
  });
  for (int64_t e = 0; e < total && e < cap; e++) {
    hij[3 * e] = r[e].h;
    hij[3 * e + 1] = r[e].i;
    hij[3 * e + 2] = r[e].j;
    vals[2 * e] = r[e].a;
    vals[2 * e + 1] = r[e].b;
  }
  return CUEVA_OK;
}

extern "C" int cueva_upload_cons(cueva_ctx* c, const double* cons5) {
  if (!c || !cons5) return CUEVA_E_ARG;
  return copy_planes_h2d(c, c->P.c, 5, cons5);
}

// device plane of (group, index) as cueva_download_plane numbers them, or null
static double* plane_of(cueva_ctx* c, int group, int index) {
  if (index < 0) return nullptr;
  if (group == 0 && index < P_N) return c->P.p[index];
  if (group == 1 && index < C_N) return c->P.c[index];
  if (group == 2 && index < S_N) return c->P.s[index];
  if (group == 3 && index < FX_N) return c->P.fx[index];
  if (group == 4 && index < FY_N) return c->P.fy[index];
  if (group >= 5 && group < 5 + MAXL && index < R_N) return c->P.rhs[group - 5][index];
  return nullptr;
}

extern "C" int cueva_plane_range(cueva_ctx* c, int group, int index, double out[2]) {
  if (!c || !out) return CUEVA_E_ARG;
  const double* a = plane_of(c, group, index);
  if (!a) return CUEVA_E_ARG;
  double* res = c->red + (size_t)RED_BLOCKS * G_N;
  LAUNCH(K_OTHER, k_range_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), c->g, a, c->red);
  LAUNCH(K_OTHER, k_range_final, dim3(1), dim3(RED_THREADS), (const double*)c->red, res);
  HIPCHK(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CUEVA_OK;
}

extern "C" int cueva_download_plane(cueva_ctx* c, int group, int index, double* out) {
  if (!c || !out) return CUEVA_E_ARG;
  double* src = plane_of(c, group, index);
  if (!src) return CUEVA_E_ARG;
  int rc = copy_planes_d2h(c, &src, 1, out);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return CUEVA_OK;
}

extern "C" int cueva_set_profiling(cueva_ctx* c, int on) {
  if (!c) return CUEVA_E_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->prof = on ? 1 : 0;
  c->nev = 0;
  return CUEVA_OK;
}

extern "C" int cueva_kernel_times(cueva_ctx* c, double ms[6], int64_t launches[6]) {
  if (!c || !ms || !launches) return CUEVA_E_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < K_NCLS; k++) { ms[k] = 0.0; launches[k] = 0; }
  for (int e = 0; e < c->nev; e++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev[2 * e], c->ev[2 * e + 1]));
    ms[c->evcls[e]] += (double)t;
    launches[c->evcls[e]] += 1;
  }
  c->nev = 0;
  return CUEVA_OK;
}

extern "C" int cueva_time_steps(cueva_ctx* c, int nsteps, double* ms) {
  if (!c || !ms || nsteps < 0) return CUEVA_E_ARG;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, c->stream));
  int rc = cueva_step(c, nsteps);
  HIPCHK(hipEventRecord(b, c->stream));
  HIPCHK(hipEventSynchronize(b));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, a, b));
  hipEventDestroy(a);
  hipEventDestroy(b);
  *ms = (double)t;
  return rc;
}

extern "C" int cueva_reduce_globals(cueva_ctx* c, double out[37]) {
  if (!c || !out) return CUEVA_E_ARG;
  double* res = c->red + (size_t)RED_BLOCKS * G_N;
  LAUNCH(K_OTHER, k_globals_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), c->g, c->P, c->S, c->red);
  LAUNCH(K_OTHER, k_globals_final, dim3(1), dim3(RED_THREADS), (const double*)c->red, res);
  HIPCHK(hipMemcpyAsync(out, res, G_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CUEVA_OK;
}

// ---------------------------------------------------------------------------
// The global diagnostics without a context and without allocating: the
// block partial sums of k_globals_partial go to a fixed device array of this
// file (one copy per GPU, held until the process ends) instead of the
// context's reduction buffer.  Its size does not depend on the grid (the
// partial kernel strides over any grid with RED_BLOCKS blocks), and every
// call writes all of it before k_globals_final reads it, so nothing needs
// clearing.  Calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[(size_t)RED_BLOCKS * G_N];
constexpr int MAX_DEVICES = 64;
static void* scratch_ptr[MAX_DEVICES];
static std::once_flag scratch_once[MAX_DEVICES];
static hipError_t scratch_err[MAX_DEVICES];

// The device address of this GPU's copy of the array (the array's name is
// only a host-side stand-in), looked up once per device with it current.
static double* scratch_of(int dev) {
  std::call_once(scratch_once[dev], [dev]() {
    void* ptr = nullptr;
    scratch_err[dev] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    scratch_ptr[dev] = scratch_err[dev] == hipSuccess ? ptr : nullptr;
  });
  return (double*)scratch_ptr[dev];
}

extern "C" int cueva_reduce_globals_noalloc(const cueva_params* p, const double* prims14, const double* cons5,
                                            double* out, void* stream) {
  if (!p || !prims14 || !cons5 || !out) return CUEVA_E_ARG;
  int rc = check_params(p);
  if (rc) return rc;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return CUEVA_E_ARG;
  double* partial = scratch_of(dev);
  if (!partial) return CUEVA_E_HIP;
  // the caller's planes in the host layout (-6:imax+6, -6:jmax+6), x fastest:
  // with pitch imax+13, Geom::at(i, j) - 2 is the element's index in a plane
  Geom g;
  memset(&g, 0, sizeof(g));
  g.imax = p->imax;
  g.jmax = p->jmax;
  g.pitch = p->imax + 13;
  g.rows = p->jmax + 13;
  g.plane = (int64_t)g.pitch * g.rows;
  g.bcx = p->bc_x;
  g.bcy = p->bc_y;
  Planes P;
  memset(&P, 0, sizeof(P));
  for (int k = 0; k < P_N; k++) P.p[k] = (double*)(prims14 + (size_t)k * (size_t)g.plane) - 2;
  for (int k = 0; k < 5; k++) P.c[k] = (double*)(cons5 + (size_t)k * (size_t)g.plane) - 2;
  Scheme S;
  memset(&S, 0, sizeof(S));
  double d[5];
  cueva_derived(p, d);
  S.dx = d[0];
  S.dy = d[1];
  S.gamma = p->gamma;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_globals_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, g, P, S, partial);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_globals_final, dim3(1), dim3(RED_THREADS), 0, s, (const double*)partial, out);
  HIPCHK(hipGetLastError());
  return CUEVA_OK;
}

extern "C" int cueva_max_speed(cueva_ctx* c, double* vmax) {
  if (!c || !vmax) return CUEVA_E_ARG;
  double* res = c->red + (size_t)RED_BLOCKS * G_N;
  LAUNCH(K_OTHER, k_speed_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), c->g, c->P, c->S, c->red);
  LAUNCH(K_OTHER, k_speed_final, dim3(1), dim3(RED_THREADS), (const double*)c->red, res);
  HIPCHK(hipMemcpyAsync(vmax, res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CUEVA_OK;
}
