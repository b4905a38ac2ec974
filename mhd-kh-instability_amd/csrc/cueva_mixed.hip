// FP32 face kernels of the mixed form: reconstruction, Riemann solve and
// flux differences in FP32, built in their own translation unit where a
// product may fuse into the sum of the same expression (the FP64 path is
// built without contraction, as the reference build evaluates products and
// sums separately; the FP64 sources and update compiled here turn it off in
// their own bodies).
#include <hip/hip_runtime.h>

#include "cueva_faces.hpp"

namespace cueva {

template <int ORD>
static hipError_t launch_ord(dim3 grid, dim3 blk, hipStream_t s, const Geom& g, const Planes& P, const Scheme& S,
                             const StatusRef& st) {
  if (S.flux_solver == 1)
    hipLaunchKernelGGL((k_faces<ORD, 1, float>), grid, blk, 0, s, g, P, S, st);
  else if (S.flux_solver == 4)
    hipLaunchKernelGGL((k_faces<ORD, 4, float>), grid, blk, 0, s, g, P, S, st);
  else if (S.flux_solver == 3)
    hipLaunchKernelGGL((k_faces<ORD, 3, float>), grid, blk, 0, s, g, P, S, st);
  else if (S.flux_solver == 2)
    hipLaunchKernelGGL((k_faces<ORD, 2, float>), grid, blk, 0, s, g, P, S, st);
  else if (S.flux_solver == 7)
    hipLaunchKernelGGL((k_faces<ORD, 7, float>), grid, blk, 0, s, g, P, S, st);
  else if (S.flux_solver == 6)
    hipLaunchKernelGGL((k_faces<ORD, 6, float>), grid, blk, 0, s, g, P, S, st);
  else
    hipLaunchKernelGGL((k_faces<ORD, 5, float>), grid, blk, 0, s, g, P, S, st);
  return hipGetLastError();
}

// one launch of the FP32 face kernel over the face pairs [-1,imax] x [-1,jmax]
hipError_t launch_faces_fp32(int rec_order, dim3 grid, dim3 blk, hipStream_t s, const Geom& g, const Planes& P,
                             const Scheme& S, const StatusRef& st) {
  if (S.flux_solver >= 1 && S.flux_solver <= 4)
    return launch_ord<5>(grid, blk, s, g, P, S, st);  // the MP order does not enter
  switch (rec_order) {
    case 5: return launch_ord<5>(grid, blk, s, g, P, S, st);
    case 7: return launch_ord<7>(grid, blk, s, g, P, S, st);
    default: return launch_ord<9>(grid, blk, s, g, P, S, st);
  }
}

// the FP32 marching face kernel of (rec_order, solver, stage kind)
const void* march_kernel_fp32(int rec_order, int solver, bool final_stage, bool c2p) {
  return march_kernel<float>(rec_order, solver, final_stage, c2p);
}

}  // namespace cueva
