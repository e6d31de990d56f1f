// sweep_seg_x_f32.hip -- the FP32 segment x sweep (k_sweep<ORD, float, 0, MODE>, sweep_core.h) in a
// translation unit of its own, built with -mllvm -disable-machine-licm (porturb_amd/build.py): with
// the flat-stencil branch of the limiter it then fits 72 VGPRs without spills, 7 waves per SIMD
// (71 VGPRs; with loop-invariant code motion on it needs 84, 5 waves, or spills at the cap).
#include "sweep_core.h"

namespace pu {

template <int ORD>
static void launch_ord(int mode, dim3 grid, dim3 block, hipStream_t s, const float *prim, const uint8_t *immf,
                       const uint8_t *any10, float *tend, const SweepParams &P, const FuseArgs<float> &FA) {
  if (mode == 2)
    hipLaunchKernelGGL((k_sweep<ORD, float, 0, 2>), grid, block, 0, s, prim, immf, any10, tend, P, FA);
  else if (mode == 1)
    hipLaunchKernelGGL((k_sweep<ORD, float, 0, 1>), grid, block, 0, s, prim, immf, any10, tend, P, FA);
  else
    hipLaunchKernelGGL((k_sweep<ORD, float, 0, 0>), grid, block, 0, s, prim, immf, any10, tend, P, FA);
}

int launch_seg_x_f32(int ord, int mode, dim3 grid, dim3 block, hipStream_t s, const float *prim,
                     const uint8_t *immf, const uint8_t *any10, float *tend, const SweepParams &P,
                     const FuseArgs<float> &FA) {
  if (mode < 0 || mode > 2) return fail(PU_ERR_ARG, "sweep mode must be 0, 1 or 2");
  switch (ord) {
    case 3: launch_ord<3>(mode, grid, block, s, prim, immf, any10, tend, P, FA); break;
    case 5: launch_ord<5>(mode, grid, block, s, prim, immf, any10, tend, P, FA); break;
    case 7: launch_ord<7>(mode, grid, block, s, prim, immf, any10, tend, P, FA); break;
    case 9: launch_ord<9>(mode, grid, block, s, prim, immf, any10, tend, P, FA); break;
    case 11: launch_ord<11>(mode, grid, block, s, prim, immf, any10, tend, P, FA); break;
    default: return fail(PU_ERR_UNSUPPORTED, "unsupported ord");
  }
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

}  // namespace pu
