// sweep_pline_f32.hip -- the FP32 whole-line y sweep (k_sweep_pline<ORD, float, 1, 1>, sweep_core.h)
// in a translation unit of its own, built with -mllvm -disable-machine-licm (porturb_amd/build.py)
// and at most 72 VGPRs: 7 waves per SIMD like the segment kernel it replaces.
#include "sweep_core.h"

namespace pu {

int launch_pline_y_f32(int ord, dim3 grid, dim3 block, hipStream_t s, const float *prim, const uint8_t *immf,
                       const uint8_t *any10, float *tend, const SweepParams &P, const FuseArgs<float> &FA) {
  switch (ord) {
    case 3: hipLaunchKernelGGL((k_sweep_pline<3, float, 1, 1>), grid, block, 0, s, prim, immf, any10, tend, P, FA); break;
    case 5: hipLaunchKernelGGL((k_sweep_pline<5, float, 1, 1>), grid, block, 0, s, prim, immf, any10, tend, P, FA); break;
    case 7: hipLaunchKernelGGL((k_sweep_pline<7, float, 1, 1>), grid, block, 0, s, prim, immf, any10, tend, P, FA); break;
    case 9: hipLaunchKernelGGL((k_sweep_pline<9, float, 1, 1>), grid, block, 0, s, prim, immf, any10, tend, P, FA); break;
    case 11: hipLaunchKernelGGL((k_sweep_pline<11, float, 1, 1>), grid, block, 0, s, prim, immf, any10, tend, P, FA); break;
    default: return fail(PU_ERR_UNSUPPORTED, "unsupported ord");
  }
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

}  // namespace pu
