// sweep_zwall_f32.hip -- the FP32 whole-column z sweep (k_sweep_zwall<ORD, float>, sweep_core.h) in
// a translation unit of its own, built with -mllvm -disable-machine-licm (porturb_amd/build.py):
// without it the compiler hoists the limiter's constants out of the chunk loop into 91 VGPRs
// (5 waves per SIMD); with it the kernel needs 77 (6 waves) and ran 2.2 ms/step faster at
// 1024x1024x128 WENO-9.
#include "sweep_core.h"

namespace pu {

int launch_zwall_f32(int ord, dim3 grid, dim3 block, hipStream_t s, const float *prim, const uint8_t *immf,
                     const uint8_t *any10, float *tend, const SweepParams &P) {
  switch (ord) {
    case 3: hipLaunchKernelGGL((k_sweep_zwall<3, float>), grid, block, 0, s, prim, immf, any10, tend, P); break;
    case 5: hipLaunchKernelGGL((k_sweep_zwall<5, float>), grid, block, 0, s, prim, immf, any10, tend, P); break;
    case 7: hipLaunchKernelGGL((k_sweep_zwall<7, float>), grid, block, 0, s, prim, immf, any10, tend, P); break;
    case 9: hipLaunchKernelGGL((k_sweep_zwall<9, float>), grid, block, 0, s, prim, immf, any10, tend, P); break;
    case 11: hipLaunchKernelGGL((k_sweep_zwall<11, float>), grid, block, 0, s, prim, immf, any10, tend, P); break;
    default: return fail(PU_ERR_UNSUPPORTED, "unsupported ord");
  }
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

}  // namespace pu
