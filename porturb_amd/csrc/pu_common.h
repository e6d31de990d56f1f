// pu_common.h -- internal definitions shared by the HIP kernels and the C ABI of libporturb.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <type_traits>
#include <string>
#include <map>
#include <vector>
#include "porturb.h"

namespace pu {

// Conserved-state indices (dynamics_rk.h:33-38) and the pressure slot of the reconstructed
// fields (dynamics_rk.h:496).
constexpr int idR = 0, idU = 1, idV = 2, idW = 3, idT = 4, idP = 5, NUM_STATE = 5;

// Reconstruction halo of the primitive-field array: hs cells for the stencils of interior cells
// plus one more for the neighbour-cell reconstruction each line segment needs at its ends.
__host__ __device__ constexpr int prim_halo(int hs) { return hs + 1; }

// x position of interior cell 0 in the primitive-field and flag arrays, and the alignment of their
// rows (elements): rows start on 128-byte boundaries for float and double and the interior of
// every row starts on one, so the row segments the kernels read and write map onto whole cache
// lines.  PRIM_OX >= prim_halo(hs) for every order.
constexpr int PRIM_OX = 32, PRIM_ROW_ALIGN = 32;

struct Grid {
  int nx, ny, nz, hs, hp;          // hp = prim_halo(hs)
  int ox;                          // x offset of the interior (PRIM_OX); y/z offsets are hp
  long long nxh, nyh, nzh;         // extents of the primitive-field array (nxh: padded row pitch)
  long long n3() const { return (long long)nx * ny * nz; }
  long long nh() const { return nxh * nyh * nzh; }
};

// Timing of one kernel family with HIP events around each launch.
struct KernelTimer {
  double total_ms = 0;
  long long count = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

}  // namespace pu

struct pu_ctx {
  pu_config cfg;
  hipStream_t stream = nullptr;
  int device = 0;                  // the HIP device current when the context was created
  pu::Grid g{};
  int nv = 0;                      // 5 + num_tracers  (conserved variables)
  int nf = 0;                      // nv + 1           (reconstructed fields incl. pressure)
  bool dycore_ready = false;
  bool les_ready = false;
  // sweeps through the segment kernel (environment variable PORTURB_SEGMENT_SWEEPS set when the
  // context is created): the whole-line kernels are checked against it
  bool segment_sweeps[3] = {false, false, false};   // x, y, z (PORTURB_SEGMENT_DIRS narrows it)
  // hydrostatic background, FP64 (host + device copies)
  std::vector<double> hy_cells[3];     // dens, theta, pressure: [nz+2hs]
  std::vector<double> hy_edges[3];     // [nz+1]
  double *d_hy = nullptr;              // device: [3*(nz+2hs)] cells then [3*(nz+1)] edges
  // immersed data
  uint8_t *d_imm_flag = nullptr;       // (prop_halos > 0), halo'd with the primitive halo
  uint8_t *d_any10 = nullptr;          // dycore_any_immersed10 [n3]
  int     *d_tau = nullptr;            // dycore_immersed_tau [n3]
  double  *d_relax = nullptr;          // pow(prop, immersed_power)/tau [n3]
  uint8_t *d_imm_les = nullptr;        // immersed_proportion_halos (hs=1) as 0 / partial / 1 codes
  double  *d_imm_les_val = nullptr;    // immersed_proportion_halos values [nz+2][ny+2][nx+2]
  // dycore work arrays (element type follows cfg.precision)
  void *d_U = nullptr, *d_Us = nullptr;        // [nv][n3]
  void *d_Ulo = nullptr, *d_Uslo = nullptr;    // FP32 only: compensation of rho, rho*theta [2][n3]
  void *d_tend = nullptr;                      // [nv][n3]
  void *d_prim = nullptr;                      // [nf][nzh][nyh][nxh]
  // LES work arrays
  double *d_les = nullptr;                     // primitive closure state, halo 1
  // learned subgrid closure (attached network, its per-cell outputs and the static mask of cells
  // whose level and 3x3x3 neighbourhood allow it)
  const pu_mlp *sgs_model = nullptr;
  float   *d_sgs_tau = nullptr;                // [3][n3][4]: per face direction d, (tau_0d, tau_1d, tau_2d, q_d)
  uint8_t *d_sgs_flags = nullptr;              // [n3]
  uint8_t *d_sgs_valid = nullptr;              // [n3]
  float   *d_sgs_fmul = nullptr;               // [108]
  bool     sgs_valid_ready = false;
  // surface fluxes: halo'd immersed roughness and temperature boxes
  double  *d_sfc_rough = nullptr, *d_sfc_temp = nullptr;
  bool     sfc_ready = false;
  // horizontal means per level (forcing modules): [3][nz] means + [3][nz][16] partial sums
  double  *d_col = nullptr;
  // reductions
  double *d_red = nullptr;                     // scratch for reductions
  // timing
  hipEvent_t t0 = nullptr, t1 = nullptr;
  bool kernel_timing = false;
  std::map<std::string, pu::KernelTimer> timers;
};

namespace pu {

void set_error(const std::string &msg);
int  fail(int code, const std::string &msg);

#define PU_HIP_CHECK(expr)                                                                  \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      return ::pu::fail(PU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

// XCD-aware block order.  The dispatcher hands workgroups to the 8 XCDs of an MI355X round-robin
// by linear workgroup id, and each XCD has its own L2; blocks that share cache lines (x-neighbours)
// should therefore run on the same XCD.  This maps the hardware id so that consecutive logical
// ids land on one XCD: hardware id h (XCD h % 8, slot h / 8) -> logical id
// start(XCD) + slot, a bijection on [0, nblocks).
__device__ __forceinline__ unsigned xcd_swizzle(unsigned h, unsigned nblocks) {
  const unsigned nx = 8, q = nblocks / nx, r = nblocks % nx;
  const unsigned xcd = h % nx, slot = h / nx;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Logical 3-D block index of the current workgroup after the swizzle (x fastest).
__device__ __forceinline__ void swizzled_block(unsigned &bx, unsigned &by, unsigned &bz) {
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned h = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned l = xcd_swizzle(h, nb);
  bx = l % gridDim.x;
  by = (l / gridDim.x) % gridDim.y;
  bz = l / (gridDim.x * gridDim.y);
}

// Reciprocal of a positive, normal argument: v_rcp_f32 in FP32; in FP64 v_rcp_f64 refined by two
// Newton steps (about 1 ulp; every call site divides by a positive quantity >= 1e-20).
template <class T> __device__ __forceinline__ T recip(T x) {
  if constexpr (std::is_same<T, float>::value) {
    return __builtin_amdgcn_rcpf(x);
  } else {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
  }
}

// Tracer pointer bundles passed by value to kernels.
struct TracerPtrs { double *p[PU_MAX_TRACERS]; };
struct TracerPtrsC { const double *p[PU_MAX_TRACERS]; };

// Kernel-timing helpers: record events around a launch when timing is enabled.
struct LaunchTimer {
  pu_ctx *ctx; const char *name; hipEvent_t a = nullptr, b = nullptr;
  LaunchTimer(pu_ctx *c, const char *n) : ctx(c), name(n) {
    if (ctx->kernel_timing) {
      hipEventCreate(&a); hipEventCreate(&b);
      hipEventRecord(a, ctx->stream);
    }
  }
  ~LaunchTimer() {
    if (ctx->kernel_timing) {
      hipEventRecord(b, ctx->stream);
      ctx->timers[name].pending.emplace_back(a, b);
    }
  }
};

void collect_timers(pu_ctx *ctx);

// Launchers implemented in the kernel translation units.
int launch_prep(pu_ctx *ctx, const void *U, const void *Ulo, hipStream_t s);
int launch_c2d_prep(pu_ctx *ctx, const pu_coupler_state *cs, hipStream_t s);
// Sweeps z, y, x.  stage < 0: tendencies into d_tend; stage 0..2: the x sweep also applies the
// stage combination (sources, SSPRK3 stage, clip, nrelax immersed relaxations) for every cell it
// finishes
int launch_sweeps(pu_ctx *ctx, hipStream_t s, int stage = -1, double dt = 0., int nrelax = 0);
int launch_convert_c2d(pu_ctx *ctx, const pu_coupler_state *cs, int relax, hipStream_t s);
int launch_convert_d2c(pu_ctx *ctx, pu_coupler_state *cs, hipStream_t s);
int launch_tend_export(pu_ctx *ctx, double *tend, hipStream_t s);
int launch_census(pu_ctx *ctx, hipStream_t s, unsigned long long *counts);
int dycore_init_device(pu_ctx *ctx, const pu_coupler_state *cs);
int les_apply_device(pu_ctx *ctx, pu_coupler_state *cs, double dt);
int les_prim_device(pu_ctx *ctx, const pu_coupler_state *cs);   // closure box d_les only
// learned closure: prediction and flags into ctx->d_sgs_tau / d_sgs_flags from ctx->d_les
int sgs_infer_device(pu_ctx *ctx, double dt);
int les_init_device(pu_ctx *ctx, const pu_coupler_state *cs);
int fill_device(double *ptr, double value, size_t count);
int copy_device(void *dst, const void *src, size_t bytes);
// k_copy of bytes (a multiple of 16, 16-byte aligned pointers) on stream s, not waited for
int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
int les_set_immersed_halos_device(pu_ctx *ctx, const double *halos);
int diagnostics_device(pu_ctx *ctx, const pu_coupler_state *cs, double *mass, double *maxw, double *maxabsw);
int case_init_device(const pu_config *cfg, int case_id, unsigned long long seed, pu_coupler_state *cs,
                     double *imm, hipStream_t s);

}  // namespace pu
