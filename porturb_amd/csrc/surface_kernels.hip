// surface_kernels.hip -- modules::apply_surface_fluxes (model/modules/surface_flux.h:9-126) on gfx950.
//
// Log-law drag against every immersed neighbour of a cell (the ground below the first level and
// the buildings): for a neighbour across a face normal to direction d with roughness z0,
// c = kappa^2 / ln^2((delta_d/2 + z0)/z0), kappa = 0.4, and the two tangential velocities u_t get
// du_t/dt = -c u_t |u_tangential| / delta_d; the temperature relaxes towards the neighbour's
// temperature T0 the same way when T0 != 0.  Forward Euler on u, v, w, T of the coupler.
//
// The halo'd roughness/temperature boxes are built once (pu_surface_fluxes_init) the way
// custom_modules::sc_init builds immersed_{roughness,temp}_halos (sc_init.h:728-767): x/y periodic,
// the bottom halo takes the surface roughness/temperature, the top halo 0; the immersed proportion
// box is the one the LES closure uses (bottom 1, top 0).
#include "pu_common.h"
#include <cmath>

namespace pu {

__device__ __forceinline__ int wraps(int i, int n) { i %= n; return i < 0 ? i + n : i; }

// halo'd box [nz+2][ny+2][nx+2] of a 3-D field with the surface value below and 0 above
__global__ __launch_bounds__(256) void k_sfc_halo(const double *__restrict__ f3, const double *__restrict__ f2, int nx,
                                                  int ny, int nz, double *__restrict__ out) {
  const long long nxh = nx + 2, nyh = ny + 2, nh = nxh * nyh * (nz + 2);
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= nh) return;
  const int ih = (int)(id % nxh), jh = (int)((id / nxh) % nyh), kh = (int)(id / (nxh * nyh));
  const int i = wraps(ih - 1, nx), j = wraps(jh - 1, ny), k = kh - 1;
  double v;
  if (k < 0) v = f2[(long long)j * nx + i];
  else if (k >= nz) v = 0.;
  else v = f3[((long long)k * ny + j) * nx + i];
  out[id] = v;
}

struct SfcParams {
  int nx, ny, nz;
  double dx, dy, dz, dt;
  const double *prop, *rough, *temp;    // halo'd boxes
};

// grid: (blocks over one level's nx*ny cells, nz levels): no 64-bit division for the indices
__global__ __launch_bounds__(256) void k_surface_fluxes(double *uu, double *vv, double *ww, double *tt, SfcParams P) {
  const int nxy = P.nx * P.ny;
  const int p = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
  if (p >= nxy) return;
  const int j = p / P.nx, i = p - j * P.nx;
  const long long id = (long long)k * nxy + p;
  const long long nxh = P.nx + 2, nyh = P.ny + 2;
  const long long c = ((long long)(k + 1) * nyh + (j + 1)) * nxh + (i + 1);
  // a cell without an immersed face neighbour has zero tendencies: it is left as it is (u + dt*0
  // is u), so most of the domain reads only the immersed proportions and writes nothing
  const long long sz = nxh * nyh;
  if (!(P.prop[c - 1] > 0 || P.prop[c + 1] > 0 || P.prop[c - nxh] > 0 || P.prop[c + nxh] > 0 ||
        P.prop[c - sz] > 0 || P.prop[c + sz] > 0))
    return;
  const double u = uu[id], v = vv[id], w = ww[id], T = tt[id];
  const double vk = 0.40;
  double tu = 0, tv = 0, tw = 0, tT = 0;
  // one immersed neighbour across a face normal to `dir` (0 x, 1 y, 2 z)
  auto face = [&](long long nb, int dir) {
    if (!(P.prop[nb] > 0)) return;
    const double delta = dir == 0 ? P.dx : (dir == 1 ? P.dy : P.dz);
    const double z0 = P.rough[nb];
    const double lg = log((delta / 2 + z0) / z0);
    const double cd = vk * vk / (lg * lg);
    const double T0 = P.temp[nb];
    if (dir == 0) {
      const double mag = sqrt(v * v + w * w);
      tv += -cd * (v - 0) * mag / delta;
      tw += -cd * (w - 0) * mag / delta;
      if (T0 != 0) tT += -cd * (T - T0) * mag / delta;
    } else if (dir == 1) {
      const double mag = sqrt(u * u + w * w);
      tu += -cd * (u - 0) * mag / delta;
      tw += -cd * (w - 0) * mag / delta;
      if (T0 != 0) tT += -cd * (T - T0) * mag / delta;
    } else {
      const double mag = sqrt(u * u + v * v);
      tu += -cd * (u - 0) * mag / delta;
      tv += -cd * (v - 0) * mag / delta;
      if (T0 != 0) tT += -cd * (T - T0) * mag / delta;
    }
  };
  face(c - 1, 0);
  face(c + 1, 0);
  face(c - nxh, 1);
  face(c + nxh, 1);
  face(c - nxh * nyh, 2);
  face(c + nxh * nyh, 2);
  uu[id] = u + P.dt * tu;
  vv[id] = v + P.dt * tv;
  ww[id] = w + P.dt * tw;
  tt[id] = T + P.dt * tT;
}

}  // namespace pu

using namespace pu;

extern "C" {

int pu_surface_fluxes_init(pu_ctx *ctx, const double *immersed_roughness, const double *immersed_temp,
                           const double *surface_roughness, const double *surface_temp) {
  if (!ctx || !immersed_roughness || !immersed_temp || !surface_roughness || !surface_temp)
    return fail(PU_ERR_ARG, "pu_surface_fluxes_init: null argument");
  const pu_config &c = ctx->cfg;
  const long long nh = (long long)(c.nx + 2) * (c.ny + 2) * (c.nz + 2);
  if (!ctx->d_sfc_rough) {
    PU_HIP_CHECK(hipMalloc(&ctx->d_sfc_rough, sizeof(double) * nh));
    PU_HIP_CHECK(hipMalloc(&ctx->d_sfc_temp, sizeof(double) * nh));
  }
  const dim3 grd((unsigned)((nh + 255) / 256));
  hipLaunchKernelGGL(k_sfc_halo, grd, dim3(256), 0, ctx->stream, immersed_roughness, surface_roughness, c.nx, c.ny,
                     c.nz, ctx->d_sfc_rough);
  PU_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sfc_halo, grd, dim3(256), 0, ctx->stream, immersed_temp, surface_temp, c.nx, c.ny, c.nz,
                     ctx->d_sfc_temp);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  ctx->sfc_ready = true;
  return PU_OK;
}

int pu_surface_fluxes_apply(pu_ctx *ctx, pu_coupler_state *cs, double dt_phys) {
  if (!ctx || !cs || !cs->uvel || !cs->vvel || !cs->wvel || !cs->temp) return fail(PU_ERR_ARG, "pu_surface_fluxes_apply: null argument");
  if (!ctx->sfc_ready) return fail(PU_ERR_STATE, "pu_surface_fluxes_init has not been called");
  if (!ctx->dycore_ready && !ctx->les_ready)
    return fail(PU_ERR_STATE, "pu_les_closure_init or pu_dycore_init has not been called (immersed halos)");
  const pu_config &c = ctx->cfg;
  SfcParams P;
  P.nx = c.nx; P.ny = c.ny; P.nz = c.nz;
  P.dx = c.xlen / c.nx; P.dy = c.ylen / c.ny; P.dz = c.zlen / c.nz; P.dt = dt_phys;
  P.prop = ctx->d_imm_les_val; P.rough = ctx->d_sfc_rough; P.temp = ctx->d_sfc_temp;
  {
    LaunchTimer lt(ctx, "surface");
    hipLaunchKernelGGL(k_surface_fluxes, dim3((unsigned)((c.nx * c.ny + 255) / 256), (unsigned)c.nz), dim3(256), 0,
                       ctx->stream, cs->uvel, cs->vvel, cs->wvel, cs->temp, P);
    PU_HIP_CHECK(hipGetLastError());
  }
  if (ctx->kernel_timing) collect_timers(ctx);
  return PU_OK;
}

}  // extern "C"
