// column_kernels.hip -- the column-mean forcing modules on gfx950:
//   modules::sponge_layer              (model/modules/sponge_layer.h:7-60)
//   modules::geostrophic_wind_forcing  (model/modules/geostrophic_wind_forcing.h:7-36)
//   the reductions and the uniform update of modules::uniform_pg_wind_forcing and
//   uniform_pg_wind_forcing_height     (model/modules/uniform_pg_wind_forcing.h:5-80)
//
// Horizontal means per level are deterministic two-pass reductions.  Pass 1 gives every
// (field, level, chunk) one 256-thread block that sums its 1/COL_CHUNKS share of the level (each
// thread a strided run, then a butterfly inside the wave and a fixed pairing of the four waves);
// pass 2 adds a level's chunk sums in chunk order and scales by 1/(nx ny).  The means stay on the
// device for the sponge and the geostrophic forcing; the uniform pressure-gradient forcing takes
// them to the host, where the reference keeps them as coupler options.  The relaxations are one
// streaming pass over the affected levels.
#include "pu_common.h"
#include <cmath>
#include <mutex>

namespace pu {

constexpr int COL_CHUNKS = 16, COL_MAXF = 3;

struct ColFields {
  const double *f[COL_MAXF];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// grid (COL_CHUNKS, nlev, nfld): partial sums of field f over level k0 + blockIdx.y
__global__ __launch_bounds__(256) void k_level_partial(ColFields F, long long plane, int k0, int nlev, double *part) {
  const int chunk = blockIdx.x, kl = blockIdx.y, f = blockIdx.z;
  const long long b = plane * chunk / COL_CHUNKS, e = plane * (chunk + 1) / COL_CHUNKS;
  const double *base = f == 0 ? F.f[0] : (f == 1 ? F.f[1] : F.f[2]);
  const double *p = base + (long long)(k0 + kl) * plane;
  double s = 0;
#pragma unroll 4
  for (long long id = b + threadIdx.x; id < e; id += 256) s += p[id];
  s = wave_sum(s);
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[((long long)f * nlev + kl) * COL_CHUNKS + chunk] = (sw[0] + sw[1]) + (sw[2] + sw[3]);
}

// one thread per (field, level): chunk sums in order, times 1/(nx ny)
__global__ __launch_bounds__(256) void k_level_final(const double *part, int nrows, double scale, double *col) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nrows) return;
  double s = 0;
#pragma unroll
  for (int c = 0; c < COL_CHUNKS; c++) s += part[(long long)r * COL_CHUNKS + c];
  col[r] = s * scale;
}

struct SpongeArgs {
  double *u, *v, *w, *T;
  const double *col;            // [3][nlev]: means of u, v, T over levels k1 ..
  long long plane;
  int k1, nlev;
  double dz, z1, z2, dt, time_scale;
};

// grid (ceil(plane/256), nlev): relaxation of the levels above z1 towards their means, w to 0
// (sponge_layer.h:48-58)
__global__ __launch_bounds__(256) void k_sponge(SpongeArgs A) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  const int kl = blockIdx.y, k = A.k1 + kl;
  if (id >= A.plane) return;
  const double z = (k + 0.5) * A.dz;
  if (!(z > A.z1)) return;
  const double factor = pow((z - A.z1) / (A.z2 - A.z1), 3.0) * A.dt / A.time_scale;
  const long long g = (long long)k * A.plane + id;
  A.u[g] = factor * A.col[kl] + (1 - factor) * A.u[g];
  A.v[g] = factor * A.col[A.nlev + kl] + (1 - factor) * A.v[g];
  A.w[g] = factor * 0 + (1 - factor) * A.w[g];
  A.T[g] = factor * A.col[2 * A.nlev + kl] + (1 - factor) * A.T[g];
}

// grid (ceil(plane/256), nz): u += dt f (<v>(k) - v_g), v += -dt f (<u>(k) - u_g)
// (geostrophic_wind_forcing.h:31-34)
__global__ __launch_bounds__(256) void k_geostrophic(double *u, double *v, const double *col, int nz, long long plane,
                                                     double dt, double fcor, double u_g, double v_g) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (id >= plane) return;
  const long long g = (long long)k * plane + id;
  u[g] += dt * (fcor * (col[nz + k] - v_g));
  v[g] += dt * (-fcor * (col[k] - u_g));
}

__global__ __launch_bounds__(256) void k_add_wind(double *u, double *v, long long n3, double du, double dv) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= n3) return;
  u[id] += du;
  v[id] += dv;
}

// Horizontal means of nfld fields over levels [k0, k0 + nlev) into ctx->d_col ([nfld][nlev]).
static int level_means(pu_ctx *ctx, ColFields F, int nfld, int k0, int nlev) {
  const pu_config &c = ctx->cfg;
  const long long plane = (long long)c.nx * c.ny;
  if (!ctx->d_col) PU_HIP_CHECK(hipMalloc(&ctx->d_col, sizeof(double) * (size_t)COL_MAXF * c.nz * (COL_CHUNKS + 1)));
  double *part = ctx->d_col + (size_t)COL_MAXF * c.nz;
  hipLaunchKernelGGL(k_level_partial, dim3(COL_CHUNKS, nlev, nfld), dim3(256), 0, ctx->stream, F, plane, k0, nlev, part);
  PU_HIP_CHECK(hipGetLastError());
  const int nrows = nfld * nlev;
  hipLaunchKernelGGL(k_level_final, dim3((nrows + 255) / 256), dim3(256), 0, ctx->stream, part, nrows,
                     1.0 / ((double)c.nx * c.ny), ctx->d_col);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

// Partial sums of the allocation-free level means (pu_level_mean_wind_noalloc): [2][max levels]
// [COL_CHUNKS] doubles, 256 KiB on every GPU the library runs on, held until the process ends.
// In host code this name is a host-side stand-in, not a device address: kernels get the address
// hipGetSymbolAddress returns for the current device (scratch_device).  k_level_partial writes
// every slot it is asked for, so nothing is cleared between calls; calls that overlap on one GPU
// share the array and race.
constexpr int NOALLOC_SLOTS = 2 * PU_NOALLOC_MAX_LEVELS * COL_CHUNKS;
static __device__ __align__(256) double scratch[NOALLOC_SLOTS];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the context's device, looked up once per device with that device
// current (the lookup loads the code object; later calls only read the table).
static int scratch_device(const pu_ctx *ctx, double **out) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= MAX_DEVICES) return fail(PU_ERR_ARG, "scratch: device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    PU_HIP_CHECK(hipGetDevice(&cur));
    if (cur != dev) PU_HIP_CHECK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) PU_HIP_CHECK(hipSetDevice(cur));
    if (e != hipSuccess || !p) return fail(PU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PU_OK;
}

}  // namespace pu

using namespace pu;

extern "C" {

int pu_sponge_layer(pu_ctx *ctx, pu_coupler_state *cs, double dt, double time_scale, double top_prop) {
  if (!ctx || !cs || !cs->uvel || !cs->vvel || !cs->wvel || !cs->temp) return fail(PU_ERR_ARG, "pu_sponge_layer: null argument");
  if (!(time_scale > 0) || !(top_prop > 0) || !(top_prop <= 1))
    return fail(PU_ERR_ARG, "pu_sponge_layer: time_scale must be > 0 and top_prop in (0, 1]");
  const pu_config &c = ctx->cfg;
  const double dz = c.zlen / c.nz;
  const double z1 = (1 - top_prop) * c.zlen, z2 = c.zlen;
  // first level of the reference's column (sponge_layer.h:34); a negative index would address
  // below the domain and holds no cell with z > z1, so it starts at level 0 here
  int k1 = (int)std::floor(z1 / dz - 0.5);
  if (k1 < 0) k1 = 0;
  if (k1 >= c.nz) return PU_OK;
  const int nlev = c.nz - k1;
  ColFields F{{cs->uvel, cs->vvel, cs->temp}};
  int rc = level_means(ctx, F, 3, k1, nlev);
  if (rc != PU_OK) return rc;
  SpongeArgs A;
  A.u = cs->uvel; A.v = cs->vvel; A.w = cs->wvel; A.T = cs->temp;
  A.col = ctx->d_col;
  A.plane = (long long)c.nx * c.ny;
  A.k1 = k1; A.nlev = nlev;
  A.dz = dz; A.z1 = z1; A.z2 = z2; A.dt = dt; A.time_scale = time_scale;
  {
    LaunchTimer lt(ctx, "sponge");
    hipLaunchKernelGGL(k_sponge, dim3((unsigned)((A.plane + 255) / 256), nlev), dim3(256), 0, ctx->stream, A);
    PU_HIP_CHECK(hipGetLastError());
  }
  if (ctx->kernel_timing) collect_timers(ctx);
  return PU_OK;
}

int pu_geostrophic_wind_forcing(pu_ctx *ctx, pu_coupler_state *cs, double dt, double lat_g, double u_g, double v_g) {
  if (!ctx || !cs || !cs->uvel || !cs->vvel) return fail(PU_ERR_ARG, "pu_geostrophic_wind_forcing: null argument");
  const pu_config &c = ctx->cfg;
  const double fcor = 2 * 7.2921e-5 * std::sin(lat_g / 180 * M_PI);
  ColFields F{{cs->uvel, cs->vvel, nullptr}};
  int rc = level_means(ctx, F, 2, 0, c.nz);
  if (rc != PU_OK) return rc;
  const long long plane = (long long)c.nx * c.ny;
  {
    LaunchTimer lt(ctx, "geostrophic");
    hipLaunchKernelGGL(k_geostrophic, dim3((unsigned)((plane + 255) / 256), c.nz), dim3(256), 0, ctx->stream, cs->uvel,
                       cs->vvel, (const double *)ctx->d_col, c.nz, plane, dt, fcor, u_g, v_g);
    PU_HIP_CHECK(hipGetLastError());
  }
  if (ctx->kernel_timing) collect_timers(ctx);
  return PU_OK;
}

int pu_level_mean_wind(pu_ctx *ctx, const pu_coupler_state *cs, int k_begin, int k_end, double *u_mean, double *v_mean) {
  if (!ctx || !cs || !cs->uvel || !cs->vvel || !u_mean || !v_mean) return fail(PU_ERR_ARG, "pu_level_mean_wind: null argument");
  const pu_config &c = ctx->cfg;
  if (k_begin < 0 || k_end > c.nz || k_begin >= k_end)
    return fail(PU_ERR_ARG, "pu_level_mean_wind: levels [" + std::to_string(k_begin) + ", " + std::to_string(k_end) +
                                ") outside [0, " + std::to_string(c.nz) + ")");
  const int nlev = k_end - k_begin;
  ColFields F{{cs->uvel, cs->vvel, nullptr}};
  int rc = level_means(ctx, F, 2, k_begin, nlev);
  if (rc != PU_OK) return rc;
  PU_HIP_CHECK(hipMemcpyAsync(u_mean, ctx->d_col, sizeof(double) * nlev, hipMemcpyDeviceToHost, ctx->stream));
  PU_HIP_CHECK(hipMemcpyAsync(v_mean, ctx->d_col + nlev, sizeof(double) * nlev, hipMemcpyDeviceToHost, ctx->stream));
  PU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return PU_OK;
}

int pu_level_mean_wind_noalloc(pu_ctx *ctx, const pu_coupler_state *cs, int k_begin, int k_end, double *u_mean,
                               double *v_mean) {
  if (!ctx || !cs || !cs->uvel || !cs->vvel || !u_mean || !v_mean)
    return fail(PU_ERR_ARG, "pu_level_mean_wind_noalloc: null argument");
  const pu_config &c = ctx->cfg;
  if (k_begin < 0 || k_end > c.nz || k_begin >= k_end)
    return fail(PU_ERR_ARG, "pu_level_mean_wind_noalloc: levels [" + std::to_string(k_begin) + ", " +
                                std::to_string(k_end) + ") outside [0, " + std::to_string(c.nz) + ")");
  const int nlev = k_end - k_begin;
  if (nlev > PU_NOALLOC_MAX_LEVELS)
    return fail(PU_ERR_ARG, "pu_level_mean_wind_noalloc: " + std::to_string(nlev) + " levels exceed the maximum of " +
                                std::to_string(PU_NOALLOC_MAX_LEVELS) + " (PU_NOALLOC_MAX_LEVELS)");
  double *part = nullptr;
  int rc = scratch_device(ctx, &part);
  if (rc != PU_OK) return rc;
  // the kernels and sum order of pu_level_mean_wind: partial sums per (field, level, chunk), then
  // each level's chunks in order; field 0 (u) goes to u_mean, field 1 (v) to v_mean
  const long long plane = (long long)c.nx * c.ny;
  ColFields F{{cs->uvel, cs->vvel, nullptr}};
  hipLaunchKernelGGL(k_level_partial, dim3(COL_CHUNKS, nlev, 2), dim3(256), 0, ctx->stream, F, plane, k_begin, nlev,
                     part);
  PU_HIP_CHECK(hipGetLastError());
  const double scale = 1.0 / ((double)c.nx * c.ny);
  hipLaunchKernelGGL(k_level_final, dim3((nlev + 255) / 256), dim3(256), 0, ctx->stream, (const double *)part, nlev,
                     scale, u_mean);
  PU_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_level_final, dim3((nlev + 255) / 256), dim3(256), 0, ctx->stream,
                     (const double *)(part + (size_t)nlev * COL_CHUNKS), nlev, scale, v_mean);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int pu_add_uniform_wind(pu_ctx *ctx, pu_coupler_state *cs, double du, double dv) {
  if (!ctx || !cs || !cs->uvel || !cs->vvel) return fail(PU_ERR_ARG, "pu_add_uniform_wind: null argument");
  const long long n3 = ctx->g.n3();
  hipLaunchKernelGGL(k_add_wind, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, ctx->stream, cs->uvel, cs->vvel, n3,
                     du, dv);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

}  // extern "C"
