// init_kernels.hip -- Dynamics_Euler_Stratified_WenoFV::init on the device, FP64 reductions for
// diagnostics, and the seeded synthetic initial states of the benchmark cases.
#include "pu_common.h"
#include <cmath>
#include <vector>

namespace pu {

__device__ __forceinline__ int wrapi_(int i, int n) { i %= n; return i < 0 ? i + n : i; }

// ------------------------------------------------------------------------------------------------
// Hydrostatic column means (dynamics_rk.h:961-984), accumulated in FP64 (the reference uses float
// running sums).  One block per level; deterministic tree reduction.
// ------------------------------------------------------------------------------------------------
struct HydroParams {
  int nx, ny, nz, ntr, idWV;
  double R_d, R_v, C0, gamma;
  unsigned int adds_mass_mask;
};

// Compensated sum (Knuth's TwoSum of every addition, the rounding errors summed apart): the level
// sums below are accurate to about one rounding whatever the order of their terms, so the
// hydrostatic profiles do not depend on how the cells are split over threads (a plain running sum
// of a level's 2^20 equal densities drifts by ~1e-13, which the split into hydrostatic and
// perturbation pressure turns into ~1e-8 Pa).
struct CSum {
  double s = 0, c = 0;
  __device__ void add(double x) {
    const double t = s + x, bb = t - s;
    c += (s - (t - bb)) + (x - bb);
    s = t;
  }
  __device__ void add(const CSum &o) { add(o.s); c += o.c; }
  __device__ double value() const { return s + c; }
};

__global__ __launch_bounds__(256) void k_hydro_means(const double *rd_, const double *tt, TracerPtrsC trc,
                                                     double *out, HydroParams P) {
  __shared__ double sr[2][256], st[2][256], sp[2][256];
  int k = blockIdx.x;
  long long plane = (long long)P.nx * P.ny;
  CSum ar, at, ap;
  for (long long q = threadIdx.x; q < plane; q += blockDim.x) {
    long long id = (long long)k * plane + q;
    double rho_d = rd_[id], temp = tt[id];
    double rho_v = trc.p[P.idWV][id];
    double press = rho_d * P.R_d * temp + rho_v * P.R_v * temp;
    double rho = rho_d;
    for (int tr = 0; tr < P.ntr; tr++) if ((P.adds_mass_mask >> tr) & 1u) rho += trc.p[tr][id];
    double rt = pow(press / P.C0, 1. / P.gamma);         // rho*theta
    ar.add(rho); at.add(rt / rho); ap.add(P.C0 * pow(rt, P.gamma));
  }
  const int t = threadIdx.x;
  sr[0][t] = ar.s; sr[1][t] = ar.c; st[0][t] = at.s; st[1][t] = at.c; sp[0][t] = ap.s; sp[1][t] = ap.c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      CSum a{sr[0][t], sr[1][t]}, b{st[0][t], st[1][t]}, c{sp[0][t], sp[1][t]};
      a.add(CSum{sr[0][t + w], sr[1][t + w]});
      b.add(CSum{st[0][t + w], st[1][t + w]});
      c.add(CSum{sp[0][t + w], sp[1][t + w]});
      sr[0][t] = a.s; sr[1][t] = a.c; st[0][t] = b.s; st[1][t] = b.c; sp[0][t] = c.s; sp[1][t] = c.c;
    }
    __syncthreads();
  }
  if (t == 0) {
    out[k] = CSum{sr[0][0], sr[1][0]}.value();
    out[P.nz + k] = CSum{st[0][0], st[1][0]}.value();
    out[2 * P.nz + k] = CSum{sp[0][0], sp[1][0]}.value();
  }
}

// ------------------------------------------------------------------------------------------------
// Immersed precompute
// ------------------------------------------------------------------------------------------------
struct ImmParams {
  int nx, ny, nz, hs, hp, ox;
  long long nxh, nyh, nzh, n3;
  double power;
};

// dycore_immersed_proportion_halos value at halo'd coordinates with halo hs (dynamics_rk.h:1010-1034):
// x/y periodic, z halos 1.  Coordinates outside the halo'd array are clamped onto it.
__device__ __forceinline__ double imm_halo_value(const double *prop, int kh, int jh, int ih, const ImmParams &P) {
  int nzh = P.nz + 2 * P.hs, nyh = P.ny + 2 * P.hs, nxh = P.nx + 2 * P.hs;
  kh = kh < 0 ? 0 : (kh > nzh - 1 ? nzh - 1 : kh);
  jh = jh < 0 ? 0 : (jh > nyh - 1 ? nyh - 1 : jh);
  ih = ih < 0 ? 0 : (ih > nxh - 1 ? nxh - 1 : ih);
  if (kh < P.hs || kh >= P.hs + P.nz) return 1.;
  return prop[((long long)(kh - P.hs) * P.ny + wrapi_(jh - P.hs, P.ny)) * P.nx + wrapi_(ih - P.hs, P.nx)];
}

// flag array on the primitive halo (hp): prop > 0, z halos immersed
__global__ __launch_bounds__(256) void k_imm_flag(const double *prop, uint8_t *flag, ImmParams P) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long nh = P.nxh * P.nyh * P.nzh;
  if (id >= nh) return;
  int ih = (int)(id % P.nxh), jh = (int)((id / P.nxh) % P.nyh), kh = (int)(id / (P.nxh * P.nyh));
  int k = kh - P.hp, j = jh - P.hp, i = ih - P.ox;
  double v;
  if (k < 0 || k >= P.nz) v = 1.;
  else v = prop[((long long)k * P.ny + wrapi_(j, P.ny)) * P.nx + wrapi_(i, P.nx)];
  flag[id] = v > 0 ? 1 : 0;
}

// immersed tau (dynamics_rk.h:389-425) and relaxation multiplier pow(prop, power)/tau (:429)
__global__ __launch_bounds__(256) void k_tau_relax(const double *prop, int *tau, double *relax, ImmParams P) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.n3) return;
  int i = (int)(id % P.nx), j = (int)((id / P.nx) % P.ny), k = (int)(id / ((long long)P.nx * P.ny));
  bool ni[4] = {false, false, false, false};
  for (int rad = 1; rad <= 3; rad++) {
    bool found = false;
    for (int kk = -rad; kk <= rad && !found; kk++)
      for (int jj = -rad; jj <= rad && !found; jj++)
        for (int ii = -rad; ii <= rad; ii++)
          if (imm_halo_value(prop, P.hs + k + kk, P.hs + j + jj, P.hs + i + ii, P) == 0) { found = true; break; }
    ni[rad] = found;
  }
  int t = ni[1] ? 8 : ni[2] ? 4 : ni[3] ? 2 : 1;
  tau[id] = t;
  relax[id] = pow(prop[id], P.power) / t;
}

// dycore_any_immersed10 (dynamics_rk.h:1127-1143) as three separable max filters of (prop > 0);
// x/y periodic, z halos 0.
__global__ __launch_bounds__(256) void k_any_x(const double *prop, uint8_t *out, ImmParams P, int R) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.n3) return;
  int i = (int)(id % P.nx);
  long long row = id - i;
  uint8_t m = 0;
  for (int d = -R; d <= R; d++) m |= prop[row + wrapi_(i + d, P.nx)] > 0;
  out[id] = m;
}
__global__ __launch_bounds__(256) void k_any_y(const uint8_t *in, uint8_t *out, ImmParams P, int R) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.n3) return;
  int i = (int)(id % P.nx), j = (int)((id / P.nx) % P.ny);
  long long k = id / ((long long)P.nx * P.ny);
  uint8_t m = 0;
  for (int d = -R; d <= R; d++) m |= in[(k * P.ny + wrapi_(j + d, P.ny)) * P.nx + i];
  out[id] = m;
}
__global__ __launch_bounds__(256) void k_any_z(const uint8_t *in, uint8_t *out, ImmParams P, int R) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.n3) return;
  long long plane = (long long)P.nx * P.ny;
  int k = (int)(id / plane);
  long long q = id - (long long)k * plane;
  uint8_t m = 0;
  for (int d = -R; d <= R; d++) {
    int k2 = k + d;
    if (k2 < 0 || k2 >= P.nz) continue;
    m |= in[(long long)k2 * plane + q];
  }
  out[id] = m;
}

// immersed_proportion_halos for the LES closure (sc_init.h:728-767): halo 1, x/y periodic,
// bottom z halo 1, top z halo 0.
// immersed_proportion_halos (values) and its codes for the closure's tile: 0 (no immersed part),
// 1 (partly immersed), 2 (fully immersed)
__global__ __launch_bounds__(256) void k_imm_les(const double *prop, double *out, uint8_t *code, int nx, int ny,
                                                 int nz) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long nxh = nx + 2, nyh = ny + 2, nzh = nz + 2;
  if (id >= nxh * nyh * nzh) return;
  int ih = (int)(id % nxh), jh = (int)((id / nxh) % nyh), kh = (int)(id / (nxh * nyh));
  double v;
  if (kh == 0) v = 1.;
  else if (kh == nz + 1) v = 0.;
  else v = prop[((long long)(kh - 1) * ny + wrapi_(jh - 1, ny)) * nx + wrapi_(ih - 1, nx)];
  out[id] = v;
  code[id] = v == 1. ? 2 : (v > 0. ? 1 : 0);
}

// A caller's immersed_proportion_halos entry in place of the one built from immersed_proportion:
// values copied, codes rebuilt.
__global__ __launch_bounds__(256) void k_imm_les_set(const double *src, double *out, uint8_t *code, long long n) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const double v = src[id];
  out[id] = v;
  code[id] = v == 1. ? 2 : (v > 0. ? 1 : 0);
}

int les_set_immersed_halos_device(pu_ctx *ctx, const double *halos) {
  const pu::Grid &g = ctx->g;
  long long nl = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  hipLaunchKernelGGL(k_imm_les_set, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, ctx->stream, halos,
                     ctx->d_imm_les_val, ctx->d_imm_les, nl);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int les_init_device(pu_ctx *ctx, const pu_coupler_state *cs) {
  const pu::Grid &g = ctx->g;
  long long nl = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  hipLaunchKernelGGL(k_imm_les, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, ctx->stream, cs->immersed_proportion,
                     ctx->d_imm_les_val, ctx->d_imm_les, g.nx, g.ny, g.nz);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return PU_OK;
}

__global__ __launch_bounds__(256) void k_fill(double *p, double v, long long n) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) p[id] = v;
}

// plain streaming copy (16-byte vectors, grid-stride); also the known-size kernel used to
// calibrate the HBM traffic counters
__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long long n16) {
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n16; id += (long long)gridDim.x * blockDim.x)
    dst[id] = src[id];
}

int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
  // one 16-byte vector per thread up to 2^24 blocks (64 GiB), grid-stride beyond: a 2 GiB copy
  // moved 6.1 TB/s this way against 5.1 TB/s with the grid capped at 65536 blocks
  long long n16 = (long long)(bytes / 16);
  long long blocks = (n16 + 255) / 256;
  if (blocks > (1LL << 24)) blocks = 1LL << 24;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n16);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int copy_device(void *dst, const void *src, size_t bytes) {
  if (!bytes) return PU_OK;
  if ((bytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16)) {
    PU_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return PU_OK;
  }
  int rc = launch_copy(dst, src, bytes, nullptr);
  if (rc) return rc;
  PU_HIP_CHECK(hipDeviceSynchronize());
  return PU_OK;
}

int fill_device(double *ptr, double value, size_t count) {
  if (!count) return PU_OK;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, nullptr, ptr, value, (long long)count);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipDeviceSynchronize());
  return PU_OK;
}

// Hydrostatic edges from the cell profiles (dynamics_rk.h:1145-1186: 4th-order interpolation in
// log space for density and pressure, linear for theta; 2nd order below ord 5) and the upload of
// cells and edges to the device.
static int upload_hydrostatics(pu_ctx *ctx) {
  const pu_config &c = ctx->cfg;
  const int hs = ctx->g.hs, nz = ctx->g.nz;
  hipStream_t s = ctx->stream;
  const std::vector<double> &r = ctx->hy_cells[0], &t = ctx->hy_cells[1], &p = ctx->hy_cells[2];
  std::vector<double> &de = ctx->hy_edges[0], &te = ctx->hy_edges[1], &pe = ctx->hy_edges[2];
  de.assign(nz + 1, 0.); te.assign(nz + 1, 0.); pe.assign(nz + 1, 0.);
  for (int k = 0; k <= nz; k++) {
    if (c.ord < 5) {
      de[k] = std::exp(0.5 * std::log(r[hs + k - 1]) + 0.5 * std::log(r[hs + k]));
      te[k] = 0.5 * t[hs + k - 1] + 0.5 * t[hs + k];
      pe[k] = std::exp(0.5 * std::log(p[hs + k - 1]) + 0.5 * std::log(p[hs + k]));
    } else {
      de[k] = std::exp(-1. / 12. * std::log(r[hs + k - 2]) + 7. / 12. * std::log(r[hs + k - 1]) +
                       7. / 12. * std::log(r[hs + k]) - 1. / 12. * std::log(r[hs + k + 1]));
      te[k] = -1. / 12. * t[hs + k - 2] + 7. / 12. * t[hs + k - 1] + 7. / 12. * t[hs + k] - 1. / 12. * t[hs + k + 1];
      pe[k] = std::exp(-1. / 12. * std::log(p[hs + k - 2]) + 7. / 12. * std::log(p[hs + k - 1]) +
                       7. / 12. * std::log(p[hs + k]) - 1. / 12. * std::log(p[hs + k + 1]));
    }
  }
  std::vector<double> all;
  for (int w = 0; w < 3; w++) all.insert(all.end(), ctx->hy_cells[w].begin(), ctx->hy_cells[w].end());
  for (int w = 0; w < 3; w++) all.insert(all.end(), ctx->hy_edges[w].begin(), ctx->hy_edges[w].end());
  PU_HIP_CHECK(hipMemcpyAsync(ctx->d_hy, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice, s));
  PU_HIP_CHECK(hipStreamSynchronize(s));
  return PU_OK;
}

// Immersed data from immersed_proportion: the flag array of the halo'd box
// (dycore_immersed_proportion_halos > 0, dynamics_rk.h:1010-1030), tau and the relaxation
// multiplier (:389-425), any_immersed10 (:1127-1143) and the closure's halo'd proportion
// (sc_init.h:728-767).
static int build_immersed(pu_ctx *ctx, const double *prop) {
  const pu::Grid &g = ctx->g;
  const pu_config &c = ctx->cfg;
  hipStream_t s = ctx->stream;
  long long n3 = g.n3();
  ImmParams P;
  P.nx = g.nx; P.ny = g.ny; P.nz = g.nz; P.hs = g.hs; P.hp = g.hp; P.ox = g.ox;
  P.nxh = g.nxh; P.nyh = g.nyh; P.nzh = g.nzh; P.n3 = n3; P.power = c.immersed_power;
  long long nh = g.nh();
  hipLaunchKernelGGL(k_imm_flag, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, prop, ctx->d_imm_flag, P);
  PU_HIP_CHECK(hipGetLastError());
  dim3 grd((unsigned)((n3 + 255) / 256)), blk(256);
  hipLaunchKernelGGL(k_tau_relax, grd, blk, 0, s, prop, ctx->d_tau, ctx->d_relax, P);
  PU_HIP_CHECK(hipGetLastError());
  uint8_t *tmp1 = nullptr, *tmp2 = nullptr;
  PU_HIP_CHECK(hipMalloc(&tmp1, n3));
  PU_HIP_CHECK(hipMalloc(&tmp2, n3));
  hipLaunchKernelGGL(k_any_x, grd, blk, 0, s, prop, tmp1, P, 10);
  hipLaunchKernelGGL(k_any_y, grd, blk, 0, s, tmp1, tmp2, P, 10);
  hipLaunchKernelGGL(k_any_z, grd, blk, 0, s, tmp2, ctx->d_any10, P, 10);
  PU_HIP_CHECK(hipGetLastError());
  long long nl = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  hipLaunchKernelGGL(k_imm_les, dim3((unsigned)((nl + 255) / 256)), blk, 0, s, prop, ctx->d_imm_les_val,
                     ctx->d_imm_les, g.nx, g.ny, g.nz);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipStreamSynchronize(s));
  hipFree(tmp1); hipFree(tmp2);
  return PU_OK;
}

int dycore_init_device(pu_ctx *ctx, const pu_coupler_state *cs) {
  const pu::Grid &g = ctx->g;
  const pu_config &c = ctx->cfg;
  hipStream_t s = ctx->stream;
  long long n3 = g.n3();
  // hydrostatic means
  {
    HydroParams P;
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz; P.ntr = c.num_tracers; P.idWV = c.idWV;
    P.R_d = c.R_d; P.R_v = c.R_v; P.C0 = c.C0; P.gamma = c.gamma_d; P.adds_mass_mask = 0;
    for (int tr = 0; tr < c.num_tracers; tr++) if (c.tracer_adds_mass[tr]) P.adds_mass_mask |= 1u << tr;
    TracerPtrsC tp;
    for (int tr = 0; tr < PU_MAX_TRACERS; tr++) tp.p[tr] = tr < c.num_tracers ? cs->tracers[tr] : nullptr;
    hipLaunchKernelGGL(k_hydro_means, dim3(g.nz), dim3(256), 0, s, cs->density_dry, cs->temp, tp, ctx->d_red, P);
    PU_HIP_CHECK(hipGetLastError());
    std::vector<double> h(3 * (size_t)g.nz);
    PU_HIP_CHECK(hipMemcpyAsync(h.data(), ctx->d_red, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    PU_HIP_CHECK(hipStreamSynchronize(s));
    int hs = g.hs, nz = g.nz, nzh = nz + 2 * hs;
    double gamma = c.gamma_d, C0 = c.C0, grav = c.grav, dz = c.zlen / nz;
    double rn = 1. / ((double)g.nx * (double)g.ny);
    std::vector<double> &r = ctx->hy_cells[0], &t = ctx->hy_cells[1], &p = ctx->hy_cells[2];
    r.assign(nzh, 0.); t.assign(nzh, 0.); p.assign(nzh, 0.);
    for (int k = 0; k < nz; k++) { r[hs + k] = h[k] * rn; t[hs + k] = h[nz + k] * rn; p[hs + k] = h[2 * nz + k] * rn; }
    // isentropic halo extrapolation (dynamics_rk.h:985-1008)
    for (int kk = 0; kk < hs; kk++) {
      {
        int k0 = hs, k = k0 - 1 - kk;
        double rho0 = r[k0], theta0 = t[k0];
        double rho0_gm1 = std::pow(rho0, gamma - 1), theta0_g = std::pow(theta0, gamma);
        r[k] = std::pow(rho0_gm1 + grav * (gamma - 1) * dz * (kk + 1) / (gamma * C0 * theta0_g), 1. / (gamma - 1));
        t[k] = theta0;
        p[k] = C0 * std::pow(r[k] * theta0, gamma);
      }
      {
        int k0 = hs + nz - 1, k = k0 + 1 + kk;
        double rho0 = r[k0], theta0 = t[k0];
        double rho0_gm1 = std::pow(rho0, gamma - 1), theta0_g = std::pow(theta0, gamma);
        r[k] = std::pow(rho0_gm1 - grav * (gamma - 1) * dz * (kk + 1) / (gamma * C0 * theta0_g), 1. / (gamma - 1));
        t[k] = theta0;
        p[k] = C0 * std::pow(r[k] * theta0, gamma);
      }
    }
    int rc = upload_hydrostatics(ctx);
    if (rc != PU_OK) return rc;
  }
  return build_immersed(ctx, cs->immersed_proportion);
}

// ------------------------------------------------------------------------------------------------
// Diagnostics: FP64 mass, max wind, max |w| (coupler.h:507-530)
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag(const double *rd_, const double *uu, const double *vv,
                                              const double *ww, TracerPtrsC trc, unsigned int amask, int ntr,
                                              long long n3, double *part) {
  __shared__ double sm[256], smw[256], sma[256];
  double m = 0, mw = 0, ma = 0;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n3; id += (long long)gridDim.x * blockDim.x) {
    double rho = rd_[id];
    for (int tr = 0; tr < ntr; tr++) if ((amask >> tr) & 1u) rho += trc.p[tr][id];
    m += rho;
    double u = uu[id], v = vv[id], w = ww[id];
    mw = fmax(mw, sqrt(u * u + v * v + w * w));
    ma = fmax(ma, fabs(w));
  }
  sm[threadIdx.x] = m; smw[threadIdx.x] = mw; sma[threadIdx.x] = ma;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sm[threadIdx.x] += sm[threadIdx.x + w];
      smw[threadIdx.x] = fmax(smw[threadIdx.x], smw[threadIdx.x + w]);
      sma[threadIdx.x] = fmax(sma[threadIdx.x], sma[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[3 * blockIdx.x] = sm[0]; part[3 * blockIdx.x + 1] = smw[0]; part[3 * blockIdx.x + 2] = sma[0]; }
}

int diagnostics_device(pu_ctx *ctx, const pu_coupler_state *cs, double *mass, double *maxw, double *maxabsw) {
  const pu_config &c = ctx->cfg;
  long long n3 = ctx->g.n3();
  const int nb = 1024;
  TracerPtrsC tp;
  unsigned amask = 0;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) tp.p[tr] = tr < c.num_tracers ? cs->tracers[tr] : nullptr;
  for (int tr = 0; tr < c.num_tracers; tr++) if (c.tracer_adds_mass[tr]) amask |= 1u << tr;
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, ctx->stream, cs->density_dry, cs->uvel, cs->vvel, cs->wvel,
                     tp, amask, c.num_tracers, n3, ctx->d_red);
  PU_HIP_CHECK(hipGetLastError());
  std::vector<double> h(3 * nb);
  PU_HIP_CHECK(hipMemcpyAsync(h.data(), ctx->d_red, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  PU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  double m = 0, mw = 0, ma = 0;
  for (int b = 0; b < nb; b++) { m += h[3 * b]; mw = std::fmax(mw, h[3 * b + 1]); ma = std::fmax(ma, h[3 * b + 2]); }
  double dv = (c.xlen / c.nx) * (c.ylen / c.ny) * (c.zlen / c.nz);
  if (mass) *mass = m * dv;
  if (maxw) *maxw = mw;
  if (maxabsw) *maxabsw = ma;
  return PU_OK;
}

// ------------------------------------------------------------------------------------------------
// Synthetic initial states (seeded, counter-based so every cell is independent of launch shape)
// ------------------------------------------------------------------------------------------------
__host__ __device__ inline unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// uniform [0,1) from (seed, stream, index)
__host__ __device__ inline double urand(unsigned long long seed, unsigned long long stream, unsigned long long idx) {
  unsigned long long h = splitmix64(seed ^ splitmix64(stream * 0x632BE59BD9B4E019ull + splitmix64(idx)));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

struct CaseParams {
  int nx, ny, nz, ntr, case_id, idWV, idTKE;
  double dx, dy, dz, xlen, ylen, zlen;
  double R_d, cp_d, C0, gamma, grav, p0;
  unsigned long long seed;
};

// Isentropic background with constant theta0 (profiles.h:12-21, hydro_const_theta).
__device__ inline void hydro_const_theta(double z, const CaseParams &P, double theta0, double &r, double &t) {
  t = theta0;
  double exner = 1. - P.grav * z / (P.cp_d * theta0);
  double p = P.p0 * pow(exner, P.cp_d / P.R_d);
  double rt = pow(p / P.C0, 1. / P.gamma);
  r = rt / t;
}


// Piecewise-constant-theta isentropic background: theta0 below z_inv, theta0 + dtheta above
// (Exner pressure integrated exactly across the jump; profiles.h:12-21 for one layer).
__device__ inline void hydro_two_layer(double z, const CaseParams &P, double theta0, double dtheta, double z_inv,
                                       double &r, double &t) {
  double exner;
  if (z <= z_inv) { t = theta0; exner = 1. - P.grav * z / (P.cp_d * theta0); }
  else {
    t = theta0 + dtheta;
    exner = 1. - P.grav * z_inv / (P.cp_d * theta0) - P.grav * (z - z_inv) / (P.cp_d * t);
  }
  double p = P.p0 * pow(exner, P.cp_d / P.R_d);
  double rt = pow(p / P.C0, 1. / P.gamma);
  r = rt / t;
}

__global__ __launch_bounds__(256) void k_case(CaseParams P, double *rd_, double *uu, double *vv, double *ww,
                                              double *tt, TracerPtrs trc, double *imm) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long n3 = (long long)P.nx * P.ny * P.nz;
  if (id >= n3) return;
  int i = (int)(id % P.nx), j = (int)((id / P.nx) % P.ny), k = (int)(id / ((long long)P.nx * P.ny));
  double x = (i + 0.5) * P.dx, y = (j + 0.5) * P.dy, z = (k + 0.5) * P.dz;
  double rho_d = 0, theta = 300, u = 0, v = 0, w = 0, immersed = 0;
  double passive[2] = {0, 0};
  if (P.case_id == PU_CASE_BUBBLE) {
    // dry warm bubble: +2 K cos^2 perturbation of radius 2 km centred at (xlen/2, ylen/2, 2 km)
    double hr, ht;
    hydro_const_theta(z, P, 300., hr, ht);
    rho_d = hr;
    double dist = sqrt(((x - P.xlen / 2) / 2000.) * ((x - P.xlen / 2) / 2000.) +
                       ((y - P.ylen / 2) / 2000.) * ((y - P.ylen / 2) / 2000.) +
                       ((z - 2000.) / 2000.) * ((z - 2000.) / 2000.)) * M_PI / 2.;
    theta = ht + (dist <= M_PI / 2. ? 2. * cos(dist) * cos(dist) : 0.);
  } else if (P.case_id == PU_CASE_ABL_URBAN) {
    // neutral ABL: log law u = (u*/kappa) ln((z+z0)/z0), u* = 0.4, kappa = 0.4, z0 = 0.1 m;
    // +/-0.5 K seeded theta noise below 50 m; 8x8 cuboid buildings (half-tile footprints,
    // heights uniform in [10,40] m); passive tracers: lowest 50 m, and the first x-tile.
    double hr, ht;
    hydro_const_theta(z, P, 300., hr, ht);
    rho_d = hr;
    const double ustar = 0.4, kappa = 0.4, z0 = 0.1;
    u = ustar / kappa * log((z + z0) / z0);
    if (z < 50.) theta = 300. + (urand(P.seed, 1, (unsigned long long)id) - 0.5);
    double tx = P.xlen / 8., ty = P.ylen / 8.;
    int bi = (int)(x / tx), bj = (int)(y / ty);
    if (bi > 7) bi = 7;
    if (bj > 7) bj = 7;
    double xc = (bi + 0.5) * tx, yc = (bj + 0.5) * ty;
    double hb = 10. + 30. * urand(P.seed, 2, (unsigned long long)(bj * 8 + bi));
    if (fabs(x - xc) < tx / 4. && fabs(y - yc) < ty / 4. && z < hb) immersed = 1.;
    if (immersed > 0) u = 0;
    passive[0] = z < 50. ? 1. : 0.;
    passive[1] = x < tx ? 1. : 0.;
  } else {
    // wind-farm ABL: log law through 10 m/s at 90 m hub height, +5 K capping inversion at 800 m,
    // seeded +/-0.5 K theta and +/-0.5 m/s u perturbations below 200 m
    double hr, ht;
    hydro_two_layer(z, P, 300., 5., 800., hr, ht);
    rho_d = hr;
    theta = ht;
    const double z0 = 0.1;
    u = 10. * log((z + z0) / z0) / log((90. + z0) / z0);
    if (z < 200.) {
      theta += urand(P.seed, 1, (unsigned long long)id) - 0.5;
      u += urand(P.seed, 3, (unsigned long long)id) - 0.5;
    }
    passive[0] = z < 100. ? 1. : 0.;
    passive[1] = 0.;
  }
  rd_[id] = rho_d;
  uu[id] = u; vv[id] = v; ww[id] = w;
  // dry air only: p = rho_d R_d T with p = C0 (rho_d theta)^gamma
  tt[id] = P.C0 * pow(rho_d * theta, P.gamma) / (rho_d * P.R_d);
  int np = 0;
  for (int tr = 0; tr < P.ntr; tr++) {
    double val = 0;
    if (tr == P.idWV) val = 0;
    else if (tr == P.idTKE) val = 0.1;                 // LES_Closure::init sets TKE = 0.1 (les_closure.h:25)
    else { val = np < 2 ? passive[np] * rho_d : 0.; np++; }
    trc.p[tr][id] = val;
  }
  if (imm) imm[id] = immersed;
}

int case_init_device(const pu_config *cfg, int case_id, unsigned long long seed, pu_coupler_state *cs,
                     double *imm, hipStream_t s) {
  CaseParams P;
  P.nx = cfg->nx; P.ny = cfg->ny; P.nz = cfg->nz; P.ntr = cfg->num_tracers; P.case_id = case_id;
  P.idWV = cfg->idWV; P.idTKE = cfg->idTKE;
  P.xlen = cfg->xlen; P.ylen = cfg->ylen; P.zlen = cfg->zlen;
  P.dx = cfg->xlen / cfg->nx; P.dy = cfg->ylen / cfg->ny; P.dz = cfg->zlen / cfg->nz;
  P.R_d = cfg->R_d; P.cp_d = cfg->cp_d; P.C0 = cfg->C0; P.gamma = cfg->gamma_d; P.grav = cfg->grav; P.p0 = cfg->p0;
  P.seed = seed;
  TracerPtrs tp;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) tp.p[tr] = tr < cfg->num_tracers ? cs->tracers[tr] : nullptr;
  long long n3 = (long long)cfg->nx * cfg->ny * cfg->nz;
  hipLaunchKernelGGL(k_case, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, P, cs->density_dry, cs->uvel,
                     cs->vvel, cs->wvel, cs->temp, tp, imm);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipStreamSynchronize(s));
  return PU_OK;
}

}  // namespace pu

// ------------------------------------------------------------------------------------------------
// Output fields and restart of the dycore (the output / restart hooks of dynamics_rk.h:1192-1241)
// ------------------------------------------------------------------------------------------------
namespace pu {

struct OutParams {
  int nx, ny, nz, hs, ntr, idWV;
  long long n3;
  double R_d, R_v, C0, gamma;
  unsigned int adds_mass_mask;
  const double *hy;       // cells: dens [nz+2hs], theta [nz+2hs], pressure [nz+2hs]
};

// density_pert, theta_pert, pressure_pert from the coupler state converted as
// convert_coupler_to_dynamics does (dynamics_rk.h:1304-1352, 1213-1231)
__global__ __launch_bounds__(256) void k_dycore_output(const double *__restrict__ rd_, const double *__restrict__ tt,
                                                       TracerPtrsC trc, double *dp, double *tp, double *pp, OutParams P) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= P.n3) return;
  const int k = (int)(id / ((long long)P.nx * P.ny));
  const double rho_d = rd_[id], temp = tt[id];
  const double rho_v = trc.p[P.idWV][id];
  const double press = rho_d * P.R_d * temp + rho_v * P.R_v * temp;
  double rho = rho_d;
  for (int tr = 0; tr < P.ntr; tr++) if ((P.adds_mass_mask >> tr) & 1u) rho += trc.p[tr][id];
  const double theta = pow(press / P.C0, 1. / P.gamma) / rho;
  const double rt = rho * theta;
  const int nzh = P.nz + 2 * P.hs;
  dp[id] = rho - P.hy[P.hs + k];
  tp[id] = rt / rho - P.hy[nzh + P.hs + k];
  pp[id] = P.C0 * pow(rt, P.gamma) - P.hy[2 * nzh + P.hs + k];
}

}  // namespace pu

using namespace pu;

extern "C" {

int pu_dycore_output_fields(pu_ctx *ctx, const pu_coupler_state *cs, double *density_pert, double *theta_pert,
                            double *pressure_pert) {
  if (!ctx || !cs || !cs->density_dry || !cs->temp || !density_pert || !theta_pert || !pressure_pert)
    return fail(PU_ERR_ARG, "pu_dycore_output_fields: null argument");
  if (!ctx->dycore_ready) return fail(PU_ERR_STATE, "pu_dycore_init has not been called");
  const pu_config &c = ctx->cfg;
  OutParams P;
  P.nx = c.nx; P.ny = c.ny; P.nz = c.nz; P.hs = ctx->g.hs; P.ntr = c.num_tracers; P.idWV = c.idWV;
  P.n3 = ctx->g.n3();
  P.R_d = c.R_d; P.R_v = c.R_v; P.C0 = c.C0; P.gamma = c.gamma_d;
  P.adds_mass_mask = 0;
  for (int tr = 0; tr < c.num_tracers; tr++) if (c.tracer_adds_mass[tr]) P.adds_mass_mask |= 1u << tr;
  P.hy = ctx->d_hy;
  TracerPtrsC tp;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) tp.p[tr] = tr < c.num_tracers ? cs->tracers[tr] : nullptr;
  hipLaunchKernelGGL(k_dycore_output, dim3((unsigned)((P.n3 + 255) / 256)), dim3(256), 0, ctx->stream, cs->density_dry,
                     cs->temp, tp, density_pert, theta_pert, pressure_pert, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int pu_dycore_restart(pu_ctx *ctx, const pu_coupler_state *cs, const double *dens_c, const double *theta_c,
                      const double *pres_c) {
  if (!ctx || !cs || !cs->immersed_proportion || !dens_c || !theta_c || !pres_c)
    return fail(PU_ERR_ARG, "pu_dycore_restart: null argument");
  if (!ctx->dycore_ready) return fail(PU_ERR_STATE, "pu_dycore_init has not been called");
  const int nzh = ctx->g.nz + 2 * ctx->g.hs;
  ctx->hy_cells[0].assign(dens_c, dens_c + nzh);
  ctx->hy_cells[1].assign(theta_c, theta_c + nzh);
  ctx->hy_cells[2].assign(pres_c, pres_c + nzh);
  int rc = upload_hydrostatics(ctx);
  if (rc != PU_OK) return rc;
  ctx->sgs_valid_ready = false;          // the learned closure's mask follows the immersed cells
  return build_immersed(ctx, cs->immersed_proportion);
}

}  // extern "C"
