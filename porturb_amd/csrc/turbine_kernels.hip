// turbine_kernels.hip -- modules::WindmillActuators::apply as actuator disks on gfx950
// (model/modules/windmill_actuators_yaw.h:523-1043, without blades, floating motions and immersed
// hub/tower material, the configuration of the reference's wind-farm drivers,
// windfarm_ensemble.cpp:106-110).
//
// Per turbine and call:
//   1. sums of u and v over the cells whose centres lie in the hub-centred cube of half-width R
//      (the blade radius) -> upstream direction (:622-653);
//   2. the disk projection: the rotor-frame point cloud (20 points per cell width) is tilted,
//      yawed and binned into cells (round(x/dx - 0.5)), weighted by the radial thrust shape times
//      the axial projection shape, once at the disk and once 4 R upstream (:654-716).  One thread
//      walks one axial line of the cloud and adds runs of points that land in the same cell with
//      one atomic; the weights are 2^-b fixed point in 64-bit integers (b <= 40, chosen so that
//      no sum can exceed 2^62: TURB_FIX_MAX_BITS), so the sums are exact and independent of the
//      order of the atomics (the reference adds floats atomically);
//   3. the weight totals (exact integer sums), the normalised weights, and the sampled inflow
//      sum(w_samp u), sum(w_samp v) in a fixed reduction order (:834-897);
//   4. on the host: C_T, C_P, power from the table at the inertially filtered inflow, the inertial
//      filter, the traces (:898-965);
//   5. the thrust and TKE tendencies of the disk cells with the normal-wind limiter (:966-1018) into
//      a per-turbine buffer; the yaw controller (:1020-1029).
// After all turbines the tendencies are added to u, v and TKE (:1035-1041).  Every device buffer
// covers a box around the disk (or its sampling image): the cloud lies within
// Rb = sqrt((xr + |overhang|)^2 + (R (1 + decay))^2) of its centre.
#include "pu_common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace pu {

// Fixed-point scale of the cloud weights: 2^fix_bits with fix_bits = min(40, 62 - ceil(log2(points))),
// points = num_x num_y num_z.  Every point's weight is at most 1 (thrust_shape2 and proj_shape_1d
// peak at 1), so a disk's total, and every cell's partial sum, stays below 2^62 in the unsigned
// 64-bit sums.  (A rotor of 60 m on 5 m cells has 7.4e7 points: at 2^40 its total would wrap.)
constexpr int TURB_FIX_MAX_BITS = 40;

struct TurbBox {
  int i0, j0, k0, ni, nj, nk;                     // cell-index origin and extents (clipped to the domain)
  __host__ __device__ long long n() const { return (long long)ni * nj * nk; }
};

struct CloudArgs {
  int nx, ny, nz, num_x, num_y, num_z;
  double dx, dy, dz, xr, R, rad, hub_radius, decayloc, overhang;
  double cos_yaw, sin_yaw, cos_tlt, sin_tlt, base_x, base_y, hub_height, up_x, up_y;
  double fix;                                     // fixed-point scale (2^fix_bits)
  TurbBox pbox, sbox;
  unsigned long long *acc_p, *acc_s;
  unsigned int *err;
};

__device__ __forceinline__ double thrust_shape2(double x, double x2, double x3, double x4) {
#pragma clang fp contract(off)
  // DefaultThrustShape2 with a = 1 (:306-317)
  if (x < x3) return -((x * x) - 2 * x * x2) / (x2 * x2);
  if (x < x4) {
    const double v3 = -((x3 * x3) - 2 * x3 * x2) / (x2 * x2);
    return (v3 * x * x - 2 * v3 * x * x4 + v3 * x4 * x4) / (x3 * x3 - 2 * x3 * x4 + x4 * x4);
  }
  return 0;
}

__device__ __forceinline__ double proj_shape_1d(double x, double xr) {
#pragma clang fp contract(off)
  const double term = 1 - (x / xr) * (x / xr);
  if (term <= 0) return 0;
  return term * term;
}

__device__ __forceinline__ void flush(unsigned long long *acc, const TurbBox &b, int ti, int tj, int tk,
                                      unsigned long long q, unsigned int *err) {
  if (q == 0) return;
  const int bi = ti - b.i0, bj = tj - b.j0, bk = tk - b.k0;
  if (bi < 0 || bi >= b.ni || bj < 0 || bj >= b.nj || bk < 0 || bk >= b.nk) { atomicAdd(err, 1u); return; }
  atomicAdd(acc + ((long long)bk * b.nj + bj) * b.ni + bi, q);
}

// one thread per (j, k) line of the rotor-frame cloud; i runs along the rotor axis
__global__ __launch_bounds__(256) void k_turbine_cloud(CloudArgs A) {
#pragma clang fp contract(off)
  const long long line = (long long)blockIdx.x * 256 + threadIdx.x;
  if (line >= (long long)A.num_y * A.num_z) return;
  const int j = (int)(line % A.num_y), k = (int)(line / A.num_y);
  const double y = -A.R + (2 * A.R * j) / (A.num_y - 1);
  const double z = -A.R + (2 * A.R * k) / (A.num_z - 1);
  const double rloc = sqrt(y * y + z * z);
  if (!(rloc <= A.R)) return;
  const double radial = rloc <= A.hub_radius ? 0. : thrust_shape2(rloc / A.rad, 0.7, 1, 1 + A.decayloc);
  int cp[3] = {-1, -1, -1}, cs[3] = {-1, -1, -1};
  unsigned long long qp = 0, qs = 0;
  for (int i = 0; i < A.num_x; i++) {
    const double x = -A.xr + (2 * A.xr * i) / (A.num_x - 1);
    const double shp = radial * proj_shape_1d(x, A.xr);
    const unsigned long long q = (unsigned long long)llrint(shp * A.fix);
    const double x1 = A.cos_tlt * (x + A.overhang) + A.sin_tlt * z;
    const double y1 = y;
    const double z1 = -A.sin_tlt * (x + A.overhang) + A.cos_tlt * z;
    double xp = A.base_x + A.cos_yaw * x1 - A.sin_yaw * y1;
    double yp = A.base_y + A.sin_yaw * x1 + A.cos_yaw * y1;
    const double zp = A.hub_height + z1;
    int ti = (int)round(xp / A.dx - 0.5), tj = (int)round(yp / A.dy - 0.5), tk = (int)round(zp / A.dz - 0.5);
    if (ti >= 0 && ti < A.nx && tj >= 0 && tj < A.ny && tk >= 0 && tk < A.nz) {
      if (ti != cp[0] || tj != cp[1] || tk != cp[2]) {
        if (cp[0] >= 0) flush(A.acc_p, A.pbox, cp[0], cp[1], cp[2], qp, A.err);
        cp[0] = ti; cp[1] = tj; cp[2] = tk; qp = 0;
      }
      qp += q;
    }
    xp += A.up_x;
    yp += A.up_y;
    ti = (int)round(xp / A.dx - 0.5); tj = (int)round(yp / A.dy - 0.5); tk = (int)round(zp / A.dz - 0.5);
    if (ti >= 0 && ti < A.nx && tj >= 0 && tj < A.ny && tk >= 0 && tk < A.nz) {
      if (ti != cs[0] || tj != cs[1] || tk != cs[2]) {
        if (cs[0] >= 0) flush(A.acc_s, A.sbox, cs[0], cs[1], cs[2], qs, A.err);
        cs[0] = ti; cs[1] = tj; cs[2] = tk; qs = 0;
      }
      qs += q;
    }
  }
  if (cp[0] >= 0) flush(A.acc_p, A.pbox, cp[0], cp[1], cp[2], qp, A.err);
  if (cs[0] >= 0) flush(A.acc_s, A.sbox, cs[0], cs[1], cs[2], qs, A.err);
}

constexpr int TURB_RED_BLOCKS = 64;

__device__ __forceinline__ long long box_global(const TurbBox &b, long long e, int nx, int ny) {
  const int bi = (int)(e % b.ni), bj = (int)((e / b.ni) % b.nj), bk = (int)(e / ((long long)b.ni * b.nj));
  return ((long long)(b.k0 + bk) * ny + (b.j0 + bj)) * nx + (b.i0 + bi);
}

// sum over the block; every thread gets it (the leading barrier lets a second call reuse sm)
template <class T> __device__ T block_sum(T v, T *sm) {
  __syncthreads();
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  return sm[0];
}

// partial sums of u and v over a box (the disk cube of step 1)
__global__ __launch_bounds__(256) void k_box_uv(const double *u, const double *v, TurbBox b, int nx, int ny, double *part) {
  __shared__ double sm[256];
  double su = 0, sv = 0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < b.n(); e += (long long)gridDim.x * 256) {
    const long long g = box_global(b, e, nx, ny);
    su += u[g];
    sv += v[g];
  }
  su = block_sum(su, sm);
  sv = block_sum(sv, sm);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = su; part[2 * blockIdx.x + 1] = sv; }
}

// partial integer sums of both accumulators
__global__ __launch_bounds__(256) void k_acc_tot(const unsigned long long *ap, long long np, const unsigned long long *as,
                                                 long long ns, unsigned long long *part) {
  __shared__ unsigned long long sm[256];
  unsigned long long sp = 0, ss = 0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < np; e += (long long)gridDim.x * 256) sp += ap[e];
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ns; e += (long long)gridDim.x * 256) ss += as[e];
  sp = block_sum(sp, sm);
  ss = block_sum(ss, sm);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = sp; part[2 * blockIdx.x + 1] = ss; }
}

// sampling: normalised weights into samp_weight_tot, partial sums of w u and w v
__global__ __launch_bounds__(256) void k_samp(const unsigned long long *as, TurbBox b, double tot, const double *u,
                                              const double *v, int nx, int ny, double *samp_tot, double *part) {
  __shared__ double sm[256];
  double s0 = 0, s1 = 0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < b.n(); e += (long long)gridDim.x * 256) {
    if (as[e] == 0) continue;
    const double w = (double)as[e] / tot;
    const long long g = box_global(b, e, nx, ny);
    if (samp_tot) samp_tot[g] += w;
    s0 += w * u[g];
    s1 += w * v[g];
  }
  s0 = block_sum(s0, sm);
  s1 = block_sum(s1, sm);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = s0; part[2 * blockIdx.x + 1] = s1; }
}

struct ThrustArgs {
  const unsigned long long *ap;
  TurbBox b;
  int nx, ny;
  double tot, turb_factor, C_T, C_TKE, mag0, cos_yaw, sin_yaw, dt, one_minus_blade_wt;
  int apply_thrust;
  const double *rho_d, *u, *v;
  double *proj_tot;
  double *tend;       // [3][box]
};

// disk cells: projection weight into proj_weight_tot and the thrust / TKE tendencies (:966-1018)
__global__ __launch_bounds__(256) void k_thrust(ThrustArgs A) {
#pragma clang fp contract(off)
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = A.b.n();
  if (e >= n) return;
  double t_u = 0, t_v = 0, t_tke = 0;
  if (A.ap[e] > 0) {
    const double w = (double)A.ap[e] / A.tot;
    const long long g = box_global(A.b, e, A.nx, A.ny);
    if (A.proj_tot) A.proj_tot[g] += w;
    const double wt = w * A.turb_factor;
    const double un = A.u[g] * A.cos_yaw + A.v[g] * A.sin_yaw;
    if (A.apply_thrust && un > 0) {
      t_u = -0.5 * A.C_T * A.mag0 * A.mag0 * A.cos_yaw * wt;
      t_v = -0.5 * A.C_T * A.mag0 * A.mag0 * A.sin_yaw * wt;
      t_tke = 0.5 * A.rho_d[g] * A.C_TKE * A.mag0 * A.mag0 * A.mag0 * wt * A.one_minus_blade_wt;
      const double un_new = (A.u[g] + A.dt * t_u) * A.cos_yaw + (A.v[g] + A.dt * t_v) * A.sin_yaw;
      if (un_new < 0) {
        const double t_old = fabs(t_u * A.cos_yaw + t_v * A.sin_yaw);
        const double t_new = un;
        t_u *= (t_new / t_old);
        t_v *= (t_new / t_old);
        t_tke *= (t_new / t_old) * (t_new / t_old);
      }
    }
  }
  A.tend[e] = t_u;
  A.tend[n + e] = t_v;
  A.tend[2 * n + e] = t_tke;
}

__global__ __launch_bounds__(256) void k_turbine_update(const double *tend, TurbBox b, int nx, int ny, double dt,
                                                        double *u, double *v, double *tke) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = b.n();
  if (e >= n) return;
  const double a = tend[e], c = tend[n + e], d = tend[2 * n + e];
  if (a == 0 && c == 0 && d == 0) return;
  const long long g = box_global(b, e, nx, ny);
  u[g] += dt * a;
  v[g] += dt * c;
  tke[g] += dt * d;
}

__global__ __launch_bounds__(256) void k_zero2(double *a, double *b, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (a) a[i] = 0;
  if (b) b[i] = 0;
}

}  // namespace pu

using namespace pu;

struct pu_turbines {
  pu_ctx *ctx = nullptr;
  std::vector<double> velmag, thrust_coef, power_coef, power_mw;
  double hub_height = 0, rad = 0, max_yaw_speed = 0, overhang = 0, hub_radius = 0, tilt = 0;
  pu_turbine_opts opts{};
  struct Turb {
    double base_x, base_y, yaw, u_inertial = 0, v_inertial = 0, yaw_uavg = 0, yaw_vavg = 0;
    bool apply_thrust;
    pu_turbine_status st{};
    std::vector<double> trace[PU_TURB_NTRACE];
    TurbBox pbox{};
  };
  std::vector<Turb> t;
  long long box_cap = 0;
  int num_x = 0, num_y = 0, num_z = 0;
  double xr = 0, R = 0, decayloc = 0, Rb = 0;
  double fix = 0;                           // fixed-point scale of the cloud weights
  unsigned long long *d_acc = nullptr;      // [2][box_cap]
  double *d_tend = nullptr;                 // [nturb][3][box_cap]
  double *d_part = nullptr;                 // [2 * TURB_RED_BLOCKS]
  unsigned long long *d_upart = nullptr;    // [2 * TURB_RED_BLOCKS]
  unsigned int *d_err = nullptr;
};

namespace {

double interp(const std::vector<double> &ref_umag, const std::vector<double> &ref_var, double umag) {   // :1050-1063
  const int imax = (int)ref_umag.size() - 1;
  if (umag < ref_umag[0] || umag > ref_umag[imax]) return 0;
  int i = 0;
  while (umag > ref_umag[i]) i++;
  if (i > 0) i--;
  const double fac = (ref_umag[i + 1] - umag) / (ref_umag[i + 1] - ref_umag[i]);
  return fac * ref_var[i] + (1 - fac) * ref_var[i + 1];
}

// cell-index box (clipped) around a centre; lo/hi by the same rounding the cloud binning uses
TurbBox make_box(const pu_config &c, double cx, double cy, double cz, double Rb) {
  const double dx = c.xlen / c.nx, dy = c.ylen / c.ny, dz = c.zlen / c.nz;
  auto range = [](double lo, double hi, double d, int n, int &i0, int &ni) {
    int a = (int)std::floor(lo / d - 0.5) - 1, b = (int)std::ceil(hi / d - 0.5) + 1;
    a = std::max(a, 0);
    b = std::min(b, n - 1);
    i0 = a;
    ni = std::max(0, b - a + 1);
  };
  TurbBox b;
  range(cx - Rb, cx + Rb, dx, c.nx, b.i0, b.ni);
  range(cy - Rb, cy + Rb, dy, c.ny, b.j0, b.nj);
  range(cz - Rb, cz + Rb, dz, c.nz, b.k0, b.nk);
  if (b.ni == 0 || b.nj == 0 || b.nk == 0) b.ni = b.nj = b.nk = 0;
  return b;
}

int nblocks(long long n) { return (int)std::max<long long>(1, std::min<long long>(TURB_RED_BLOCKS, (n + 255) / 256)); }

}  // namespace

extern "C" {

int pu_turbines_create(pu_ctx *ctx, const pu_turbine_ref *ref, int nturb, const double *x, const double *y,
                       const unsigned char *apply_thrust, const pu_turbine_opts *opts, pu_turbines **out) {
  if (!ctx || !ref || !out || nturb < 0 || (nturb > 0 && (!x || !y)) || !ref->velocity_magnitude ||
      !ref->thrust_coef || !ref->power_coef || !ref->power_megawatts)
    return fail(PU_ERR_ARG, "pu_turbines_create: null argument");
  if (ref->ntable < 2) return fail(PU_ERR_ARG, "pu_turbines_create: the turbine table needs at least 2 rows");
  for (int i = 0; i + 1 < ref->ntable; i++)
    if (!(ref->velocity_magnitude[i + 1] > ref->velocity_magnitude[i]))
      return fail(PU_ERR_ARG, "pu_turbines_create: velocity_magnitude must increase");
  if (!(ref->blade_radius > 0)) return fail(PU_ERR_ARG, "pu_turbines_create: blade_radius must be > 0");
  if (ctx->cfg.idTKE < 0) return fail(PU_ERR_ARG, "pu_turbines_create: the TKE tracer is required");
  const pu_config &c = ctx->cfg;
  auto *T = new pu_turbines();
  T->ctx = ctx;
  T->velmag.assign(ref->velocity_magnitude, ref->velocity_magnitude + ref->ntable);
  T->thrust_coef.assign(ref->thrust_coef, ref->thrust_coef + ref->ntable);
  T->power_coef.assign(ref->power_coef, ref->power_coef + ref->ntable);
  T->power_mw.assign(ref->power_megawatts, ref->power_megawatts + ref->ntable);
  T->hub_height = ref->hub_height; T->rad = ref->blade_radius; T->max_yaw_speed = ref->max_yaw_speed;
  T->overhang = ref->overhang; T->hub_radius = ref->hub_radius; T->tilt = ref->shaft_tilt;
  if (opts) T->opts = *opts;
  const double dx = c.xlen / c.nx;
  const double decay = 3 * dx / T->rad;
  T->decayloc = decay + 0.02;
  T->xr = std::max(5., 5 * dx);
  T->R = T->rad * (1 + T->decayloc);
  T->num_x = (int)std::ceil(20 / dx * T->xr * 2);
  T->num_y = (int)std::ceil(20 / dx * T->rad * (1 + T->decayloc) * 2);
  T->num_z = T->num_y;
  {
    const double points = (double)T->num_x * T->num_y * T->num_z;
    const int bits = std::min(TURB_FIX_MAX_BITS, 62 - (int)std::ceil(std::log2(points)));
    if (bits < 16) { delete T; return fail(PU_ERR_ARG, "pu_turbines_create: rotor cloud too fine for the fixed-point sums"); }
    T->fix = std::ldexp(1.0, bits);
  }
  T->Rb = std::sqrt((T->xr + std::fabs(T->overhang)) * (T->xr + std::fabs(T->overhang)) + T->R * T->R);
  const double dy = c.ylen / c.ny, dz = c.zlen / c.nz;
  T->box_cap = (long long)(2 * (int)std::ceil(T->Rb / dx) + 5) * (2 * (int)std::ceil(T->Rb / dy) + 5) *
               (2 * (int)std::ceil(T->Rb / dz) + 5);
  for (int i = 0; i < nturb; i++) {
    pu_turbines::Turb tb;
    tb.base_x = x[i]; tb.base_y = y[i];
    tb.yaw = opts ? opts->initial_yaw : 0.;
    tb.apply_thrust = apply_thrust ? apply_thrust[i] != 0 : true;
    T->t.push_back(tb);
  }
  auto cleanup = [&](int rc) { pu_turbines_destroy(T); return rc; };
  hipError_t e;
  if ((e = hipMalloc(&T->d_acc, sizeof(unsigned long long) * 2 * T->box_cap)) != hipSuccess ||
      (e = hipMalloc(&T->d_tend, sizeof(double) * 3 * T->box_cap * std::max(1, nturb))) != hipSuccess ||
      (e = hipMalloc(&T->d_part, sizeof(double) * 2 * TURB_RED_BLOCKS)) != hipSuccess ||
      (e = hipMalloc(&T->d_upart, sizeof(unsigned long long) * 2 * TURB_RED_BLOCKS)) != hipSuccess ||
      (e = hipMalloc(&T->d_err, sizeof(unsigned int))) != hipSuccess)
    return cleanup(fail(PU_ERR_HIP, std::string("pu_turbines_create: ") + hipGetErrorString(e)));
  *out = T;
  return PU_OK;
}

int pu_turbines_apply(pu_turbines *T, pu_coupler_state *cs, double dt, double *proj_weight_tot, double *samp_weight_tot) {
  if (!T || !cs || !cs->density_dry || !cs->uvel || !cs->vvel) return fail(PU_ERR_ARG, "pu_turbines_apply: null argument");
  pu_ctx *ctx = T->ctx;
  const pu_config &c = ctx->cfg;
  double *tke = cs->tracers[c.idTKE];
  if (!tke) return fail(PU_ERR_ARG, "pu_turbines_apply: the TKE tracer pointer is null");
  hipStream_t s = ctx->stream;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const double dx = c.xlen / nx, dy = c.ylen / ny, dz = c.zlen / nz;
  const long long n3 = (long long)nx * ny * nz;
  if (proj_weight_tot || samp_weight_tot) {
    hipLaunchKernelGGL(k_zero2, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, proj_weight_tot, samp_weight_tot, n3);
    PU_HIP_CHECK(hipGetLastError());
  }
  const double cos_tlt = std::cos(T->tilt), sin_tlt = std::sin(T->tilt);
  const double dxloc = dx * (63 / T->rad);
  double blade_wt = 1;
  if (dxloc > 2) { const double xx = (std::log2(dxloc) - 1) / 3; blade_wt = -2 * xx * xx * xx + 3 * xx * xx; }
  const double turb_factor = M_PI * T->rad * T->rad / (dx * dy * dz);
  std::vector<double> hp(2 * TURB_RED_BLOCKS);
  std::vector<unsigned long long> hu(2 * TURB_RED_BLOCKS);
  LaunchTimer lt(ctx, "turbines");
  for (size_t it = 0; it < T->t.size(); it++) {
    auto &tb = T->t[it];
    const double cos_yaw = std::cos(tb.yaw), sin_yaw = std::sin(tb.yaw);
    const double rad = T->rad;
    // 1. disk cube sums (cell centres within [centre - R, centre + R])
    TurbBox cube{0, 0, 0, 0, 0, 0};
    auto span = [](double lo, double hi, double d, int n, int &a, int &m) {
      a = -1; m = 0;
      for (int i = 0; i < n; i++) {
        const double xc = (i + 0.5) * d;
        if (xc >= lo && xc <= hi) { if (a < 0) a = i; m++; }
      }
      if (a < 0) a = 0;
    };
    span(tb.base_x - rad, tb.base_x + rad, dx, nx, cube.i0, cube.ni);
    span(tb.base_y - rad, tb.base_y + rad, dy, ny, cube.j0, cube.nj);
    span(T->hub_height - rad, T->hub_height + rad, dz, nz, cube.k0, cube.nk);
    double su = 0, sv = 0;
    if (cube.n() > 0) {
      const int nb = nblocks(cube.n());
      hipLaunchKernelGGL(k_box_uv, dim3(nb), dim3(256), 0, s, cs->uvel, cs->vvel, cube, nx, ny, T->d_part);
      PU_HIP_CHECK(hipGetLastError());
      PU_HIP_CHECK(hipMemcpyAsync(hp.data(), T->d_part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s));
      PU_HIP_CHECK(hipStreamSynchronize(s));
      for (int b = 0; b < nb; b++) { su += hp[2 * b]; sv += hp[2 * b + 1]; }
    }
    const double upstream_dir = T->opts.have_upstream_dir ? T->opts.upstream_dir : std::atan2(sv, su);
    const double up_x = -4 * rad * std::cos(upstream_dir), up_y = -4 * rad * std::sin(upstream_dir);
    // 2. cloud projection into the disk and sampling boxes
    const double hx = tb.base_x, hy = tb.base_y;
    TurbBox pbox = make_box(c, hx, hy, T->hub_height, T->Rb);
    TurbBox sbox = make_box(c, hx + up_x, hy + up_y, T->hub_height, T->Rb);
    if (pbox.n() > T->box_cap || sbox.n() > T->box_cap) return fail(PU_ERR_STATE, "pu_turbines_apply: turbine box overflow");
    unsigned long long *acc_p = T->d_acc, *acc_s = T->d_acc + T->box_cap;
    PU_HIP_CHECK(hipMemsetAsync(T->d_acc, 0, sizeof(unsigned long long) * 2 * T->box_cap, s));
    PU_HIP_CHECK(hipMemsetAsync(T->d_err, 0, sizeof(unsigned int), s));
    CloudArgs A;
    A.nx = nx; A.ny = ny; A.nz = nz; A.num_x = T->num_x; A.num_y = T->num_y; A.num_z = T->num_z;
    A.dx = dx; A.dy = dy; A.dz = dz; A.xr = T->xr; A.R = T->R; A.rad = rad; A.hub_radius = T->hub_radius;
    A.decayloc = T->decayloc; A.overhang = T->overhang;
    A.cos_yaw = cos_yaw; A.sin_yaw = sin_yaw; A.cos_tlt = cos_tlt; A.sin_tlt = sin_tlt;
    A.base_x = hx; A.base_y = hy; A.hub_height = T->hub_height; A.up_x = up_x; A.up_y = up_y;
    A.fix = T->fix;
    A.pbox = pbox; A.sbox = sbox; A.acc_p = acc_p; A.acc_s = acc_s; A.err = T->d_err;
    const long long lines = (long long)T->num_y * T->num_z;
    hipLaunchKernelGGL(k_turbine_cloud, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, s, A);
    PU_HIP_CHECK(hipGetLastError());
    // 3. totals (exact), sampled inflow
    unsigned long long tp = 0, ts = 0;
    {
      const int nb = nblocks(std::max(pbox.n(), sbox.n()));
      hipLaunchKernelGGL(k_acc_tot, dim3(nb), dim3(256), 0, s, acc_p, pbox.n(), acc_s, sbox.n(), T->d_upart);
      PU_HIP_CHECK(hipGetLastError());
      unsigned int err = 0;
      PU_HIP_CHECK(hipMemcpyAsync(hu.data(), T->d_upart, sizeof(unsigned long long) * 2 * nb, hipMemcpyDeviceToHost, s));
      PU_HIP_CHECK(hipMemcpyAsync(&err, T->d_err, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
      PU_HIP_CHECK(hipStreamSynchronize(s));
      if (err) return fail(PU_ERR_STATE, "pu_turbines_apply: cloud point outside its box");
      for (int b = 0; b < nb; b++) { tp += hu[2 * b]; ts += hu[2 * b + 1]; }
    }
    double s0 = 0, s1 = 0;
    if (ts > 0) {
      const int nb = nblocks(sbox.n());
      hipLaunchKernelGGL(k_samp, dim3(nb), dim3(256), 0, s, acc_s, sbox, (double)ts, cs->uvel, cs->vvel, nx, ny,
                         samp_weight_tot, T->d_part);
      PU_HIP_CHECK(hipGetLastError());
      PU_HIP_CHECK(hipMemcpyAsync(hp.data(), T->d_part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s));
      PU_HIP_CHECK(hipStreamSynchronize(s));
      for (int b = 0; b < nb; b++) { s0 += hp[2 * b]; s1 += hp[2 * b + 1]; }
    }
    // 4. coefficients and inertial filter (host)
    const double instant_u0 = s0 * cos_yaw, instant_v0 = s1 * sin_yaw;
    const double instant_mag0 = std::max(0., instant_u0 + instant_v0);
    const double inertial_mag0 = std::max(0., tb.u_inertial + tb.v_inertial);
    double C_T = interp(T->velmag, T->thrust_coef, inertial_mag0);
    double C_P = interp(T->velmag, T->power_coef, inertial_mag0);
    const double pwr = interp(T->velmag, T->power_mw, inertial_mag0);
    if (inertial_mag0 > 1.e-10) {
      if (!T->opts.orig_C_T) {
        const double a = std::max(0., std::min(1., 1 - C_P / (C_T + 1.e-10)));
        C_T = 4 * a * (1 - a);
      }
      C_P = std::min(C_T, pwr * 1.e6 / (0.5 * 1.2 * M_PI * rad * rad * inertial_mag0 * inertial_mag0 * inertial_mag0));
    } else {
      C_T = 0;
      C_P = 0;
    }
    const double inertial_tau = 30;
    tb.u_inertial = instant_u0 * dt / inertial_tau + (inertial_tau - dt) / inertial_tau * tb.u_inertial;
    tb.v_inertial = instant_v0 * dt / inertial_tau + (inertial_tau - dt) / inertial_tau * tb.v_inertial;
    tb.trace[PU_TURB_POWER].push_back(pwr);
    tb.trace[PU_TURB_YAW].push_back(tb.yaw);
    tb.trace[PU_TURB_U_SAMP].push_back(s0);
    tb.trace[PU_TURB_V_SAMP].push_back(s1);
    tb.trace[PU_TURB_CP].push_back(C_P);
    tb.trace[PU_TURB_CT].push_back(C_T);
    // 5. tendencies of the disk cells
    ThrustArgs B;
    B.ap = acc_p; B.b = pbox; B.nx = nx; B.ny = ny;
    B.tot = (double)tp; B.turb_factor = turb_factor; B.C_T = C_T; B.C_TKE = 0.25 * (C_T - C_P); B.mag0 = instant_mag0;
    B.cos_yaw = cos_yaw; B.sin_yaw = sin_yaw; B.dt = dt; B.one_minus_blade_wt = 1 - blade_wt;
    B.apply_thrust = tb.apply_thrust && tp > 0;
    B.rho_d = cs->density_dry; B.u = cs->uvel; B.v = cs->vvel; B.proj_tot = proj_weight_tot;
    B.tend = T->d_tend + (long long)it * 3 * T->box_cap;
    tb.pbox = pbox;
    if (pbox.n() > 0 && tp > 0) {
      hipLaunchKernelGGL(k_thrust, dim3((unsigned)((pbox.n() + 255) / 256)), dim3(256), 0, s, B);
      PU_HIP_CHECK(hipGetLastError());
    } else {
      tb.pbox.ni = tb.pbox.nj = tb.pbox.nk = 0;
    }
    tb.st.yaw = tb.yaw;
    tb.st.u_samp_inertial = tb.u_inertial; tb.st.v_samp_inertial = tb.v_inertial;
    tb.st.power = pwr; tb.st.ct = C_T; tb.st.cp = C_P; tb.st.u_samp = s0; tb.st.v_samp = s1;
    tb.st.upstream_u = su; tb.st.upstream_v = sv; tb.st.instant_mag0 = instant_mag0;
    tb.st.nsteps = (int)tb.trace[PU_TURB_POWER].size();
    // yaw controller (:81-103, 1020-1029)
    if (!T->opts.fixed_yaw) {
      const double tau = 60;
      tb.yaw_uavg = (tau - dt) / tau * tb.yaw_uavg + dt / tau * su;
      tb.yaw_vavg = (tau - dt) / tau * tb.yaw_vavg + dt / tau * sv;
      const double dir_upwind = std::atan2(tb.yaw_vavg, tb.yaw_uavg);
      double diff = dir_upwind - tb.yaw;
      if (diff > M_PI) diff -= 2 * M_PI;
      if (diff < -M_PI) diff += 2 * M_PI;
      double tend = diff / dt;
      tend = tend > 0 ? std::min(T->max_yaw_speed, tend) : std::max(-T->max_yaw_speed, tend);
      tb.yaw = tb.yaw + dt * tend;
    }
  }
  // tendencies of all turbines, computed from the same velocities, added afterwards (:1035-1041)
  for (size_t it = 0; it < T->t.size(); it++) {
    const TurbBox &b = T->t[it].pbox;
    if (b.n() == 0) continue;
    hipLaunchKernelGGL(k_turbine_update, dim3((unsigned)((b.n() + 255) / 256)), dim3(256), 0, s,
                       T->d_tend + (long long)it * 3 * T->box_cap, b, nx, ny, dt, cs->uvel, cs->vvel, tke);
    PU_HIP_CHECK(hipGetLastError());
  }
  return PU_OK;
}

int pu_turbines_status(pu_turbines *T, int iturb, pu_turbine_status *out) {
  if (!T || !out || iturb < 0 || iturb >= (int)T->t.size()) return fail(PU_ERR_ARG, "pu_turbines_status: bad argument");
  *out = T->t[iturb].st;
  out->yaw = T->t[iturb].yaw;
  out->u_samp_inertial = T->t[iturb].u_inertial;
  out->v_samp_inertial = T->t[iturb].v_inertial;
  return PU_OK;
}

int pu_turbines_trace(pu_turbines *T, int iturb, int which, double *out) {
  if (!T || !out || iturb < 0 || iturb >= (int)T->t.size() || which < 0 || which >= PU_TURB_NTRACE)
    return fail(PU_ERR_ARG, "pu_turbines_trace: bad argument");
  auto &v = T->t[iturb].trace[which];
  std::copy(v.begin(), v.end(), out);
  return PU_OK;
}

int pu_turbines_destroy(pu_turbines *T) {
  if (!T) return PU_OK;
  if (T->d_acc) hipFree(T->d_acc);
  if (T->d_tend) hipFree(T->d_tend);
  if (T->d_part) hipFree(T->d_part);
  if (T->d_upart) hipFree(T->d_upart);
  if (T->d_err) hipFree(T->d_err);
  delete T;
  return PU_OK;
}

}  // extern "C"
