// dycore_kernels.hip -- the per-stage kernels of the WENO finite-volume dycore step on gfx950.
//
// One SSPRK3 stage (dynamics_rk.h:279-367 with compute_tendencies :498-732) is four launches:
//   prep      : conserved state -> reconstructed primitive fields [rho', u, v, w, theta', p', q_l]
//               with the x/y periodic/precursor halos and z wall/periodic halos filled in place
//               (dynamics_rk.h:547-558 + coupler.h:985-1165 + dynamics_rk.h:736-804);
//   sweep x/y/z: WENO reconstruction of every field for every cell of a line segment, the
//               acoustic upwind Riemann flux at each face and the flux divergence, with the face
//               states exchanged through LDS -- no edge arrays in HBM (:572-728, :808-912);
//   the sweeps run z, y, x; the x sweep, as it finishes each cell's tendencies, also applies the
//   gravity/Coriolis sources, the SSPRK stage combination, positive-tracer clipping and the
//   immersed-boundary relaxation (:308-362, :370-473), so the finished tendencies never return to
//   HBM.
// Layout: every field SoA [var][z][y][x], x fastest.  The state is double (PU_FP64) or float with a
// float compensation term for rho and rho*theta (PU_FP32).
#include "sweep_core.h"

namespace pu {

struct PrepParams {
  int nx, ny, nz, hs, hp, ox, ntr, bc_x, bc_y, bc_z;
  long long n3, nxh, nyh, nzh, nh;
  double C0, gamma;
  const double *hy_r, *hy_t, *hy_p;     // cells [nz+2hs]
};

// ------------------------------------------------------------------------------------------------
// prep: primitive fields with halos (cross-shaped: the halo corners are never read)
// ------------------------------------------------------------------------------------------------
// Stores of the primitive fields and of the coupler state, non-temporal (PU_NT_STORE=0: ordinary
// stores): the fields are far larger than the caches and the next kernel reads them from HBM
// either way.  Measured at 1024x1024x128 WENO-9 FP64 (k_prep only): the step 186.4 -> 185.2 ms.
#ifndef PU_NT_STORE
#define PU_NT_STORE 1
#endif
template <class ST>
__device__ __forceinline__ void put_prim(ST *p, ST v) {
  if (PU_NT_STORE) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <class ST>
__global__ __launch_bounds__(256) void k_prep(const ST *__restrict__ U, const ST *__restrict__ Ulo,
                                              ST *__restrict__ prim, PrepParams P) {
  int ih = blockIdx.x * blockDim.x + threadIdx.x;
  int jh = blockIdx.y * blockDim.y + threadIdx.y;
  int kh = blockIdx.z;
  if (ih >= P.nxh || jh >= P.nyh || kh >= P.nzh) return;
  int i = ih - P.ox, j = jh - P.hp, k = kh - P.hp;
  if (i < -P.hp || i >= P.nx + P.hp) return;          // row padding
  bool ox = i < 0 || i >= P.nx, oy = j < 0 || j >= P.ny, oz = k < 0 || k >= P.nz;
  if ((int)ox + (int)oy + (int)oz >= 2) return;
  int is = i, js = j, ks = k;
  bool zwall = false;
  if (ox) is = P.bc_x == PU_BC_PERIODIC ? wrapi(i, P.nx) : clampi(i, 0, P.nx - 1);
  if (oy) js = P.bc_y == PU_BC_PERIODIC ? wrapi(j, P.ny) : clampi(j, 0, P.ny - 1);
  if (oz) {
    if (P.bc_z == PU_BC_PERIODIC) ks = wrapi(k, P.nz);
    else { ks = clampi(k, 0, P.nz - 1); zwall = true; }
  }
  long long sid = ((long long)ks * P.ny + js) * P.nx + is;
  long long hid = ((long long)kh * P.nyh + jh) * P.nxh + ih;
  double rho = load_cons(U, Ulo, idR, sid, P.n3);
  double rt = load_cons(U, Ulo, idT, sid, P.n3);
  int kc = P.hs + ks;
  double r_r = 1. / rho;
  put_prim(prim + (long long)idP * P.nh + hid, (ST)(P.C0 * pow(rt, P.gamma) - P.hy_p[kc]));
  put_prim(prim + (long long)idR * P.nh + hid, (ST)(rho - P.hy_r[kc]));
  put_prim(prim + (long long)idU * P.nh + hid, (ST)((double)U[(long long)idU * P.n3 + sid] * r_r));
  put_prim(prim + (long long)idV * P.nh + hid, (ST)((double)U[(long long)idV * P.n3 + sid] * r_r));
  put_prim(prim + (long long)idW * P.nh + hid, zwall ? (ST)0 : (ST)((double)U[(long long)idW * P.n3 + sid] * r_r));
  put_prim(prim + (long long)idT * P.nh + hid, zwall ? (ST)0 : (ST)(rt * r_r - P.hy_t[kc]));
  for (int tr = 0; tr < P.ntr; tr++)
    put_prim(prim + (long long)(NUM_STATE + 1 + tr) * P.nh + hid, (ST)((double)U[(long long)(NUM_STATE + tr) * P.n3 + sid] * r_r));
}

// ------------------------------------------------------------------------------------------------
// coupler <-> dynamics conversions (dynamics_rk.h:1251-1352), with the first relaxation fused
// ------------------------------------------------------------------------------------------------
struct ConvParams {
  int nx, ny, nz, hs, ntr, idWV, relax;
  long long n3;
  double R_d, R_v, C0, gamma;
  unsigned int adds_mass_mask, positive_mask;
  const double *hy_r, *hy_t, *relax_mult;
};

// One cell of the coupler -> dynamics conversion (dynamics_rk.h:1251-1304): the conserved state
// (rho, rho u, rho v, rho w, rho theta, rho q) of coupler cell id at level k into x[], with the
// opening immersed relaxation when P.relax.  All loads come before any store of the caller (the
// stores may alias the tracer arrays).
__device__ __forceinline__ void c2d_cell(const double *__restrict__ rd_, const double *__restrict__ uu,
                                         const double *__restrict__ vv, const double *__restrict__ ww,
                                         const double *__restrict__ tt, const TracerPtrs &trc, long long id, int k,
                                         const ConvParams &P, double *x) {
  double rho_d = rd_[id], u = uu[id], v = vv[id], w = ww[id], temp = tt[id];
  double m[PU_MAX_TRACERS];
#pragma unroll
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) m[tr] = tr < P.ntr ? trc.p[tr][id] : 0.;
  double rho_v = trc.p[P.idWV][id];
  double press = rho_d * P.R_d * temp + rho_v * P.R_v * temp;
  double rho = rho_d;
#pragma unroll
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) if (tr < P.ntr && ((P.adds_mass_mask >> tr) & 1u)) rho += m[tr];
  double theta = pow(press / P.C0, 1. / P.gamma) / rho;
  double vals[NUM_STATE] = {rho, rho * u, rho * v, rho * w, rho * theta};
  double mult = P.relax ? P.relax_mult[id] : 0.;
  double tr_r = P.hy_r[P.hs + k], tr_rt = P.hy_r[P.hs + k] * P.hy_t[P.hs + k];
#pragma unroll
  for (int l = 0; l < NUM_STATE; l++) {
    double xv = vals[l];
    if (P.relax) {
      double target = l == idR ? tr_r : (l == idT ? tr_rt : 0.);
      xv = xv + (target - xv) * mult;
    }
    x[l] = xv;
  }
#pragma unroll
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) {
    double xv = m[tr];
    if (tr < P.ntr && P.relax) {
      xv = xv + (0. - xv) * mult;
      if ((P.positive_mask >> tr) & 1u) xv = fmax(0., xv);
    }
    x[NUM_STATE + tr] = xv;
  }
}

template <class ST>
__global__ __launch_bounds__(256) void k_c2d(const double *__restrict__ rd_, const double *__restrict__ uu,
                                             const double *__restrict__ vv, const double *__restrict__ ww,
                                             const double *__restrict__ tt, TracerPtrs trc, ST *U, ST *Ulo,
                                             ConvParams P) {
  // grid: (blocks over one level's nx*ny cells, nz levels)
  const int nxy = P.nx * P.ny;
  const int p = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (p >= nxy) return;
  const long long id = (long long)k * nxy + p;
  double x[NUM_STATE + PU_MAX_TRACERS];
  c2d_cell(rd_, uu, vv, ww, tt, trc, id, k, P, x);
#pragma unroll
  for (int l = 0; l < NUM_STATE + PU_MAX_TRACERS; l++)
    if (l < NUM_STATE + P.ntr) store_cons(U, Ulo, l, id, P.n3, x[l]);
}

// The value of conserved field l as load_cons reads it back after store_cons (FP32 storage: the
// float, plus the float compensation term for rho and rho theta).
template <class ST> __device__ __forceinline__ double stored_cons(int l, double v) {
  if constexpr (std::is_same<ST, float>::value) {
    const float hi = (float)v;
    if (l == idR || l == idT) return (double)hi + (double)(float)(v - (double)hi);
    return (double)hi;
  } else {
    return v;
  }
}

// The first stage's prep fused with the coupler -> dynamics conversion: every point of the
// primitive-field box converts its source cell's coupler state (as k_c2d) and derives the primitive
// fields from the conserved values as stored (as k_prep); interior points also store the
// conserved state.  Saves reading the conserved state back (one pass over nv fields).
template <class ST>
__global__ __launch_bounds__(256) void k_c2d_prep(const double *__restrict__ rd_, const double *__restrict__ uu,
                                                  const double *__restrict__ vv, const double *__restrict__ ww,
                                                  const double *__restrict__ tt, TracerPtrs trc, ST *U, ST *Ulo,
                                                  ST *__restrict__ prim, PrepParams P, ConvParams C) {
  int ih = blockIdx.x * blockDim.x + threadIdx.x;
  int jh = blockIdx.y * blockDim.y + threadIdx.y;
  int kh = blockIdx.z;
  if (ih >= P.nxh || jh >= P.nyh || kh >= P.nzh) return;
  int i = ih - P.ox, j = jh - P.hp, k = kh - P.hp;
  if (i < -P.hp || i >= P.nx + P.hp) return;          // row padding
  bool ox = i < 0 || i >= P.nx, oy = j < 0 || j >= P.ny, oz = k < 0 || k >= P.nz;
  if ((int)ox + (int)oy + (int)oz >= 2) return;
  int is = i, js = j, ks = k;
  bool zwall = false;
  if (ox) is = P.bc_x == PU_BC_PERIODIC ? wrapi(i, P.nx) : clampi(i, 0, P.nx - 1);
  if (oy) js = P.bc_y == PU_BC_PERIODIC ? wrapi(j, P.ny) : clampi(j, 0, P.ny - 1);
  if (oz) {
    if (P.bc_z == PU_BC_PERIODIC) ks = wrapi(k, P.nz);
    else { ks = clampi(k, 0, P.nz - 1); zwall = true; }
  }
  long long sid = ((long long)ks * P.ny + js) * P.nx + is;
  long long hid = ((long long)kh * P.nyh + jh) * P.nxh + ih;
  double x[NUM_STATE + PU_MAX_TRACERS];
  c2d_cell(rd_, uu, vv, ww, tt, trc, sid, ks, C, x);
  if (!ox && !oy && !oz) {
#pragma unroll
    for (int l = 0; l < NUM_STATE + PU_MAX_TRACERS; l++)
      if (l < NUM_STATE + P.ntr) store_cons(U, Ulo, l, sid, P.n3, x[l]);
  }
  double rho = stored_cons<ST>(idR, x[idR]);
  double rt = stored_cons<ST>(idT, x[idT]);
  int kc = P.hs + ks;
  double r_r = 1. / rho;
  put_prim(prim + (long long)idP * P.nh + hid, (ST)(P.C0 * pow(rt, P.gamma) - P.hy_p[kc]));
  put_prim(prim + (long long)idR * P.nh + hid, (ST)(rho - P.hy_r[kc]));
  put_prim(prim + (long long)idU * P.nh + hid, (ST)(stored_cons<ST>(idU, x[idU]) * r_r));
  put_prim(prim + (long long)idV * P.nh + hid, (ST)(stored_cons<ST>(idV, x[idV]) * r_r));
  put_prim(prim + (long long)idW * P.nh + hid, zwall ? (ST)0 : (ST)(stored_cons<ST>(idW, x[idW]) * r_r));
  put_prim(prim + (long long)idT * P.nh + hid, zwall ? (ST)0 : (ST)(rt * r_r - P.hy_t[kc]));
#pragma unroll
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++)
    if (tr < P.ntr)
      put_prim(prim + (long long)(NUM_STATE + 1 + tr) * P.nh + hid, (ST)(stored_cons<ST>(NUM_STATE + tr, x[NUM_STATE + tr]) * r_r));
}

template <class ST>
__global__ __launch_bounds__(256) void k_d2c(const ST *__restrict__ U, const ST *__restrict__ Ulo,
                                             double *rd_, double *uu, double *vv, double *ww, double *tt,
                                             TracerPtrs trc, ConvParams P) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.n3) return;
  double rho = load_cons(U, Ulo, idR, id, P.n3);
  double u = load_cons(U, Ulo, idU, id, P.n3) / rho;
  double v = load_cons(U, Ulo, idV, id, P.n3) / rho;
  double w = load_cons(U, Ulo, idW, id, P.n3) / rho;
  double theta = load_cons(U, Ulo, idT, id, P.n3) / rho;
  // all tracer loads before the first store
  double m[PU_MAX_TRACERS];
#pragma unroll
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) m[tr] = tr < P.ntr ? load_cons(U, Ulo, NUM_STATE + tr, id, P.n3) : 0.;
  double press = P.C0 * pow(rho * theta, P.gamma);
  double rho_v = load_cons(U, Ulo, NUM_STATE + P.idWV, id, P.n3);
  double rho_d = rho;
#pragma unroll
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++)
    if (tr < P.ntr && ((P.adds_mass_mask >> tr) & 1u)) rho_d -= m[tr];
  double temp = press / (rho_d * P.R_d + rho_v * P.R_v);
  put_prim(rd_ + id, rho_d); put_prim(uu + id, u); put_prim(vv + id, v); put_prim(ww + id, w); put_prim(tt + id, temp);
#pragma unroll
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++)
    if (tr < P.ntr) put_prim(trc.p[tr] + id, m[tr]);
}

template <class ST>
__global__ __launch_bounds__(256) void k_tend_export(const ST *__restrict__ tend, const ST *U, const ST *Ulo,
                                                     double *out, CombineParams P) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.n3) return;
  int k = (int)(id / ((long long)P.nx * P.ny));
  const int nv = NUM_STATE + P.ntr;
  double rho_in = load_cons(U, Ulo, idR, id, P.n3);
  for (int l = 0; l < nv; l++) {
    double tn = (double)tend[(long long)l * P.n3 + id];
    if (l == idV && P.sim2d) tn = 0.;
    if (l == idW && P.enable_gravity) tn += -P.grav * (rho_in - P.hy_r[P.hs + k]);
    if (P.coriolis && l == idU) tn += P.fcor * load_cons(U, Ulo, idV, id, P.n3);
    if (P.coriolis && l == idV) tn -= P.fcor * load_cons(U, Ulo, idU, id, P.n3);
    out[(long long)l * P.n3 + id] = tn;
  }
}

// ------------------------------------------------------------------------------------------------
// Reconstruction census of one sweep direction: for every interior cell and every field the sweep
// reconstructs (rho, the normal velocity and p' first, then the transverse velocities, theta' and
// the tracers, with the sweeps' zero-derivative clamps), the branch the limiter takes -- counts of
// [limiter, flat stencil, linear] added to counts[0..2].  Measures the work the sweeps execute on a
// given state.  grid: (blocks over a level's nx*ny cells, nz levels).
// ------------------------------------------------------------------------------------------------
template <int ORD, class T, int DIR>
__global__ __launch_bounds__(256) void k_flat_census(const T *__restrict__ prim, const uint8_t *__restrict__ immf,
                                                     const uint8_t *__restrict__ any10, SweepParams P,
                                                     unsigned long long *counts) {
  __shared__ unsigned long long sc[3];
  if (threadIdx.x < 3) sc[threadIdx.x] = 0;
  __syncthreads();
  const int nxy = P.nx * P.ny;
  const int p = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  unsigned c[3] = {0u, 0u, 0u};
  if (p < nxy) {
    const int i = p % P.nx, j = p / P.nx;
    const long long sstride = DIR == 0 ? 1 : (DIR == 1 ? P.nxh : P.nxh * P.nyh);
    const long long base = ((long long)(k + P.hp) * P.nyh + (j + P.hp)) * P.nxh + (i + P.ox);
    const long long cid = (long long)k * nxy + p;
    StencilFlags<ORD> sf;
    sf.setup(immf, base, sstride, any10[cid] != 0, P.weno_all);
    constexpr int ln = 1 + DIR, lt1 = DIR == 0 ? idV : idU, lt2 = DIR == 2 ? idV : idW;
    const int ls[5] = {idR, ln, idP, lt1, lt2};
    for (int f = 0; f < 5; f++) c[sf.template branch<T>(prim + (long long)ls[f] * P.nh + base, sstride, f >= 2)]++;
    c[sf.template branch<T>(prim + (long long)idT * P.nh + base, sstride, false)]++;
    for (int tr = 0; tr < P.ntr; tr++)
      c[sf.template branch<T>(prim + (long long)(NUM_STATE + 1 + tr) * P.nh + base, sstride, false)]++;
  }
  for (int q = 0; q < 3; q++) if (c[q]) atomicAdd(&sc[q], (unsigned long long)c[q]);
  __syncthreads();
  if (threadIdx.x < 3 && sc[threadIdx.x]) atomicAdd(&counts[threadIdx.x], sc[threadIdx.x]);
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static inline const double *hy_cells(pu_ctx *ctx, int w) { return ctx->d_hy + (size_t)w * (ctx->g.nz + 2 * ctx->g.hs); }
static inline const double *hy_edges(pu_ctx *ctx, int w) {
  return ctx->d_hy + (size_t)3 * (ctx->g.nz + 2 * ctx->g.hs) + (size_t)w * (ctx->g.nz + 1);
}

template <class ST>
static int prep_t(pu_ctx *ctx, const void *U, const void *Ulo, hipStream_t s) {
  const pu::Grid &g = ctx->g;
  PrepParams P;
  P.nx = g.nx; P.ny = g.ny; P.nz = g.nz; P.hs = g.hs; P.hp = g.hp; P.ox = g.ox; P.ntr = ctx->cfg.num_tracers;
  P.bc_x = ctx->cfg.bc_x; P.bc_y = ctx->cfg.bc_y; P.bc_z = ctx->cfg.bc_z;
  P.n3 = g.n3(); P.nxh = g.nxh; P.nyh = g.nyh; P.nzh = g.nzh; P.nh = g.nh();
  P.C0 = ctx->cfg.C0; P.gamma = ctx->cfg.gamma_d;
  P.hy_r = hy_cells(ctx, 0); P.hy_t = hy_cells(ctx, 1); P.hy_p = hy_cells(ctx, 2);
  dim3 blk(64, 4, 1), grd((unsigned)((g.nxh + 63) / 64), (unsigned)((g.nyh + 3) / 4), (unsigned)g.nzh);
  LaunchTimer lt(ctx, "prep");
  hipLaunchKernelGGL(k_prep<ST>, grd, blk, 0, s, (const ST *)U, (const ST *)Ulo, (ST *)ctx->d_prim, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int launch_prep(pu_ctx *ctx, const void *U, const void *Ulo, hipStream_t s) {
  return ctx->cfg.precision == PU_FP32 ? prep_t<float>(ctx, U, Ulo, s) : prep_t<double>(ctx, U, Ulo, s);
}

static CombineParams combine_params(pu_ctx *ctx);

static SweepParams sweep_params(pu_ctx *ctx, int dir) {
  const pu::Grid &g = ctx->g;
  SweepParams P;
  P.nx = g.nx; P.ny = g.ny; P.nz = g.nz; P.hs = g.hs; P.hp = g.hp; P.ox = g.ox; P.ntr = ctx->cfg.num_tracers;
  P.bc_along = dir == 0 ? ctx->cfg.bc_x : (dir == 1 ? ctx->cfg.bc_y : ctx->cfg.bc_z);
  if (dir == 2 && P.bc_along == PU_BC_SOLID_WALL) P.bc_along = PU_BC_PRECURSOR;   // non-periodic
  P.wall_z = ctx->cfg.bc_z == PU_BC_SOLID_WALL;
  P.weno_all = ctx->cfg.weno_all;
  P.n3 = g.n3(); P.nxh = g.nxh; P.nyh = g.nyh; P.nzh = g.nzh; P.nh = g.nh();
  double d = dir == 0 ? ctx->cfg.xlen / g.nx : (dir == 1 ? ctx->cfg.ylen / g.ny : ctx->cfg.zlen / g.nz);
  P.rd = 1. / d;
  P.hy_rc = hy_cells(ctx, 0); P.hy_tc = hy_cells(ctx, 1);
  P.hy_re = hy_edges(ctx, 0); P.hy_te = hy_edges(ctx, 1);
  return P;
}

// mode: 0 first sweep (writes d_tend), 1 adds to d_tend, 2 adds and combines stage `stage`
template <int ORD, class T, int DIR>
static int sweep_one(pu_ctx *ctx, hipStream_t s, int mode, int stage, double dt, int nrelax) {
  const pu::Grid &g = ctx->g;
  SweepParams P = sweep_params(ctx, DIR);
  static const char *names[3] = {"weno_x", "weno_y", "weno_z"};
  LaunchTimer lt(ctx, names[DIR]);
  constexpr int BX = SegShape<DIR>::BX, BY = SegShape<DIR>::BY, NC = SegShape<DIR>::BS - 2;
  dim3 blk(BX, BY, 1), grd;
  if (DIR == 0) grd = dim3((unsigned)((g.nx + NC - 1) / NC), (unsigned)((g.ny + BY - 1) / BY), (unsigned)g.nz);
  else if (DIR == 1) grd = dim3((unsigned)((g.nx + BX - 1) / BX), (unsigned)((g.ny + NC - 1) / NC), (unsigned)g.nz);
  else grd = dim3((unsigned)((g.nx + BX - 1) / BX), (unsigned)g.ny, (unsigned)((g.nz + NC - 1) / NC));
  FuseArgs<T> FA{};
  const T *prim = (const T *)ctx->d_prim;
  T *tend = (T *)ctx->d_tend;
  constexpr bool line_kernel =
      DIR <= 1 && ((PU_PLINE >> ((DIR == 0 ? 1 : 0) + (std::is_same<T, float>::value ? 2 : 0))) & 1);
  if constexpr (line_kernel) {
    if (!ctx->segment_sweeps[DIR] && mode >= 1 && P.bc_along == PU_BC_PERIODIC && (DIR == 0 ? g.nx : g.ny) >= 2) {
      constexpr int D = DIR <= 1 ? DIR : 0;
      dim3 lb(LineShape<D>::BX, LineShape<D>::BY, 1);
      dim3 lg(DIR == 0 ? (unsigned)((g.ny + lb.y - 1) / lb.y) : (unsigned)((g.nx + lb.x - 1) / lb.x), (unsigned)g.nz, 1);
      if constexpr (std::is_same<T, float>::value && DIR == 1) {
        // FP32 y lines: their own translation unit (sweep_pline_f32.hip); y never combines the stage
        if (mode != 1) return fail(PU_ERR_ARG, "the y sweep does not combine the stage");
        const int rc = launch_pline_y_f32(ORD, lg, lb, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
        if (rc) return rc;
      } else if (mode == 2) {
        FA.U = (T *)ctx->d_U; FA.Ulo = (T *)ctx->d_Ulo; FA.Us = (T *)ctx->d_Us; FA.Uslo = (T *)ctx->d_Uslo;
        FA.C = combine_params(ctx);
        FA.C.stage = stage; FA.C.dt = dt; FA.C.nrelax = nrelax;
        hipLaunchKernelGGL((k_sweep_pline<ORD, T, D, 2>), lg, lb, 0, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
      } else {
        hipLaunchKernelGGL((k_sweep_pline<ORD, T, D, 1>), lg, lb, 0, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
      }
      PU_HIP_CHECK(hipGetLastError());
      return PU_OK;
    }
  }
  if (PU_ZWALL && !ctx->segment_sweeps[2] && DIR == 2 && mode == 0 && ctx->cfg.bc_z == PU_BC_SOLID_WALL) {
    dim3 zb(ZW_BX, ZW_BS, 1), zg((unsigned)((g.nx + ZW_BX - 1) / ZW_BX), (unsigned)g.ny, 1);
    if constexpr (std::is_same<T, float>::value) {
      const int rc = launch_zwall_f32(ORD, zg, zb, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL((k_sweep_zwall<ORD, T>), zg, zb, 0, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P);
    }
  } else if (mode == 2) {
    FA.U = (T *)ctx->d_U; FA.Ulo = (T *)ctx->d_Ulo; FA.Us = (T *)ctx->d_Us; FA.Uslo = (T *)ctx->d_Uslo;
    FA.C = combine_params(ctx);
    FA.C.stage = stage; FA.C.dt = dt; FA.C.nrelax = nrelax;
    if constexpr (std::is_same<T, float>::value && DIR == 0)
      return launch_seg_x_f32(ORD, 2, grd, blk, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
    else
      hipLaunchKernelGGL((k_sweep<ORD, T, DIR, 2>), grd, blk, 0, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
  } else if constexpr (std::is_same<T, float>::value && DIR == 0) {
    // FP32 x segments: their own translation unit (sweep_seg_x_f32.hip)
    return launch_seg_x_f32(ORD, mode, grd, blk, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
  } else if (mode == 1) {
    hipLaunchKernelGGL((k_sweep<ORD, T, DIR, 1>), grd, blk, 0, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
  } else {
    hipLaunchKernelGGL((k_sweep<ORD, T, DIR, 0>), grd, blk, 0, s, prim, ctx->d_imm_flag, ctx->d_any10, tend, P, FA);
  }
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

// z, then y, then x (the x sweep finishes the tendencies and, for stage >= 0, combines the stage)
template <int ORD, class T>
static int sweeps_t(pu_ctx *ctx, hipStream_t s, int stage, double dt, int nrelax) {
  int rc;
  if ((rc = sweep_one<ORD, T, 2>(ctx, s, 0, stage, dt, nrelax))) return rc;
  if (ctx->g.ny > 1 && (rc = sweep_one<ORD, T, 1>(ctx, s, 1, stage, dt, nrelax))) return rc;
  if ((rc = sweep_one<ORD, T, 0>(ctx, s, stage >= 0 ? 2 : 1, stage, dt, nrelax))) return rc;
  return PU_OK;
}

int launch_sweeps(pu_ctx *ctx, hipStream_t s, int stage, double dt, int nrelax) {
  bool f32 = ctx->cfg.precision == PU_FP32;
  switch (ctx->cfg.ord) {
    case 3: return f32 ? sweeps_t<3, float>(ctx, s, stage, dt, nrelax) : sweeps_t<3, double>(ctx, s, stage, dt, nrelax);
    case 5: return f32 ? sweeps_t<5, float>(ctx, s, stage, dt, nrelax) : sweeps_t<5, double>(ctx, s, stage, dt, nrelax);
    case 7: return f32 ? sweeps_t<7, float>(ctx, s, stage, dt, nrelax) : sweeps_t<7, double>(ctx, s, stage, dt, nrelax);
    case 9: return f32 ? sweeps_t<9, float>(ctx, s, stage, dt, nrelax) : sweeps_t<9, double>(ctx, s, stage, dt, nrelax);
    case 11: return f32 ? sweeps_t<11, float>(ctx, s, stage, dt, nrelax) : sweeps_t<11, double>(ctx, s, stage, dt, nrelax);
  }
  return fail(PU_ERR_UNSUPPORTED, "unsupported ord");
}

template <int ORD, class T>
static int census_t(pu_ctx *ctx, hipStream_t s, unsigned long long *counts) {
  const pu::Grid &g = ctx->g;
  dim3 grd((unsigned)((g.nx * g.ny + 255) / 256), (unsigned)g.nz), blk(256);
  const T *prim = (const T *)ctx->d_prim;
  hipLaunchKernelGGL((k_flat_census<ORD, T, 0>), grd, blk, 0, s, prim, ctx->d_imm_flag, ctx->d_any10,
                     sweep_params(ctx, 0), counts);
  if (g.ny > 1)
    hipLaunchKernelGGL((k_flat_census<ORD, T, 1>), grd, blk, 0, s, prim, ctx->d_imm_flag, ctx->d_any10,
                       sweep_params(ctx, 1), counts + 3);
  hipLaunchKernelGGL((k_flat_census<ORD, T, 2>), grd, blk, 0, s, prim, ctx->d_imm_flag, ctx->d_any10,
                     sweep_params(ctx, 2), counts + 6);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

// census of the reconstruction branches on the primitive fields of the last prep; counts: device
// [3 directions][limiter, flat, linear], zeroed by the caller
int launch_census(pu_ctx *ctx, hipStream_t s, unsigned long long *counts) {
  bool f32 = ctx->cfg.precision == PU_FP32;
  switch (ctx->cfg.ord) {
    case 3: return f32 ? census_t<3, float>(ctx, s, counts) : census_t<3, double>(ctx, s, counts);
    case 5: return f32 ? census_t<5, float>(ctx, s, counts) : census_t<5, double>(ctx, s, counts);
    case 7: return f32 ? census_t<7, float>(ctx, s, counts) : census_t<7, double>(ctx, s, counts);
    case 9: return f32 ? census_t<9, float>(ctx, s, counts) : census_t<9, double>(ctx, s, counts);
    case 11: return f32 ? census_t<11, float>(ctx, s, counts) : census_t<11, double>(ctx, s, counts);
  }
  return fail(PU_ERR_UNSUPPORTED, "unsupported ord");
}

static CombineParams combine_params(pu_ctx *ctx) {
  CombineParams P{};
  P.nx = ctx->g.nx; P.ny = ctx->g.ny; P.nz = ctx->g.nz; P.hs = ctx->g.hs; P.ntr = ctx->cfg.num_tracers;
  P.n3 = ctx->g.n3();
  P.enable_gravity = ctx->cfg.enable_gravity;
  P.grav = ctx->cfg.grav;
  // a 2-D run (ny == 1, Coupler::is_sim2d) has no y fluxes, no Coriolis and a zero V tendency
  // (dynamics_rk.h:663,716,720-721); the y sweep is not launched
  P.sim2d = ctx->g.ny == 1;
  P.coriolis = ctx->cfg.latitude != 0 && !P.sim2d;
  P.fcor = 2 * 7.2921e-5 * sin(ctx->cfg.latitude / 180 * M_PI);
  P.hy_r = hy_cells(ctx, 0); P.hy_t = hy_cells(ctx, 1);
  P.relax = ctx->d_relax;
  P.positive_mask = 0;
  for (int tr = 0; tr < ctx->cfg.num_tracers; tr++) if (ctx->cfg.tracer_positive[tr]) P.positive_mask |= 1u << tr;
  P.stage = 0; P.nrelax = 1; P.dt = 0;
  return P;
}

static ConvParams conv_params(pu_ctx *ctx) {
  ConvParams P;
  P.nx = ctx->g.nx; P.ny = ctx->g.ny; P.nz = ctx->g.nz; P.hs = ctx->g.hs; P.ntr = ctx->cfg.num_tracers;
  P.idWV = ctx->cfg.idWV; P.n3 = ctx->g.n3();
  P.R_d = ctx->cfg.R_d; P.R_v = ctx->cfg.R_v; P.C0 = ctx->cfg.C0; P.gamma = ctx->cfg.gamma_d;
  P.adds_mass_mask = 0; P.positive_mask = 0;
  for (int tr = 0; tr < ctx->cfg.num_tracers; tr++) {
    if (ctx->cfg.tracer_adds_mass[tr]) P.adds_mass_mask |= 1u << tr;
    if (ctx->cfg.tracer_positive[tr]) P.positive_mask |= 1u << tr;
  }
  P.hy_r = hy_cells(ctx, 0); P.hy_t = hy_cells(ctx, 1); P.relax_mult = ctx->d_relax;
  P.relax = 0;
  return P;
}

int launch_convert_c2d(pu_ctx *ctx, const pu_coupler_state *cs, int relax, hipStream_t s) {
  ConvParams P = conv_params(ctx);
  P.relax = relax;
  TracerPtrs tp;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) tp.p[tr] = tr < ctx->cfg.num_tracers ? cs->tracers[tr] : nullptr;
  dim3 grd((unsigned)((ctx->g.nx * ctx->g.ny + 255) / 256), (unsigned)ctx->g.nz), blk(256);
  if (ctx->cfg.precision == PU_FP32)
    hipLaunchKernelGGL(k_c2d<float>, grd, blk, 0, s, cs->density_dry, cs->uvel, cs->vvel, cs->wvel, cs->temp, tp,
                       (float *)ctx->d_U, (float *)ctx->d_Ulo, P);
  else
    hipLaunchKernelGGL(k_c2d<double>, grd, blk, 0, s, cs->density_dry, cs->uvel, cs->vvel, cs->wvel, cs->temp, tp,
                       (double *)ctx->d_U, (double *)ctx->d_Ulo, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

template <class ST>
static int c2d_prep_t(pu_ctx *ctx, const pu_coupler_state *cs, hipStream_t s) {
  const pu::Grid &g = ctx->g;
  PrepParams P;
  P.nx = g.nx; P.ny = g.ny; P.nz = g.nz; P.hs = g.hs; P.hp = g.hp; P.ox = g.ox; P.ntr = ctx->cfg.num_tracers;
  P.bc_x = ctx->cfg.bc_x; P.bc_y = ctx->cfg.bc_y; P.bc_z = ctx->cfg.bc_z;
  P.n3 = g.n3(); P.nxh = g.nxh; P.nyh = g.nyh; P.nzh = g.nzh; P.nh = g.nh();
  P.C0 = ctx->cfg.C0; P.gamma = ctx->cfg.gamma_d;
  P.hy_r = hy_cells(ctx, 0); P.hy_t = hy_cells(ctx, 1); P.hy_p = hy_cells(ctx, 2);
  ConvParams C = conv_params(ctx);
  C.relax = 1;
  TracerPtrs tp;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) tp.p[tr] = tr < ctx->cfg.num_tracers ? cs->tracers[tr] : nullptr;
  dim3 blk(64, 4, 1), grd((unsigned)((g.nxh + 63) / 64), (unsigned)((g.nyh + 3) / 4), (unsigned)g.nzh);
  LaunchTimer lt(ctx, "prep");
  hipLaunchKernelGGL(k_c2d_prep<ST>, grd, blk, 0, s, cs->density_dry, cs->uvel, cs->vvel, cs->wvel, cs->temp, tp,
                     (ST *)ctx->d_U, (ST *)ctx->d_Ulo, (ST *)ctx->d_prim, P, C);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

// the conversion with the opening relaxation of a time step, fused with its first stage's prep
int launch_c2d_prep(pu_ctx *ctx, const pu_coupler_state *cs, hipStream_t s) {
  return ctx->cfg.precision == PU_FP32 ? c2d_prep_t<float>(ctx, cs, s) : c2d_prep_t<double>(ctx, cs, s);
}

int launch_convert_d2c(pu_ctx *ctx, pu_coupler_state *cs, hipStream_t s) {
  ConvParams P = conv_params(ctx);
  TracerPtrs tp;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) tp.p[tr] = tr < ctx->cfg.num_tracers ? cs->tracers[tr] : nullptr;
  long long n = ctx->g.n3();
  dim3 grd((unsigned)((n + 255) / 256)), blk(256);
  if (ctx->cfg.precision == PU_FP32)
    hipLaunchKernelGGL(k_d2c<float>, grd, blk, 0, s, (const float *)ctx->d_U, (const float *)ctx->d_Ulo,
                       cs->density_dry, cs->uvel, cs->vvel, cs->wvel, cs->temp, tp, P);
  else
    hipLaunchKernelGGL(k_d2c<double>, grd, blk, 0, s, (const double *)ctx->d_U, (const double *)ctx->d_Ulo,
                       cs->density_dry, cs->uvel, cs->vvel, cs->wvel, cs->temp, tp, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int launch_tend_export(pu_ctx *ctx, double *out, hipStream_t s) {
  CombineParams P = combine_params(ctx);
  long long n = ctx->g.n3();
  dim3 grd((unsigned)((n + 255) / 256)), blk(256);
  if (ctx->cfg.precision == PU_FP32)
    hipLaunchKernelGGL(k_tend_export<float>, grd, blk, 0, s, (const float *)ctx->d_tend, (const float *)ctx->d_U,
                       (const float *)ctx->d_Ulo, out, P);
  else
    hipLaunchKernelGGL(k_tend_export<double>, grd, blk, 0, s, (const double *)ctx->d_tend, (const double *)ctx->d_U,
                       (const double *)ctx->d_Ulo, out, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

}  // namespace pu
