// les_kernels.hip -- LES_Closure::apply (les_closure.h:30-323) on gfx950.
//
// Two launches per call:
//   k_les_prim   : coupler state -> primitive closure fields (rho_d, u, v, w, theta, K, q_l) on a
//                  halo-1 box, with the periodic/precursor x/y halos (coupler.h:785-851) and the
//                  z wall/periodic halos (les_closure.h:418-536) computed in place;
//   k_les_update : per cell, the six face fluxes (:92-229), the TKE source (:233-278) and the
//                  forward-Euler update (:280-320) written straight back to the coupler entries
//                  (:375-414).  Each face flux is computed by both cells that share it with the same
//                  code, so no flux arrays go to HBM.
// In the FP64 dycore mode the closure runs in double; in the FP32 mode its box, fluxes and
// tendencies are float and the tendencies are applied as increments to the double entries.
#include "pu_common.h"
#include <cmath>

namespace pu {

struct LesParams {
  int nx, ny, nz, ntr_all, nq, idTKE, bc_x, bc_y, bc_z, enable_gravity, dns;
  long long nxh, nyh, nzh, nh, n3;
  int qidx[PU_MAX_TRACERS];        // tracer index of each diffused tracer
  double dx, dy, dz, delta, grav, nu, total_mult, diss_mult, shear_mult, dt;
  double visc_max_x, visc_max_y, visc_max_z;
  double rdelta, nu_pr, rR_d;      // 1 / delta, nu / Pr (Pr = 0.7), 1 / R_d
  double rdx, rdy, rdz;            // 1 / dx, 1 / dy, 1 / dz
  double R_d, C0, gamma;
  // learned closure: per-cell stresses/heat fluxes per face direction d, [3][n3] of
  // (tau_0d, tau_1d, tau_2d, q_d), and flags [n3] (null: eddy viscosity only)
  const float4 *tau;
  const unsigned char *sflags;
};

// periodic image of an index at most one period outside [0, n) (every use here is one cell away)
__device__ __forceinline__ int wrapl(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// fields of the closure box: 0 rho_d, 1 u, 2 v, 3 w, 4 theta, 5 K, 6+q diffused tracer mixing ratios
// (values computed in FP64, stored as T: double, or float for the FP32 dycore mode)
template <class T>
__global__ __launch_bounds__(256) void k_les_prim(const double *__restrict__ rd_, const double *__restrict__ uu,
                                                  const double *__restrict__ vv, const double *__restrict__ ww,
                                                  const double *__restrict__ tt, TracerPtrsC trc,
                                                  T *__restrict__ F, LesParams P) {
  // grid: (blocks over one halo'd level's nxh*nyh points, nzh levels)
  const int nxyh = (int)(P.nxh * P.nyh);
  const int p = blockIdx.x * blockDim.x + threadIdx.x, kh = blockIdx.y;
  if (p >= nxyh) return;
  const int jh = p / (int)P.nxh, ih = p - jh * (int)P.nxh;
  const long long id = (long long)kh * nxyh + p;
  int i = ih - 1, j = jh - 1, k = kh - 1;
  bool ox = i < 0 || i >= P.nx, oy = j < 0 || j >= P.ny;
  // x/y source (periodic exchange; precursor replaces only the face halos, corners stay periodic)
  int is = i, js = j;
  if (ox && oy) { is = wrapl(i, P.nx); js = wrapl(j, P.ny); }
  else if (ox) is = P.bc_x == PU_BC_PRECURSOR ? (i < 0 ? 0 : P.nx - 1) : wrapl(i, P.nx);
  else if (oy) js = P.bc_y == PU_BC_PRECURSOR ? (j < 0 ? 0 : P.ny - 1) : wrapl(j, P.ny);
  bool zhalo = k < 0 || k >= P.nz;
  int ks = k;
  if (zhalo) ks = P.bc_z == PU_BC_PERIODIC ? wrapl(k, P.nz) : (k < 0 ? 0 : P.nz - 1);
  long long sid = ((long long)ks * P.ny + js) * P.nx + is;
  double rho_d = rd_[sid];
  double vals[6];
  vals[0] = rho_d;
  vals[1] = uu[sid];
  vals[2] = vv[sid];
  vals[3] = ww[sid];
  vals[4] = pow(rho_d * P.R_d * tt[sid] / P.C0, 1. / P.gamma) / rho_d;
  vals[5] = fmax(0., trc.p[P.idTKE][sid] / rho_d);
  if (zhalo && P.bc_z != PU_BC_PERIODIC) {
    vals[3] = 0.;
    // hydrostatic extrapolation of rho from the first interior level (les_closure.h:503-510)
    double g = P.enable_gravity ? P.grav : 0.;
    double rho0_gm1 = pow(rho_d, P.gamma - 1), theta0_g = pow(vals[4], P.gamma);
    double sgn = k < 0 ? 1. : -1.;
    vals[0] = pow(rho0_gm1 + sgn * g * (P.gamma - 1) * P.dz * 1 / (P.gamma * P.C0 * theta0_g), 1. / (P.gamma - 1));
  }
  for (int f = 0; f < 6; f++) F[(long long)f * P.nh + id] = (T)vals[f];
  for (int q = 0; q < P.nq; q++) F[(long long)(6 + q) * P.nh + id] = (T)(trc.p[P.qidx[q]][sid] / rho_d);
}

template <class T> struct FaceCommon {
  bool zero;
  T rho, visc, visc_th;
  T f[5];      // ru rv rw rt tke
};

// The closure box as the face fluxes read it: field f at box index idx, the immersed proportion,
// and the box strides.  Tile: the block's cells plus a one-cell halo staged in LDS.
// Cells of one block tile (x, y, z; LTX * LTY * LTZ = 256); PU_LES_T{X,Y,Z} override them for
// tuning builds.
#ifndef PU_LES_TX
#define PU_LES_TX 16
#define PU_LES_TY 4
#define PU_LES_TZ 4
#endif
constexpr int LTX = PU_LES_TX, LTY = PU_LES_TY, LTZ = PU_LES_TZ;
static_assert(LTX * LTY * LTZ == 256, "one cell per thread of a 256-thread block");
constexpr int LBX = LTX + 2, LBY = LTY + 2, LBZ = LTZ + 2;   // tile + one-cell halo

template <class T> struct TileBox {
  const T *sm;                               // closure fields of the tile + halo, [field][n]
  const uint8_t *icode;                      // immersed codes of the box points: 0 none, 1 partial, 2 full
  const float4 *tau;                         // learned closure: global [3][n3] per-direction stresses, or null
  long long n3;
  const unsigned char *flag;                 // learned closure: [n] flags of the box points
  static constexpr int sx = 1, sy = LBX, sz = LBX * LBY, n = LBX * LBY * LBZ;
  __device__ T S(int f, int idx) const { return sm[f * n + idx]; }
  __device__ bool imm_full(int idx) const { return icode[idx] == 2; }   // immersed proportion == 1
  __device__ bool imm_any(int idx) const { return icode[idx] != 0; }    // immersed proportion > 0
};

// Face flux on the low face of box cell c (at halo coordinates kh, jh, ih of the closure box) in
// direction DIR (les_closure.h:92-229).  T = double reproduces the reference's formulas in FP64;
// T = float evaluates them in FP32 (the FP32 dycore mode).
// gm, gc: the global cell ids of the face's low and high cells (periodic boundaries wrapped), used
// by the learned closure only.
template <int DIR, class T>
__device__ __forceinline__ void face_flux(const TileBox<T> &B, const LesParams &P, int c, int kh, int jh, int ih,
                                          long long gm, long long gc, FaceCommon<T> &fc) {
  constexpr int sx = TileBox<T>::sx, sy = TileBox<T>::sy, sz = TileBox<T>::sz;
  const int sn = DIR == 0 ? sx : (DIR == 1 ? sy : sz);            // normal stride
  const int m = c - sn;                                            // the low neighbour
  auto S = [&](int f, int idx) { return B.S(f, idx); };
  if (B.imm_full(m) || B.imm_full(c)) { fc.zero = true; return; }
  fc.zero = false;
  const T h = T(0.5);
  const T rdx = (T)P.rdx, rdy = (T)P.rdy, rdz = (T)P.rdz;
  const T r2dx = h * rdx, r2dy = h * rdy, r2dz = h * rdz;
  T du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz, dt_dz, dtn, dKn;
  T vmax;
  if (DIR == 0) {
    du_dz = h * ((S(1, m + sz) - S(1, m - sz)) * r2dz + (S(1, c + sz) - S(1, c - sz)) * r2dz);
    dw_dz = h * ((S(3, m + sz) - S(3, m - sz)) * r2dz + (S(3, c + sz) - S(3, c - sz)) * r2dz);
    dt_dz = h * ((S(4, m + sz) - S(4, m - sz)) * r2dz + (S(4, c + sz) - S(4, c - sz)) * r2dz);
    du_dy = h * ((S(1, m + sy) - S(1, m - sy)) * r2dy + (S(1, c + sy) - S(1, c - sy)) * r2dy);
    dv_dy = h * ((S(2, m + sy) - S(2, m - sy)) * r2dy + (S(2, c + sy) - S(2, c - sy)) * r2dy);
    du_dx = (S(1, c) - S(1, m)) * rdx;
    dv_dx = (S(2, c) - S(2, m)) * rdx;
    dw_dx = (S(3, c) - S(3, m)) * rdx;
    dtn = (S(4, c) - S(4, m)) * rdx;
    dKn = (S(5, c) - S(5, m)) * rdx;
    dv_dz = dw_dy = T(0);
    vmax = (T)P.visc_max_x;
  } else if (DIR == 1) {
    dv_dz = h * ((S(2, m + sz) - S(2, m - sz)) * r2dz + (S(2, c + sz) - S(2, c - sz)) * r2dz);
    dw_dz = h * ((S(3, m + sz) - S(3, m - sz)) * r2dz + (S(3, c + sz) - S(3, c - sz)) * r2dz);
    dt_dz = h * ((S(4, m + sz) - S(4, m - sz)) * r2dz + (S(4, c + sz) - S(4, c - sz)) * r2dz);
    du_dx = h * ((S(1, m + sx) - S(1, m - sx)) * r2dx + (S(1, c + sx) - S(1, c - sx)) * r2dx);
    dv_dx = h * ((S(2, m + sx) - S(2, m - sx)) * r2dx + (S(2, c + sx) - S(2, c - sx)) * r2dx);
    du_dy = (S(1, c) - S(1, m)) * rdy;
    dv_dy = (S(2, c) - S(2, m)) * rdy;
    dw_dy = (S(3, c) - S(3, m)) * rdy;
    dtn = (S(4, c) - S(4, m)) * rdy;
    dKn = (S(5, c) - S(5, m)) * rdy;
    du_dz = dw_dx = T(0);
    vmax = (T)P.visc_max_y;
  } else {
    du_dx = h * ((S(1, m + sx) - S(1, m - sx)) * r2dx + (S(1, c + sx) - S(1, c - sx)) * r2dx);
    dw_dx = h * ((S(3, m + sx) - S(3, m - sx)) * r2dx + (S(3, c + sx) - S(3, c - sx)) * r2dx);
    dv_dy = h * ((S(2, m + sy) - S(2, m - sy)) * r2dy + (S(2, c + sy) - S(2, c - sy)) * r2dy);
    dw_dy = h * ((S(3, m + sy) - S(3, m - sy)) * r2dy + (S(3, c + sy) - S(3, c - sy)) * r2dy);
    du_dz = (S(1, c) - S(1, m)) * rdz;
    dv_dz = (S(2, c) - S(2, m)) * rdz;
    dw_dz = (S(3, c) - S(3, m)) * rdz;
    dt_dz = (S(4, c) - S(4, m)) * rdz;
    dtn = dt_dz;
    dKn = (S(5, c) - S(5, m)) * rdz;
    du_dy = dv_dx = T(0);
    vmax = (T)P.visc_max_z;
  }
  const T rho = h * (S(0, m) + S(0, c));
  const T K = h * (S(5, m) + S(5, c));
  const T t = h * (S(4, m) + S(4, c));
  // les_closure.h:116-127 with the divisions as reciprocal multiplies (t, N + 1e-20 > 0) and
  // km / Pr_t = km (1 + 2 ell) / delta
  const T sK = sqrt(K);
  const T N = dt_dz >= T(0) && P.enable_gravity ? sqrt((T)P.grav * recip(t) * dt_dz) : T(0);
  const T ell = fmin(T(0.76) * sK * recip(N + T(1.e-20)), (T)P.delta);
  const T km = (T)P.total_mult * T(0.1) * ell * sK;
  const T visc_tot = P.dns ? (T)P.nu : fmin(km + (T)P.nu, h * vmax);
  const T visc_tot_th = P.dns ? (T)P.nu : fmin(km * (T(1) + T(2) * ell) * (T)P.rdelta + (T)P.nu_pr, h * vmax);
  const T t23 = T(2. / 3.);
  fc.rho = rho; fc.visc = visc_tot; fc.visc_th = visc_tot_th;
  if (DIR == 0) {
    fc.f[0] = -rho * visc_tot * (du_dx + du_dx - t23 * (du_dx + dv_dy + dw_dz));
    fc.f[1] = -rho * visc_tot * (dv_dx + du_dy);
    fc.f[2] = -rho * visc_tot * (dw_dx + du_dz);
  } else if (DIR == 1) {
    fc.f[0] = -rho * visc_tot * (du_dy + dv_dx);
    fc.f[1] = -rho * visc_tot * (dv_dy + dv_dy - t23 * (du_dx + dv_dy + dw_dz));
    fc.f[2] = -rho * visc_tot * (dw_dy + dv_dz);
  } else {
    fc.f[0] = -rho * visc_tot * (du_dz + dw_dx);
    fc.f[1] = -rho * visc_tot * (dv_dz + dw_dy);
    fc.f[2] = -rho * visc_tot * (dw_dz + dw_dz - t23 * (du_dx + dv_dy + dw_dz));
  }
  fc.f[3] = -rho * visc_tot_th * (dtn);
  fc.f[4] = -rho * visc_tot * T(2) * (dKn);
  if (B.tau) {
    // Learned closure (no reference counterpart): on a face whose two cells both hold the
    // network's flag, the momentum fluxes become rho * avg(tau_{i,DIR}) (bit 0) and the theta
    // flux rho * avg(q_DIR) (bit 1).  Box points outside the domain carry flag 0 unless the
    // boundary is periodic (then they hold the wrapped cell's values).  The network's outputs are
    // stored per face direction, so a face reads one 16-byte vector of each of its two cells.
    const int both = B.flag[m] & B.flag[c];
    if (both) {
      const float4 a = B.tau[DIR * B.n3 + gm], b = B.tau[DIR * B.n3 + gc];
      if (both & 1) {
        fc.f[0] = rho * (h * ((T)a.x + (T)b.x));
        fc.f[1] = rho * (h * ((T)a.y + (T)b.y));
        fc.f[2] = rho * (h * ((T)a.z + (T)b.z));
      }
      if (both & 2) fc.f[3] = rho * (h * ((T)a.w + (T)b.w));
    }
  }
}

// One block: an LTX x LTY x LTZ tile of cells.  The closure box values of the tile plus a one-cell halo
// (fields rho_d, u, v, w, theta, K and the diffused tracers, and the box points' immersed codes as
// bytes) are staged in LDS once; each thread then evaluates its cell's six face fluxes, the TKE
// source and the update.
// T = double: the reference's update in FP64 (les_closure.h:280-320, :404-413).  T = float: the
// tendencies in FP32, applied as increments to the FP64 coupler entries (u += dt tend_u / rho,
// T *= (1 + dt tend_theta / (rho theta))^gamma, TKE and tracer masses += dt tend), so the entries
// keep their FP64 resolution.
template <class T>
__global__ __launch_bounds__(256) void k_les_update(const T *__restrict__ F, const uint8_t *__restrict__ icode,
                                                    double *uu, double *vv, double *ww, double *tt, TracerPtrs trc,
                                                    LesParams P) {
  extern __shared__ __align__(16) unsigned char smraw[];
  T *sm = reinterpret_cast<T *>(smraw);
  const int nfl = 6 + P.nq;
  unsigned bx, by, bz;
  swizzled_block(bx, by, bz);
  const int i0 = bx * LTX, j0 = by * LTY, k0 = bz * LTZ;
  constexpr int NB = TileBox<T>::n;
  constexpr int PER = (NB + 255) / 256;         // box points per thread
  // Box point offsets of this thread (the same for every field); the loads of FB fields are all
  // issued before their LDS writes, so a block waits on memory a few times, not once per element.
  long long goff[PER];
  bool gin[PER];
#pragma unroll
  for (int e = 0; e < PER; e++) {
    const int q = threadIdx.x + 256 * e;
    const int li = q % LBX, lj = (q / LBX) % LBY, lk = q / (LBX * LBY);
    const int ih = i0 + li, jh = j0 + lj, kh = k0 + lk;       // closure-box (halo) coordinates
    gin[e] = q < NB && ih < P.nxh && jh < P.nyh && kh < P.nzh;
    goff[e] = ((long long)kh * P.nyh + jh) * P.nxh + ih;
  }
  constexpr int FB = 4;
  for (int f0 = 0; f0 < nfl; f0 += FB) {
    T buf[FB][PER];
#pragma unroll
    for (int df = 0; df < FB; df++)
#pragma unroll
      for (int e = 0; e < PER; e++) {
        const int f = f0 + df;
        buf[df][e] = T(0);
        if (f < nfl && gin[e]) buf[df][e] = F[(long long)f * P.nh + goff[e]];
      }
#pragma unroll
    for (int df = 0; df < FB; df++)
#pragma unroll
      for (int e = 0; e < PER; e++) {
        const int q = threadIdx.x + 256 * e;
        if (f0 + df < nfl && q < NB) sm[(f0 + df) * NB + q] = buf[df][e];
      }
  }
  uint8_t *sic = reinterpret_cast<uint8_t *>(sm + NB * nfl);
  {
    uint8_t ib[PER];
#pragma unroll
    for (int e = 0; e < PER; e++) ib[e] = gin[e] ? icode[goff[e]] : (uint8_t)0;
#pragma unroll
    for (int e = 0; e < PER; e++) {
      const int q = threadIdx.x + 256 * e;
      if (q < NB) sic[q] = ib[e];
    }
  }
  unsigned char *sflag = nullptr;
  if (P.sflags) {
    // the network's flags of the box points (interior cells; periodic boundaries wrap)
    sflag = sic + NB;
    unsigned char fb[PER];
#pragma unroll
    for (int e = 0; e < PER; e++) {
      const int q = threadIdx.x + 256 * e;
      int i = i0 + q % LBX - 1, j = j0 + (q / LBX) % LBY - 1, k = k0 + q / (LBX * LBY) - 1;
      bool ok = q < NB;
      // points more than one cell past the domain belong to no cell's faces (ragged last tiles)
      if (i < 0 || i >= P.nx) { if (P.bc_x == PU_BC_PERIODIC && i <= P.nx) i = wrapl(i, P.nx); else ok = false; }
      if (j < 0 || j >= P.ny) { if (P.bc_y == PU_BC_PERIODIC && j <= P.ny) j = wrapl(j, P.ny); else ok = false; }
      if (k < 0 || k >= P.nz) { if (P.bc_z == PU_BC_PERIODIC && k <= P.nz) k = wrapl(k, P.nz); else ok = false; }
      fb[e] = ok ? P.sflags[((long long)k * P.ny + j) * P.nx + i] : (unsigned char)0;
    }
#pragma unroll
    for (int e = 0; e < PER; e++) {
      const int q = threadIdx.x + 256 * e;
      if (q < NB) sflag[q] = fb[e];
    }
  }
  __syncthreads();
  const int li = threadIdx.x % LTX, lj = (threadIdx.x / LTX) % LTY, lk = threadIdx.x / (LTX * LTY);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  if (i >= P.nx || j >= P.ny || k >= P.nz) return;
  const long long id = ((long long)k * P.ny + j) * P.nx + i;
  const int ih = i + 1, jh = j + 1, kh = k + 1;
  const TileBox<T> B{sm, sic, P.sflags ? P.tau : nullptr, P.n3, sflag};
  constexpr int sx = TileBox<T>::sx, sy = TileBox<T>::sy, sz = TileBox<T>::sz;
  const int c = ((lk + 1) * LBY + (lj + 1)) * LBX + (li + 1);
  auto S = [&](int f, int idx) { return B.S(f, idx); };
  FaceCommon<T> fx0, fx1, fy0, fy1, fz0, fz1;
  // global ids of the six neighbours, read only by the learned closure (wrapped where periodic;
  // out-of-domain ones carry flag 0)
  long long gn[6] = {0, 0, 0, 0, 0, 0};
  if (P.sflags) {
    const int im = P.bc_x == PU_BC_PERIODIC ? wrapl(i - 1, P.nx) : i - 1;
    const int ip = P.bc_x == PU_BC_PERIODIC ? wrapl(i + 1, P.nx) : i + 1;
    const int jm = P.bc_y == PU_BC_PERIODIC ? wrapl(j - 1, P.ny) : j - 1;
    const int jp = P.bc_y == PU_BC_PERIODIC ? wrapl(j + 1, P.ny) : j + 1;
    const int km = P.bc_z == PU_BC_PERIODIC ? wrapl(k - 1, P.nz) : k - 1;
    const int kp = P.bc_z == PU_BC_PERIODIC ? wrapl(k + 1, P.nz) : k + 1;
    const long long row = (long long)P.nx, pln = (long long)P.nx * P.ny;
    gn[0] = id + (im - i);
    gn[1] = id + (ip - i);
    gn[2] = id + (jm - j) * row;
    gn[3] = id + (jp - j) * row;
    gn[4] = id + (km - k) * pln;
    gn[5] = id + (kp - k) * pln;
  }
  face_flux<0>(B, P, c, kh, jh, ih, gn[0], id, fx0);
  face_flux<0>(B, P, c + sx, kh, jh, ih + 1, id, gn[1], fx1);
  face_flux<1>(B, P, c, kh, jh, ih, gn[2], id, fy0);
  face_flux<1>(B, P, c + sy, kh, jh + 1, ih, id, gn[3], fy1);
  face_flux<2>(B, P, c, kh, jh, ih, gn[4], id, fz0);
  face_flux<2>(B, P, c + sz, kh + 1, jh, ih, id, gn[5], fz1);
  const T h = T(0.5);
  const T rdx = (T)P.rdx, rdy = (T)P.rdy, rdz = (T)P.rdz;
  const T r2dx = h * rdx, r2dy = h * rdy, r2dz = h * rdz;
  const T rdelta = (T)P.rdelta;
  // TKE source (les_closure.h:233-278)
  const T rho = S(0, c), K = S(5, c), t = S(4, c);
  const T dt_dz = (S(4, c + sz) - S(4, c - sz)) * r2dz;
  const T sK = sqrt(K), rt_ = recip(t);
  const T N = dt_dz >= T(0) && P.enable_gravity ? sqrt((T)P.grav * rt_ * dt_dz) : T(0);
  const T ell = fmin(T(0.76) * sK * recip(N + T(1.e-20)), (T)P.delta);
  const T km = T(0.1) * ell * sK;
  T src = T(0);
  if (!B.imm_full(c)) {
    // -(g rho km) / (t Pr_t) dt_dz with 1 / Pr_t = (1 + 2 ell) / delta
    if (P.enable_gravity) src += -((T)P.grav * rho * km) * (rt_ * (T(1) + T(2) * ell) * rdelta) * dt_dz;
    src -= (T)P.diss_mult * rho * (T(0.19) + T(0.51) * ell * rdelta) * rdelta * (K * sK);
    int im1 = B.imm_any(c - sx) ? c : c - sx, ip1 = B.imm_any(c + sx) ? c : c + sx;
    int jm1 = B.imm_any(c - sy) ? c : c - sy, jp1 = B.imm_any(c + sy) ? c : c + sy;
    int km1 = B.imm_any(c - sz) ? c : c - sz, kp1 = B.imm_any(c + sz) ? c : c + sz;
    const T du_dx = (S(1, c + sx) - S(1, c - sx)) * r2dx;
    const T dv_dx = (S(2, ip1) - S(2, im1)) * r2dx;
    const T dw_dx = (S(3, ip1) - S(3, im1)) * r2dx;
    const T du_dy = (S(1, jp1) - S(1, jm1)) * r2dy;
    const T dv_dy = (S(2, c + sy) - S(2, c - sy)) * r2dy;
    const T dw_dy = (S(3, jp1) - S(3, jm1)) * r2dy;
    const T du_dz = (S(1, kp1) - S(1, km1)) * r2dz;
    const T dv_dz = (S(2, kp1) - S(2, km1)) * r2dz;
    const T dw_dz = (S(3, c + sz) - S(3, c - sz)) * r2dz;
    const T j1_i1 = (du_dx + du_dx) * du_dx, j1_i2 = (dv_dx + du_dy) * dv_dx, j1_i3 = (dw_dx + du_dz) * dw_dx;
    const T j2_i1 = (du_dy + dv_dx) * du_dy, j2_i2 = (dv_dy + dv_dy) * dv_dy, j2_i3 = (dw_dy + dv_dz) * dw_dy;
    const T j3_i1 = (du_dz + dw_dx) * du_dz, j3_i2 = (dv_dz + dw_dy) * dv_dz, j3_i3 = (dw_dz + dw_dz) * dw_dz;
    src += (T)P.shear_mult * rho * km * (j1_i1 + j1_i2 + j1_i3 + j2_i1 + j2_i2 + j2_i3 + j3_i1 + j3_i2 + j3_i3);
  }
  auto fv = [](const FaceCommon<T> &fc, int f) { return fc.zero ? T(0) : fc.f[f]; };
  T tend[5];
  for (int f = 0; f < 5; f++)
    tend[f] = -(fv(fx1, f) - fv(fx0, f)) * rdx - (fv(fy1, f) - fv(fy0, f)) * rdy - (fv(fz1, f) - fv(fz0, f)) * rdz;
  tend[4] += src;
  const T dt = (T)P.dt;
  const T rrho = recip(rho);
  if constexpr (std::is_same<T, double>::value) {
    const double ru = S(1, c) * rho + dt * tend[0];
    const double rv = S(2, c) * rho + dt * tend[1];
    const double rw = S(3, c) * rho + dt * tend[2];
    const double rt = S(4, c) * rho + dt * tend[3];
    const double rk = fmax(0., K * rho + dt * tend[4]);
    uu[id] = ru * rrho;
    vv[id] = rv * rrho;
    ww[id] = rw * rrho;
    tt[id] = P.C0 * pow(rt, P.gamma) * (rrho * P.rR_d);
    trc.p[P.idTKE][id] = rk;
  } else {
    uu[id] += (double)(dt * tend[0] * rrho);
    vv[id] += (double)(dt * tend[1] * rrho);
    ww[id] += (double)(dt * tend[2] * rrho);
    const T x = dt * tend[3] * rrho * recip(t);                  // relative change of rho*theta
    const double T0 = tt[id];
    tt[id] = T0 + T0 * (double)expm1f((float)P.gamma * log1pf(x));
    double *tke = trc.p[P.idTKE];
    tke[id] = fmax(0., fmax(0., tke[id]) + (double)(dt * tend[4]));
  }
  // diffused tracers: flux -rho*visc*dq/dn on each open face
  for (int q = 0; q < P.nq; q++) {
    auto Q = [&](int idx) { return B.S(6 + q, idx); };
    const T qx0 = fx0.zero ? T(0) : -fx0.rho * fx0.visc * ((Q(c) - Q(c - sx)) * rdx);
    const T qx1 = fx1.zero ? T(0) : -fx1.rho * fx1.visc * ((Q(c + sx) - Q(c)) * rdx);
    const T qy0 = fy0.zero ? T(0) : -fy0.rho * fy0.visc * ((Q(c) - Q(c - sy)) * rdy);
    const T qy1 = fy1.zero ? T(0) : -fy1.rho * fy1.visc * ((Q(c + sy) - Q(c)) * rdy);
    const T qz0 = fz0.zero ? T(0) : -fz0.rho * fz0.visc * ((Q(c) - Q(c - sz)) * rdz);
    const T qz1 = fz1.zero ? T(0) : -fz1.rho * fz1.visc * ((Q(c + sz) - Q(c)) * rdz);
    const T tq = -(qx1 - qx0) * rdx - (qy1 - qy0) * rdy - (qz1 - qz0) * rdz;
    if constexpr (std::is_same<T, double>::value) trc.p[P.qidx[q]][id] = Q(c) * rho + dt * tq;
    else trc.p[P.qidx[q]][id] += (double)(dt * tq);
  }
}

static int les_params(pu_ctx *ctx, const pu_coupler_state *cs, double dt, LesParams &P, TracerPtrs &tp,
                      TracerPtrsC &tpc) {
  const pu_config &c = ctx->cfg;
  if (c.idTKE < 0 || c.idTKE >= c.num_tracers) return fail(PU_ERR_ARG, "LES closure needs a TKE tracer (idTKE)");
  P = LesParams{};
  P.nx = c.nx; P.ny = c.ny; P.nz = c.nz; P.ntr_all = c.num_tracers; P.idTKE = c.idTKE;
  P.bc_x = c.bc_x; P.bc_y = c.bc_y; P.bc_z = c.bc_z;
  P.enable_gravity = c.enable_gravity; P.dns = c.dns;
  P.nxh = c.nx + 2; P.nyh = c.ny + 2; P.nzh = c.nz + 2; P.nh = P.nxh * P.nyh * P.nzh;
  P.n3 = (long long)c.nx * c.ny * c.nz;
  P.nq = 0;
  for (int tr = 0; tr < c.num_tracers; tr++) if (c.tracer_diffuse[tr]) P.qidx[P.nq++] = tr;
  P.dx = c.xlen / c.nx; P.dy = c.ylen / c.ny; P.dz = c.zlen / c.nz;
  P.delta = std::pow(P.dx * P.dy * P.dz, 1. / 3.);
  P.rdelta = 1. / P.delta;
  P.nu_pr = c.kinematic_viscosity / 0.7;
  P.rR_d = 1. / c.R_d;
  P.rdx = 1. / P.dx; P.rdy = 1. / P.dy; P.rdz = 1. / P.dz;
  P.grav = c.grav; P.nu = c.kinematic_viscosity;
  P.total_mult = c.les_total_mult; P.diss_mult = c.les_dissipation_mult; P.shear_mult = c.les_shear_prod_mult;
  P.dt = dt;
  P.visc_max_x = 0.25 * P.dx * P.dx / dt; P.visc_max_y = 0.25 * P.dy * P.dy / dt; P.visc_max_z = 0.25 * P.dz * P.dz / dt;
  P.R_d = c.R_d; P.C0 = c.C0; P.gamma = c.gamma_d;
  P.tau = nullptr;
  P.sflags = nullptr;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) {
    tp.p[tr] = tr < c.num_tracers ? cs->tracers[tr] : nullptr;
    tpc.p[tr] = tp.p[tr];
  }
  return PU_OK;
}

template <class T>
static int launch_les_prim(pu_ctx *ctx, const pu_coupler_state *cs, const LesParams &P, const TracerPtrsC &tpc) {
  hipLaunchKernelGGL(k_les_prim<T>, dim3((unsigned)((P.nxh * P.nyh + 255) / 256), (unsigned)P.nzh), dim3(256), 0,
                     ctx->stream, cs->density_dry,
                     cs->uvel, cs->vvel, cs->wvel, cs->temp, tpc, (T *)ctx->d_les, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

template <class T>
static int launch_les_update(pu_ctx *ctx, pu_coupler_state *cs, const LesParams &P, const TracerPtrs &tp) {
  const dim3 grd((unsigned)((P.nx + LTX - 1) / LTX), (unsigned)((P.ny + LTY - 1) / LTY), (unsigned)((P.nz + LTZ - 1) / LTZ));
  // box fields (<= 114 KB for 16 tracers in FP64), the immersed codes and the network's flags
  const size_t lds = sizeof(T) * TileBox<T>::n * (6 + P.nq) + TileBox<T>::n * (P.sflags ? 2 : 1);
  if (lds > 65536)
    PU_HIP_CHECK(hipFuncSetAttribute((const void *)k_les_update<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_les_update<T>, grd, dim3(256), lds, ctx->stream, (const T *)ctx->d_les,
                     (const uint8_t *)ctx->d_imm_les, cs->uvel, cs->vvel, cs->wvel, cs->temp, tp, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int les_prim_device(pu_ctx *ctx, const pu_coupler_state *cs) {
  LesParams P;
  TracerPtrs tp;
  TracerPtrsC tpc;
  int rc = les_params(ctx, cs, 1., P, tp, tpc);
  if (rc) return rc;
  return ctx->cfg.precision == PU_FP32 ? launch_les_prim<float>(ctx, cs, P, tpc) : launch_les_prim<double>(ctx, cs, P, tpc);
}

int les_apply_device(pu_ctx *ctx, pu_coupler_state *cs, double dt) {
  LesParams P;
  TracerPtrs tp;
  TracerPtrsC tpc;
  int rc = les_params(ctx, cs, dt, P, tp, tpc);
  if (rc) return rc;
  const bool f32 = ctx->cfg.precision == PU_FP32;
  {
    LaunchTimer lt(ctx, "les");
    rc = f32 ? launch_les_prim<float>(ctx, cs, P, tpc) : launch_les_prim<double>(ctx, cs, P, tpc);
    if (rc) return rc;
  }
  if (ctx->sgs_model) {
    LaunchTimer lt(ctx, "sgs");
    rc = sgs_infer_device(ctx, dt);
    if (rc) return rc;
    P.tau = reinterpret_cast<const float4 *>(ctx->d_sgs_tau);
    P.sflags = ctx->d_sgs_flags;
  }
  {
    LaunchTimer lt(ctx, "les");
    rc = f32 ? launch_les_update<float>(ctx, cs, P, tp) : launch_les_update<double>(ctx, cs, P, tp);
    if (rc) return rc;
  }
  return PU_OK;
}

}  // namespace pu
