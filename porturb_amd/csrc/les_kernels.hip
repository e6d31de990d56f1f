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
#include "pu_common.h"
#include <cmath>

namespace pu {

struct LesParams {
  int nx, ny, nz, ntr_all, nq, idTKE, bc_x, bc_y, bc_z, enable_gravity, dns;
  long long nxh, nyh, nzh, nh, n3;
  int qidx[PU_MAX_TRACERS];        // tracer index of each diffused tracer
  double dx, dy, dz, delta, grav, nu, total_mult, diss_mult, shear_mult, dt;
  double visc_max_x, visc_max_y, visc_max_z;
  double R_d, C0, gamma;
  const double *imm;               // immersed_proportion_halos [nzh][nyh][nxh]
  // learned closure: per-cell stresses/heat fluxes [9][n3] and flags [n3] (null: eddy viscosity only)
  const float *tau;
  const unsigned char *sflags;
};

__device__ __forceinline__ int wrapl(int i, int n) { i %= n; return i < 0 ? i + n : i; }

// fields of the closure box: 0 rho_d, 1 u, 2 v, 3 w, 4 theta, 5 K, 6+q diffused tracer mixing ratios
__global__ __launch_bounds__(256) void k_les_prim(const double *__restrict__ rd_, const double *__restrict__ uu,
                                                  const double *__restrict__ vv, const double *__restrict__ ww,
                                                  const double *__restrict__ tt, TracerPtrsC trc,
                                                  double *__restrict__ F, LesParams P) {
  long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.nh) return;
  int ih = (int)(id % P.nxh), jh = (int)((id / P.nxh) % P.nyh), kh = (int)(id / (P.nxh * P.nyh));
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
  for (int f = 0; f < 6; f++) F[(long long)f * P.nh + id] = vals[f];
  for (int q = 0; q < P.nq; q++) F[(long long)(6 + q) * P.nh + id] = trc.p[P.qidx[q]][sid] / rho_d;
}

struct FaceCommon {
  bool zero;
  double rho, visc, visc_th;
  double f[5];      // ru rv rw rt tke
};

// The closure box as the face fluxes read it: field f at box index idx, the immersed proportion,
// and the box strides.  Tile: the block's cells plus a one-cell halo staged in LDS.
struct TileBox {
  const double *sm;
  int nfl;                                   // fields before the immersed-proportion slot
  static constexpr int sx = 1, sy = 18, sz = 18 * 6, n = 18 * 6 * 6;
  __device__ double S(int f, int idx) const { return sm[f * n + idx]; }
  __device__ double imm(int idx) const { return sm[nfl * n + idx]; }
};

// Face flux on the low face of box cell c (at halo coordinates kh, jh, ih of the closure box) in
// direction DIR (les_closure.h:92-229).
template <int DIR, class Box>
__device__ __forceinline__ void face_flux(const Box &B, const LesParams &P, int c, int kh, int jh, int ih,
                                          FaceCommon &fc) {
  constexpr int sx = Box::sx, sy = Box::sy, sz = Box::sz;
  const int sn = DIR == 0 ? sx : (DIR == 1 ? sy : sz);            // normal stride
  const int m = c - sn;                                            // the low neighbour
  auto S = [&](int f, int idx) { return B.S(f, idx); };
  if (B.imm(m) == 1 || B.imm(c) == 1) { fc.zero = true; return; }
  fc.zero = false;
  const double rdx = 1. / P.dx, rdy = 1. / P.dy, rdz = 1. / P.dz;
  const double r2dx = 0.5 * rdx, r2dy = 0.5 * rdy, r2dz = 0.5 * rdz;
  double du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz, dt_dz, dtn, dKn;
  double vmax;
  if (DIR == 0) {
    du_dz = 0.5 * ((S(1, m + sz) - S(1, m - sz)) * r2dz + (S(1, c + sz) - S(1, c - sz)) * r2dz);
    dw_dz = 0.5 * ((S(3, m + sz) - S(3, m - sz)) * r2dz + (S(3, c + sz) - S(3, c - sz)) * r2dz);
    dt_dz = 0.5 * ((S(4, m + sz) - S(4, m - sz)) * r2dz + (S(4, c + sz) - S(4, c - sz)) * r2dz);
    du_dy = 0.5 * ((S(1, m + sy) - S(1, m - sy)) * r2dy + (S(1, c + sy) - S(1, c - sy)) * r2dy);
    dv_dy = 0.5 * ((S(2, m + sy) - S(2, m - sy)) * r2dy + (S(2, c + sy) - S(2, c - sy)) * r2dy);
    du_dx = (S(1, c) - S(1, m)) * rdx;
    dv_dx = (S(2, c) - S(2, m)) * rdx;
    dw_dx = (S(3, c) - S(3, m)) * rdx;
    dtn = (S(4, c) - S(4, m)) * rdx;
    dKn = (S(5, c) - S(5, m)) * rdx;
    dv_dz = dw_dy = 0;
    vmax = P.visc_max_x;
  } else if (DIR == 1) {
    dv_dz = 0.5 * ((S(2, m + sz) - S(2, m - sz)) * r2dz + (S(2, c + sz) - S(2, c - sz)) * r2dz);
    dw_dz = 0.5 * ((S(3, m + sz) - S(3, m - sz)) * r2dz + (S(3, c + sz) - S(3, c - sz)) * r2dz);
    dt_dz = 0.5 * ((S(4, m + sz) - S(4, m - sz)) * r2dz + (S(4, c + sz) - S(4, c - sz)) * r2dz);
    du_dx = 0.5 * ((S(1, m + sx) - S(1, m - sx)) * r2dx + (S(1, c + sx) - S(1, c - sx)) * r2dx);
    dv_dx = 0.5 * ((S(2, m + sx) - S(2, m - sx)) * r2dx + (S(2, c + sx) - S(2, c - sx)) * r2dx);
    du_dy = (S(1, c) - S(1, m)) * rdy;
    dv_dy = (S(2, c) - S(2, m)) * rdy;
    dw_dy = (S(3, c) - S(3, m)) * rdy;
    dtn = (S(4, c) - S(4, m)) * rdy;
    dKn = (S(5, c) - S(5, m)) * rdy;
    du_dz = dw_dx = 0;
    vmax = P.visc_max_y;
  } else {
    du_dx = 0.5 * ((S(1, m + sx) - S(1, m - sx)) * r2dx + (S(1, c + sx) - S(1, c - sx)) * r2dx);
    dw_dx = 0.5 * ((S(3, m + sx) - S(3, m - sx)) * r2dx + (S(3, c + sx) - S(3, c - sx)) * r2dx);
    dv_dy = 0.5 * ((S(2, m + sy) - S(2, m - sy)) * r2dy + (S(2, c + sy) - S(2, c - sy)) * r2dy);
    dw_dy = 0.5 * ((S(3, m + sy) - S(3, m - sy)) * r2dy + (S(3, c + sy) - S(3, c - sy)) * r2dy);
    du_dz = (S(1, c) - S(1, m)) * rdz;
    dv_dz = (S(2, c) - S(2, m)) * rdz;
    dw_dz = (S(3, c) - S(3, m)) * rdz;
    dt_dz = (S(4, c) - S(4, m)) * rdz;
    dtn = dt_dz;
    dKn = (S(5, c) - S(5, m)) * rdz;
    du_dy = dv_dx = 0;
    vmax = P.visc_max_z;
  }
  double rho = 0.5 * (S(0, m) + S(0, c));
  double K = 0.5 * (S(5, m) + S(5, c));
  double t = 0.5 * (S(4, m) + S(4, c));
  double N = dt_dz >= 0 && P.enable_gravity ? sqrt(P.grav / t * dt_dz) : 0;
  double ell = fmin(0.76 * sqrt(K) / (N + 1.e-20), P.delta);
  double km = P.total_mult * 0.1 * ell * sqrt(K);
  double Pr_t = P.delta / (1 + 2 * ell);
  const double Pr = 0.7;
  double visc_tot = P.dns ? P.nu : fmin(km + P.nu, 0.5 * vmax);
  double visc_tot_th = P.dns ? P.nu : fmin(km / Pr_t + P.nu / Pr, 0.5 * vmax);
  fc.rho = rho; fc.visc = visc_tot; fc.visc_th = visc_tot_th;
  if (DIR == 0) {
    fc.f[0] = -rho * visc_tot * (du_dx + du_dx - 2. / 3. * (du_dx + dv_dy + dw_dz));
    fc.f[1] = -rho * visc_tot * (dv_dx + du_dy);
    fc.f[2] = -rho * visc_tot * (dw_dx + du_dz);
  } else if (DIR == 1) {
    fc.f[0] = -rho * visc_tot * (du_dy + dv_dx);
    fc.f[1] = -rho * visc_tot * (dv_dy + dv_dy - 2. / 3. * (du_dx + dv_dy + dw_dz));
    fc.f[2] = -rho * visc_tot * (dw_dy + dv_dz);
  } else {
    fc.f[0] = -rho * visc_tot * (du_dz + dw_dx);
    fc.f[1] = -rho * visc_tot * (dv_dz + dw_dy);
    fc.f[2] = -rho * visc_tot * (dw_dz + dw_dz - 2. / 3. * (du_dx + dv_dy + dw_dz));
  }
  fc.f[3] = -rho * visc_tot_th * (dtn);
  fc.f[4] = -rho * visc_tot * 2 * (dKn);
  if (P.sflags) {
    // Learned closure (no reference counterpart): on a face whose two cells both hold the
    // network's flag, the momentum fluxes become rho * avg(tau_{i,DIR}) (bit 0) and the theta
    // flux rho * avg(q_DIR) (bit 1).  Only the DIR coordinate of the face's cells can leave the
    // interior; it wraps for periodic boundaries, other faces keep the eddy-viscosity flux.
    int ic = ih - 1, jc = jh - 1, kc = kh - 1;
    int im = ic - (DIR == 0), jm = jc - (DIR == 1), km = kc - (DIR == 2);
    const int n = DIR == 0 ? P.nx : (DIR == 1 ? P.ny : P.nz);
    const bool per = (DIR == 0 ? P.bc_x : (DIR == 1 ? P.bc_y : P.bc_z)) == PU_BC_PERIODIC;
    int &lo = DIR == 0 ? im : (DIR == 1 ? jm : km);
    int &hi = DIR == 0 ? ic : (DIR == 1 ? jc : kc);
    bool ok = true;
    if (lo < 0) { if (per) lo += n; else ok = false; }
    if (hi >= n) { if (per) hi -= n; else ok = false; }
    if (ok) {
      const long long a = ((long long)km * P.ny + jm) * P.nx + im, b = ((long long)kc * P.ny + jc) * P.nx + ic;
      const int both = P.sflags[a] & P.sflags[b];
      if (both & 1) {
        for (int f = 0; f < 3; f++) {
          const int lo_ = f < DIR ? f : DIR, hi_ = f < DIR ? DIR : f;
          const int t = lo_ == 0 ? hi_ : (lo_ == 1 ? 2 + hi_ : 5);
          fc.f[f] = rho * (0.5 * ((double)P.tau[t * P.n3 + a] + (double)P.tau[t * P.n3 + b]));
        }
      }
      if (both & 2) fc.f[3] = rho * (0.5 * ((double)P.tau[(6 + DIR) * P.n3 + a] + (double)P.tau[(6 + DIR) * P.n3 + b]));
    }
  }
}

constexpr int LTX = 16, LTY = 4, LTZ = 4;       // cells of one block tile (x, y, z)

// One block: a 16x4x4 tile of cells.  The closure box values of the tile plus a one-cell halo
// (fields rho_d, u, v, w, theta, K, the diffused tracers and the immersed proportion) are staged in
// LDS once; each thread then evaluates its cell's six face fluxes, the TKE source and the update.
__global__ __launch_bounds__(256) void k_les_update(const double *__restrict__ F, double *uu, double *vv, double *ww,
                                                    double *tt, TracerPtrs trc, LesParams P) {
  extern __shared__ double sm[];
  const int nfl = 6 + P.nq;
  unsigned bx, by, bz;
  swizzled_block(bx, by, bz);
  const int i0 = bx * LTX, j0 = by * LTY, k0 = bz * LTZ;
  for (int p = threadIdx.x; p < TileBox::n * (nfl + 1); p += 256) {
    const int f = p / TileBox::n, q = p % TileBox::n;
    const int li = q % 18, lj = (q / 18) % 6, lk = q / (18 * 6);
    const int ih = i0 + li, jh = j0 + lj, kh = k0 + lk;       // closure-box (halo) coordinates
    double v = 0.;
    if (ih < P.nxh && jh < P.nyh && kh < P.nzh) {
      const long long g = ((long long)kh * P.nyh + jh) * P.nxh + ih;
      v = f < nfl ? F[(long long)f * P.nh + g] : P.imm[g];
    }
    sm[p] = v;
  }
  __syncthreads();
  const int li = threadIdx.x % LTX, lj = (threadIdx.x / LTX) % LTY, lk = threadIdx.x / (LTX * LTY);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  if (i >= P.nx || j >= P.ny || k >= P.nz) return;
  const long long id = ((long long)k * P.ny + j) * P.nx + i;
  const int ih = i + 1, jh = j + 1, kh = k + 1;
  const TileBox B{sm, nfl};
  constexpr int sx = TileBox::sx, sy = TileBox::sy, sz = TileBox::sz;
  const int c = ((lk + 1) * 6 + (lj + 1)) * 18 + (li + 1);
  auto S = [&](int f, int idx) { return B.S(f, idx); };
  FaceCommon fx0, fx1, fy0, fy1, fz0, fz1;
  face_flux<0>(B, P, c, kh, jh, ih, fx0);
  face_flux<0>(B, P, c + sx, kh, jh, ih + 1, fx1);
  face_flux<1>(B, P, c, kh, jh, ih, fy0);
  face_flux<1>(B, P, c + sy, kh, jh + 1, ih, fy1);
  face_flux<2>(B, P, c, kh, jh, ih, fz0);
  face_flux<2>(B, P, c + sz, kh + 1, jh, ih, fz1);
  const double rdx = 1. / P.dx, rdy = 1. / P.dy, rdz = 1. / P.dz;
  const double r2dx = 0.5 * rdx, r2dy = 0.5 * rdy, r2dz = 0.5 * rdz;
  // TKE source (les_closure.h:233-278)
  double rho = S(0, c), K = S(5, c), t = S(4, c);
  double dt_dz = (S(4, c + sz) - S(4, c - sz)) * r2dz;
  double N = dt_dz >= 0 && P.enable_gravity ? sqrt(P.grav / t * dt_dz) : 0;
  double ell = fmin(0.76 * sqrt(K) / (N + 1.e-20), P.delta);
  double km = 0.1 * ell * sqrt(K);
  double Pr_t = P.delta / (1 + 2 * ell);
  double src = 0;
  auto imm = [&](int idx) { return B.imm(idx); };
  if (imm(c) < 1) {
    if (P.enable_gravity) src += -(P.grav * rho * km) / (t * Pr_t) * dt_dz;
    src -= P.diss_mult * rho * (0.19 + 0.51 * ell / P.delta) / P.delta * (K * sqrt(K));
    int im1 = imm(c - sx) > 0 ? c : c - sx, ip1 = imm(c + sx) > 0 ? c : c + sx;
    int jm1 = imm(c - sy) > 0 ? c : c - sy, jp1 = imm(c + sy) > 0 ? c : c + sy;
    int km1 = imm(c - sz) > 0 ? c : c - sz, kp1 = imm(c + sz) > 0 ? c : c + sz;
    double du_dx = (S(1, c + sx) - S(1, c - sx)) * r2dx;
    double dv_dx = (S(2, ip1) - S(2, im1)) * r2dx;
    double dw_dx = (S(3, ip1) - S(3, im1)) * r2dx;
    double du_dy = (S(1, jp1) - S(1, jm1)) * r2dy;
    double dv_dy = (S(2, c + sy) - S(2, c - sy)) * r2dy;
    double dw_dy = (S(3, jp1) - S(3, jm1)) * r2dy;
    double du_dz = (S(1, kp1) - S(1, km1)) * r2dz;
    double dv_dz = (S(2, kp1) - S(2, km1)) * r2dz;
    double dw_dz = (S(3, c + sz) - S(3, c - sz)) * r2dz;
    double j1_i1 = (du_dx + du_dx) * du_dx, j1_i2 = (dv_dx + du_dy) * dv_dx, j1_i3 = (dw_dx + du_dz) * dw_dx;
    double j2_i1 = (du_dy + dv_dx) * du_dy, j2_i2 = (dv_dy + dv_dy) * dv_dy, j2_i3 = (dw_dy + dv_dz) * dw_dy;
    double j3_i1 = (du_dz + dw_dx) * du_dz, j3_i2 = (dv_dz + dw_dy) * dv_dz, j3_i3 = (dw_dz + dw_dz) * dw_dz;
    src += P.shear_mult * rho * km * (j1_i1 + j1_i2 + j1_i3 + j2_i1 + j2_i2 + j2_i3 + j3_i1 + j3_i2 + j3_i3);
  }
  auto fv = [](const FaceCommon &fc, int f) { return fc.zero ? 0. : fc.f[f]; };
  double tend[5];
  for (int f = 0; f < 5; f++)
    tend[f] = -(fv(fx1, f) - fv(fx0, f)) * rdx - (fv(fy1, f) - fv(fy0, f)) * rdy - (fv(fz1, f) - fv(fz0, f)) * rdz;
  tend[4] += src;
  double dt = P.dt;
  double ru = S(1, c) * rho + dt * tend[0];
  double rv = S(2, c) * rho + dt * tend[1];
  double rw = S(3, c) * rho + dt * tend[2];
  double rt = S(4, c) * rho + dt * tend[3];
  double rk = fmax(0., K * rho + dt * tend[4]);
  uu[id] = ru / rho;
  vv[id] = rv / rho;
  ww[id] = rw / rho;
  tt[id] = P.C0 * pow(rt, P.gamma) / (rho * P.R_d);
  trc.p[P.idTKE][id] = rk;
  // diffused tracers: flux -rho*visc*dq/dn on each open face
  for (int q = 0; q < P.nq; q++) {
    auto Q = [&](int idx) { return B.S(6 + q, idx); };
    double qx0 = fx0.zero ? 0. : -fx0.rho * fx0.visc * ((Q(c) - Q(c - sx)) * rdx);
    double qx1 = fx1.zero ? 0. : -fx1.rho * fx1.visc * ((Q(c + sx) - Q(c)) * rdx);
    double qy0 = fy0.zero ? 0. : -fy0.rho * fy0.visc * ((Q(c) - Q(c - sy)) * rdy);
    double qy1 = fy1.zero ? 0. : -fy1.rho * fy1.visc * ((Q(c + sy) - Q(c)) * rdy);
    double qz0 = fz0.zero ? 0. : -fz0.rho * fz0.visc * ((Q(c) - Q(c - sz)) * rdz);
    double qz1 = fz1.zero ? 0. : -fz1.rho * fz1.visc * ((Q(c + sz) - Q(c)) * rdz);
    double tq = -(qx1 - qx0) * rdx - (qy1 - qy0) * rdy - (qz1 - qz0) * rdz;
    trc.p[P.qidx[q]][id] = Q(c) * rho + dt * tq;
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
  P.grav = c.grav; P.nu = c.kinematic_viscosity;
  P.total_mult = c.les_total_mult; P.diss_mult = c.les_dissipation_mult; P.shear_mult = c.les_shear_prod_mult;
  P.dt = dt;
  P.visc_max_x = 0.25 * P.dx * P.dx / dt; P.visc_max_y = 0.25 * P.dy * P.dy / dt; P.visc_max_z = 0.25 * P.dz * P.dz / dt;
  P.R_d = c.R_d; P.C0 = c.C0; P.gamma = c.gamma_d;
  P.imm = ctx->d_imm_les_val;
  P.tau = nullptr;
  P.sflags = nullptr;
  for (int tr = 0; tr < PU_MAX_TRACERS; tr++) {
    tp.p[tr] = tr < c.num_tracers ? cs->tracers[tr] : nullptr;
    tpc.p[tr] = tp.p[tr];
  }
  return PU_OK;
}

int les_prim_device(pu_ctx *ctx, const pu_coupler_state *cs) {
  LesParams P;
  TracerPtrs tp;
  TracerPtrsC tpc;
  int rc = les_params(ctx, cs, 1., P, tp, tpc);
  if (rc) return rc;
  hipLaunchKernelGGL(k_les_prim, dim3((unsigned)((P.nh + 255) / 256)), dim3(256), 0, ctx->stream, cs->density_dry,
                     cs->uvel, cs->vvel, cs->wvel, cs->temp, tpc, ctx->d_les, P);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int les_apply_device(pu_ctx *ctx, pu_coupler_state *cs, double dt) {
  LesParams P;
  TracerPtrs tp;
  TracerPtrsC tpc;
  int rc = les_params(ctx, cs, dt, P, tp, tpc);
  if (rc) return rc;
  {
    LaunchTimer lt(ctx, "les");
    hipLaunchKernelGGL(k_les_prim, dim3((unsigned)((P.nh + 255) / 256)), dim3(256), 0, ctx->stream, cs->density_dry,
                       cs->uvel, cs->vvel, cs->wvel, cs->temp, tpc, ctx->d_les, P);
    PU_HIP_CHECK(hipGetLastError());
  }
  if (ctx->sgs_model) {
    LaunchTimer lt(ctx, "sgs");
    rc = sgs_infer_device(ctx, dt);
    if (rc) return rc;
    P.tau = ctx->d_sgs_tau;
    P.sflags = ctx->d_sgs_flags;
  }
  {
    LaunchTimer lt(ctx, "les");
    const dim3 grd((unsigned)((P.nx + LTX - 1) / LTX), (unsigned)((P.ny + LTY - 1) / LTY), (unsigned)((P.nz + LTZ - 1) / LTZ));
    const size_t lds = sizeof(double) * TileBox::n * (7 + P.nq);     // <= 119 KB for 16 tracers
    if (lds > 65536)
      PU_HIP_CHECK(hipFuncSetAttribute((const void *)k_les_update, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_les_update, grd, dim3(256), lds, ctx->stream, ctx->d_les, cs->uvel, cs->vvel, cs->wvel,
                       cs->temp, tp, P);
    PU_HIP_CHECK(hipGetLastError());
  }
  return PU_OK;
}

}  // namespace pu
