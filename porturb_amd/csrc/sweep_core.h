// sweep_core.h -- the sweep kernels (WENO limiter, Riemann flux, flux divergence and the fused
// stage combination): the segment sweep k_sweep, the whole-column z sweep k_sweep_zwall and the
// whole-line sweep k_sweep_pline.  dycore_kernels.hip launches them; the FP32 z, y and x instances
// are compiled in translation units of their own (sweep_zwall_f32.hip, sweep_pline_f32.hip,
// sweep_seg_x_f32.hip) without loop-invariant code motion of machine code: hoisting the limiter's
// constants out of the loops costs those kernels registers and waves of occupancy.
#pragma once
#include "pu_common.h"
#include "weno_ops.h"
#include <type_traits>

namespace pu {

struct SweepParams {
  int nx, ny, nz, hs, hp, ox, ntr, bc_along, wall_z, weno_all;
  long long n3, nxh, nyh, nzh, nh;
  double rd;                       // 1/dx, 1/dy or 1/dz
  const double *hy_rc, *hy_tc;     // cells [nz+2hs]
  const double *hy_re, *hy_te;     // edges [nz+1]
};

__device__ __forceinline__ int wrapi(int i, int n) { i %= n; return i < 0 ? i + n : i; }
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }


// ------------------------------------------------------------------------------------------------
// WENO limiter (WenoLimiter.h compute_limited_edges / gll_1 / gll_2 / convexify / map_rs)
// ------------------------------------------------------------------------------------------------

template <class T, int S> __device__ __forceinline__ void convexify(T *w) {
  T tot = w[0];
#pragma unroll
  for (int m = 1; m < S; m++) tot += w[m];
  if (tot > T(1.e-20)) {
    T r = recip(tot);
#pragma unroll
    for (int m = 0; m < S; m++) w[m] *= r;
  }
}

template <class T> __device__ __forceinline__ T map_rs(T w, T d) {     // k=6, m=3, s=2000
  T dl = w - d;
  T d2 = dl * dl;
  T t1 = d2 * d2 * d2;
  T ww = fma(-w, w, w);
  T t2 = ww * ww * ww;
  return d + (t1 * dl) * recip(t1 + T(2000) * t2);
}

template <class T> __device__ __forceinline__ T map_w(T w, T d) {      // ord 3
  return w * (d + d * d - T(3) * d * w + w * w) * recip(d * d + (T(1) - T(2) * d) * w);
}

// P[m] = prod_{j != m} D[j]
template <class T, int S> __device__ __forceinline__ void excl_products(const T *D, T *P) {
  T pre[S];
  pre[0] = T(1);
#pragma unroll
  for (int m = 1; m < S; m++) pre[m] = pre[m - 1] * D[m - 1];
  T suf = T(1);
#pragma unroll
  for (int m = S - 1; m >= 0; m--) { P[m] = pre[m] * suf; suf *= D[m]; }
}

// FP64 limiter with few divisions and no redundant scaling (same mathematics as the reference,
// WenoLimiter.h gll_1/gll_2, with a few roundings reordered):
//  * TV normalisation: with tot = sum(TV) > 1e-20 the reference uses TV_m/tot, so
//    d_m/((TV_m/tot)^2 + 1e-20) = tot^2 * d_m/(TV_m^2 + 1e-20*tot^2); the common factor tot^2
//    cancels in the weight normalisation, leaving D_m = TV_m^2 + 1e-20*tot^2 (1e-20 when the guard
//    fails).  imm_L/imm_R take a maximum, which commutes with the positive scaling.
//  * w_m = d_m/D_m normalised equals d_m*P_m / sum_k d_k*P_k with P_m = prod_{j!=m} D_j
//    (D >= 1e-20*tot^2 keeps products of S-1 factors far from underflow in double).
//  * map_rs returns A_m/B_m; the renormalised mapped weights are A_m*Q_m / sum A_k*Q_k with
//    Q_m = prod_{j!=m} B_j, and the final division is applied once to the face value.
// Build switch for the flat-stencil branch of the limiters (1: on).
#ifndef PU_FLAT_SHORTCUT
#define PU_FLAT_SHORTCUT 1
#endif

template <int ORD, bool FLAT>
__device__ __forceinline__ void weno_edges_fp64(const double *s, bool do_map, bool imm_L, bool imm_R,
                                                double &qL, double &qR) {
  using O = pu_weno::Ops<ORD>;
  constexpr int S = O::S;
  double TV[S];
  O::template tv<double>(s, TV);
  double tot = TV[0];
#pragma unroll
  for (int m = 1; m < S; m++) tot += TV[m];
  // Flat stencil (tot <= 1e-20, the reference's convexify guard): every D_m rounds to 1e-20, the
  // weights are the ideal ones and the map leaves them unchanged, so the limited edges are the
  // linear reconstruction to rounding.  Decided per cell (the result depends on the cell's own
  // stencil only); a wave whose cells are all flat skips the limiter.
  if constexpr (FLAT && PU_FLAT_SHORTCUT) {
    if (tot <= 1.e-20) {
      qL = O::template linL<double>(s);
      qR = O::template linR<double>(s);
      return;
    }
  }
  const double escale = tot > 1.e-20 ? 1.e-20 * (tot * tot) : 1.e-20;
  if (imm_L || imm_R) {
    double mx = TV[0];
#pragma unroll
    for (int m = 1; m < S; m++) mx = fmax(mx, TV[m]);
    if (imm_L) TV[S - 1] = mx;
    if (imm_R) TV[0] = mx;
  }
  double D[S], P[S];
#pragma unroll
  for (int m = 0; m < S; m++) D[m] = fma(TV[m], TV[m], escale);
  excl_products<double, S>(D, P);
  double wl[S], wr[S], sl = 0, sr = 0;
#pragma unroll
  for (int m = 0; m < S; m++) {
    wl[m] = O::idealL(m) * P[m]; sl += wl[m];
    wr[m] = O::idealR(m) * P[m]; sr += wr[m];
  }
  if (!do_map) {
    qL = O::template combL<double>(s, wl) * recip(sl);
    qR = O::template combR<double>(s, wr) * recip(sr);
    return;
  }
  const double rl = recip(sl), rr = recip(sr);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    double *w = side == 0 ? wl : wr;
    const double r = side == 0 ? rl : rr;
    double A[S], B[S], Q[S];
#pragma unroll
    for (int m = 0; m < S; m++) {
      const double d = side == 0 ? O::idealL(m) : O::idealR(m);
      const double x = w[m] * r;
      if constexpr (O::use_map_rs) {
        double dl = x - d, d2 = dl * dl, t1 = d2 * d2 * d2, ww = fma(-x, x, x), t2 = ww * ww * ww;
        B[m] = fma(2000., t2, t1);
        A[m] = fma(d, B[m], t1 * dl);
      } else {
        B[m] = d * d + (1. - 2. * d) * x;
        A[m] = x * (d + d * d - 3. * d * x + x * x);
      }
    }
    excl_products<double, S>(B, Q);
    double sum = 0;
#pragma unroll
    for (int m = 0; m < S; m++) { w[m] = A[m] * Q[m]; sum += w[m]; }
    if (side == 0) qL = O::template combL<double>(s, w) * recip(sum);
    else qR = O::template combR<double>(s, w) * recip(sum);
  }
}

// FLAT: flat stencils take the linear reconstruction (see weno_edges_fp64)
template <int ORD, class T, bool FLAT = true>
__device__ __forceinline__ void weno_edges(const T *s, bool do_map, bool imm_L, bool imm_R, T &qL, T &qR) {
  using O = pu_weno::Ops<ORD>;
  constexpr int S = O::S;
  if constexpr (std::is_same<T, double>::value) {
    weno_edges_fp64<ORD, FLAT>(s, do_map, imm_L, imm_R, qL, qR);
    return;
  } else {
  T TV[S];
  O::template tv<T>(s, TV);
  if constexpr (FLAT && PU_FLAT_SHORTCUT) {                              // flat stencils: as in FP64 above
    T tot = TV[0];
#pragma unroll
    for (int m = 1; m < S; m++) tot += TV[m];
    if (tot <= T(1.e-20)) {
      qL = O::template linL<T>(s);
      qR = O::template linR<T>(s);
      return;
    }
  }
  convexify<T, S>(TV);
  if (imm_L || imm_R) {
    T mx = TV[0];
#pragma unroll
    for (int m = 1; m < S; m++) mx = fmax(mx, TV[m]);
    if (imm_L) TV[S - 1] = mx;
    if (imm_R) TV[0] = mx;
  }
  T rinv[S];
#pragma unroll
  for (int m = 0; m < S; m++) rinv[m] = recip(TV[m] * TV[m] + T(1.e-20));
  T wl[S], wr[S];
#pragma unroll
  for (int m = 0; m < S; m++) { wl[m] = T(O::idealL(m)) * rinv[m]; wr[m] = T(O::idealR(m)) * rinv[m]; }
  convexify<T, S>(wl);
  convexify<T, S>(wr);
  if (do_map && ORD <= 9) {
    // map_rs(w, d) = A/B; the renormalised mapped weights are A_m Q_m / sum_k A_k Q_k with
    // Q_m = prod_{j != m} B_j, and the one division goes to the face value.  Safe in FP32 up to
    // ord 9: the weights are normalised, so B_m = (w-d)^6 + 2000 (w(1-w))^3 lies in
    // [0.99 d_m^6, 32], and with the ideal weights of ord <= 9 (>= 1/126) a product of S-1 factors
    // stays above 1e-29 and below 1e6 (ord 11 could reach 1e-42: it keeps the divisions below).
#pragma unroll
    for (int side = 0; side < 2; side++) {
      T *w = side == 0 ? wl : wr;
      T A[S], B[S], Q[S];
#pragma unroll
      for (int m = 0; m < S; m++) {
        const T d = T(side == 0 ? O::idealL(m) : O::idealR(m));
        const T x = w[m];
        if constexpr (O::use_map_rs) {
          const T dl = x - d, d2 = dl * dl, t1 = d2 * d2 * d2, ww = fma(-x, x, x), t2 = ww * ww * ww;
          B[m] = fma(T(2000), t2, t1);
          A[m] = fma(d, B[m], t1 * dl);
        } else {
          B[m] = d * d + (T(1) - T(2) * d) * x;
          A[m] = x * (d + d * d - T(3) * d * x + x * x);
        }
      }
      excl_products<T, S>(B, Q);
      T sum = T(0);
#pragma unroll
      for (int m = 0; m < S; m++) { w[m] = A[m] * Q[m]; sum += w[m]; }
      if (side == 0) qL = O::template combL<T>(s, w) * recip(sum);
      else qR = O::template combR<T>(s, w) * recip(sum);
    }
  } else if (do_map) {
#pragma unroll
    for (int m = 0; m < S; m++) {
      wl[m] = map_rs(wl[m], T(O::idealL(m)));
      wr[m] = map_rs(wr[m], T(O::idealR(m)));
    }
    // renormalisation folded into the face value (the mapped weights sum to about one)
    T sl = wl[0], sr = wr[0];
#pragma unroll
    for (int m = 1; m < S; m++) { sl += wl[m]; sr += wr[m]; }
    qL = O::template combL<T>(s, wl) * recip(sl);
    qR = O::template combR<T>(s, wr) * recip(sr);
  } else {
    qL = O::template combL<T>(s, wl);
    qR = O::template combR<T>(s, wr);
  }
  }
}

// Immersed flags of a cell's stencil (dynamics_rk.h:572-607 with modify_stencil_immersed_der0,
// :478-493): the zero-derivative clamp window [lo, hi] used for the transverse velocities and p,
// imm_L / imm_R, do_map = no immersed cell within the 10-cell box, and WENO or the linear
// reconstruction.
template <int ORD> struct StencilFlags {
  static constexpr int hs = (ORD - 1) / 2;
  int lo = 0, hi = ORD - 1;
  bool imm_L = false, imm_R = false, do_map = true, use_weno = true;

  // fbase: flag-array index of the cell; any: its any-immersed-within-10 flag
  __device__ __forceinline__ void setup(const uint8_t *__restrict__ immf, long long fbase, long long sstride, bool any,
                                        int weno_all) {
    uint8_t f[ORD];
#pragma unroll
    for (int s = 0; s < ORD; s++) f[s] = immf[fbase + (long long)(s - hs) * sstride];
    imm_L = f[hs - 1];
    imm_R = f[hs + 1];
    if (!f[hs]) {
#pragma unroll
      for (int s = ORD - 1; s > hs; s--) if (f[s]) hi = s - 1;
#pragma unroll
      for (int s = 0; s < hs; s++) if (f[s]) lo = s + 1;
    }
    do_map = !any;
    use_weno = any || weno_all;
  }

  // both interface values of the cell whose field value sits at p (stride sstride along the sweep)
  template <class T, bool FLAT = true>
  __device__ __forceinline__ void recon(const T *p, long long sstride, bool modify, T &qL, T &qR) const {
    p -= (long long)hs * sstride;
    T s[ORD];
    if (modify) {
#pragma unroll
      for (int q = 0; q < ORD; q++) s[q] = p[(long long)clampi(q, lo, hi) * sstride];
    } else {
#pragma unroll
      for (int q = 0; q < ORD; q++) s[q] = p[(long long)q * sstride];
    }
    if (use_weno) {
      weno_edges<ORD, T, FLAT>(s, do_map, imm_L, imm_R, qL, qR);
    } else {
      qL = pu_weno::Ops<ORD>::template linL<T>(s);
      qR = pu_weno::Ops<ORD>::template linR<T>(s);
    }
  }

  // The branch recon<T> takes for the same stencil: 0 the limiter, 1 the flat-stencil branch
  // (total variation <= 1e-20, evaluated as weno_edges evaluates it), 2 the linear reconstruction
  // of a cell without WENO.
  template <class T>
  __device__ __forceinline__ int branch(const T *p, long long sstride, bool modify) const {
    if (!use_weno) return 2;
    using O = pu_weno::Ops<ORD>;
    p -= (long long)hs * sstride;
    T s[ORD];
#pragma unroll
    for (int q = 0; q < ORD; q++) s[q] = p[(long long)(modify ? clampi(q, lo, hi) : q) * sstride];
    T TV[O::S];
    O::template tv<T>(s, TV);
    T tot = TV[0];
#pragma unroll
    for (int m = 1; m < O::S; m++) tot += TV[m];
    return (PU_FLAT_SHORTCUT && tot <= T(1.e-20)) ? 1 : 0;
  }
};

// acoustic upwind state of one face (dynamics_rk.h:640-651): mass flux, pressure, upwind side
template <class T> struct Upw { T ru, p; bool left; };

template <class T>
__device__ __forceinline__ Upw<T> riemann(T rL, T uL, T pL, T rR, T uR, T pR, T hr) {
  rL += hr; rR += hr;
  const T cs = T(350.);
  T ruL = uL * rL, ruR = uR * rR;
  Upw<T> u;
  u.p = T(0.5) * (pL + pR - cs * (ruR - ruL));
  u.ru = T(0.5) * (ruL + ruR - (pR - pL) * T(1. / 350.));
  u.left = u.ru > 0;
  return u;
}

// ------------------------------------------------------------------------------------------------
// z sweep between solid walls (the first sweep of a stage, writing the tendencies): one block per
// 8 columns walks the whole column in chunks of 32 cells, every thread reconstructing one cell per
// chunk.  Walls need no neighbour cells (the wall faces take the adjacent cell's own values,
// dynamics_rk.h:864-888), and the face between two chunks takes the lower chunk's top values from
// a small LDS carry, so no cell is reconstructed twice (the segment kernel reconstructs 160 cells
// for each 128-level column).  Thread (x, t) of chunk q owns cell r = 32q + t: it computes the
// face below r (r = 0: the bottom wall) and finishes the tendency of cell r-1; the thread of the
// top cell also finishes its own tendency with the top wall face, whose Riemann state with w = 0
// on both sides and equal left/right values is exactly mass flux 0 and pressure p (every other
// flux through it is 0).  Otherwise the arithmetic of k_sweep<ORD, T, 2, 0>, operand for operand.
// ------------------------------------------------------------------------------------------------
constexpr int ZW_BX = 8, ZW_BS = 32;
#ifndef PU_ZWALL
#define PU_ZWALL 1          // 0: z walls use the segment kernel too (tuning builds)
#endif
constexpr int ZW_NQ = 6 + PU_MAX_TRACERS;     // carried face values: rho, w, p, then the advected fields

template <int ORD, class T>
__global__ __launch_bounds__(ZW_BX * ZW_BS) void k_sweep_zwall(const T *__restrict__ prim,
                                                             const uint8_t *__restrict__ immf,
                                                             const uint8_t *__restrict__ any10, T *__restrict__ tend,
                                                             SweepParams P) {
  constexpr int BX = ZW_BX, BS = ZW_BS, NT = BX * BS;
  __shared__ T bufA0[NT], bufA1[NT], bufA2[NT], bufB0[NT], bufB1[NT];
  __shared__ T carQ[2][ZW_NQ][BX], carF[2][ZW_NQ][BX];     // by chunk parity

  const int tx = threadIdx.x, ty = threadIdx.y, lin = ty * BX + tx;
  unsigned bx, by, bz;
  swizzled_block(bx, by, bz);
  const int i = bx * BX + tx, j = by, nz = P.nz;
  const bool col_ok = i < P.nx;
  const long long sstride = P.nxh * P.nyh;
  const T rd = (T)P.rd;
  const int nchunk = (nz + BS - 1) / BS;
  const int nB = 3 + P.ntr;

  for (int q = 0; q < nchunk; q++) {
    const int r = q * BS + ty;
    const bool valid = col_ok && r < nz;
    const bool bot = r == 0, top = valid && r == nz - 1;
    const bool carry_in = ty == 0 && q > 0;          // face below r: values of the previous chunk
    const bool carry_out = ty == BS - 1;
    const int pin = (q + 1) & 1, pout = q & 1;
    const long long base = ((long long)(r + P.hp) * P.nyh + (j + P.hp)) * P.nxh + (i + P.ox);
    const long long cid = ((long long)r * P.ny + j) * P.nx + i;

    StencilFlags<ORD> sf;
    if (valid) sf.setup(immf, base, sstride, any10[cid] != 0, P.weno_all);
    auto recon = [&](int l, bool modify, T &qL, T &qR) {
      sf.template recon<T>(prim + (long long)l * P.nh + base, sstride, modify, qL, qR);
    };
    T hr = 0, ht = 0;
    if (valid) { hr = (T)P.hy_re[r]; ht = (T)P.hy_te[r]; }

    // ---- phase A: density, w, pressure -> acoustic upwinding ----
    T qLr = 0, qRr = 0, qLu = 0, qRu = 0, qLp = 0, qRp = 0;
    if (valid) {
      recon(idR, false, qLr, qRr);
      recon(idW, false, qLu, qRu);
      recon(idP, true, qLp, qRp);
    }
    bufA0[lin] = qRr; bufA1[lin] = qRu; bufA2[lin] = qRp;
    __syncthreads();
    Upw<T> up{T(0), T(0), true};
    T Fr = 0, Fu = 0;
    if (valid) {
      T rL, uL, pL;
      if (carry_in) { rL = carQ[pin][0][tx]; uL = carQ[pin][1][tx]; pL = carQ[pin][2][tx]; }
      else if (ty > 0) { rL = bufA0[lin - BX]; uL = bufA1[lin - BX]; pL = bufA2[lin - BX]; }
      else { rL = 0; uL = 0; pL = 0; }
      T rR = qLr, uR = qLu, pR = qLp;
      if (bot) { rL = rR; uL = 0; uR = 0; pL = pR; }
      up = riemann(rL, uL, pL, rR, uR, pR, hr);
      Fr = up.ru;
      Fu = up.ru * (up.left ? uL : uR) + up.p;
    }
    bufB0[lin] = Fr; bufB1[lin] = Fu;
    __syncthreads();
    if (valid) {
      if (r > 0) {
        const T Fr0 = ty > 0 ? bufB0[lin - BX] : carF[pin][0][tx];
        const T Fu0 = ty > 0 ? bufB1[lin - BX] : carF[pin][1][tx];
        tend[(long long)idR * P.n3 + cid - (long long)P.nx * P.ny] = -(Fr - Fr0) * rd;
        tend[(long long)idW * P.n3 + cid - (long long)P.nx * P.ny] = -(Fu - Fu0) * rd;
      }
      if (top) {
        tend[(long long)idR * P.n3 + cid] = -(T(0) - Fr) * rd;
        tend[(long long)idW * P.n3 + cid] = -(qRp - Fu) * rd;
      }
    }
    if (carry_out) {
      carQ[pout][0][tx] = qRr; carQ[pout][1][tx] = qRu; carQ[pout][2][tx] = qRp;
      carF[pout][0][tx] = Fr; carF[pout][1][tx] = Fu;
    }

    // ---- phase B: the advected fields, one at a time ----
    for (int b = 0; b < nB; b++) {
      int l, fidx;
      bool modify = false;
      if (b == 0) { l = idU; fidx = idU; modify = true; }
      else if (b == 1) { l = idV; fidx = idV; modify = true; }
      else if (b == 2) { l = idT; fidx = idT; }
      else { l = NUM_STATE + 1 + (b - 3); fidx = NUM_STATE + (b - 3); }
      T qL = 0, qR = 0;
      if (valid) recon(l, modify, qL, qR);
      bufA0[lin] = qR;
      __syncthreads();
      T F = 0;
      if (valid) {
        T vL = carry_in ? carQ[pin][3 + b][tx] : (ty > 0 ? bufA0[lin - BX] : T(0)), vR = qL;
        if (bot) vL = vR;
        if (bot && l == idT) { vL = 0; vR = 0; }
        if (l == idT) { vL += ht; vR += ht; }
        F = up.ru * (up.left ? vL : vR);
      }
      bufB0[lin] = F;
      __syncthreads();
      if (valid) {
        T *tp = tend + (long long)fidx * P.n3 + cid;
        if (r > 0) tp[-(long long)P.nx * P.ny] = -(F - (ty > 0 ? bufB0[lin - BX] : carF[pin][2 + b][tx])) * rd;
        if (top) *tp = -(T(0) - F) * rd;
      }
      if (carry_out) { carQ[pout][3 + b][tx] = qR; carF[pout][2 + b][tx] = F; }
    }
  }
}


struct CombineParams {
  int nx, ny, nz, hs, ntr, stage, nrelax, enable_gravity, coriolis;
  int sim2d;                        // ny == 1: the V tendency is zero (dynamics_rk.h:716)
  long long n3;
  double dt, grav, fcor;
  const double *hy_r, *hy_t;        // cells
  const double *relax;              // [n3]
  unsigned int positive_mask;       // bit tr set if tracer tr is positive-definite
};

// Conserved value as double: FP32 storage adds the compensation term of rho (l=0), rho*theta (l=4).
template <class ST>
__device__ __forceinline__ double load_cons(const ST *U, const ST *Ulo, int l, long long id, long long n3) {
  double v = (double)U[(long long)l * n3 + id];
  if constexpr (std::is_same<ST, float>::value) {
    if (l == idR) v += (double)Ulo[id];
    else if (l == idT) v += (double)Ulo[n3 + id];
  }
  return v;
}

template <class ST>
__device__ __forceinline__ void store_cons(ST *U, ST *Ulo, int l, long long id, long long n3, double v) {
  if constexpr (std::is_same<ST, float>::value) {
    float hi = (float)v;
    U[(long long)l * n3 + id] = hi;
    if (l == idR) Ulo[id] = (float)(v - (double)hi);
    else if (l == idT) Ulo[n3 + id] = (float)(v - (double)hi);
  } else {
    U[(long long)l * n3 + id] = v;
  }
}

// ------------------------------------------------------------------------------------------------
// SSPRK3 stage combination of one conserved field of one cell (time_step_rk_3_3,
// dynamics_rk.h:279-367): gravity and Coriolis sources, the stage's convex combination, the
// positivity clip of the last stage and nrelax immersed relaxations (enforce_immersed_boundaries,
// :370-473).  rho_in, u_in, v_in are the stage-input state of the cell, read before any field of the
// cell is overwritten (stage 1 writes Us in place).
// ------------------------------------------------------------------------------------------------
template <class ST> struct FuseArgs {
  ST *U, *Ulo, *Us, *Uslo;
  CombineParams C;
};

// un, us: the field's values in U and (stages 1, 2) Us, read by the caller ahead of time
template <class ST>
__device__ __forceinline__ void combine_field_v(const FuseArgs<ST> &F, int l, long long id, double tn, double un,
                                                double us, double rho_in, double u_in, double v_in, double mult,
                                                double tr_r, double tr_rt) {
  const CombineParams &P = F.C;
  if (l == idV && P.sim2d) tn = 0.;
  if (l == idW && P.enable_gravity) tn += -P.grav * (rho_in - tr_r);
  if (P.coriolis && l == idU) tn += P.fcor * v_in;
  if (P.coriolis && l == idV) tn -= P.fcor * u_in;
  double v;
  if (P.stage == 0) {
    v = un + P.dt * tn;
  } else if (P.stage == 1) {
    v = (3. / 4.) * un + (1. / 4.) * us + (1. / 4.) * P.dt * tn;
  } else {
    v = (1. / 3.) * un + (2. / 3.) * us + (2. / 3.) * P.dt * tn;
  }
  bool positive = l >= NUM_STATE && ((P.positive_mask >> (l - NUM_STATE)) & 1u);
  if (P.stage == 2 && positive) v = fmax(0., v);
  double target = l == idR ? tr_r : (l == idT ? tr_rt : 0.);
  for (int r = 0; r < P.nrelax; r++) {
    v = v + (target - v) * mult;
    if (positive) v = fmax(0., v);
  }
  if (P.stage == 2) store_cons(F.U, F.Ulo, l, id, P.n3, v);
  else store_cons(F.Us, F.Uslo, l, id, P.n3, v);
}

template <class ST>
__device__ __forceinline__ void stage_inputs(const FuseArgs<ST> &F, long long id, int k, double &rho_in, double &u_in,
                                             double &v_in, double &mult, double &tr_r, double &tr_rt) {
  const CombineParams &P = F.C;
  const ST *Sin = P.stage == 0 ? F.U : F.Us;
  const ST *Sinlo = P.stage == 0 ? F.Ulo : F.Uslo;
  rho_in = load_cons(Sin, Sinlo, idR, id, P.n3);
  u_in = P.coriolis ? load_cons(Sin, Sinlo, idU, id, P.n3) : 0.;
  v_in = P.coriolis ? load_cons(Sin, Sinlo, idV, id, P.n3) : 0.;
  mult = P.relax[id];
  tr_r = P.hy_r[P.hs + k];
  tr_rt = P.hy_r[P.hs + k] * P.hy_t[P.hs + k];
}

// ------------------------------------------------------------------------------------------------
// x or y sweep along a periodic line: one block walks whole lines in chunks (x: 4 rows, chunks of
// 64 cells; y: 8 columns, chunks of 32), every thread reconstructing one cell per chunk, so no
// cell is reconstructed twice (the segment kernel reconstructs 1120 cells for each 1024-cell y
// line, 1088 for x).  Thread t of chunk q owns cell r = BS q + t: it computes the face below r and
// finishes the tendency of cell r-1 (through the stage combination in MODE 2); the face between two
// chunks takes the lower chunk's top values from an LDS carry.  The periodic face between cells
// N-1 and 0 is computed last, by the thread of cell N-1, from cell 0's lower values kept in LDS;
// the tendencies of cells 0 and N-1, which need it, are finished after the last chunk.  The
// operations are those of k_sweep<ORD, T, DIR, MODE> (whose face 0 comes from the periodic
// neighbour cell -1, with the values of cell N-1), but the compiler fuses some multiply-adds
// differently: results agree with it to rounding, not bitwise.
// ------------------------------------------------------------------------------------------------
// Where it is used (bits: y FP64, x FP64, y FP32, x FP32; the others use the segment kernel).  The
// chunk loop costs registers: at 1024x1024x128 WENO-9 the y sweep in FP64 keeps its occupancy and
// runs 58.1 -> 54.7 ms/step; in FP32 it needs its own translation unit (sweep_pline_f32.hip, no
// machine-code loop-invariant motion) capped at 72 VGPRs to keep the segment kernel's 7 waves per
// SIMD, and then runs 30.8 -> 29.9 ms/step (at 6 waves 30.4).  The x sweep, which also combines the
// stage, lost occupancy in both precisions (FP64 4 -> 3, FP32 7 -> 4) and ran slower than the 6 %
// of halo reconstructions it saves.
#ifndef PU_PLINE
#define PU_PLINE 5
#endif
template <int DIR> struct LineShape;
template <> struct LineShape<0> { static constexpr int BX = 64, BY = 4, BS = 64, NA = 4; };
template <> struct LineShape<1> { static constexpr int BX = 8, BY = 32, BS = 32, NA = 8; };

// FP32: at most 72 VGPRs (7 waves per SIMD), the segment kernel's occupancy
#ifndef PU_PLINE_F32_WAVES
#define PU_PLINE_F32_WAVES 7
#endif
template <int ORD, class T, int DIR, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(std::is_same<T, float>::value ? PU_PLINE_F32_WAVES : 1)))
void k_sweep_pline(const T *__restrict__ prim, const uint8_t *__restrict__ immf,
                                                     const uint8_t *__restrict__ any10, T *__restrict__ tend,
                                                     SweepParams P, FuseArgs<T> FA) {
  using SH = LineShape<DIR>;
  constexpr int BX = SH::BX, BY = SH::BY, BS = SH::BS, NA = SH::NA, NT = BX * BY;
  constexpr int nbs = DIR == 0 ? 1 : BX;
  __shared__ T bufA0[NT], bufA1[NT], bufA2[NT], bufB0[NT], bufB1[NT];
  __shared__ T carQ[2][ZW_NQ][NA], carF[2][ZW_NQ][NA];          // by chunk parity
  __shared__ T firstQ[ZW_NQ][NA], firstF[ZW_NQ][NA];             // cell 0's lower values, face 1's fluxes
  __shared__ T lastF[ZW_NQ][NA], wrapF[ZW_NQ][NA], wrapU[2][NA]; // faces N-1 and N (= 0)

  const int tx = threadIdx.x, ty = threadIdx.y, lin = ty * BX + tx;
  const int t = DIR == 0 ? tx : ty, a = DIR == 0 ? ty : tx;
  unsigned bx, by, bz;
  swizzled_block(bx, by, bz);
  const int k = by;
  const int ia = DIR == 0 ? 0 : bx * BX + tx, ja = DIR == 0 ? bx * BY + ty : 0;   // across position
  const int N = DIR == 0 ? P.nx : P.ny;
  const bool across_ok = DIR == 0 ? ja < P.ny : ia < P.nx;
  const long long sstride = DIR == 0 ? 1 : P.nxh;
  const long long cstride = DIR == 0 ? 1 : P.nx;
  const T rd = (T)P.rd;
  const T hr = (T)P.hy_rc[P.hs + k], ht = (T)P.hy_tc[P.hs + k];
  const int nchunk = (N + BS - 1) / BS;
  const int nB = 3 + P.ntr;
  constexpr int ln = 1 + DIR;
  constexpr int lt1 = DIR == 0 ? idV : idU;
  auto field_of = [&](int b, int &l, int &fidx, bool &modify) {
    modify = b < 2;
    if (b == 0) { l = lt1; fidx = lt1; }
    else if (b == 1) { l = idW; fidx = idW; }
    else if (b == 2) { l = idT; fidx = idT; }
    else { l = NUM_STATE + 1 + (b - 3); fidx = NUM_STATE + (b - 3); }
  };
  auto cell_id = [&](int r) {
    const int i = DIR == 0 ? r : ia, j = DIR == 0 ? ja : r;
    return ((long long)k * P.ny + j) * P.nx + i;
  };

  // tendency of one field of cell `cell`: MODE 1 adds it to `tend`, MODE 2 finishes it through the
  // stage combination with the cell's stage inputs S
  // the values a cell's finished tendency needs are loaded before the field's reconstruction, in
  // both precisions (FP32 y sweep: 27.1 -> 26.3 ms per step at 1024x1024x128, the register cap of
  // its translation unit unchanged)
  constexpr bool PREFETCH = true;
  struct Pre { T t; double u, us; };
  struct StageIn { double rho, u, v, mult, trr, trt; };
  auto load_pre = [&](int fidx, long long cell, bool on) {
    Pre p{T(0), 0., 0.};
    if (on) {
      p.t = tend[(long long)fidx * P.n3 + cell];
      if constexpr (MODE == 2) {
        p.u = load_cons(FA.U, FA.Ulo, fidx, cell, P.n3);
        if (FA.C.stage > 0) p.us = load_cons(FA.Us, FA.Uslo, fidx, cell, P.n3);
      }
    }
    return p;
  };
  auto load_stage = [&](long long cell, bool on) {
    StageIn S{0., 0., 0., 0., 0., 0.};
    if constexpr (MODE == 2) {
      if (on) stage_inputs(FA, cell, k, S.rho, S.u, S.v, S.mult, S.trr, S.trt);
    }
    return S;
  };
  auto finish = [&](int fidx, long long cell, T v, const Pre &pf, const StageIn &S) {
    if constexpr (MODE == 2) {
      combine_field_v(FA, fidx, cell, (double)(pf.t + v), pf.u, pf.us, S.rho, S.u, S.v, S.mult, S.trr, S.trt);
    } else {
      tend[(long long)fidx * P.n3 + cell] = pf.t + v;
    }
  };

  for (int q = 0; q < nchunk; q++) {
    const int r = q * BS + t;
    const bool valid = across_ok && r < N;
    const bool carry_in = t == 0 && q > 0, carry_out = t == BS - 1;
    const bool first = valid && r == 0, second = valid && r == 1, wrapper = valid && r == N - 1;
    const bool has_face = valid && r >= 1, has_tend = valid && r >= 2;
    const int pin = (q + 1) & 1, pout = q & 1;
    const long long cid = cell_id(r), eid = cid - cstride;   // own cell, tendency cell r-1
    const int i = DIR == 0 ? r : ia, j = DIR == 0 ? ja : r;
    const long long base = ((long long)(k + P.hp) * P.nyh + (j + P.hp)) * P.nxh + (i + P.ox);

    StencilFlags<ORD> sf;
    if (valid) sf.setup(immf, base, sstride, any10[cid] != 0, P.weno_all);
    auto recon = [&](int l, bool modify, T &qL, T &qR) {
      sf.template recon<T>(prim + (long long)l * P.nh + base, sstride, modify, qL, qR);
    };
    const StageIn S = load_stage(eid, has_tend);
    auto prefetch = [&](int fidx) { return PREFETCH ? load_pre(fidx, eid, has_tend) : Pre{T(0), 0., 0.}; };
    auto pre_at_use = [&](int fidx, const Pre &p0) { return PREFETCH ? p0 : load_pre(fidx, eid, has_tend); };

    // ---- phase A: density, normal velocity, pressure -> acoustic upwinding ----
    T qLr = 0, qRr = 0, qLu = 0, qRu = 0, qLp = 0, qRp = 0;
    const Pre preR = prefetch(idR), preU = prefetch(ln);
    if (valid) {
      recon(idR, false, qLr, qRr);
      recon(ln, false, qLu, qRu);
      recon(idP, true, qLp, qRp);
    }
    bufA0[lin] = qRr; bufA1[lin] = qRu; bufA2[lin] = qRp;
    if (first) { firstQ[0][a] = qLr; firstQ[1][a] = qLu; firstQ[2][a] = qLp; }
    __syncthreads();
    Upw<T> up{T(0), T(0), true};
    T Fr = 0, Fu = 0;
    if (has_face) {
      T rL, uL, pL;
      if (carry_in) { rL = carQ[pin][0][a]; uL = carQ[pin][1][a]; pL = carQ[pin][2][a]; }
      else { rL = bufA0[lin - nbs]; uL = bufA1[lin - nbs]; pL = bufA2[lin - nbs]; }
      up = riemann(rL, uL, pL, qLr, qLu, qLp, hr);
      Fr = up.ru;
      Fu = up.ru * (up.left ? uL : qLu) + up.p;
    }
    if (wrapper) {
      const T uR = firstQ[1][a];
      const Upw<T> w = riemann(qRr, qRu, qRp, firstQ[0][a], uR, firstQ[2][a], hr);
      wrapU[0][a] = w.ru; wrapU[1][a] = w.left ? T(1) : T(0);
      wrapF[0][a] = w.ru; wrapF[1][a] = w.ru * (w.left ? qRu : uR) + w.p;
      lastF[0][a] = Fr; lastF[1][a] = Fu;
    }
    bufB0[lin] = Fr; bufB1[lin] = Fu;
    if (second) { firstF[0][a] = Fr; firstF[1][a] = Fu; }
    __syncthreads();
    if (has_tend) {
      const T Fr0 = carry_in ? carF[pin][0][a] : bufB0[lin - nbs];
      const T Fu0 = carry_in ? carF[pin][1][a] : bufB1[lin - nbs];
      finish(idR, eid, -(Fr - Fr0) * rd, pre_at_use(idR, preR), S);
      finish(ln, eid, -(Fu - Fu0) * rd, pre_at_use(ln, preU), S);
    }
    if (carry_out) {
      carQ[pout][0][a] = qRr; carQ[pout][1][a] = qRu; carQ[pout][2][a] = qRp;
      carF[pout][0][a] = Fr; carF[pout][1][a] = Fu;
    }

    // ---- phase B: the advected fields, one at a time ----
    for (int b = 0; b < nB; b++) {
      int l, fidx;
      bool modify;
      field_of(b, l, fidx, modify);
      T qL = 0, qR = 0;
      const Pre pf = prefetch(fidx);
      if (valid) recon(l, modify, qL, qR);
      bufA0[lin] = qR;
      if (first) firstQ[3 + b][a] = qL;
      __syncthreads();
      T F = 0;
      if (has_face) {
        T vL = carry_in ? carQ[pin][3 + b][a] : bufA0[lin - nbs], vR = qL;
        if (l == idT) { vL += ht; vR += ht; }
        F = up.ru * (up.left ? vL : vR);
      }
      if (wrapper) {
        T vL = qR, vR = firstQ[3 + b][a];
        if (l == idT) { vL += ht; vR += ht; }
        wrapF[2 + b][a] = wrapU[0][a] * (wrapU[1][a] != T(0) ? vL : vR);
        lastF[2 + b][a] = F;
      }
      bufB0[lin] = F;
      if (second) firstF[2 + b][a] = F;
      __syncthreads();
      if (has_tend) finish(fidx, eid, -(F - (carry_in ? carF[pin][2 + b][a] : bufB0[lin - nbs])) * rd, pre_at_use(fidx, pf), S);
      if (carry_out) { carQ[pout][3 + b][a] = qR; carF[pout][2 + b][a] = F; }
    }
  }
  __syncthreads();

  // ---- cells 0 (thread t = 0) and N-1 (t = 1): the periodic face closes their tendencies ----
  if (across_ok && t < 2) {
    const int slot_lo = t == 0 ? 0 : 1;                 // which face arrays bound the cell
    const long long cell = cell_id(t == 0 ? 0 : N - 1);
    const StageIn S = load_stage(cell, true);
    auto close = [&](int slot, int fidx) {
      const T Fhi = slot_lo == 0 ? firstF[slot][a] : wrapF[slot][a];
      const T Flo = slot_lo == 0 ? wrapF[slot][a] : lastF[slot][a];
      finish(fidx, cell, -(Fhi - Flo) * rd, load_pre(fidx, cell, true), S);
    };
    close(0, idR);
    close(1, ln);
    for (int b = 0; b < nB; b++) {
      int l, fidx;
      bool modify;
      field_of(b, l, fidx, modify);
      close(2 + b, fidx);
    }
  }
}

// ------------------------------------------------------------------------------------------------
// sweep along one direction: reconstruction + Riemann + divergence, face states through LDS.
// A 256-thread block covers line segments of BS-2 cells along the sweep plus one neighbour cell at
// each end (the cell whose reconstruction the segment's end faces need); x-contiguous threads keep
// the loads of y and z sweeps coalesced.  Whole-line blocks (no neighbour cells) were measured
// slower: 1024-thread blocks leave one block per CU, idle at every barrier.
// ------------------------------------------------------------------------------------------------
// Block shapes (BX across, BS along the sweep; BY = BS for y/z).  PU_SEG_{Y,Z}_{BX,BS} override
// them at compile time for tuning builds.
#ifndef PU_SEG_Y_BX
#define PU_SEG_Y_BX 8
#endif
#ifndef PU_SEG_Y_BS
#define PU_SEG_Y_BS 32
#endif
#ifndef PU_SEG_Z_BX
#define PU_SEG_Z_BX 8
#endif
#ifndef PU_SEG_Z_BS
#define PU_SEG_Z_BS 32
#endif
template <int DIR> struct SegShape;
template <> struct SegShape<0> { static constexpr int BX = 64, BY = 4, BS = 64; };
template <> struct SegShape<1> { static constexpr int BX = PU_SEG_Y_BX, BY = PU_SEG_Y_BS, BS = PU_SEG_Y_BS; };
template <> struct SegShape<2> { static constexpr int BX = PU_SEG_Z_BX, BY = PU_SEG_Z_BS, BS = PU_SEG_Z_BS; };

// MODE: 0 = first sweep of the stage (writes the tendencies), 1 = adds to them, 2 = last sweep:
// each cell's final tendency goes straight into the stage combination (combine_field_v) instead of
// back to `tend`.  The stage runs z (0), y (1), x (2): the last sweep is the one whose rows are
// 64 cells wide, so the stage-state reads and writes of the combination are coalesced.
// FP32 x sweep: at most 64 VGPRs (8 waves per SIMD; it fits without spills, and ran 10.75 against
// 11.21 ms per launch at 72 VGPRs and 7 waves over 205 steps of the 1024x1024x128 urban ABL)
#ifndef PU_SEG_X_F32_WAVES
#define PU_SEG_X_F32_WAVES 8
#endif
template <int ORD, class T, int DIR, int MODE>
__global__ __launch_bounds__(SegShape<DIR>::BX * SegShape<DIR>::BY)
__attribute__((amdgpu_waves_per_eu((std::is_same<T, float>::value && DIR == 0) ? PU_SEG_X_F32_WAVES : 1))) void k_sweep(const T *__restrict__ prim, const uint8_t *__restrict__ immf,
                                               const uint8_t *__restrict__ any10, T *__restrict__ tend,
                                               SweepParams P, FuseArgs<T> FA) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int BX = SegShape<DIR>::BX, BY = SegShape<DIR>::BY, BS = SegShape<DIR>::BS;
  constexpr int NC = BS - 2;
  constexpr int NT = BX * BY;
  __shared__ T bufA0[NT], bufA1[NT], bufA2[NT], bufB0[NT], bufB1[NT];

  const int tx = threadIdx.x, ty = threadIdx.y, lin = ty * BX + tx;
  const int t = DIR == 0 ? tx : ty;
  constexpr int nbs = DIR == 0 ? 1 : BX;
  unsigned bx, by, bz;
  swizzled_block(bx, by, bz);
  int i, j, k, c, nal;
  if (DIR == 0) { c = bx * NC - 1 + tx; i = c; j = by * BY + ty; k = bz; nal = P.nx; }
  else if (DIR == 1) { i = bx * BX + tx; c = by * NC - 1 + ty; j = c; k = bz; nal = P.ny; }
  else { i = bx * BX + tx; j = by; c = bz * NC - 1 + ty; k = c; nal = P.nz; }
  const bool across_ok = DIR == 0 ? (j < P.ny) : (i < P.nx);
  const bool in_range = across_ok && c >= -1 && c <= nal;
  const bool periodic = P.bc_along == PU_BC_PERIODIC;
  const bool ghost = c < 0 || c >= nal;
  const bool do_recon = in_range && (!ghost || periodic);
  const bool has_face = in_range && t >= 1 && c >= 0;
  const bool has_tend = in_range && t >= 1 && t <= NC && c < nal;
  const bool bface_lo = !periodic && c == 0;
  const bool bface_hi = !periodic && c == nal;
  const bool wallz = DIR == 2 && P.wall_z;

  const long long sstride = DIR == 0 ? 1 : (DIR == 1 ? P.nxh : P.nxh * P.nyh);
  const long long base = ((long long)(k + P.hp) * P.nyh + (j + P.hp)) * P.nxh + (i + P.ox);

  StencilFlags<ORD> sf;
  if (do_recon) {
    // A periodic neighbour cell past either end is the wrapped interior cell: its flags come from
    // that cell's own position (the z halos of the flag array hold "immersed", not a periodic copy,
    // dynamics_rk.h:1031-1034; the reference takes that face state from the wrapped cell's
    // reconstruction through edge_exchange, :890-908).
    const int cw = wrapi(c, nal);
    int ci = i, cj = j, ck = k;
    if (DIR == 0) ci = cw; else if (DIR == 1) cj = cw; else ck = cw;
    sf.setup(immf, base + (long long)(cw - c) * sstride, sstride, any10[((long long)ck * P.ny + cj) * P.nx + ci] != 0,
             P.weno_all);
  }
  auto recon = [&](int l, bool modify, T &qL, T &qR) {
    sf.template recon<T>(prim + (long long)l * P.nh + base, sstride, modify, qL, qR);
  };

  const long long cid = ((long long)k * P.ny + j) * P.nx + i;
  const T rd = (T)P.rd;
  double f_rho = 0, f_u = 0, f_v = 0, f_mult = 0, f_trr = 0, f_trt = 0;
  if constexpr (MODE == 2) {
    if (has_tend) stage_inputs(FA, cid, k, f_rho, f_u, f_v, f_mult, f_trr, f_trt);
  }
  // FP64: the values a cell's finished tendency is combined with are loaded before the field's
  // reconstruction, so their latency hides behind it (measured 3.5 ms/step faster at cfg 2).  FP32:
  // loaded when needed -- the extra registers cost the FP32 kernel one wave of occupancy (slower).
  constexpr bool PREFETCH = std::is_same<T, double>::value;
  struct Pre { T t; double u, us; };
  auto load_pre = [&](int fidx) {
    Pre p{T(0), 0., 0.};
    if (has_tend) {
      if constexpr (MODE >= 1) p.t = tend[(long long)fidx * P.n3 + cid];
      if constexpr (MODE == 2) {
        p.u = load_cons(FA.U, FA.Ulo, fidx, cid, P.n3);
        if (FA.C.stage > 0) p.us = load_cons(FA.Us, FA.Uslo, fidx, cid, P.n3);
      }
    }
    return p;
  };
  auto prefetch = [&](int fidx) { return PREFETCH ? load_pre(fidx) : Pre{T(0), 0., 0.}; };
  auto accumulate = [&](int fidx, T v, const Pre &pf0) {
    T *tp = tend + (long long)fidx * P.n3 + cid;
    const Pre pf = PREFETCH ? pf0 : load_pre(fidx);
    if constexpr (MODE == 2) {
      const T tsum = pf.t + v;
      combine_field_v(FA, fidx, cid, (double)tsum, pf.u, pf.us, f_rho, f_u, f_v, f_mult, f_trr, f_trt);
    } else if constexpr (MODE == 1) {
      *tp = pf.t + v;
    } else {
      *tp = v;
    }
  };

  // hydrostatic add-back: cell values for x/y faces, edge values for z faces (:642-645, :686-689)
  T hr = 0, ht = 0;
  if (has_face) {
    if (DIR == 2) { hr = (T)P.hy_re[c]; ht = (T)P.hy_te[c]; }
    else { hr = (T)P.hy_rc[hs + k]; ht = (T)P.hy_tc[hs + k]; }
  }

  // ---- phase A: density, normal velocity, pressure -> acoustic upwinding ----
  constexpr int ln = 1 + DIR;
  T qLr = 0, qRr = 0, qLu = 0, qRu = 0, qLp = 0, qRp = 0;
  const Pre preR = prefetch(idR), preU = prefetch(ln);
  if (do_recon) {
    recon(idR, false, qLr, qRr);
    recon(ln, false, qLu, qRu);
    recon(idP, true, qLp, qRp);
  }
  bufA0[lin] = qRr; bufA1[lin] = qRu; bufA2[lin] = qRp;
  __syncthreads();
  Upw<T> up{T(0), T(0), true};
  T Fr = 0, Fu = 0;
  if (has_face) {
    T rL = bufA0[lin - nbs], uL = bufA1[lin - nbs], pL = bufA2[lin - nbs];
    T rR = qLr, uR = qLu, pR = qLp;
    if (bface_lo) { rL = rR; uL = uR; pL = pR; }
    if (bface_hi) { rR = rL; uR = uL; pR = pL; }
    if (wallz && (bface_lo || bface_hi)) { uL = 0; uR = 0; }     // W = 0 on both sides at walls
    up = riemann(rL, uL, pL, rR, uR, pR, hr);
    Fr = up.ru;
    Fu = up.ru * (up.left ? uL : uR) + up.p;
  }
  bufB0[lin] = Fr; bufB1[lin] = Fu;
  __syncthreads();
  if (has_tend) {
    accumulate(idR, -(bufB0[lin + nbs] - bufB0[lin]) * rd, preR);
    accumulate(ln, -(bufB1[lin + nbs] - bufB1[lin]) * rd, preU);
  }

  // ---- phase B: the advected fields, one at a time ----
  constexpr int lt1 = DIR == 0 ? idV : idU;
  constexpr int lt2 = DIR == 2 ? idV : idW;
  const int nB = 3 + P.ntr;
  for (int b = 0; b < nB; b++) {
    int l, fidx;
    bool modify = false;
    if (b == 0) { l = lt1; fidx = lt1; modify = true; }
    else if (b == 1) { l = lt2; fidx = lt2; modify = true; }
    else if (b == 2) { l = idT; fidx = idT; }
    else { l = NUM_STATE + 1 + (b - 3); fidx = NUM_STATE + (b - 3); }
    T qL = 0, qR = 0;
    const Pre pf = prefetch(fidx);
    if (do_recon) recon(l, modify, qL, qR);
    bufA0[lin] = qR;
    __syncthreads();
    T F = 0;
    if (has_face) {
      T vL = bufA0[lin - nbs], vR = qL;
      if (bface_lo) vL = vR;
      if (bface_hi) vR = vL;
      if (wallz && (bface_lo || bface_hi) && l == idT) { vL = 0; vR = 0; }
      if (l == idT) { vL += ht; vR += ht; }
      F = up.ru * (up.left ? vL : vR);
    }
    bufB0[lin] = F;
    __syncthreads();
    if (has_tend) accumulate(fidx, -(bufB0[lin + nbs] - bufB0[lin]) * rd, pf);
  }
}

// FP32 whole-column z sweep (sweep_zwall_f32.hip)
int launch_zwall_f32(int ord, dim3 grid, dim3 block, hipStream_t s, const float *prim, const uint8_t *immf,
                     const uint8_t *any10, float *tend, const SweepParams &P);
// FP32 whole-line y sweep (sweep_pline_f32.hip), MODE 1
int launch_pline_y_f32(int ord, dim3 grid, dim3 block, hipStream_t s, const float *prim, const uint8_t *immf,
                       const uint8_t *any10, float *tend, const SweepParams &P, const FuseArgs<float> &FA);
// FP32 segment x sweep (sweep_seg_x_f32.hip), MODE 0, 1 or 2
int launch_seg_x_f32(int ord, int mode, dim3 grid, dim3 block, hipStream_t s, const float *prim,
                     const uint8_t *immf, const uint8_t *any10, float *tend, const SweepParams &P,
                     const FuseArgs<float> &FA);

}  // namespace pu
