// sweep_core.h -- device code shared by the sweep kernels of dycore_kernels.hip and by
// sweep_zwall_f32.hip (the FP32 whole-column z sweep, compiled in its own translation unit without
// loop-invariant code motion of machine code: hoisting the limiter's constants out of the chunk
// loop costs that kernel registers and a wave of occupancy).
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
template <int ORD>
__device__ __forceinline__ void weno_edges_fp64(const double *s, bool do_map, bool imm_L, bool imm_R,
                                                double &qL, double &qR) {
  using O = pu_weno::Ops<ORD>;
  constexpr int S = O::S;
  double TV[S];
  O::template tv<double>(s, TV);
  double tot = TV[0];
#pragma unroll
  for (int m = 1; m < S; m++) tot += TV[m];
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

template <int ORD, class T>
__device__ __forceinline__ void weno_edges(const T *s, bool do_map, bool imm_L, bool imm_R, T &qL, T &qR) {
  using O = pu_weno::Ops<ORD>;
  constexpr int S = O::S;
  if constexpr (std::is_same<T, double>::value) {
    weno_edges_fp64<ORD>(s, do_map, imm_L, imm_R, qL, qR);
    return;
  } else {
  T TV[S];
  O::template tv<T>(s, TV);
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
  template <class T>
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
      weno_edges<ORD, T>(s, do_map, imm_L, imm_R, qL, qR);
    } else {
      qL = pu_weno::Ops<ORD>::template linL<T>(s);
      qR = pu_weno::Ops<ORD>::template linR<T>(s);
    }
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


// FP32 whole-column z sweep (sweep_zwall_f32.hip)
int launch_zwall_f32(int ord, dim3 grid, dim3 block, hipStream_t s, const float *prim, const uint8_t *immf,
                     const uint8_t *any10, float *tend, const SweepParams &P);

}  // namespace pu
