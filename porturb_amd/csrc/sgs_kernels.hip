// sgs_kernels.hip -- learned subgrid closure on gfx950 (no counterpart in the reference).
//
// A per-cell network 108 -> 64 -> 64 -> 9 (ReLU) predicts the subgrid stresses and heat fluxes of
// a cell from the differences of (u, v, w, theta) over its 3x3x3 neighbourhood.  The network runs
// on the matrix cores: v_mfma_f32_16x16x16_bf16, BF16 operands, FP32 accumulation.
//
// Matrix-core layout used throughout (one wave, 16x16x16):
//   A operand: lane l holds A[row = l%16][k = 4*(l/16) .. +3]
//   B operand: lane l holds B[k = 4*(l/16) .. +3][col = l%16]
//   C/D      : lane l holds D[row = 4*(l/16) .. +3][col = l%16]
// The layers are evaluated transposed, D = W * A^T (rows = units, columns = 16 samples), so a
// layer's D fragment for units 16m..16m+15 is exactly the B fragment of K-block m of the next
// layer: the activations of inference never leave registers.
//
// Kernels:
//   k_mlp_train<TRAIN>  64 samples per block: forward, and with TRAIN the backward pass and the
//                       block's partial gradient (weight gradients with K = the block's samples);
//   k_mlp_reduce        sums the block partials in a fixed order, optional Adam update;
//   k_sgs_filter        box filter of a fine run (training targets);
//   k_sgs_features      training inputs of listed cells;
//   k_sgs_valid         static mask of cells whose level and neighbourhood allow the network;
//   k_sgs_infer         inputs + network + fallback rule for every cell of the closure box.
#include "pu_common.h"
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <vector>

namespace pu {

constexpr int NIN = PU_SGS_NIN, NINP = 112, NH = PU_SGS_NHIDDEN, NOUT = PU_SGS_NOUT, NOUTP = 16;
// padded device parameter layout (fp32 master weights and their bf16 copy)
constexpr int P_W1 = 0, P_B1 = P_W1 + NH * NINP, P_W2 = P_B1 + NH, P_B2 = P_W2 + NH * NH, P_W3 = P_B2 + NH,
              P_B3 = P_W3 + NOUTP * NH, P_N = P_B3 + NOUTP;
constexpr int SAMPLES_PER_BLOCK = 64;
// Fallback rule: resolved strains and theta gradients below this fraction of the cell's speed
// (theta) per cell width are rounding noise and count as zero (the same constant as oracle/sgs.c).
constexpr double SGS_NOISE_REL = 2e-6;

typedef short bfx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_bits(float x) { return __builtin_bit_cast(short, (__bf16)x); }

__device__ __forceinline__ fx4 mfma16(bfx4 a, bfx4 b, fx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bfx4 ld4(const short *p) { return *reinterpret_cast<const bfx4 *>(p); }

// gfx950 v_mfma_f32_16x16x32_bf16: lane l holds A[row l%16][k = 8*(l/16) .. +7],
// B[k = 8*(l/16) .. +7][col l%16]; C/D as for 16x16x16.
typedef short bfx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ fx4 mfma32(bfx8 a, bfx8 b, fx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                  0, 0, 0);
}
__device__ __forceinline__ bfx8 ld8(const short *p) { return *reinterpret_cast<const bfx8 *>(p); }
__device__ __forceinline__ bfx8 cat8(bfx4 lo, bfx4 hi) {
  return bfx8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// Input order of the 64-unit layers in the 16x16x32 inference: position k of a 32-wide K block
// kb holds hidden unit 16 (2 kb + (k%8 >= 4)) + 4 (k%32 / 8) + k%4, so the two 16-unit output
// fragments of the previous layer, side by side, are exactly the lane's 8 B-operand values.
__host__ __device__ constexpr int unit_of_k(int k) {
  return 16 * (2 * (k / 32) + ((k % 8) >= 4 ? 1 : 0)) + 4 * ((k % 32) / 8) + (k % 4);
}

// Hidden-layer activation: ReLU, and its derivative (0 at x = 0).
__device__ __forceinline__ float relu(float x) { return fmaxf(x, 0.f); }
__device__ __forceinline__ float relu_deriv(float x) { return x > 0.f ? 1.f : 0.f; }

// sum over the 16 lanes that share l/16 (one column group of a D fragment)
__device__ __forceinline__ float sum16(float v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}

}  // namespace pu

struct pu_mlp {
  float *d_master = nullptr;         // [P_N]
  short *d_wbf = nullptr;            // [P_N] bf16 copy of the master weights
  float *d_m = nullptr, *d_v = nullptr;
  float *d_partial = nullptr;        // [max_blocks][P_N]
  float *d_ploss = nullptr;          // [max_blocks]
  float *d_grad = nullptr;           // [P_N]
  int max_blocks = 0;
  int t = 0;                         // Adam step count
  float feat_scale[PU_SGS_NIN];
  float xmax = 1e30f;
  float out_mean[PU_SGS_NOUT], out_scale[PU_SGS_NOUT];
};

namespace pu {

// ------------------------------------------------------------------------------------------
// training: forward / backward of 64 samples per block
// ------------------------------------------------------------------------------------------
struct TrainArgs {
  const short *X;       // [N][NINP] bf16
  const float *T;       // [N][NOUT]
  const int *idx;       // [n] rows, or null
  int n;
  const short *wbf;
  const float *master;
  float *partial;       // [nblocks][P_N]
  float *ploss;         // [nblocks]
  float *Y;             // forward output [n][NOUT] (forward-only launch)
  float scale;          // 2 / (n * NOUT)
};

template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_train(TrainArgs A) {
  __shared__ __align__(16) short sX[SAMPLES_PER_BLOCK * NINP];
  __shared__ __align__(16) short sA1[SAMPLES_PER_BLOCK * NH], sA2[SAMPLES_PER_BLOCK * NH];
  __shared__ __align__(16) short sD1[SAMPLES_PER_BLOCK * NH], sD2[SAMPLES_PER_BLOCK * NH];
  __shared__ __align__(16) short sDY[SAMPLES_PER_BLOCK * NOUTP];
  __shared__ float sBias[4][NH + NH + NOUTP];
  __shared__ float sLoss[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, r16 = lane & 15, g = lane >> 4;
  const long long base = (long long)blockIdx.x * SAMPLES_PER_BLOCK;
  // stage the block's inputs (rows beyond n are zero); all loads issued before the LDS writes
  constexpr int PER = SAMPLES_PER_BLOCK * (NINP / 4) / 256;
  uint2 xin[PER];
#pragma unroll
  for (int e = 0; e < PER; e++) {
    const int c = tid + 256 * e;
    const int s = c / (NINP / 4), q = c % (NINP / 4);
    const long long gs = base + s;
    xin[e] = make_uint2(0u, 0u);
    if (gs < A.n) {
      const long long row = A.idx ? (long long)A.idx[gs] : gs;
      xin[e] = *reinterpret_cast<const uint2 *>(A.X + row * NINP + 4 * q);
    }
  }
#pragma unroll
  for (int e = 0; e < PER; e++) {
    const int c = tid + 256 * e;
    *reinterpret_cast<uint2 *>(sX + (c / (NINP / 4)) * NINP + 4 * (c % (NINP / 4))) = xin[e];
  }
  __syncthreads();
  const int my = 16 * w + r16;                    // this lane's sample column within the block
  const long long gsample = base + my;
  const bool valid = gsample < A.n;
  const long long row = valid ? (A.idx ? (long long)A.idx[gsample] : gsample) : 0;

  // ---- forward ----
  float h1[4][4], h2[4][4];
  for (int mt = 0; mt < 4; mt++) {
    fx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < NINP / 16; kb++)
      acc = mfma16(ld4(A.wbf + P_W1 + (16 * mt + r16) * NINP + 16 * kb + 4 * g), ld4(sX + my * NINP + 16 * kb + 4 * g), acc);
    for (int r = 0; r < 4; r++) {
      const int u = 16 * mt + 4 * g + r;
      const float h = acc[r] + A.master[P_B1 + u];
      h1[mt][r] = h;
      sA1[my * NH + u] = bf16_bits(relu(h));
    }
  }
  __syncthreads();
  for (int mt = 0; mt < 4; mt++) {
    fx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < NH / 16; kb++)
      acc = mfma16(ld4(A.wbf + P_W2 + (16 * mt + r16) * NH + 16 * kb + 4 * g), ld4(sA1 + my * NH + 16 * kb + 4 * g), acc);
    for (int r = 0; r < 4; r++) {
      const int u = 16 * mt + 4 * g + r;
      const float h = acc[r] + A.master[P_B2 + u];
      h2[mt][r] = h;
      sA2[my * NH + u] = bf16_bits(relu(h));
    }
  }
  __syncthreads();
  fx4 yacc = {0.f, 0.f, 0.f, 0.f};
  for (int kb = 0; kb < NH / 16; kb++)
    yacc = mfma16(ld4(A.wbf + P_W3 + r16 * NH + 16 * kb + 4 * g), ld4(sA2 + my * NH + 16 * kb + 4 * g), yacc);
  // lane holds output o = 4g + r of sample `my`
  if (!TRAIN) {
    for (int r = 0; r < 4; r++) {
      const int o = 4 * g + r;
      if (o < NOUT && valid) A.Y[gsample * NOUT + o] = yacc[r] + A.master[P_B3 + o];
    }
    return;
  }

  // ---- backward ----
  float lossp = 0.f, dyv[4];
  for (int r = 0; r < 4; r++) {
    const int o = 4 * g + r;
    float d = 0.f;
    if (o < NOUT && valid) {
      const float e = (yacc[r] + A.master[P_B3 + o]) - A.T[row * NOUT + o];
      lossp += e * e;
      d = A.scale * e;
    }
    dyv[r] = d;
    sDY[my * NOUTP + o] = bf16_bits(d);
  }
  for (int r = 0; r < 4; r++) {
    const float v = sum16(dyv[r]);
    if (r16 == 0) sBias[w][2 * NH + 4 * g + r] = v;
  }
  {
    float v = sum16(lossp);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (lane == 0) sLoss[w] = v;
  }
  __syncthreads();
  // dA2 = W3^T dy, dH2 = dA2 * relu'(h2)
  for (int mt = 0; mt < 4; mt++) {
    bfx4 a;
    for (int q = 0; q < 4; q++) a[q] = A.wbf[P_W3 + (4 * g + q) * NH + 16 * mt + r16];
    fx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(a, ld4(sDY + my * NOUTP + 4 * g), acc);
    for (int r = 0; r < 4; r++) {
      const int u = 16 * mt + 4 * g + r;
      const float d = acc[r] * relu_deriv(h2[mt][r]);
      sD2[my * NH + u] = bf16_bits(d);
      const float v = sum16(d);
      if (r16 == 0) sBias[w][NH + u] = v;
    }
  }
  __syncthreads();
  // dA1 = W2^T dH2, dH1 = dA1 * relu'(h1)
  for (int mt = 0; mt < 4; mt++) {
    fx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < NH / 16; kb++) {
      bfx4 a;
      for (int q = 0; q < 4; q++) a[q] = A.wbf[P_W2 + (16 * kb + 4 * g + q) * NH + 16 * mt + r16];
      acc = mfma16(a, ld4(sD2 + my * NH + 16 * kb + 4 * g), acc);
    }
    for (int r = 0; r < 4; r++) {
      const int u = 16 * mt + 4 * g + r;
      const float d = acc[r] * relu_deriv(h1[mt][r]);
      sD1[my * NH + u] = bf16_bits(d);
      const float v = sum16(d);
      if (r16 == 0) sBias[w][u] = v;
    }
  }
  __syncthreads();
  // weight gradients, K = the block's 64 samples: dW[out][in] = sum_s dH[s][out] * act[s][in]
  float *out = A.partial + (size_t)blockIdx.x * P_N;
  auto wgrad = [&](const short *dH, int ldh, int mt, const short *act, int lda, int nt, float *dst, int ldd) {
    fx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < SAMPLES_PER_BLOCK / 16; kb++) {
      bfx4 a, b;
      for (int q = 0; q < 4; q++) {
        const int s = 16 * kb + 4 * g + q;
        a[q] = dH[s * ldh + 16 * mt + r16];
        b[q] = act[s * lda + 16 * nt + r16];
      }
      acc = mfma16(a, b, acc);
    }
    for (int r = 0; r < 4; r++) dst[(16 * mt + 4 * g + r) * ldd + 16 * nt + r16] = acc[r];
  };
  wgrad(sDY, NOUTP, 0, sA2, NH, w, out + P_W3, NH);                           // dW3: 4 tiles, one per wave
  for (int nt = 0; nt < 4; nt++) wgrad(sD2, NH, w, sA1, NH, nt, out + P_W2, NH);       // dW2: 16 tiles
  for (int nt = 0; nt < NINP / 16; nt++) wgrad(sD1, NH, w, sX, NINP, nt, out + P_W1, NINP);  // dW1: 28 tiles
  for (int u = tid; u < NH + NH + NOUTP; u += 256) {
    const float v = sBias[0][u] + sBias[1][u] + sBias[2][u] + sBias[3][u];
    const int dst = u < NH ? P_B1 + u : (u < 2 * NH ? P_B2 + (u - NH) : P_B3 + (u - 2 * NH));
    out[dst] = v;
  }
  if (tid == 0) A.ploss[blockIdx.x] = sLoss[0] + sLoss[1] + sLoss[2] + sLoss[3];
}

// Sum of the block partials (fixed order), optional Adam update of master weights and bf16 copy.
struct ReduceArgs {
  const float *partial;
  int nblocks;
  float *grad;          // [P_N] out (may be null)
  const float *gin;     // given gradient instead of partials (may be null)
  float *master, *m, *v;
  short *wbf;
  int update;
  float lr, b1, b2, eps, c1, c2;
};

__global__ __launch_bounds__(256) void k_mlp_reduce(ReduceArgs A) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P_N) return;
  float gsum;
  if (A.gin) gsum = A.gin[p];
  else {
    gsum = 0.f;
    for (int b = 0; b < A.nblocks; b++) gsum += A.partial[(size_t)b * P_N + p];
  }
  if (A.grad) A.grad[p] = gsum;
  if (!A.update) return;
  // Adam (Kingma & Ba 2015, Algorithm 1), each product and sum rounded separately
  const float m = __fadd_rn(__fmul_rn(A.b1, A.m[p]), __fmul_rn(1.f - A.b1, gsum));
  const float v = __fadd_rn(__fmul_rn(A.b2, A.v[p]), __fmul_rn(1.f - A.b2, __fmul_rn(gsum, gsum)));
  A.m[p] = m;
  A.v[p] = v;
  const float mh = m / A.c1, vh = v / A.c2;
  const float np = __fsub_rn(A.master[p], __fmul_rn(A.lr, mh) / __fadd_rn(sqrtf(vh), A.eps));
  A.master[p] = np;
  A.wbf[p] = bf16_bits(np);
}

__global__ __launch_bounds__(256) void k_mlp_refresh_bf16(const float *master, short *wbf) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < P_N) wbf[p] = bf16_bits(master[p]);
}

__global__ __launch_bounds__(256) void k_pack_inputs(const float *X, long long n, short *Xbf) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= n * NINP) return;
  const long long s = id / NINP;
  const int f = (int)(id % NINP);
  Xbf[id] = f < NIN ? bf16_bits(X[s * NIN + f]) : (short)0;
}

// ------------------------------------------------------------------------------------------
// training data: box filter and inputs
// ------------------------------------------------------------------------------------------
struct FilterArgs {
  const double *rho, *u, *v, *w, *T, *imm;
  int nx, ny, nz, r, nxc, nyc, nzc;
  long long n3, n3c, nxh, nyh, nh;
  double R_d, C0, inv_gamma;
  double *les, *immc;
  float *targets;
};

__global__ __launch_bounds__(256) void k_sgs_filter(FilterArgs A) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= A.n3c) return;
  const int I = (int)(id % A.nxc), J = (int)((id / A.nxc) % A.nyc), K = (int)(id / ((long long)A.nxc * A.nyc));
  double m[5] = {0, 0, 0, 0, 0}, uu[6] = {0, 0, 0, 0, 0, 0}, ut[3] = {0, 0, 0}, im = 0;
  for (int a = 0; a < A.r; a++)
    for (int b = 0; b < A.r; b++)
      for (int e = 0; e < A.r; e++) {
        const long long f = ((long long)(K * A.r + a) * A.ny + (J * A.r + b)) * A.nx + (I * A.r + e);
        const double rho = A.rho[f], u = A.u[f], v = A.v[f], w = A.w[f];
        const double th = pow(rho * A.R_d * A.T[f] / A.C0, A.inv_gamma) / rho;
        m[0] += rho; m[1] += u; m[2] += v; m[3] += w; m[4] += th;
        uu[0] += u * u; uu[1] += u * v; uu[2] += u * w; uu[3] += v * v; uu[4] += v * w; uu[5] += w * w;
        ut[0] += u * th; ut[1] += v * th; ut[2] += w * th;
        im = fmax(im, A.imm[f]);
      }
  const double inv = 1. / ((double)A.r * A.r * A.r);
  for (int f = 0; f < 5; f++) m[f] *= inv;
  for (int f = 0; f < 6; f++) uu[f] *= inv;
  for (int f = 0; f < 3; f++) ut[f] *= inv;
  const double tau[6] = {uu[0] - m[1] * m[1], uu[1] - m[1] * m[2], uu[2] - m[1] * m[3],
                         uu[3] - m[2] * m[2], uu[4] - m[2] * m[3], uu[5] - m[3] * m[3]};
  const double q[3] = {ut[0] - m[1] * m[4], ut[1] - m[2] * m[4], ut[2] - m[3] * m[4]};
  const long long h = ((long long)(K + 1) * A.nyh + (J + 1)) * A.nxh + (I + 1);
  for (int f = 0; f < 5; f++) A.les[f * A.nh + h] = m[f];
  A.les[5 * A.nh + h] = 0.5 * (tau[0] + tau[3] + tau[5]);
  A.immc[h] = im;
  for (int f = 0; f < 6; f++) A.targets[f * A.n3c + id] = (float)tau[f];
  for (int f = 0; f < 3; f++) A.targets[(6 + f) * A.n3c + id] = (float)q[f];
}

// halo of the filtered box: x/y periodic (corners wrap both), z copies with w = 0, immersed 1
__global__ __launch_bounds__(256) void k_sgs_filter_halo(FilterArgs A) {
  const long long h = (long long)blockIdx.x * 256 + threadIdx.x;
  if (h >= A.nh) return;
  const int ih = (int)(h % A.nxh), jh = (int)((h / A.nxh) % A.nyh), kh = (int)(h / (A.nxh * A.nyh));
  const int i = ih - 1, j = jh - 1, k = kh - 1;
  if (k >= 0 && k < A.nzc && j >= 0 && j < A.nyc && i >= 0 && i < A.nxc) return;
  const int ks = k < 0 ? 0 : (k >= A.nzc ? A.nzc - 1 : k);
  const int js = (j % A.nyc + A.nyc) % A.nyc, is = (i % A.nxc + A.nxc) % A.nxc;
  const long long src = ((long long)(ks + 1) * A.nyh + (js + 1)) * A.nxh + (is + 1);
  for (int f = 0; f < 6; f++) A.les[f * A.nh + h] = A.les[f * A.nh + src];
  if (k != ks) {
    A.les[3 * A.nh + h] = 0.;
    A.immc[h] = 1.;
  } else {
    A.immc[h] = A.immc[src];
  }
}

// One input of a cell: neighbour n of the 3x3x3 box, quantity q (u, v, w, theta); the difference
// is formed in the box's precision, rounded to float once.
template <class T>
__device__ __forceinline__ float sgs_input(const T *__restrict__ F, long long nh, long long c, long long nb, int q,
                                           float fm) {
  const T *Fq = F + (long long)(1 + q) * nh;
  return (float)(Fq[nb] - Fq[c]) * fm;
}

struct FeatArgs {
  const double *les;
  int nx, ny, nz;
  long long nxh, nyh, nh;
  const float *fmul;
  const long long *cells;
  long long n;
  short *Xbf;
};

__global__ __launch_bounds__(256) void k_sgs_features(FeatArgs A) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= A.n * 28) return;
  const long long s = t / 28;
  const int n = (int)(t % 28);            // neighbour slot (27 = padding)
  const long long id = A.cells[s];
  const int i = (int)(id % A.nx), j = (int)((id / A.nx) % A.ny), k = (int)(id / ((long long)A.nx * A.ny));
  const long long c = ((long long)(k + 1) * A.nyh + (j + 1)) * A.nxh + (i + 1);
  short out[4] = {0, 0, 0, 0};
  if (n < 27) {
    const int dz = n / 9 - 1, dy = (n / 3) % 3 - 1, dx = n % 3 - 1;
    const long long nb = c + dz * A.nxh * A.nyh + dy * A.nxh + dx;
    for (int q = 0; q < 4; q++) out[q] = bf16_bits(sgs_input(A.les, A.nh, c, nb, q, A.fmul[4 * n + q]));
  }
  short *dst = A.Xbf + s * NINP + 4 * n;
  for (int q = 0; q < 4; q++) dst[q] = out[q];
}

// ------------------------------------------------------------------------------------------
// inference inside the closure
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sgs_valid(const double *__restrict__ imm, int nx, int ny, int nz,
                                                   unsigned char *__restrict__ valid) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n3 = (long long)nx * ny * nz;
  if (id >= n3) return;
  const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((long long)nx * ny));
  const long long nxh = nx + 2, nyh = ny + 2;
  bool ok = k >= 1 && k <= nz - 2;
  for (int dz = -1; dz <= 1 && ok; dz++)
    for (int dy = -1; dy <= 1 && ok; dy++)
      for (int dx = -1; dx <= 1; dx++)
        if (imm[((long long)(k + 1 + dz) * nyh + (j + 1 + dy)) * nxh + (i + 1 + dx)] > 0) { ok = false; break; }
  valid[id] = ok ? 1 : 0;
}

template <class T> struct InferArgs {
  const T *F;
  long long nxh, nyh, nh, n3;
  int nx, ny, nz;
  const unsigned char *valid;
  const short *wbf;
  const float *master;
  const float *fmul;
  float xmax;
  float out_mean[NOUT], out_scale[NOUT];
  double r2dx, r2dy, r2dz, nu_max;
  float4 *tau;          // [3][n3]: per face direction d, (tau_0d, tau_1d, tau_2d, q_d) of each cell
  unsigned char *flags;
};
// One group = 128 consecutive cells of an x row (i0..i0+127 at j, k); a block loops over groups.
// The group's 3x3 neighbourhood rows of (u, v, w, theta), 130 x 3 x 3 points, are staged in LDS
// with row-contiguous loads; each wave then takes 32 cells as two 16-cell tiles that share every
// weight fragment it reads from LDS: inputs from LDS, the network on the matrix cores, the outputs
// stored from the lanes that hold them in the last layer's D fragments, and (lanes 0-31, one cell
// each) the fallback rule.
constexpr int GX = 128, GBX = GX + 2, GBN = GBX * 3 * 3;

// Waves per SIMD the inference kernel is built for (3: up to 168 VGPRs, no spills).
#ifndef PU_SGS_INFER_WAVES
#define PU_SGS_INFER_WAVES 3
#endif

template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PU_SGS_INFER_WAVES))) void k_sgs_infer(InferArgs<T> A) {
  constexpr int K1 = 128;                       // layer-1 inputs: 4 K blocks of 32 (108 used)
  __shared__ __align__(16) short sW1[NH * NINP];   // row stride 112: block 3 reads 16 columns
  __shared__ __align__(16) short sW2[NH * NH];  // columns in unit_of_k order
  __shared__ __align__(16) short sW3[NOUTP * NH];
  __shared__ __align__(16) float sB[NH + NH + NOUTP];
  __shared__ __align__(16) float sMS[2 * NOUTP];   // output mean, output scale (16 slots each)
  __shared__ float sFm[K1];
  __shared__ T sQ[4 * GBN];                     // u, v, w, theta of the group's neighbourhood
  const int tid = threadIdx.x, lane = tid & 63, r16 = lane & 15, g = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);   // the wave (uniform: row addresses are scalar)
  for (int p = tid; p < NH * NINP / 4; p += 256)
    reinterpret_cast<uint2 *>(sW1)[p] = reinterpret_cast<const uint2 *>(A.wbf + P_W1)[p];
  for (int p = tid; p < NH * NH; p += 256) sW2[p] = A.wbf[P_W2 + (p / NH) * NH + unit_of_k(p % NH)];
  for (int p = tid; p < NOUTP * NH; p += 256) sW3[p] = A.wbf[P_W3 + (p / NH) * NH + unit_of_k(p % NH)];
  for (int p = tid; p < NH + NH + NOUTP; p += 256)
    sB[p] = p < NH ? A.master[P_B1 + p] : (p < 2 * NH ? A.master[P_B2 + p - NH] : A.master[P_B3 + p - 2 * NH]);
  for (int p = tid; p < K1; p += 256) sFm[p] = p < NIN ? A.fmul[p] : 0.f;
  if (tid < NOUTP) {
    sMS[tid] = tid < NOUT ? A.out_mean[tid] : 0.f;
    sMS[NOUTP + tid] = tid < NOUT ? A.out_scale[tid] : 0.f;
  }
  // group bookkeeping in 32 bits (the host checks that the group count fits)
  const unsigned ngx = (unsigned)((A.nx + GX - 1) / GX);
  const unsigned ngroups = ngx * (unsigned)A.ny * (unsigned)A.nz;
  for (unsigned grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const unsigned rest = grp / ngx;
    const int i0 = (int)(grp - rest * ngx) * GX, j = (int)(rest % (unsigned)A.ny), k = (int)(rest / (unsigned)A.ny);
    const int lxmax = (int)(A.nxh - i0);        // box x positions inside the array
    // Staging: the box is 36 rows (quantity q, z offset, y offset) of 130 points; wave w loads rows
    // w, w+4, ... with one lane per x point (points lane, 64 + lane; lanes 0, 1 also 128, 129), so
    // each row's address is wave-uniform.  All loads are issued before any LDS write.
    constexpr int RPW = 36 / 4;
    T buf[RPW], buf2[RPW], buf3[RPW];
#pragma unroll
    for (int rr = 0; rr < RPW; rr++) {
      const int r = w + 4 * rr;
      const int q = r / 9, lz = (r / 3) % 3, ly = r % 3;
      const T *row = A.F + (long long)(1 + q) * A.nh + ((long long)(k + lz) * A.nyh + (j + ly)) * A.nxh + i0;
      buf[rr] = lane < lxmax ? row[lane] : T(0);
      buf2[rr] = 64 + lane < lxmax ? row[64 + lane] : T(0);
      buf3[rr] = (lane < 2 && 128 + lane < lxmax) ? row[128 + lane] : T(0);
    }
    __syncthreads();                            // the previous group is done with sQ
#pragma unroll
    for (int rr = 0; rr < RPW; rr++) {
      const int r = w + 4 * rr;
      sQ[r * GBX + lane] = buf[rr];
      sQ[r * GBX + 64 + lane] = buf2[rr];
      if (lane < 2) sQ[r * GBX + 128 + lane] = buf3[rr];
    }
    __syncthreads();
    // inputs of the two tiles: tile t holds cells 32 w + 16 t + r16; K block kb, lane group g ->
    // neighbours 8 kb + 2 g and 8 kb + 2 g + 1
    bfx8 xb[2][4];
    float mxt[2] = {0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int c = (1 * 3 + 1) * GBX + 32 * w + 16 * t + r16 + 1;   // the cell's index in the staged box
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        bfx8 v = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int n = 8 * kb + 2 * g + h;
          if (n < 27) {
            const int dz = n / 9 - 1, dy = (n / 3) % 3 - 1, dx = n % 3 - 1;
            const int nb = c + (dz * 3 + dy) * GBX + dx;
#pragma unroll
            for (int q = 0; q < 4; q++) {
              const float x = (float)(sQ[q * GBN + nb] - sQ[q * GBN + c]) * sFm[4 * n + q];
              mxt[t] = fmaxf(mxt[t], fabsf(x));
              v[4 * h + q] = bf16_bits(x);
            }
          }
        }
        xb[t][kb] = v;
      }
    }
    // the largest |input| of each tile's cell r16: the four lane groups hold a quarter of its inputs each
#pragma unroll
    for (int t = 0; t < 2; t++) {
      mxt[t] = fmaxf(mxt[t], __shfl_xor(mxt[t], 16));
      mxt[t] = fmaxf(mxt[t], __shfl_xor(mxt[t], 32));
    }
    // layer 1: D tiles of units 16 mt .. 16 mt + 15, one weight fragment read for both tiles
    bfx4 a1[2][4], a2[2][4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const fx4 bias = *reinterpret_cast<const fx4 *>(sB + 16 * mt + 4 * g);   // the bias starts the sum
      fx4 acc0 = bias, acc1 = bias;
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        // inputs 112..127 do not exist (their B values are 0): lanes 32..63 of block 3 use a zero A
        const bfx8 wa = (kb == 3 && g >= 2) ? bfx8{0, 0, 0, 0, 0, 0, 0, 0} : ld8(sW1 + (16 * mt + r16) * NINP + 32 * kb + 8 * g);
        acc0 = mfma32(wa, xb[0][kb], acc0);
        acc1 = mfma32(wa, xb[1][kb], acc1);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        a1[0][mt][r] = bf16_bits(relu(acc0[r]));
        a1[1][mt][r] = bf16_bits(relu(acc1[r]));
      }
    }
    // layer 2: the B fragment of K block kb2 is (a1[2 kb2], a1[2 kb2 + 1]) in unit_of_k order
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const fx4 bias = *reinterpret_cast<const fx4 *>(sB + NH + 16 * mt + 4 * g);
      fx4 acc0 = bias, acc1 = bias;
#pragma unroll
      for (int kb = 0; kb < 2; kb++) {
        const bfx8 wa = ld8(sW2 + (16 * mt + r16) * NH + 32 * kb + 8 * g);
        acc0 = mfma32(wa, cat8(a1[0][2 * kb], a1[0][2 * kb + 1]), acc0);
        acc1 = mfma32(wa, cat8(a1[1][2 * kb], a1[1][2 * kb + 1]), acc1);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        a2[0][mt][r] = bf16_bits(relu(acc0[r]));
        a2[1][mt][r] = bf16_bits(relu(acc1[r]));
      }
    }
    const fx4 b3 = *reinterpret_cast<const fx4 *>(sB + 2 * NH + 4 * g);
    fx4 y0 = b3, y1 = b3;
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      const bfx8 wa = ld8(sW3 + r16 * NH + 32 * kb + 8 * g);
      y0 = mfma32(wa, cat8(a2[0][2 * kb], a2[0][2 * kb + 1]), y0);
      y1 = mfma32(wa, cat8(a2[1][2 * kb], a2[1][2 * kb + 1]), y1);
    }
    // Lane (r16, g) holds outputs 4 g .. 4 g + 3 of cell r16 of each tile: denormalise them there.
    const fx4 om = *reinterpret_cast<const fx4 *>(sMS + 4 * g), os = *reinterpret_cast<const fx4 *>(sMS + NOUTP + 4 * g);
    float t40[4], t41[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      t40[r] = __fadd_rn(om[r], __fmul_rn(os[r], y0[r]));
      t41[r] = __fadd_rn(om[r], __fmul_rn(os[r], y1[r]));
    }
    const long long idrow = ((long long)k * A.ny + j) * A.nx;
    // lane L < 32 takes cell 32 w + L (tile L / 16, cell L % 16): it gathers that cell's 9 outputs
    // from the lanes that hold them, stores them per face direction and applies the fallback rule
    const int src = r16;
    float t[NOUT];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float a0 = __shfl(t40[r], src), a1v = __shfl(t41[r], src);
      const float b0 = __shfl(t40[r], src + 16), b1v = __shfl(t41[r], src + 16);
      t[r] = g == 0 ? a0 : a1v;
      t[4 + r] = g == 0 ? b0 : b1v;
    }
    {
      const float c0 = __shfl(t40[0], src + 32), c1 = __shfl(t41[0], src + 32);
      t[8] = g == 0 ? c0 : c1;
    }
    const float mxl = g == 0 ? mxt[0] : mxt[1];
    const int xl = 32 * w + lane;                // this lane's cell (lanes 0..31)
    const int i = i0 + xl;
    if (lane < 32 && i < A.nx) {
      const long long id = idrow + i;
      // outputs t = (tau11, tau12, tau13, tau22, tau23, tau33, q1, q2, q3)
      A.tau[id] = make_float4(t[0], t[1], t[2], t[6]);
      A.tau[A.n3 + id] = make_float4(t[1], t[3], t[4], t[7]);
      A.tau[2 * A.n3 + id] = make_float4(t[2], t[4], t[5], t[8]);
      const int c = (1 * 3 + 1) * GBX + xl + 1;
      unsigned char fl = 0;
      if (A.valid[id] && mxl <= A.xmax) {
        const T *U = sQ, *V = sQ + GBN, *W = sQ + 2 * GBN, *TH = sQ + 3 * GBN;
        const int sy = GBX, sz = 3 * GBX;
        const double ux = (U[c + 1] - U[c - 1]) * A.r2dx, uy = (U[c + sy] - U[c - sy]) * A.r2dy, uz = (U[c + sz] - U[c - sz]) * A.r2dz;
        const double vx = (V[c + 1] - V[c - 1]) * A.r2dx, vy = (V[c + sy] - V[c - sy]) * A.r2dy, vz = (V[c + sz] - V[c - sz]) * A.r2dz;
        const double wx = (W[c + 1] - W[c - 1]) * A.r2dx, wy = (W[c + sy] - W[c - sy]) * A.r2dy, wz = (W[c + sz] - W[c - sz]) * A.r2dz;
        const double tx = (TH[c + 1] - TH[c - 1]) * A.r2dx, ty = (TH[c + sy] - TH[c - sy]) * A.r2dy, tz = (TH[c + sz] - TH[c - sz]) * A.r2dz;
        const double s11 = ux, s22 = vy, s33 = wz, s12 = 0.5 * (uy + vx), s13 = 0.5 * (uz + wx), s23 = 0.5 * (vz + wy);
        const double eps = -(t[0] * s11 + t[3] * s22 + t[5] * s33 + 2 * (t[1] * s12 + t[2] * s13 + t[4] * s23));
        const double ss = s11 * s11 + s22 * s22 + s33 * s33 + 2 * (s12 * s12 + s13 * s13 + s23 * s23);
        const double chi = -(t[6] * tx + t[7] * ty + t[8] * tz);
        const double gg = tx * tx + ty * ty + tz * tz;
        // A strain (theta gradient) at the rounding level of the cell's own values bounds nothing:
        // with a zero one any stress (flux) gives eps = 0 (chi = 0), and rounding-level differences
        // would decide the flag.  Below SGS_NOISE_REL of the cell's speed (theta) per cell width
        // they count as zero.
        const double rmax = fmax(A.r2dx, fmax(A.r2dy, A.r2dz));
        const double smin = SGS_NOISE_REL * (fabs((double)U[c]) + fabs((double)V[c]) + fabs((double)W[c])) * rmax;
        const double gmin = SGS_NOISE_REL * fabs((double)TH[c]) * rmax;
        if (ss > smin * smin && eps >= 0 && eps <= A.nu_max * 2 * ss) fl |= 1;
        if (gg > gmin * gmin && chi >= 0 && chi <= A.nu_max * gg) fl |= 2;
      }
      A.flags[id] = fl;
    }
  }
}

static void to_padded(const float *c, std::vector<float> &p) {
  p.assign(P_N, 0.f);
  const float *W1 = c, *b1 = W1 + NH * NIN, *W2 = b1 + NH, *b2 = W2 + NH * NH, *W3 = b2 + NH, *b3 = W3 + NOUT * NH;
  for (int o = 0; o < NH; o++) for (int i = 0; i < NIN; i++) p[P_W1 + o * NINP + i] = W1[o * NIN + i];
  for (int o = 0; o < NH; o++) p[P_B1 + o] = b1[o];
  for (int o = 0; o < NH; o++) for (int i = 0; i < NH; i++) p[P_W2 + o * NH + i] = W2[o * NH + i];
  for (int o = 0; o < NH; o++) p[P_B2 + o] = b2[o];
  for (int o = 0; o < NOUT; o++) for (int i = 0; i < NH; i++) p[P_W3 + o * NH + i] = W3[o * NH + i];
  for (int o = 0; o < NOUT; o++) p[P_B3 + o] = b3[o];
}

static void to_compact(const float *p, float *c) {
  float *W1 = c, *b1 = W1 + NH * NIN, *W2 = b1 + NH, *b2 = W2 + NH * NH, *W3 = b2 + NH, *b3 = W3 + NOUT * NH;
  for (int o = 0; o < NH; o++) for (int i = 0; i < NIN; i++) W1[o * NIN + i] = p[P_W1 + o * NINP + i];
  for (int o = 0; o < NH; o++) b1[o] = p[P_B1 + o];
  for (int o = 0; o < NH; o++) for (int i = 0; i < NH; i++) W2[o * NH + i] = p[P_W2 + o * NH + i];
  for (int o = 0; o < NH; o++) b2[o] = p[P_B2 + o];
  for (int o = 0; o < NOUT; o++) for (int i = 0; i < NH; i++) W3[o * NH + i] = p[P_W3 + o * NH + i];
  for (int o = 0; o < NOUT; o++) b3[o] = p[P_B3 + o];
}

static int mlp_ensure_blocks(pu_mlp *m, int nblocks) {
  if (nblocks <= m->max_blocks) return PU_OK;
  if (m->d_partial) hipFree(m->d_partial);
  if (m->d_ploss) hipFree(m->d_ploss);
  m->d_partial = nullptr;
  m->d_ploss = nullptr;
  m->max_blocks = 0;
  PU_HIP_CHECK(hipMalloc(&m->d_partial, sizeof(float) * (size_t)nblocks * P_N));
  PU_HIP_CHECK(hipMalloc(&m->d_ploss, sizeof(float) * (size_t)nblocks));
  m->max_blocks = nblocks;
  return PU_OK;
}

static int mlp_batch(pu_mlp *m, const short *Xbf, const float *T, const int *idx, int n, float *loss) {
  if (!m || !Xbf || !T || n <= 0) return fail(PU_ERR_ARG, "pu_mlp: null argument or empty batch");
  const int nb = (n + SAMPLES_PER_BLOCK - 1) / SAMPLES_PER_BLOCK;
  int rc = mlp_ensure_blocks(m, nb);
  if (rc) return rc;
  TrainArgs A{Xbf, T, idx, n, m->d_wbf, m->d_master, m->d_partial, m->d_ploss, nullptr, 2.f / ((float)n * NOUT)};
  hipLaunchKernelGGL(k_mlp_train<true>, dim3(nb), dim3(256), 0, 0, A);
  PU_HIP_CHECK(hipGetLastError());
  if (loss) {
    std::vector<float> pl(nb);
    PU_HIP_CHECK(hipMemcpy(pl.data(), m->d_ploss, sizeof(float) * nb, hipMemcpyDeviceToHost));
    double s = 0;
    for (float v : pl) s += v;
    *loss = (float)(s / ((double)n * NOUT));
  }
  return PU_OK;
}

static int mlp_reduce(pu_mlp *m, int nblocks, const float *gin, float *grad_out, bool update, float lr) {
  ReduceArgs R{};
  R.partial = m->d_partial; R.nblocks = nblocks; R.grad = grad_out; R.gin = gin;
  R.master = m->d_master; R.m = m->d_m; R.v = m->d_v; R.wbf = m->d_wbf;
  R.update = update ? 1 : 0;
  R.lr = lr; R.b1 = 0.9f; R.b2 = 0.999f; R.eps = 1e-8f;
  if (update) {
    m->t += 1;
    R.c1 = 1.f - powf(R.b1, (float)m->t);
    R.c2 = 1.f - powf(R.b2, (float)m->t);
  }
  hipLaunchKernelGGL(k_mlp_reduce, dim3((P_N + 255) / 256), dim3(256), 0, 0, R);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

// The network's outputs as tau [9][n3] (tau11, tau12, tau13, tau22, tau23, tau33, q1, q2, q3) from
// the per-direction layout the closure update reads.
__global__ __launch_bounds__(256) void k_tau_unpack(const float4 *__restrict__ t4, long long n3, float *__restrict__ tau) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n3) return;
  const float4 a = t4[id], b = t4[n3 + id], c = t4[2 * n3 + id];
  tau[id] = a.x; tau[n3 + id] = a.y; tau[2 * n3 + id] = a.z;
  tau[3 * n3 + id] = b.y; tau[4 * n3 + id] = b.z; tau[5 * n3 + id] = c.z;
  tau[6 * n3 + id] = a.w; tau[7 * n3 + id] = b.w; tau[8 * n3 + id] = c.w;
}

template <class T>
static int sgs_infer_t(pu_ctx *ctx, double dt) {
  const pu_config &c = ctx->cfg;
  const pu_mlp *m = ctx->sgs_model;
  const long long n3 = (long long)c.nx * c.ny * c.nz;
  if (!ctx->sgs_valid_ready) {
    hipLaunchKernelGGL(k_sgs_valid, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, ctx->stream, ctx->d_imm_les_val,
                       c.nx, c.ny, c.nz, ctx->d_sgs_valid);
    PU_HIP_CHECK(hipGetLastError());
    ctx->sgs_valid_ready = true;
  }
  const double dx = c.xlen / c.nx, dy = c.ylen / c.ny, dz = c.zlen / c.nz;
  InferArgs<T> A{};
  A.F = (const T *)ctx->d_les;
  A.nxh = c.nx + 2; A.nyh = c.ny + 2; A.nh = A.nxh * A.nyh * (c.nz + 2); A.n3 = n3;
  A.nx = c.nx; A.ny = c.ny; A.nz = c.nz;
  A.valid = ctx->d_sgs_valid;
  A.wbf = m->d_wbf; A.master = m->d_master; A.fmul = ctx->d_sgs_fmul;
  A.xmax = m->xmax;
  for (int o = 0; o < NOUT; o++) { A.out_mean[o] = m->out_mean[o]; A.out_scale[o] = m->out_scale[o]; }
  A.r2dx = 1. / (2 * dx); A.r2dy = 1. / (2 * dy); A.r2dz = 1. / (2 * dz);
  const double vmx = 0.25 * dx * dx / dt, vmy = 0.25 * dy * dy / dt, vmz = 0.25 * dz * dz / dt;
  A.nu_max = 0.5 * std::fmin(vmx, std::fmin(vmy, vmz));
  A.tau = reinterpret_cast<float4 *>(ctx->d_sgs_tau); A.flags = ctx->d_sgs_flags;
  const long long ngroups = (long long)((c.nx + GX - 1) / GX) * c.ny * c.nz;
  if (ngroups >= (1LL << 32)) return fail(PU_ERR_UNSUPPORTED, "learned closure: grid too large for the inference kernel");
  const unsigned grid = (unsigned)std::min<long long>(ngroups, 256LL * 16);
  hipLaunchKernelGGL(k_sgs_infer<T>, dim3(grid), dim3(256), 0, ctx->stream, A);
  PU_HIP_CHECK(hipGetLastError());
  return PU_OK;
}

int sgs_infer_device(pu_ctx *ctx, double dt) {
  return ctx->cfg.precision == PU_FP32 ? sgs_infer_t<float>(ctx, dt) : sgs_infer_t<double>(ctx, dt);
}

}  // namespace pu

using namespace pu;

extern "C" {

int pu_mlp_create(pu_mlp **out) {
  if (!out) return fail(PU_ERR_ARG, "null output pointer");
  *out = nullptr;
  pu_mlp *m = new pu_mlp();
  for (int f = 0; f < NIN; f++) m->feat_scale[f] = 1.f;
  for (int o = 0; o < NOUT; o++) { m->out_mean[o] = 0.f; m->out_scale[o] = 1.f; }
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&m->d_master, sizeof(float) * P_N)) != hipSuccess || (e = hipMalloc(&m->d_wbf, sizeof(short) * P_N)) != hipSuccess ||
      (e = hipMalloc(&m->d_m, sizeof(float) * P_N)) != hipSuccess || (e = hipMalloc(&m->d_v, sizeof(float) * P_N)) != hipSuccess ||
      (e = hipMalloc(&m->d_grad, sizeof(float) * P_N)) != hipSuccess) {
    pu_mlp_destroy(m);
    return fail(PU_ERR_HIP, std::string("pu_mlp_create: ") + hipGetErrorString(e));
  }
  hipMemset(m->d_master, 0, sizeof(float) * P_N);
  hipMemset(m->d_wbf, 0, sizeof(short) * P_N);
  hipMemset(m->d_m, 0, sizeof(float) * P_N);
  hipMemset(m->d_v, 0, sizeof(float) * P_N);
  *out = m;
  return PU_OK;
}

int pu_mlp_destroy(pu_mlp *m) {
  if (!m) return PU_OK;
  hipDeviceSynchronize();
  void *ptrs[] = {m->d_master, m->d_wbf, m->d_m, m->d_v, m->d_partial, m->d_ploss, m->d_grad};
  for (void *p : ptrs) if (p) hipFree(p);
  delete m;
  return PU_OK;
}

int pu_mlp_set_params(pu_mlp *m, const float *params) {
  if (!m || !params) return fail(PU_ERR_ARG, "pu_mlp_set_params: null argument");
  std::vector<float> p;
  to_padded(params, p);
  PU_HIP_CHECK(hipMemcpy(m->d_master, p.data(), sizeof(float) * P_N, hipMemcpyHostToDevice));
  PU_HIP_CHECK(hipMemset(m->d_m, 0, sizeof(float) * P_N));
  PU_HIP_CHECK(hipMemset(m->d_v, 0, sizeof(float) * P_N));
  m->t = 0;
  hipLaunchKernelGGL(k_mlp_refresh_bf16, dim3((P_N + 255) / 256), dim3(256), 0, 0, m->d_master, m->d_wbf);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipDeviceSynchronize());
  return PU_OK;
}

int pu_mlp_get_params(const pu_mlp *m, float *params) {
  if (!m || !params) return fail(PU_ERR_ARG, "pu_mlp_get_params: null argument");
  std::vector<float> p(P_N);
  PU_HIP_CHECK(hipDeviceSynchronize());
  PU_HIP_CHECK(hipMemcpy(p.data(), m->d_master, sizeof(float) * P_N, hipMemcpyDeviceToHost));
  to_compact(p.data(), params);
  return PU_OK;
}

int pu_mlp_set_normalisation(pu_mlp *m, const float *feat_scale, float xmax, const float *out_mean, const float *out_scale) {
  if (!m || !feat_scale || !out_mean || !out_scale) return fail(PU_ERR_ARG, "pu_mlp_set_normalisation: null argument");
  if (!(xmax >= 0)) return fail(PU_ERR_ARG, "pu_mlp_set_normalisation: xmax must be >= 0");
  std::memcpy(m->feat_scale, feat_scale, sizeof(m->feat_scale));
  std::memcpy(m->out_mean, out_mean, sizeof(m->out_mean));
  std::memcpy(m->out_scale, out_scale, sizeof(m->out_scale));
  m->xmax = xmax;
  return PU_OK;
}

int pu_mlp_get_normalisation(const pu_mlp *m, float *feat_scale, float *xmax, float *out_mean, float *out_scale) {
  if (!m) return fail(PU_ERR_ARG, "pu_mlp_get_normalisation: null model");
  if (feat_scale) std::memcpy(feat_scale, m->feat_scale, sizeof(m->feat_scale));
  if (xmax) *xmax = m->xmax;
  if (out_mean) std::memcpy(out_mean, m->out_mean, sizeof(m->out_mean));
  if (out_scale) std::memcpy(out_scale, m->out_scale, sizeof(m->out_scale));
  return PU_OK;
}

// File: "PUSGS001", int32 nin, nhidden, nout, nparam, then float32 feat_scale[nin], xmax,
// out_mean[nout], out_scale[nout], params[nparam] (little endian).
int pu_mlp_save(const pu_mlp *m, const char *path) {
  if (!m || !path) return fail(PU_ERR_ARG, "pu_mlp_save: null argument");
  std::vector<float> p(PU_SGS_NPARAM);
  int rc = pu_mlp_get_params(m, p.data());
  if (rc) return rc;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(PU_ERR_ARG, std::string("pu_mlp_save: cannot open ") + path);
  const int hdr[4] = {NIN, NH, NOUT, PU_SGS_NPARAM};
  bool ok = std::fwrite("PUSGS001", 1, 8, f) == 8 && std::fwrite(hdr, sizeof(int), 4, f) == 4 &&
            std::fwrite(m->feat_scale, sizeof(float), NIN, f) == (size_t)NIN && std::fwrite(&m->xmax, sizeof(float), 1, f) == 1 &&
            std::fwrite(m->out_mean, sizeof(float), NOUT, f) == (size_t)NOUT &&
            std::fwrite(m->out_scale, sizeof(float), NOUT, f) == (size_t)NOUT &&
            std::fwrite(p.data(), sizeof(float), p.size(), f) == p.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? PU_OK : fail(PU_ERR_ARG, std::string("pu_mlp_save: write failed: ") + path);
}

int pu_mlp_load(const char *path, pu_mlp **out) {
  if (!path || !out) return fail(PU_ERR_ARG, "pu_mlp_load: null argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(PU_ERR_ARG, std::string("pu_mlp_load: cannot open ") + path);
  char magic[8];
  int hdr[4];
  float fs[NIN], xmax, om[NOUT], os[NOUT];
  std::vector<float> p(PU_SGS_NPARAM);
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "PUSGS001", 8) == 0 && std::fread(hdr, sizeof(int), 4, f) == 4 &&
            hdr[0] == NIN && hdr[1] == NH && hdr[2] == NOUT && hdr[3] == PU_SGS_NPARAM &&
            std::fread(fs, sizeof(float), NIN, f) == (size_t)NIN && std::fread(&xmax, sizeof(float), 1, f) == 1 &&
            std::fread(om, sizeof(float), NOUT, f) == (size_t)NOUT && std::fread(os, sizeof(float), NOUT, f) == (size_t)NOUT &&
            std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
  std::fclose(f);
  if (!ok) return fail(PU_ERR_ARG, std::string("pu_mlp_load: not a 108-64-64-9 network file: ") + path);
  pu_mlp *m = nullptr;
  int rc = pu_mlp_create(&m);
  if (rc) return rc;
  if ((rc = pu_mlp_set_params(m, p.data())) || (rc = pu_mlp_set_normalisation(m, fs, xmax, om, os))) {
    pu_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return PU_OK;
}

int pu_mlp_pack_inputs(const float *X, long long n, unsigned short *Xbf) {
  if (!X || !Xbf || n <= 0) return fail(PU_ERR_ARG, "pu_mlp_pack_inputs: null argument or n <= 0");
  const long long tot = n * NINP;
  hipLaunchKernelGGL(k_pack_inputs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, X, n, (short *)Xbf);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipDeviceSynchronize());
  return PU_OK;
}

int pu_mlp_forward(pu_mlp *m, const unsigned short *Xbf, const int *idx, int n, float *Y) {
  if (!m || !Xbf || !Y || n <= 0) return fail(PU_ERR_ARG, "pu_mlp_forward: null argument or n <= 0");
  const int nb = (n + SAMPLES_PER_BLOCK - 1) / SAMPLES_PER_BLOCK;
  TrainArgs A{(const short *)Xbf, nullptr, idx, n, m->d_wbf, m->d_master, nullptr, nullptr, Y, 0.f};
  hipLaunchKernelGGL(k_mlp_train<false>, dim3(nb), dim3(256), 0, 0, A);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipDeviceSynchronize());
  return PU_OK;
}

int pu_mlp_loss_grad(pu_mlp *m, const unsigned short *Xbf, const float *T, const int *idx, int n, float *grad, float *loss) {
  if (!grad) return fail(PU_ERR_ARG, "pu_mlp_loss_grad: null gradient");
  int rc = mlp_batch(m, (const short *)Xbf, T, idx, n, loss);
  if (rc) return rc;
  rc = mlp_reduce(m, (n + SAMPLES_PER_BLOCK - 1) / SAMPLES_PER_BLOCK, nullptr, m->d_grad, false, 0.f);
  if (rc) return rc;
  std::vector<float> p(P_N), c(PU_SGS_NPARAM);
  PU_HIP_CHECK(hipMemcpy(p.data(), m->d_grad, sizeof(float) * P_N, hipMemcpyDeviceToHost));
  to_compact(p.data(), c.data());
  PU_HIP_CHECK(hipMemcpy(grad, c.data(), sizeof(float) * PU_SGS_NPARAM, hipMemcpyHostToDevice));
  return PU_OK;
}

int pu_mlp_train_step(pu_mlp *m, const unsigned short *Xbf, const float *T, const int *idx, int n, float lr, float *loss) {
  int rc = mlp_batch(m, (const short *)Xbf, T, idx, n, loss);
  if (rc) return rc;
  return mlp_reduce(m, (n + SAMPLES_PER_BLOCK - 1) / SAMPLES_PER_BLOCK, nullptr, nullptr, true, lr);
}

int pu_mlp_adam_apply(pu_mlp *m, const float *grad, float lr) {
  if (!m || !grad) return fail(PU_ERR_ARG, "pu_mlp_adam_apply: null argument");
  std::vector<float> c(PU_SGS_NPARAM), p;
  PU_HIP_CHECK(hipMemcpy(c.data(), grad, sizeof(float) * PU_SGS_NPARAM, hipMemcpyDeviceToHost));
  to_padded(c.data(), p);
  PU_HIP_CHECK(hipMemcpy(m->d_grad, p.data(), sizeof(float) * P_N, hipMemcpyHostToDevice));
  int rc = mlp_reduce(m, 0, m->d_grad, nullptr, true, lr);
  if (rc) return rc;
  PU_HIP_CHECK(hipDeviceSynchronize());
  return PU_OK;
}

int pu_sgs_filter(pu_ctx *ctx, const pu_coupler_state *cs, int r, double *les, double *imm, float *targets) {
  if (!ctx || !cs || !les || !imm || !targets) return fail(PU_ERR_ARG, "pu_sgs_filter: null argument");
  const pu_config &c = ctx->cfg;
  if (r < 1 || c.nx % r || c.ny % r || c.nz % r) return fail(PU_ERR_ARG, "pu_sgs_filter: r must divide nx, ny, nz");
  if (!cs->density_dry || !cs->uvel || !cs->vvel || !cs->wvel || !cs->temp || !cs->immersed_proportion)
    return fail(PU_ERR_ARG, "pu_sgs_filter: coupler state needs density_dry, uvel, vvel, wvel, temp, immersed_proportion");
  FilterArgs A{};
  A.rho = cs->density_dry; A.u = cs->uvel; A.v = cs->vvel; A.w = cs->wvel; A.T = cs->temp; A.imm = cs->immersed_proportion;
  A.nx = c.nx; A.ny = c.ny; A.nz = c.nz; A.r = r;
  A.nxc = c.nx / r; A.nyc = c.ny / r; A.nzc = c.nz / r;
  A.n3 = (long long)c.nx * c.ny * c.nz;
  A.n3c = (long long)A.nxc * A.nyc * A.nzc;
  A.nxh = A.nxc + 2; A.nyh = A.nyc + 2; A.nh = A.nxh * A.nyh * (A.nzc + 2);
  A.R_d = c.R_d; A.C0 = c.C0; A.inv_gamma = 1. / c.gamma_d;
  A.les = les; A.immc = imm; A.targets = targets;
  hipLaunchKernelGGL(k_sgs_filter, dim3((unsigned)((A.n3c + 255) / 256)), dim3(256), 0, ctx->stream, A);
  PU_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sgs_filter_halo, dim3((unsigned)((A.nh + 255) / 256)), dim3(256), 0, ctx->stream, A);
  PU_HIP_CHECK(hipGetLastError());
  PU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return PU_OK;
}

int pu_sgs_features(const double *les, int nx, int ny, int nz, const float *fmul, const long long *cells, long long n,
                    unsigned short *Xbf) {
  if (!les || !fmul || !cells || !Xbf || n <= 0 || nx < 1 || ny < 1 || nz < 1)
    return fail(PU_ERR_ARG, "pu_sgs_features: null argument or empty");
  float *d_fm = nullptr;
  PU_HIP_CHECK(hipMalloc(&d_fm, sizeof(float) * NIN));
  hipError_t e = hipMemcpy(d_fm, fmul, sizeof(float) * NIN, hipMemcpyHostToDevice);
  if (e != hipSuccess) { hipFree(d_fm); return fail(PU_ERR_HIP, hipGetErrorString(e)); }
  FeatArgs A{les, nx, ny, nz, nx + 2, ny + 2, (long long)(nx + 2) * (ny + 2) * (nz + 2), d_fm, cells, n, (short *)Xbf};
  const long long tot = n * 28;
  hipLaunchKernelGGL(k_sgs_features, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, A);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  hipFree(d_fm);
  if (e != hipSuccess) return fail(PU_ERR_HIP, std::string("pu_sgs_features: ") + hipGetErrorString(e));
  return PU_OK;
}

int pu_les_closure_set_model(pu_ctx *ctx, const pu_mlp *m) {
  if (!ctx) return fail(PU_ERR_ARG, "null context");
  if (!m) { ctx->sgs_model = nullptr; return PU_OK; }
  const pu_config &c = ctx->cfg;
  if (c.idTKE < 0) return fail(PU_ERR_ARG, "the learned closure needs the LES closure (idTKE)");
  const size_t n3 = (size_t)c.nx * c.ny * c.nz;
  if (!ctx->d_sgs_tau) {
    PU_HIP_CHECK(hipMalloc(&ctx->d_sgs_tau, sizeof(float) * 12 * n3));
    PU_HIP_CHECK(hipMalloc(&ctx->d_sgs_flags, n3));
    PU_HIP_CHECK(hipMalloc(&ctx->d_sgs_valid, n3));
    PU_HIP_CHECK(hipMalloc(&ctx->d_sgs_fmul, sizeof(float) * NIN));
    ctx->sgs_valid_ready = false;
  }
  const double delta = std::pow((c.xlen / c.nx) * (c.ylen / c.ny) * (c.zlen / c.nz), 1. / 3.);
  float fm[NIN];
  for (int f = 0; f < NIN; f++) fm[f] = (float)((double)m->feat_scale[f] / delta);
  PU_HIP_CHECK(hipMemcpy(ctx->d_sgs_fmul, fm, sizeof(fm), hipMemcpyHostToDevice));
  ctx->sgs_model = m;
  return PU_OK;
}

int pu_les_closure_sgs(pu_ctx *ctx, const pu_coupler_state *cs, double dt_phys, float *tau, unsigned char *flags) {
  if (!ctx || !cs) return fail(PU_ERR_ARG, "null argument");
  if (!ctx->sgs_model) return fail(PU_ERR_STATE, "no network attached (pu_les_closure_set_model)");
  if (!ctx->dycore_ready && !ctx->les_ready)
    return fail(PU_ERR_STATE, "pu_les_closure_init or pu_dycore_init has not been called (immersed halos)");
  if (!(dt_phys > 0)) return fail(PU_ERR_ARG, "dt_phys must be positive");
  int rc = les_prim_device(ctx, cs);
  if (rc) return rc;
  rc = sgs_infer_device(ctx, dt_phys);
  if (rc) return rc;
  const size_t n3 = (size_t)ctx->cfg.nx * ctx->cfg.ny * ctx->cfg.nz;
  if (tau) {
    hipLaunchKernelGGL(k_tau_unpack, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const float4 *>(ctx->d_sgs_tau), (long long)n3, tau);
    PU_HIP_CHECK(hipGetLastError());
  }
  if (flags) PU_HIP_CHECK(hipMemcpyAsync(flags, ctx->d_sgs_flags, n3, hipMemcpyDeviceToDevice, ctx->stream));
  PU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return PU_OK;
}

}  // extern "C"
