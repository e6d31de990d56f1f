/*
 * sgs.c -- CPU oracle for the learned subgrid closure.  TEST INFRASTRUCTURE ONLY.
 *
 * The learned closure has no counterpart in the reference: it replaces the eddy-viscosity face
 * fluxes of LES_Closure (model/modules/les_closure.h:92-229) for momentum and potential
 * temperature with fluxes of a per-cell network prediction of the subgrid stresses.  The nearest
 * prior art in the reference is build/postproc/compare_les.py:11-60 (explicit vs closure SGS
 * tendencies).  Everything here is the specification the HIP kernels are checked against:
 *
 *   oc_sgs_filter      box filter of a fine run onto an r-times coarser grid: filtered
 *                      (rho, u, v, w, theta), subgrid stresses tau_ij = <u_i u_j> - <u_i><u_j>
 *                      and heat fluxes q_j = <u_j theta> - <u_j><theta>;
 *   oc_sgs_features    the 108 inputs of one cell: for the 27 cells of its 3x3x3 neighbourhood
 *                      (n = (dz+1)*9 + (dy+1)*3 + (dx+1)) the differences of (u, v, w, theta)
 *                      to the centre cell, times a per-input multiplier fmul (= 1/(delta*sigma));
 *   oc_mlp_forward     108 -> 64 -> 64 -> 9, ReLU on the hidden layers;
 *   oc_mlp_loss_grad   mean-squared loss over n samples and its exact gradient (backpropagation);
 *   oc_mlp_adam        one Adam step (Kingma & Ba 2015, Algorithm 1, bias-corrected);
 *   oc_sgs_infer       prediction + fallback decision for every interior cell of a halo-1
 *                      closure box (the box LES_Closure builds, les_closure.h:328-536).
 *
 * With bf16 != 0 every matrix-product operand (inputs, weights, activations, back-propagated
 * deltas) is rounded to bfloat16 (round to nearest even) first, as the HIP kernels feed the
 * matrix cores; sums are accumulated in double here (float on the GPU).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NIN OC_MLP_NIN
#define NH OC_MLP_NH
#define NOUT OC_MLP_NOUT
/* fallback rule: strains / theta gradients below this fraction of the cell's speed / theta per
 * cell width count as zero (the same constant as the product's inference kernel) */
#define SGS_NOISE_REL 2e-6

/* offsets of the blocks in the parameter vector */
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NH * NIN)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NOUT * NH)

float oc_bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;   /* inf / nan unchanged */
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&x, &u, 4);
  return x;
}

static inline float rnd(float x, int bf16) { return bf16 ? oc_bf16_round(x) : x; }

/* ReLU and its derivative (0 at x = 0, as torch's backward takes it) */
double oc_relu(double x) { return x > 0. ? x : 0.; }

double oc_relu_deriv(double x) { return x > 0. ? 1. : 0.; }

/* One sample through the network.  Keeps the pre-activations (float, as the GPU holds them) and
 * the rounded activations for the backward pass. */
typedef struct {
  float x[NIN];            /* rounded inputs */
  float h1[NH], a1[NH];    /* pre-activation (float) and rounded activation */
  float h2[NH], a2[NH];
  float y[NOUT];
} fwd_t;

static void forward_one(const float *P, const float *xin, int bf16, fwd_t *f) {
  for (int i = 0; i < NIN; i++) f->x[i] = rnd(xin[i], bf16);
  for (int o = 0; o < NH; o++) {
    double s = 0;
    for (int i = 0; i < NIN; i++) s += (double)rnd(P[OFF_W1 + o * NIN + i], bf16) * f->x[i];
    f->h1[o] = (float)s + P[OFF_B1 + o];
    f->a1[o] = rnd((float)oc_relu(f->h1[o]), bf16);
  }
  for (int o = 0; o < NH; o++) {
    double s = 0;
    for (int i = 0; i < NH; i++) s += (double)rnd(P[OFF_W2 + o * NH + i], bf16) * f->a1[i];
    f->h2[o] = (float)s + P[OFF_B2 + o];
    f->a2[o] = rnd((float)oc_relu(f->h2[o]), bf16);
  }
  for (int o = 0; o < NOUT; o++) {
    double s = 0;
    for (int i = 0; i < NH; i++) s += (double)rnd(P[OFF_W3 + o * NH + i], bf16) * f->a2[i];
    f->y[o] = (float)s + P[OFF_B3 + o];
  }
}

void oc_mlp_forward(const float *P, const float *X, long n, int bf16, float *Y) {
#pragma omp parallel for
  for (long s = 0; s < n; s++) {
    fwd_t f;
    forward_one(P, X + s * NIN, bf16, &f);
    for (int o = 0; o < NOUT; o++) Y[s * NOUT + o] = f.y[o];
  }
}

/* loss = sum_{s,o} (y_so - t_so)^2 / (n * NOUT); G = d loss / d P (double, same layout as P). */
double oc_mlp_loss_grad(const float *P, const float *X, const float *T, long n, int bf16, double *G) {
  memset(G, 0, sizeof(double) * OC_MLP_NPARAM);
  double loss = 0;
  const double scale = 2. / ((double)n * NOUT);
  for (long s = 0; s < n; s++) {
    fwd_t f;
    forward_one(P, X + s * NIN, bf16, &f);
    float dy[NOUT], dyr[NOUT];
    for (int o = 0; o < NOUT; o++) {
      double e = (double)f.y[o] - T[s * NOUT + o];
      loss += e * e;
      dy[o] = (float)(scale * e);
      dyr[o] = rnd(dy[o], bf16);
    }
    /* layer 3 */
    for (int o = 0; o < NOUT; o++) {
      G[OFF_B3 + o] += dy[o];
      for (int i = 0; i < NH; i++) G[OFF_W3 + o * NH + i] += (double)dyr[o] * f.a2[i];
    }
    float dh2[NH], dh2r[NH];
    for (int i = 0; i < NH; i++) {
      double s3 = 0;
      for (int o = 0; o < NOUT; o++) s3 += (double)rnd(P[OFF_W3 + o * NH + i], bf16) * dyr[o];
      dh2[i] = (float)((double)(float)s3 * oc_relu_deriv(f.h2[i]));
      dh2r[i] = rnd(dh2[i], bf16);
    }
    /* layer 2 */
    for (int o = 0; o < NH; o++) {
      G[OFF_B2 + o] += dh2[o];
      for (int i = 0; i < NH; i++) G[OFF_W2 + o * NH + i] += (double)dh2r[o] * f.a1[i];
    }
    float dh1[NH], dh1r[NH];
    for (int i = 0; i < NH; i++) {
      double s2 = 0;
      for (int o = 0; o < NH; o++) s2 += (double)rnd(P[OFF_W2 + o * NH + i], bf16) * dh2r[o];
      dh1[i] = (float)((double)(float)s2 * oc_relu_deriv(f.h1[i]));
      dh1r[i] = rnd(dh1[i], bf16);
    }
    /* layer 1 */
    for (int o = 0; o < NH; o++) {
      G[OFF_B1 + o] += dh1[o];
      for (int i = 0; i < NIN; i++) G[OFF_W1 + o * NIN + i] += (double)dh1r[o] * f.x[i];
    }
  }
  return loss / ((double)n * NOUT);
}

void oc_mlp_adam(float *P, const float *G, float *m, float *v, long np, int t, float lr, float b1, float b2,
                 float eps) {
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  for (long i = 0; i < np; i++) {
    float g = G[i];
    m[i] = b1 * m[i] + (1.f - b1) * g;
    v[i] = b2 * v[i] + (1.f - b2) * (g * g);
    float mh = m[i] / c1, vh = v[i] / c2;
    P[i] = P[i] - lr * mh / (sqrtf(vh) + eps);
  }
}

/* ----------------------------------------------------------------------------------------- */
/* filter, features, inference                                                                */
/* ----------------------------------------------------------------------------------------- */

static inline int wrapi(int i, int n) { i %= n; return i < 0 ? i + n : i; }

int oc_sgs_filter(const oc_config *c, const double *cp, const double *imm_fine, int r, double *les,
                  double *imm_c, float *targets) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  if (r < 1 || nx % r || ny % r || nz % r) return -1;
  int nxc = nx / r, nyc = ny / r, nzc = nz / r;
  size_t n3 = (size_t)nx * ny * nz, n3c = (size_t)nxc * nyc * nzc;
  size_t nxh = nxc + 2, nyh = nyc + 2, nzh = nzc + 2, nh = nxh * nyh * nzh;
  double inv = 1. / ((double)r * r * r);
#define HC(k, j, i) ((((size_t)(k)) * nyh + (j)) * nxh + (i))
#pragma omp parallel for collapse(2)
  for (int K = 0; K < nzc; K++) for (int J = 0; J < nyc; J++) for (int I = 0; I < nxc; I++) {
    double m[5] = {0, 0, 0, 0, 0}, uu[6] = {0, 0, 0, 0, 0, 0}, ut[3] = {0, 0, 0}, im = 0;
    for (int a = 0; a < r; a++) for (int b = 0; b < r; b++) for (int e = 0; e < r; e++) {
      size_t id = ((size_t)(K * r + a) * ny + (J * r + b)) * nx + (I * r + e);
      double rho = cp[id], u = cp[n3 + id], v = cp[2 * n3 + id], w = cp[3 * n3 + id];
      double th = pow(rho * c->R_d * cp[4 * n3 + id] / c->C0, 1. / c->gamma_d) / rho;
      m[0] += rho; m[1] += u; m[2] += v; m[3] += w; m[4] += th;
      uu[0] += u * u; uu[1] += u * v; uu[2] += u * w; uu[3] += v * v; uu[4] += v * w; uu[5] += w * w;
      ut[0] += u * th; ut[1] += v * th; ut[2] += w * th;
      if (imm_fine[id] > im) im = imm_fine[id];
    }
    for (int f = 0; f < 5; f++) m[f] *= inv;
    for (int f = 0; f < 6; f++) uu[f] *= inv;
    for (int f = 0; f < 3; f++) ut[f] *= inv;
    double tau[6] = {uu[0] - m[1] * m[1], uu[1] - m[1] * m[2], uu[2] - m[1] * m[3],
                     uu[3] - m[2] * m[2], uu[4] - m[2] * m[3], uu[5] - m[3] * m[3]};
    double q[3] = {ut[0] - m[1] * m[4], ut[1] - m[2] * m[4], ut[2] - m[3] * m[4]};
    size_t h = HC(K + 1, J + 1, I + 1), id = ((size_t)K * nyc + J) * nxc + I;
    for (int f = 0; f < 5; f++) les[f * nh + h] = m[f];
    les[5 * nh + h] = 0.5 * (tau[0] + tau[3] + tau[5]);
    imm_c[h] = im;
    for (int f = 0; f < 6; f++) targets[f * n3c + id] = (float)tau[f];
    for (int f = 0; f < 3; f++) targets[(6 + f) * n3c + id] = (float)q[f];
  }
  /* halos: x/y periodic (corners wrap both), z: copy of the boundary level with w = 0; immersed
   * z halos = 1 */
  for (size_t kh = 0; kh < nzh; kh++) for (size_t jh = 0; jh < nyh; jh++) for (size_t ih = 0; ih < nxh; ih++) {
    int k = (int)kh - 1, j = (int)jh - 1, i = (int)ih - 1;
    int interior = k >= 0 && k < nzc && j >= 0 && j < nyc && i >= 0 && i < nxc;
    if (interior) continue;
    int ks = k < 0 ? 0 : (k >= nzc ? nzc - 1 : k);
    size_t hs = HC(ks + 1, wrapi(j, nyc) + 1, wrapi(i, nxc) + 1), h = HC(kh, jh, ih);
    for (int f = 0; f < 6; f++) les[f * nh + h] = les[f * nh + hs];
    if (k != ks) { les[3 * nh + h] = 0; imm_c[h] = 1; }
    else imm_c[h] = imm_c[hs];
  }
#undef HC
  return 0;
}

/* features of one cell at halo coordinates (kh, jh, ih); returns max |x| */
static float features_one(const double *les, size_t nxh, size_t nyh, size_t nh, int kh, int jh, int ih,
                          const float *fmul, float *x) {
  size_t c = ((size_t)kh * nyh + jh) * nxh + ih;
  float mx = 0;
  for (int dz = -1; dz <= 1; dz++) for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
    int n = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
    size_t nb = ((size_t)(kh + dz) * nyh + (jh + dy)) * nxh + (ih + dx);
    for (int q = 0; q < 4; q++) {
      const double *F = les + (size_t)(1 + q) * nh;
      float v = (float)(F[nb] - F[c]) * fmul[4 * n + q];
      x[4 * n + q] = v;
      float a = fabsf(v);
      if (a > mx) mx = a;
    }
  }
  return mx;
}

void oc_sgs_fmul(const float *feat_scale, double delta, float *fmul) {
  for (int f = 0; f < NIN; f++) fmul[f] = (float)((double)feat_scale[f] / delta);
}

int oc_sgs_features(const double *les, int nx, int ny, int nz, const float *fmul, const long *cells, long n,
                    float *X) {
  size_t nxh = nx + 2, nyh = ny + 2, nh = nxh * nyh * (size_t)(nz + 2);
#pragma omp parallel for
  for (long s = 0; s < n; s++) {
    long id = cells[s];
    int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((long)nx * ny));
    features_one(les, nxh, nyh, nh, k + 1, j + 1, i + 1, fmul, X + s * NIN);
  }
  return 0;
}

/* Prediction and fallback flags for every interior cell of a halo-1 closure box.
 * flags bit 0: momentum fluxes from the network; bit 1: heat fluxes from the network.
 * A cell falls back to the eddy-viscosity closure (bit cleared) when
 *   - it is in the first or last level, or any of its 27 neighbours has immersed proportion > 0;
 *   - any normalised input exceeds M->xmax in magnitude (out of the training range);
 *   - momentum: the resolved strain is at the rounding level of the cell's speed (S_ij S_ij <=
 *     (SGS_NOISE_REL (|u| + |v| + |w|) / (2 min(dx, dy, dz)))^2: nothing bounds the stresses), or
 *     the predicted dissipation eps = -tau_ij S_ij is negative, or exceeds nu_max * 2 S_ij S_ij (an
 *     effective viscosity above the closure's stability cap nu_max);
 *   - heat: the resolved theta gradient is at the rounding level of theta (|grad theta| <=
 *     SGS_NOISE_REL theta / (2 min(dx, dy, dz))), or chi = -q_j dtheta/dx_j is negative or exceeds
 *     nu_max |grad theta|^2.
 * tau: [9][nz][ny][nx] (tau11 tau12 tau13 tau22 tau23 tau33 q1 q2 q3, physical units). */
void oc_sgs_infer(const oc_sgs_model *M, const double *les, const double *imm, int nx, int ny, int nz, double dx,
                  double dy, double dz, double nu_max, int bf16, float *tau, unsigned char *flags) {
  size_t nxh = nx + 2, nyh = ny + 2, nh = nxh * nyh * (size_t)(nz + 2), n3 = (size_t)nx * ny * nz;
  double delta = pow(dx * dy * dz, 1. / 3.);
  float fmul[NIN];
  oc_sgs_fmul(M->feat_scale, delta, fmul);
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    size_t id = ((size_t)k * ny + j) * nx + i;
    int kh = k + 1, jh = j + 1, ih = i + 1;
    size_t c = ((size_t)kh * nyh + jh) * nxh + ih;
    float x[NIN];
    float mx = features_one(les, nxh, nyh, nh, kh, jh, ih, fmul, x);
    fwd_t f;
    forward_one(M->params, x, bf16, &f);
    float t[9];
    for (int o = 0; o < 9; o++) t[o] = M->out_mean[o] + M->out_scale[o] * f.y[o];
    for (int o = 0; o < 9; o++) tau[o * n3 + id] = t[o];
    int ok = k >= 1 && k <= nz - 2 && mx <= M->xmax;
    for (int dzz = -1; dzz <= 1 && ok; dzz++) for (int dyy = -1; dyy <= 1 && ok; dyy++)
      for (int dxx = -1; dxx <= 1; dxx++)
        if (imm[((size_t)(kh + dzz) * nyh + (jh + dyy)) * nxh + (ih + dxx)] > 0) { ok = 0; break; }
    unsigned char fl = 0;
    if (ok) {
      const double *U = les + nh, *V = les + 2 * nh, *W = les + 3 * nh, *TH = les + 4 * nh;
      size_t sx = 1, sy = nxh, sz = nxh * nyh;
      double ux = (U[c + sx] - U[c - sx]) / (2 * dx), uy = (U[c + sy] - U[c - sy]) / (2 * dy), uz = (U[c + sz] - U[c - sz]) / (2 * dz);
      double vx = (V[c + sx] - V[c - sx]) / (2 * dx), vy = (V[c + sy] - V[c - sy]) / (2 * dy), vz = (V[c + sz] - V[c - sz]) / (2 * dz);
      double wx = (W[c + sx] - W[c - sx]) / (2 * dx), wy = (W[c + sy] - W[c - sy]) / (2 * dy), wz = (W[c + sz] - W[c - sz]) / (2 * dz);
      double tx = (TH[c + sx] - TH[c - sx]) / (2 * dx), ty = (TH[c + sy] - TH[c - sy]) / (2 * dy), tz = (TH[c + sz] - TH[c - sz]) / (2 * dz);
      double s11 = ux, s22 = vy, s33 = wz, s12 = 0.5 * (uy + vx), s13 = 0.5 * (uz + wx), s23 = 0.5 * (vz + wy);
      double eps = -(t[0] * s11 + t[3] * s22 + t[5] * s33 + 2 * (t[1] * s12 + t[2] * s13 + t[4] * s23));
      double ss = s11 * s11 + s22 * s22 + s33 * s33 + 2 * (s12 * s12 + s13 * s13 + s23 * s23);
      double chi = -(t[6] * tx + t[7] * ty + t[8] * tz);
      double gg = tx * tx + ty * ty + tz * tz;
      double rmax = fmax(1 / (2 * dx), fmax(1 / (2 * dy), 1 / (2 * dz)));
      double smin = SGS_NOISE_REL * (fabs(U[c]) + fabs(V[c]) + fabs(W[c])) * rmax;
      double gmin = SGS_NOISE_REL * fabs(TH[c]) * rmax;
      if (ss > smin * smin && eps >= 0 && eps <= nu_max * 2 * ss) fl |= 1;
      if (gg > gmin * gmin && chi >= 0 && chi <= nu_max * gg) fl |= 2;
    }
    flags[id] = fl;
  }
}
