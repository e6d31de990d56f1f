/*
 * oracle.c -- FP64 CPU restatement of portUrb's WENO-FV dycore step and TKE LES closure.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline); never linked into the
 * product.  Each function cites the reference lines it restates.  Every `float` of the reference
 * hot path is promoted to double here (the reference's reconstruction, Riemann solver,
 * hydrostatic arrays and coefs_to_tv are float; dynamics_rk.h:384-385,525-570,638-695,961-966,
 * TransformMatrices.h:105,429).
 *
 * Deliberate, documented differences from the as-written reference:
 *  - hydrostatic column means are accumulated in double with compensated (TwoSum) additions
 *    (reference: float running sums, dynamics_rk.h:961-975);
 *  - the immersed tau box scan (dynamics_rk.h:389-425) clamps indices into the halo'd array
 *    (the reference reads out of bounds for ord 5, whose halo is 2 but the scan radius is 3).
 */
#include "oracle.h"

/* s + x with the rounding error of the addition added to *c (Knuth's TwoSum) */
static void two_sum_add(double *s, double *c, double x) {
  double t = *s + x, bb = t - *s;
  *c += (*s - (t - bb)) + (x - bb);
  *s = t;
}

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXORD 11
#define MAXS   6

/* ----------------------------------------------------------------------------------------- */
/* Coefficient derivation (numerical, long double).                                          */
/* Definitions: helpers/sage/transformation_matrices.sage:19-27 (stencil -> coefs),           */
/* :140-146 with poly_utils.sage:23-26 (TV with degree truncation), WenoLimiter ideal weights.*/
/* ----------------------------------------------------------------------------------------- */

typedef struct {
  int ord, hs, N, S;
  double C[MAXS][MAXS][MAXS];     /* [m][n][i]: a_n of sub-stencil m = sum_i C s_{m+i} */
  double faceL[MAXS][MAXORD];     /* p_m(-1/2) weights over the full stencil */
  double faceR[MAXS][MAXORD];
  double idealL[MAXS], idealR[MAXS];
  double linL[MAXORD], linR[MAXORD];
  double tv[MAXS][MAXS];          /* upper triangular: TV = sum_{i<=j} tv[i][j] a_i a_j */
} weno_tab;

static weno_tab g_tabs[MAXORD + 1];
static int      g_tab_ready[MAXORD + 1];

static void ld_invert(int n, long double A[][MAXORD], long double Ainv[][MAXORD]) {
  long double M[MAXORD][2 * MAXORD];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) M[i][j] = j < n ? A[i][j] : (long double)(j - n == i);
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++) if (fabsl(M[r][c]) > fabsl(M[p][c])) p = r;
    for (int j = 0; j < 2 * n; j++) { long double t = M[c][j]; M[c][j] = M[p][j]; M[p][j] = t; }
    long double piv = M[c][c];
    for (int j = 0; j < 2 * n; j++) M[c][j] /= piv;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      long double f = M[r][c];
      for (int j = 0; j < 2 * n; j++) M[r][j] -= f * M[c][j];
    }
  }
  for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) Ainv[i][j] = M[i][j + n];
}

/* Polynomial coefficient map for cell averages at integer offsets `off` (unit cells). */
static void ld_poly_from_avgs(int n, const int *off, long double Cinv[][MAXORD]) {
  long double A[MAXORD][MAXORD];
  for (int r = 0; r < n; r++) {
    long double lo = off[r] - 0.5L, hi = off[r] + 0.5L;
    long double plo = lo, phi = hi;
    for (int c = 0; c < n; c++) {
      A[r][c] = (phi - plo) / (c + 1);
      plo *= lo; phi *= hi;
    }
  }
  ld_invert(n, A, Cinv);
}

static const weno_tab *get_tab(int ord) {
  if (ord < 3 || ord > MAXORD || (ord % 2) == 0) return NULL;
  if (g_tab_ready[ord]) return &g_tabs[ord];
  weno_tab *t = &g_tabs[ord];
  memset(t, 0, sizeof(*t));
  t->ord = ord; t->hs = (ord - 1) / 2; t->N = t->hs + 1; t->S = t->hs + 1;
  int N = t->N, S = t->S, hs = t->hs;
  long double faceL[MAXS][MAXORD] = {{0}}, faceR[MAXS][MAXORD] = {{0}};
  for (int m = 0; m < S; m++) {
    int off[MAXORD];
    for (int r = 0; r < N; r++) off[r] = m + r - hs;
    long double Ci[MAXORD][MAXORD];
    ld_poly_from_avgs(N, off, Ci);
    for (int n = 0; n < N; n++) for (int i = 0; i < N; i++) t->C[m][n][i] = (double)Ci[n][i];
    for (int i = 0; i < N; i++) {
      long double vl = 0, vr = 0, xl = 1, xr = 1;
      for (int n = 0; n < N; n++) { vl += Ci[n][i] * xl; vr += Ci[n][i] * xr; xl *= -0.5L; xr *= 0.5L; }
      faceL[m][m + i] = vl; faceR[m][m + i] = vr;
    }
  }
  /* full-stencil polynomial -> linear face values (dynamics_rk.h:539-545) */
  {
    int off[MAXORD];
    for (int r = 0; r < ord; r++) off[r] = r - hs;
    long double Ci[MAXORD][MAXORD];
    ld_poly_from_avgs(ord, off, Ci);
    long double linL[MAXORD], linR[MAXORD];
    for (int i = 0; i < ord; i++) {
      long double vl = 0, vr = 0, xl = 1, xr = 1;
      for (int n = 0; n < ord; n++) { vl += Ci[n][i] * xl; vr += Ci[n][i] * xr; xl *= -0.5L; xr *= 0.5L; }
      linL[i] = vl; linR[i] = vr;
      t->linL[i] = (double)vl; t->linR[i] = (double)vr;
    }
    /* ideal weights by forward substitution over the first S cells */
    long double dL[MAXS], dR[MAXS];
    for (int c = 0; c < S; c++) {
      long double aL = linL[c], aR = linR[c];
      for (int m = 0; m < c; m++) { aL -= dL[m] * faceL[m][c]; aR -= dR[m] * faceR[m][c]; }
      dL[c] = aL / faceL[c][c]; dR[c] = aR / faceR[c][c];
    }
    for (int m = 0; m < S; m++) { t->idealL[m] = (double)dL[m]; t->idealR[m] = (double)dR[m]; }
  }
  for (int m = 0; m < S; m++)
    for (int i = 0; i < ord; i++) { t->faceL[m][i] = (double)faceL[m][i]; t->faceR[m][i] = (double)faceR[m][i]; }
  /* TV quadratic form with truncation of each squared derivative to degree < N */
  {
    long double tv[MAXS][MAXS] = {{0}};
    for (int l = 1; l < N; l++) {
      for (int n1 = l; n1 < N; n1++) {
        for (int n2 = l; n2 < N; n2++) {
          int q = (n1 - l) + (n2 - l);
          if (q >= N || (q % 2)) continue;
          long double c1 = 1, c2 = 1;
          for (int f = n1 - l + 1; f <= n1; f++) c1 *= f;
          for (int f = n2 - l + 1; f <= n2; f++) c2 *= f;
          long double integ = 2.0L * powl(0.5L, q + 1) / (q + 1);
          int a = n1 < n2 ? n1 : n2, b = n1 < n2 ? n2 : n1;
          tv[a][b] += c1 * c2 * integ;
        }
      }
    }
    for (int a = 0; a < N; a++) for (int b = 0; b < N; b++) t->tv[a][b] = (double)tv[a][b];
  }
  g_tab_ready[ord] = 1;
  return t;
}

int oc_weno_tables(int ord, double *C, double *faceL, double *faceR, double *idealL,
                   double *idealR, double *linL, double *linR, double *tv) {
  const weno_tab *t = get_tab(ord);
  if (!t) return -1;
  int N = t->N, S = t->S;
  for (int m = 0; m < S; m++) for (int n = 0; n < N; n++) for (int i = 0; i < N; i++)
    C[(m * N + n) * N + i] = t->C[m][n][i];
  for (int m = 0; m < S; m++) for (int i = 0; i < ord; i++) {
    faceL[m * ord + i] = t->faceL[m][i]; faceR[m * ord + i] = t->faceR[m][i];
  }
  for (int m = 0; m < S; m++) { idealL[m] = t->idealL[m]; idealR[m] = t->idealR[m]; }
  for (int i = 0; i < ord; i++) { linL[i] = t->linL[i]; linR[i] = t->linR[i]; }
  for (int a = 0; a < N; a++) for (int b = 0; b < N; b++) tv[a * N + b] = t->tv[a][b];
  return 0;
}

/* ----------------------------------------------------------------------------------------- */
/* Limiter: WenoLimiter.h:10-96 (convexify, map, map_rs) and compute_limited_edges/gll_1/gll_2  */
/* (:116-180 ord 3, :186-260 ord 5, :264-345 ord 7, :349-437 ord 9, :441-537 ord 11).         */
/* ----------------------------------------------------------------------------------------- */

static void convexify(int n, double *w) {              /* WenoLimiter.h:10-55 */
  double tot = 0;
  for (int i = 0; i < n; i++) tot += w[i];
  if (tot > 1.e-20) for (int i = 0; i < n; i++) w[i] /= tot;
}

static double map_w(double w, double d) {              /* WenoLimiter.h:64 */
  return w * (d + d * d - 3 * d * w + w * w) / (d * d + (1 - 2 * d) * w);
}

static double map_rs(double w, double d) {             /* WenoLimiter.h:76-82, k=6 m=3 s=2000 */
  double term1 = w - d;
  for (int i = 1; i < 6; i++) term1 *= w - d;
  double term2 = w * (1 - w);
  for (int i = 1; i < 3; i++) term2 *= w * (1 - w);
  return d + (term1 * (w - d)) / (term1 + 2000. * term2);
}

/* Final (normalised) weights of one face: w_m = d_m/(TV_m^2 + 1e-20), convexified, mapped by
 * map_rs (map for ord 3) and convexified again when do_map (WenoLimiter.h gll_1/gll_2). */
static void face_weights(const weno_tab *t, const double *TV, const double *ideal, int do_map, double *w) {
  int S = t->S;
  for (int m = 0; m < S; m++) w[m] = ideal[m] / (TV[m] * TV[m] + 1.e-20);
  convexify(S, w);
  if (do_map) {
    for (int m = 0; m < S; m++) w[m] = t->ord == 3 ? map_w(w[m], ideal[m]) : map_rs(w[m], ideal[m]);
    convexify(S, w);
  }
}

double oc_weno_map_rs(double w, double d) { return map_rs(w, d); }

static double gll_face(const weno_tab *t, const double *TV, const double *ideal,
                       const double (*face)[MAXORD], int do_map, const double *s) {
  int S = t->S;
  double w[MAXS];
  face_weights(t, TV, ideal, do_map, w);
  double q = 0;
  for (int m = 0; m < S; m++) {
    double v = 0;
    for (int i = 0; i < t->ord; i++) v += face[m][i] * s[i];
    q += v * w[m];
  }
  return q;
}

/* Smoothness indicators of the S sub-stencils (coefs_to_tv of sten_to_coefs), convexified, with
 * the imm_L / imm_R replacement by the maximum (WenoLimiter.h compute_limited_edges). */
static void stencil_tv(const weno_tab *t, const double *s, int imm_L, int imm_R, double *TV) {
  int S = t->S, N = t->N;
  for (int m = 0; m < S; m++) {
    double a[MAXS];
    for (int n = 0; n < N; n++) {
      double v = 0;
      for (int i = 0; i < N; i++) v += t->C[m][n][i] * s[m + i];
      a[n] = v;
    }
    double tv = 0;
    for (int i = 1; i < N; i++) for (int j = i; j < N; j++) tv += t->tv[i][j] * a[i] * a[j];
    TV[m] = tv;
  }
  convexify(S, TV);
  if (imm_L) { double mx = TV[0]; for (int m = 1; m < S; m++) mx = fmax(mx, TV[m]); TV[S - 1] = mx; }
  if (imm_R) { double mx = TV[0]; for (int m = 1; m < S; m++) mx = fmax(mx, TV[m]); TV[0] = mx; }
}

static void limited_edges(const weno_tab *t, const double *s, int do_map, int imm_L, int imm_R,
                          double *qL, double *qR) {
  double TV[MAXS];
  stencil_tv(t, s, imm_L, imm_R, TV);
  *qL = gll_face(t, TV, t->idealL, (const double (*)[MAXORD])t->faceL, do_map, s);
  *qR = gll_face(t, TV, t->idealR, (const double (*)[MAXORD])t->faceR, do_map, s);
}

int oc_weno_face_weights(int ord, const double *stencil, int do_map, int imm_L, int imm_R,
                         double *wL, double *wR) {
  const weno_tab *t = get_tab(ord);
  if (!t) return -1;
  double TV[MAXS];
  stencil_tv(t, stencil, imm_L, imm_R, TV);
  face_weights(t, TV, t->idealL, do_map, wL);
  face_weights(t, TV, t->idealR, do_map, wR);
  return 0;
}

int oc_weno_limited_edges(int ord, const double *stencil, int do_map, int imm_L, int imm_R,
                          double *qL, double *qR) {
  const weno_tab *t = get_tab(ord);
  if (!t) return -1;
  limited_edges(t, stencil, do_map, imm_L, imm_R, qL, qR);
  return 0;
}

/* dynamics_rk.h:478-493 */
static void modify_stencil_immersed_der0(int ord, double *st, const unsigned char *imm) {
  int hs = (ord - 1) / 2;
  if (imm[hs]) return;
  for (int i2 = hs + 1; i2 < ord; i2++) {
    if (imm[i2]) { for (int i3 = i2; i3 < ord; i3++) st[i3] = st[i2 - 1]; break; }
  }
  for (int i2 = hs - 1; i2 >= 0; i2--) {
    if (imm[i2]) { for (int i3 = i2; i3 >= 0; i3--) st[i3] = st[i2 + 1]; break; }
  }
}

/* ----------------------------------------------------------------------------------------- */
/* Dycore                                                                                    */
/* ----------------------------------------------------------------------------------------- */

enum { idR = 0, idU = 1, idV = 2, idW = 3, idT = 4, idP = 5, NUM_STATE = 5 };

struct oc_dycore {
  oc_config cfg;
  int hs;
  double *hy_dens_cells, *hy_theta_cells, *hy_pressure_cells;   /* [nz+2hs] */
  double *hy_dens_edges, *hy_theta_edges, *hy_pressure_edges;   /* [nz+1]   */
  double *imm_halo;          /* dycore_immersed_proportion_halos [nz+2hs][ny+2hs][nx+2hs] */
  unsigned char *any10;      /* dycore_any_immersed10 [nz][ny][nx] */
  int *tau;                  /* dycore_immersed_tau [nz][ny][nx], built on first relax */
  int tau_ready;
};

#define H3(a, k, j, i) (a)[((size_t)(k) * (size_t)(ny + 2 * hs) + (size_t)(j)) * (size_t)(nx + 2 * hs) + (size_t)(i)]
#define I3(a, k, j, i) (a)[((size_t)(k) * (size_t)ny + (size_t)(j)) * (size_t)nx + (size_t)(i)]

static void set_threads(const oc_config *c) {
#ifdef _OPENMP
  if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#else
  (void)c;
#endif
}

double oc_dycore_compute_time_step(const oc_config *cfg) {     /* dynamics_rk.h:57-64 */
  double dx = cfg->xlen / cfg->nx, dy = cfg->ylen / cfg->ny, dz = cfg->zlen / cfg->nz;
  double maxwave = 350 + 100;
  return cfg->cfl * fmin(fmin(dx, dy), dz) / maxwave;
}

/* convert_coupler_to_dynamics (dynamics_rk.h:1304-1352): interior of halo'd state/tracers */
static void convert_c2d(const oc_config *c, int hs, const double *cp, double *state, double *tracers) {
  int nx = c->nx, ny = c->ny, nz = c->nz, ntr = c->num_tracers;
  size_t n3 = (size_t)nx * ny * nz;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
  const double *rho_d = cp, *uu = cp + n3, *vv = cp + 2 * n3, *ww = cp + 3 * n3, *tt = cp + 4 * n3;
  const double *trc = cp + 5 * n3;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    size_t id = ((size_t)k * ny + j) * nx + i;
    double rd = rho_d[id], u = uu[id], v = vv[id], w = ww[id], temp = tt[id];
    double rv = trc[(size_t)c->idWV * n3 + id];
    double press = rd * c->R_d * temp + rv * c->R_v * temp;
    double rho = rd;
    for (int tr = 0; tr < ntr; tr++) if (c->tracer_adds_mass[tr]) rho += trc[(size_t)tr * n3 + id];
    double theta = pow(press / c->C0, 1. / c->gamma_d) / rho;
    H3(state + idR * nh, hs + k, hs + j, hs + i) = rho;
    H3(state + idU * nh, hs + k, hs + j, hs + i) = rho * u;
    H3(state + idV * nh, hs + k, hs + j, hs + i) = rho * v;
    H3(state + idW * nh, hs + k, hs + j, hs + i) = rho * w;
    H3(state + idT * nh, hs + k, hs + j, hs + i) = rho * theta;
    for (int tr = 0; tr < ntr; tr++) H3(tracers + tr * nh, hs + k, hs + j, hs + i) = trc[(size_t)tr * n3 + id];
  }
}

/* convert_dynamics_to_coupler (dynamics_rk.h:1251-1299) */
static void convert_d2c(const oc_config *c, int hs, const double *state, const double *tracers, double *cp) {
  int nx = c->nx, ny = c->ny, nz = c->nz, ntr = c->num_tracers;
  size_t n3 = (size_t)nx * ny * nz;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    size_t id = ((size_t)k * ny + j) * nx + i;
    double rho = H3(state + idR * nh, hs + k, hs + j, hs + i);
    double u = H3(state + idU * nh, hs + k, hs + j, hs + i) / rho;
    double v = H3(state + idV * nh, hs + k, hs + j, hs + i) / rho;
    double w = H3(state + idW * nh, hs + k, hs + j, hs + i) / rho;
    double theta = H3(state + idT * nh, hs + k, hs + j, hs + i) / rho;
    double press = c->C0 * pow(rho * theta, c->gamma_d);
    double rho_v = H3(tracers + c->idWV * nh, hs + k, hs + j, hs + i);
    double rho_d = rho;
    for (int tr = 0; tr < ntr; tr++) if (c->tracer_adds_mass[tr]) rho_d -= H3(tracers + tr * nh, hs + k, hs + j, hs + i);
    double temp = press / (rho_d * c->R_d + rho_v * c->R_v);
    cp[id] = rho_d; cp[n3 + id] = u; cp[2 * n3 + id] = v; cp[3 * n3 + id] = w; cp[4 * n3 + id] = temp;
    for (int tr = 0; tr < ntr; tr++) cp[(5 + (size_t)tr) * n3 + id] = H3(tracers + tr * nh, hs + k, hs + j, hs + i);
  }
}

/* Periodic wrap of an index into [0,n). */
static inline int wrap(int i, int n) { i %= n; return i < 0 ? i + n : i; }

oc_dycore *oc_dycore_create(const oc_config *cfg, const double *coupler, const double *immersed_prop) {
  const weno_tab *t = get_tab(cfg->ord);
  if (!t || cfg->num_tracers > OC_MAX_TRACERS || cfg->idWV < 0) return NULL;
  set_threads(cfg);
  oc_dycore *d = (oc_dycore *)calloc(1, sizeof(oc_dycore));
  d->cfg = *cfg;
  int hs = d->hs = t->hs;
  int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz, ntr = cfg->num_tracers;
  double gamma = cfg->gamma_d, C0 = cfg->C0, grav = cfg->grav, dz = cfg->zlen / nz;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
  /* init (dynamics_rk.h:957-984): convert, then horizontal means per level (double sums) */
  double *state = (double *)calloc(NUM_STATE * nh, sizeof(double));
  double *tracers = (double *)calloc((size_t)(ntr > 0 ? ntr : 1) * nh, sizeof(double));
  convert_c2d(cfg, hs, coupler, state, tracers);
  int nzh = nz + 2 * hs;
  double *r = d->hy_dens_cells = (double *)calloc(nzh, sizeof(double));
  double *tt = d->hy_theta_cells = (double *)calloc(nzh, sizeof(double));
  double *p = d->hy_pressure_cells = (double *)calloc(nzh, sizeof(double));
#pragma omp parallel for
  for (int k = hs; k < hs + nz; k++) {
    /* compensated sums (TwoSum of each addition, the rounding errors summed apart): accurate to
     * about one rounding in any order, where a running sum of a level's equal densities drifts by
     * ~n eps */
    double sr = 0, st = 0, sp = 0, cr = 0, ct = 0, cp = 0;
    for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      double rho = H3(state + idR * nh, k, hs + j, hs + i), rt = H3(state + idT * nh, k, hs + j, hs + i);
      two_sum_add(&sr, &cr, rho);
      two_sum_add(&st, &ct, rt / rho);
      two_sum_add(&sp, &cp, C0 * pow(rt, gamma));
    }
    double rn = 1. / ((double)nx * (double)ny);
    r[k] = (sr + cr) * rn; tt[k] = (st + ct) * rn; p[k] = (sp + cp) * rn;
  }
  /* isentropic halo extrapolation (dynamics_rk.h:985-1008) */
  for (int kk = 0; kk < hs; kk++) {
    {
      int k0 = hs, k = k0 - 1 - kk;
      double rho0 = r[k0], theta0 = tt[k0];
      double rho0_gm1 = pow(rho0, gamma - 1), theta0_g = pow(theta0, gamma);
      r[k] = pow(rho0_gm1 + grav * (gamma - 1) * dz * (kk + 1) / (gamma * C0 * theta0_g), 1. / (gamma - 1));
      tt[k] = theta0;
      p[k] = C0 * pow(r[k] * theta0, gamma);
    }
    {
      int k0 = hs + nz - 1, k = k0 + 1 + kk;
      double rho0 = r[k0], theta0 = tt[k0];
      double rho0_gm1 = pow(rho0, gamma - 1), theta0_g = pow(theta0, gamma);
      r[k] = pow(rho0_gm1 - grav * (gamma - 1) * dz * (kk + 1) / (gamma * C0 * theta0_g), 1. / (gamma - 1));
      tt[k] = theta0;
      p[k] = C0 * pow(r[k] * theta0, gamma);
    }
  }
  /* hydrostatic edges (dynamics_rk.h:1145-1186) */
  d->hy_dens_edges = (double *)calloc(nz + 1, sizeof(double));
  d->hy_theta_edges = (double *)calloc(nz + 1, sizeof(double));
  d->hy_pressure_edges = (double *)calloc(nz + 1, sizeof(double));
  for (int k = 0; k <= nz; k++) {
    if (cfg->ord < 5) {
      d->hy_dens_edges[k] = exp(0.5 * log(r[hs + k - 1]) + 0.5 * log(r[hs + k]));
      d->hy_theta_edges[k] = 0.5 * tt[hs + k - 1] + 0.5 * tt[hs + k];
      d->hy_pressure_edges[k] = exp(0.5 * log(p[hs + k - 1]) + 0.5 * log(p[hs + k]));
    } else {
      d->hy_dens_edges[k] = exp(-1. / 12. * log(r[hs + k - 2]) + 7. / 12. * log(r[hs + k - 1]) +
                                7. / 12. * log(r[hs + k]) - 1. / 12. * log(r[hs + k + 1]));
      d->hy_theta_edges[k] = -1. / 12. * tt[hs + k - 2] + 7. / 12. * tt[hs + k - 1] +
                             7. / 12. * tt[hs + k] - 1. / 12. * tt[hs + k + 1];
      d->hy_pressure_edges[k] = exp(-1. / 12. * log(p[hs + k - 2]) + 7. / 12. * log(p[hs + k - 1]) +
                                    7. / 12. * log(p[hs + k]) - 1. / 12. * log(p[hs + k + 1]));
    }
  }
  /* dycore_immersed_proportion_halos: periodic x/y halos, z halos = 1 (dynamics_rk.h:1010-1034) */
  d->imm_halo = (double *)calloc(nh, sizeof(double));
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz + 2 * hs; k++) for (int j = 0; j < ny + 2 * hs; j++) for (int i = 0; i < nx + 2 * hs; i++) {
    double v;
    if (k < hs || k >= hs + nz) v = 1;
    else v = I3(immersed_prop, k - hs, wrap(j - hs, ny), wrap(i - hs, nx));
    H3(d->imm_halo, k, j, i) = v;
  }
  /* dycore_any_immersed10: any prop>0 within a 21^3 box, x/y periodic, z halos 0 (:1127-1143) */
  d->any10 = (unsigned char *)calloc((size_t)nx * ny * nz, 1);
  {
    int R = 10;
    /* separable: max over x, then y, then z of the (prop>0) indicator */
    unsigned char *a = (unsigned char *)calloc((size_t)nx * ny * nz, 1);
    unsigned char *b = (unsigned char *)calloc((size_t)nx * ny * nz, 1);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      unsigned char m = 0;
      for (int ii = -R; ii <= R; ii++) m |= I3(immersed_prop, k, j, wrap(i + ii, nx)) > 0;
      I3(a, k, j, i) = m;
    }
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      unsigned char m = 0;
      for (int jj = -R; jj <= R; jj++) m |= I3(a, k, wrap(j + jj, ny), i);
      I3(b, k, j, i) = m;
    }
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      unsigned char m = 0;
      for (int kk = -R; kk <= R; kk++) {
        int k2 = k + kk;
        if (k2 < 0 || k2 >= nz) continue;           /* z halos hold 0 */
        m |= I3(b, k2, j, i);
      }
      d->any10[((size_t)k * ny + j) * nx + i] = m;
    }
    free(a); free(b);
  }
  d->tau = (int *)calloc((size_t)nx * ny * nz, sizeof(int));
  free(state); free(tracers);
  return d;
}

void oc_dycore_destroy(oc_dycore *d) {
  if (!d) return;
  free(d->hy_dens_cells); free(d->hy_theta_cells); free(d->hy_pressure_cells);
  free(d->hy_dens_edges); free(d->hy_theta_edges); free(d->hy_pressure_edges);
  free(d->imm_halo); free(d->any10); free(d->tau);
  free(d);
}

int oc_dycore_hydrostatics(const oc_dycore *d, double *dc, double *tc, double *pc,
                           double *de, double *te, double *pe) {
  int nzh = d->cfg.nz + 2 * d->hs;
  memcpy(dc, d->hy_dens_cells, nzh * sizeof(double));
  memcpy(tc, d->hy_theta_cells, nzh * sizeof(double));
  memcpy(pc, d->hy_pressure_cells, nzh * sizeof(double));
  memcpy(de, d->hy_dens_edges, (d->cfg.nz + 1) * sizeof(double));
  memcpy(te, d->hy_theta_edges, (d->cfg.nz + 1) * sizeof(double));
  memcpy(pe, d->hy_pressure_edges, (d->cfg.nz + 1) * sizeof(double));
  return 0;
}

/* immersed tau (dynamics_rk.h:389-425), indices clamped into the halo'd array */
static void build_tau(oc_dycore *d) {
  int nx = d->cfg.nx, ny = d->cfg.ny, nz = d->cfg.nz, hs = d->hs;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    int ni[4] = {0, 0, 0, 0};
    for (int rad = 1; rad <= 3; rad++) {
      for (int kk = -rad; kk <= rad; kk++) for (int jj = -rad; jj <= rad; jj++) for (int ii = -rad; ii <= rad; ii++) {
        int k2 = hs + k + kk, j2 = hs + j + jj, i2 = hs + i + ii;
        if (k2 < 0) k2 = 0; if (k2 > nz + 2 * hs - 1) k2 = nz + 2 * hs - 1;
        if (j2 < 0) j2 = 0; if (j2 > ny + 2 * hs - 1) j2 = ny + 2 * hs - 1;
        if (i2 < 0) i2 = 0; if (i2 > nx + 2 * hs - 1) i2 = nx + 2 * hs - 1;
        if (H3(d->imm_halo, k2, j2, i2) == 0) ni[rad] = 1;
      }
    }
    int tau = ni[1] ? 8 : ni[2] ? 4 : ni[3] ? 2 : 1;
    d->tau[((size_t)k * ny + j) * nx + i] = tau;
  }
  d->tau_ready = 1;
}

int oc_dycore_immersed_masks(oc_dycore *d, unsigned char *any10, int *tau) {
  if (!d->tau_ready) build_tau(d);
  size_t n = (size_t)d->cfg.nx * d->cfg.ny * d->cfg.nz;
  memcpy(any10, d->any10, n);
  memcpy(tau, d->tau, n * sizeof(int));
  return 0;
}

/* enforce_immersed_boundaries (dynamics_rk.h:370-473) */
static void enforce_immersed(oc_dycore *d, double *state, double *tracers) {
  if (!d->tau_ready) build_tau(d);
  const oc_config *c = &d->cfg;
  int nx = c->nx, ny = c->ny, nz = c->nz, hs = d->hs, ntr = c->num_tracers;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    double mult = pow(H3(d->imm_halo, hs + k, hs + j, hs + i), c->immersed_power) /
                  d->tau[((size_t)k * ny + j) * nx + i];
    double *v;
    v = &H3(state + idR * nh, hs + k, hs + j, hs + i); *v = *v + (d->hy_dens_cells[hs + k] - *v) * mult;
    v = &H3(state + idU * nh, hs + k, hs + j, hs + i); *v = *v + (0 - *v) * mult;
    v = &H3(state + idV * nh, hs + k, hs + j, hs + i); *v = *v + (0 - *v) * mult;
    v = &H3(state + idW * nh, hs + k, hs + j, hs + i); *v = *v + (0 - *v) * mult;
    v = &H3(state + idT * nh, hs + k, hs + j, hs + i);
    *v = *v + (d->hy_dens_cells[hs + k] * d->hy_theta_cells[hs + k] - *v) * mult;
    for (int tr = 0; tr < ntr; tr++) {
      v = &H3(tracers + tr * nh, hs + k, hs + j, hs + i);
      *v = *v + (0 - *v) * mult;
      if (c->tracer_positive[tr]) *v = fmax(0., *v);
    }
  }
}

/* compute_tendencies (dynamics_rk.h:498-732) incl. halo_boundary_conditions (:736-804) and
 * edge_exchange (:808-912), single rank (periodic exchanges wrap onto itself). */
static void compute_tendencies(oc_dycore *d, const double *state, double *state_tend,
                               const double *tracers, double *tracers_tend) {
  const oc_config *c = &d->cfg;
  const weno_tab *t = get_tab(c->ord);
  int nx = c->nx, ny = c->ny, nz = c->nz, hs = d->hs, ntr = c->num_tracers, ord = c->ord;
  int nf = NUM_STATE + ntr + 1;
  int sim2d = ny == 1;
  double dx = c->xlen / nx, dy = c->ylen / ny, dz = c->zlen / nz;
  double r_dx = 1. / dx, r_dy = 1. / dy, r_dz = 1. / dz;
  double C0 = c->C0, gamma = c->gamma_d, grav = c->grav;
  double fcor = 2 * 7.2921e-5 * sin(c->latitude / 180 * M_PI);
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
  size_t n3 = (size_t)nx * ny * nz;
  const double *hyr = d->hy_dens_cells, *hyt = d->hy_theta_cells, *hyp = d->hy_pressure_cells;

  double *fl = (double *)calloc((size_t)nf * nh, sizeof(double));
#define FL(l, k, j, i) H3(fl + (size_t)(l) * nh, k, j, i)
  /* fields_loc (:550-558) */
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    FL(idP, hs + k, hs + j, hs + i) = C0 * pow(H3(state + idT * nh, hs + k, hs + j, hs + i), gamma) - hyp[hs + k];
    double r_r = 1. / H3(state + idR * nh, hs + k, hs + j, hs + i);
    FL(idR, hs + k, hs + j, hs + i) = H3(state + idR * nh, hs + k, hs + j, hs + i);
    for (int l = 1; l < NUM_STATE; l++) FL(l, hs + k, hs + j, hs + i) = H3(state + l * nh, hs + k, hs + j, hs + i) * r_r;
    for (int l = 0; l < ntr; l++) FL(NUM_STATE + 1 + l, hs + k, hs + j, hs + i) = H3(tracers + l * nh, hs + k, hs + j, hs + i) * r_r;
    FL(idR, hs + k, hs + j, hs + i) -= hyr[hs + k];
    FL(idT, hs + k, hs + j, hs + i) -= hyt[hs + k];
  }
  /* halo_exchange_x / _y (coupler.h:985-1021, 1128-1165): faces only, periodic self-exchange */
#pragma omp parallel for collapse(2)
  for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int ii = 0; ii < hs; ii++) {
    FL(l, hs + k, hs + j, ii) = FL(l, hs + k, hs + j, nx + ii);
    FL(l, hs + k, hs + j, nx + hs + ii) = FL(l, hs + k, hs + j, hs + ii);
  }
#pragma omp parallel for collapse(2)
  for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int jj = 0; jj < hs; jj++) for (int i = 0; i < nx; i++) {
    FL(l, hs + k, jj, hs + i) = FL(l, hs + k, ny + jj, hs + i);
    FL(l, hs + k, ny + hs + jj, hs + i) = FL(l, hs + k, hs + jj, hs + i);
  }
  /* halo_boundary_conditions (:736-804) */
  if (c->bc_x == OC_BC_PRECURSOR) {
    for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int ii = 0; ii < hs; ii++) {
      FL(l, hs + k, hs + j, ii) = FL(l, hs + k, hs + j, hs + 0);
      FL(l, hs + k, hs + j, hs + nx + ii) = FL(l, hs + k, hs + j, hs + nx - 1);
    }
  }
  if (c->bc_y == OC_BC_PRECURSOR) {
    for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int jj = 0; jj < hs; jj++) for (int i = 0; i < nx; i++) {
      FL(l, hs + k, jj, hs + i) = FL(l, hs + k, hs + 0, hs + i);
      FL(l, hs + k, hs + ny + jj, hs + i) = FL(l, hs + k, hs + ny - 1, hs + i);
    }
  }
  if (c->bc_z == OC_BC_SOLID_WALL) {
    for (int l = 0; l < nf; l++) for (int kk = 0; kk < hs; kk++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      if (l == idW || l == idT) {
        FL(l, kk, hs + j, hs + i) = 0;
        FL(l, hs + nz + kk, hs + j, hs + i) = 0;
      } else {
        FL(l, kk, hs + j, hs + i) = FL(l, hs + 0, hs + j, hs + i);
        FL(l, hs + nz + kk, hs + j, hs + i) = FL(l, hs + nz - 1, hs + j, hs + i);
      }
    }
  } else {
    for (int l = 0; l < nf; l++) for (int kk = 0; kk < hs; kk++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      FL(l, kk, hs + j, hs + i) = FL(l, nz + kk, hs + j, hs + i);
      FL(l, hs + nz + kk, hs + j, hs + i) = FL(l, hs + kk, hs + j, hs + i);
    }
  }

  /* limits arrays (:566-570) */
  size_t nlx = (size_t)nz * ny * (nx + 1), nly = (size_t)nz * (ny + 1) * nx, nlz = (size_t)(nz + 1) * ny * nx;
  double *lx = (double *)calloc(2 * (size_t)nf * nlx, sizeof(double));
  double *ly = (double *)calloc(2 * (size_t)nf * nly, sizeof(double));
  double *lz = (double *)calloc(2 * (size_t)nf * nlz, sizeof(double));
#define LX(s, l, k, j, i) lx[(((size_t)(s) * nf + (l)) * nz + (k)) * (size_t)ny * (nx + 1) + (size_t)(j) * (nx + 1) + (i)]
#define LY(s, l, k, j, i) ly[(((size_t)(s) * nf + (l)) * nz + (k)) * (size_t)(ny + 1) * nx + (size_t)(j) * nx + (i)]
#define LZ(s, l, k, j, i) lz[(((size_t)(s) * nf + (l)) * (nz + 1) + (k)) * (size_t)ny * nx + (size_t)(j) * nx + (i)]
#define IMMH(k, j, i) H3(d->imm_halo, k, j, i)

  /* WENO sweeps (:572-625) */
  for (int dir = 0; dir < 3; dir++) {
#pragma omp parallel for collapse(2)
    for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      double st[MAXORD];
      unsigned char imm[MAXORD];
      for (int s = 0; s < ord; s++) {
        int kk = hs + k, jj = hs + j, ii = hs + i;
        if (dir == 0) ii = i + s; else if (dir == 1) jj = j + s; else kk = k + s;
        imm[s] = IMMH(kk, jj, ii) > 0;
        st[s] = FL(l, kk, jj, ii);
      }
      int tl1 = dir == 0 ? idV : idU, tl2 = dir == 2 ? idV : idW;
      if (l == tl1 || l == tl2 || l == idP) modify_stencil_immersed_der0(ord, st, imm);
      int any = d->any10[((size_t)k * ny + j) * nx + i];
      double qL, qR;
      if (any || c->weno_all) {
        limited_edges(t, st, !any, imm[hs - 1], imm[hs + 1], &qL, &qR);
      } else {
        qL = 0; qR = 0;
        for (int s = 0; s < ord; s++) { qL += t->linL[s] * st[s]; qR += t->linR[s] * st[s]; }
      }
      if (dir == 0) { LX(1, l, k, j, i) = qL; LX(0, l, k, j, i + 1) = qR; }
      else if (dir == 1) { LY(1, l, k, j, i) = qL; LY(0, l, k, j + 1, i) = qR; }
      else { LZ(1, l, k, j, i) = qL; LZ(0, l, k + 1, j, i) = qR; }
    }
  }
  /* edge_exchange (:808-912) */
  for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) {
    double w0 = LX(1, l, k, j, 0), e0 = LX(0, l, k, j, nx);
    LX(0, l, k, j, 0) = e0;
    LX(1, l, k, j, nx) = w0;
  }
  for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) {
    double s0 = LY(1, l, k, 0, i), n0 = LY(0, l, k, ny, i);
    LY(0, l, k, 0, i) = n0;
    LY(1, l, k, ny, i) = s0;
  }
  if (c->bc_x == OC_BC_PRECURSOR) {
    for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) {
      LX(0, l, k, j, 0) = LX(1, l, k, j, 0);
      LX(1, l, k, j, nx) = LX(0, l, k, j, nx);
    }
  }
  if (c->bc_y == OC_BC_PRECURSOR) {
    for (int l = 0; l < nf; l++) for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) {
      LY(0, l, k, 0, i) = LY(1, l, k, 0, i);
      LY(1, l, k, ny, i) = LY(0, l, k, ny, i);
    }
  }
  if (c->bc_z == OC_BC_SOLID_WALL) {
    for (int l = 0; l < nf; l++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      if (l == idW || l == idT) {
        LZ(0, l, 0, j, i) = 0; LZ(1, l, 0, j, i) = 0; LZ(0, l, nz, j, i) = 0; LZ(1, l, nz, j, i) = 0;
      } else {
        LZ(0, l, 0, j, i) = LZ(1, l, 0, j, i);
        LZ(1, l, nz, j, i) = LZ(0, l, nz, j, i);
      }
    }
  } else {
    for (int l = 0; l < nf; l++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      LZ(0, l, 0, j, i) = LZ(0, l, nz, j, i);
      LZ(1, l, nz, j, i) = LZ(1, l, 0, j, i);
    }
  }

  /* Riemann (:636-707): acoustic upwinding with constant cs, fluxes in place into slot 0 */
  const double cs = 350.;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz + 1; k++) for (int j = 0; j < ny + 1; j++) for (int i = 0; i < nx + 1; i++) {
    if (j < ny && k < nz) {
      LX(0, idR, k, j, i) += hyr[hs + k]; LX(1, idR, k, j, i) += hyr[hs + k];
      LX(0, idT, k, j, i) += hyt[hs + k]; LX(1, idT, k, j, i) += hyt[hs + k];
      double r_L = LX(0, idR, k, j, i), r_R = LX(1, idR, k, j, i);
      double ru_L = LX(0, idU, k, j, i) * r_L, ru_R = LX(1, idU, k, j, i) * r_R;
      double p_L = LX(0, idP, k, j, i), p_R = LX(1, idP, k, j, i);
      double p_upw = 0.5 * (p_L + p_R - cs * (ru_R - ru_L));
      double ru_upw = 0.5 * (ru_L + ru_R - (p_R - p_L) / cs);
      int ind = ru_upw > 0 ? 0 : 1;
      LX(0, idR, k, j, i) = ru_upw;
      LX(0, idU, k, j, i) = ru_upw * LX(ind, idU, k, j, i) + p_upw;
      LX(0, idV, k, j, i) = ru_upw * LX(ind, idV, k, j, i);
      LX(0, idW, k, j, i) = ru_upw * LX(ind, idW, k, j, i);
      LX(0, idT, k, j, i) = ru_upw * LX(ind, idT, k, j, i);
      for (int tr = 0; tr < ntr; tr++) LX(0, NUM_STATE + tr, k, j, i) = ru_upw * LX(ind, NUM_STATE + 1 + tr, k, j, i);
    }
    if (i < nx && k < nz && !sim2d) {
      LY(0, idR, k, j, i) += hyr[hs + k]; LY(1, idR, k, j, i) += hyr[hs + k];
      LY(0, idT, k, j, i) += hyt[hs + k]; LY(1, idT, k, j, i) += hyt[hs + k];
      double r_L = LY(0, idR, k, j, i), r_R = LY(1, idR, k, j, i);
      double rv_L = LY(0, idV, k, j, i) * r_L, rv_R = LY(1, idV, k, j, i) * r_R;
      double p_L = LY(0, idP, k, j, i), p_R = LY(1, idP, k, j, i);
      double p_upw = 0.5 * (p_L + p_R - cs * (rv_R - rv_L));
      double rv_upw = 0.5 * (rv_L + rv_R - (p_R - p_L) / cs);
      int ind = rv_upw > 0 ? 0 : 1;
      LY(0, idR, k, j, i) = rv_upw;
      LY(0, idU, k, j, i) = rv_upw * LY(ind, idU, k, j, i);
      LY(0, idV, k, j, i) = rv_upw * LY(ind, idV, k, j, i) + p_upw;
      LY(0, idW, k, j, i) = rv_upw * LY(ind, idW, k, j, i);
      LY(0, idT, k, j, i) = rv_upw * LY(ind, idT, k, j, i);
      for (int tr = 0; tr < ntr; tr++) LY(0, NUM_STATE + tr, k, j, i) = rv_upw * LY(ind, NUM_STATE + 1 + tr, k, j, i);
    }
    if (i < nx && j < ny) {
      LZ(0, idR, k, j, i) += d->hy_dens_edges[k]; LZ(1, idR, k, j, i) += d->hy_dens_edges[k];
      LZ(0, idT, k, j, i) += d->hy_theta_edges[k]; LZ(1, idT, k, j, i) += d->hy_theta_edges[k];
      double r_L = LZ(0, idR, k, j, i), r_R = LZ(1, idR, k, j, i);
      double rw_L = LZ(0, idW, k, j, i) * r_L, rw_R = LZ(1, idW, k, j, i) * r_R;
      double p_L = LZ(0, idP, k, j, i), p_R = LZ(1, idP, k, j, i);
      double p_upw = 0.5 * (p_L + p_R - cs * (rw_R - rw_L));
      double rw_upw = 0.5 * (rw_L + rw_R - (p_R - p_L) / cs);
      int ind = rw_upw > 0 ? 0 : 1;
      LZ(0, idR, k, j, i) = rw_upw;
      LZ(0, idU, k, j, i) = rw_upw * LZ(ind, idU, k, j, i);
      LZ(0, idV, k, j, i) = rw_upw * LZ(ind, idV, k, j, i);
      LZ(0, idW, k, j, i) = rw_upw * LZ(ind, idW, k, j, i) + p_upw;
      LZ(0, idT, k, j, i) = rw_upw * LZ(ind, idT, k, j, i);
      for (int tr = 0; tr < ntr; tr++) LZ(0, NUM_STATE + tr, k, j, i) = rw_upw * LZ(ind, NUM_STATE + 1 + tr, k, j, i);
    }
  }
  /* divergence + gravity + Coriolis (:709-728) */
  int mx = NUM_STATE > ntr ? NUM_STATE : ntr;
#pragma omp parallel for collapse(2)
  for (int l = 0; l < mx; l++) for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    size_t id = ((size_t)k * ny + j) * nx + i;
    double ydiv_s = sim2d ? 0 : 1;
    if (l < NUM_STATE) {
      double tnd = -(LX(0, l, k, j, i + 1) - LX(0, l, k, j, i)) * r_dx
                   - ydiv_s * (LY(0, l, k, j + 1, i) - LY(0, l, k, j, i)) * r_dy
                   - (LZ(0, l, k + 1, j, i) - LZ(0, l, k, j, i)) * r_dz;
      if (l == idV && sim2d) tnd = 0;
      if (l == idW && c->enable_gravity) tnd += -grav * (H3(state + idR * nh, hs + k, hs + j, hs + i) - hyr[hs + k]);
      if (c->latitude != 0 && !sim2d && l == idU) tnd += fcor * H3(state + idV * nh, hs + k, hs + j, hs + i);
      if (c->latitude != 0 && !sim2d && l == idV) tnd -= fcor * H3(state + idU * nh, hs + k, hs + j, hs + i);
      state_tend[(size_t)l * n3 + id] = tnd;
    }
    if (l < ntr) {
      tracers_tend[(size_t)l * n3 + id] =
          -(LX(0, NUM_STATE + l, k, j, i + 1) - LX(0, NUM_STATE + l, k, j, i)) * r_dx
          - ydiv_s * (LY(0, NUM_STATE + l, k, j + 1, i) - LY(0, NUM_STATE + l, k, j, i)) * r_dy
          - (LZ(0, NUM_STATE + l, k + 1, j, i) - LZ(0, NUM_STATE + l, k, j, i)) * r_dz;
    }
  }
  free(fl); free(lx); free(ly); free(lz);
#undef FL
#undef LX
#undef LY
#undef LZ
#undef IMMH
}

/* time_step_rk_3_3 (dynamics_rk.h:279-367) */
static void ssprk3(oc_dycore *d, double *state, double *tracers, double dt) {
  const oc_config *c = &d->cfg;
  int nx = c->nx, ny = c->ny, nz = c->nz, hs = d->hs, ntr = c->num_tracers;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
  size_t n3 = (size_t)nx * ny * nz;
  double *state_tmp = (double *)calloc(NUM_STATE * nh, sizeof(double));
  double *tracers_tmp = (double *)calloc((size_t)(ntr > 0 ? ntr : 1) * nh, sizeof(double));
  double *state_tend = (double *)calloc(NUM_STATE * n3, sizeof(double));
  double *tracers_tend = (double *)calloc((size_t)(ntr > 0 ? ntr : 1) * n3, sizeof(double));
  int nv = NUM_STATE + ntr;

  enforce_immersed(d, state, tracers);
  for (int stage = 0; stage < 3; stage++) {
    const double *si = stage == 0 ? state : state_tmp;
    const double *ti = stage == 0 ? tracers : tracers_tmp;
    compute_tendencies(d, si, state_tend, ti, tracers_tend);
#pragma omp parallel for collapse(2)
    for (int l = 0; l < nv; l++) for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      size_t id = ((size_t)k * ny + j) * nx + i;
      int isstate = l < NUM_STATE;
      int ll = isstate ? l : l - NUM_STATE;
      double *un = isstate ? state + ll * nh : tracers + ll * nh;
      double *ut = isstate ? state_tmp + ll * nh : tracers_tmp + ll * nh;
      double tn = isstate ? state_tend[ll * n3 + id] : tracers_tend[ll * n3 + id];
      double *pn = &H3(un, hs + k, hs + j, hs + i), *pt = &H3(ut, hs + k, hs + j, hs + i);
      if (stage == 0) {
        *pt = *pn + dt * tn;
      } else if (stage == 1) {
        *pt = (3. / 4.) * *pn + (1. / 4.) * *pt + (1. / 4.) * dt * tn;
      } else {
        *pn = (1. / 3.) * *pn + (2. / 3.) * *pt + (2. / 3.) * dt * tn;
        if (!isstate && c->tracer_positive[ll]) *pn = fmax(0., *pn);
      }
    }
    if (stage < 2) enforce_immersed(d, state_tmp, tracers_tmp);
    else enforce_immersed(d, state, tracers);
  }
  free(state_tmp); free(tracers_tmp); free(state_tend); free(tracers_tend);
}

int oc_dycore_time_step(oc_dycore *d, double *coupler, double dt_phys) {
  const oc_config *c = &d->cfg;
  set_threads(c);
  int nx = c->nx, ny = c->ny, nz = c->nz, hs = d->hs, ntr = c->num_tracers;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
  double *state = (double *)calloc(NUM_STATE * nh, sizeof(double));
  double *tracers = (double *)calloc((size_t)(ntr > 0 ? ntr : 1) * nh, sizeof(double));
  convert_c2d(c, hs, coupler, state, tracers);
  double dt_dyn = oc_dycore_compute_time_step(c);
  int ncycles = (int)ceil(dt_phys / dt_dyn);
  dt_dyn = dt_phys / ncycles;
  for (int icycle = 0; icycle < ncycles; icycle++) ssprk3(d, state, tracers, dt_dyn);
  convert_d2c(c, hs, state, tracers, coupler);
  free(state); free(tracers);
  return 0;
}

int oc_dycore_tendencies(oc_dycore *d, const double *coupler, double *tend) {
  const oc_config *c = &d->cfg;
  set_threads(c);
  int nx = c->nx, ny = c->ny, nz = c->nz, hs = d->hs, ntr = c->num_tracers;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
  size_t n3 = (size_t)nx * ny * nz;
  double *state = (double *)calloc(NUM_STATE * nh, sizeof(double));
  double *tracers = (double *)calloc((size_t)(ntr > 0 ? ntr : 1) * nh, sizeof(double));
  convert_c2d(c, hs, coupler, state, tracers);
  compute_tendencies(d, state, tend, tracers, tend + NUM_STATE * n3);
  free(state); free(tracers);
  return 0;
}

/* The dycore's output fields (the write hook at dynamics_rk.h:1192-1233): density_pert =
 * rho - hy_dens_cells, theta_pert = rho theta / rho - hy_theta_cells, pressure_pert =
 * C0 (rho theta)^gamma - hy_pressure_cells, with the state from convert_coupler_to_dynamics. */
int oc_dycore_output_fields(const oc_dycore *d, const double *cp, double *dp, double *tp, double *pp) {
  const oc_config *c = &d->cfg;
  int nx = c->nx, ny = c->ny, nz = c->nz, hs = d->hs;
  size_t nh = (size_t)(nx + 2 * hs) * (ny + 2 * hs) * (nz + 2 * hs);
  double *state = (double *)calloc(5 * nh, sizeof(double));
  double *tracers = (double *)calloc((size_t)(c->num_tracers > 0 ? c->num_tracers : 1) * nh, sizeof(double));
  convert_c2d(c, hs, cp, state, tracers);
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    double r = H3(state + idR * nh, hs + k, hs + j, hs + i), rt = H3(state + idT * nh, hs + k, hs + j, hs + i);
    I3(dp, k, j, i) = r - d->hy_dens_cells[hs + k];
    I3(tp, k, j, i) = rt / r - d->hy_theta_cells[hs + k];
    I3(pp, k, j, i) = c->C0 * pow(rt, c->gamma_d) - d->hy_pressure_cells[hs + k];
  }
  free(state); free(tracers);
  return 0;
}

double oc_total_mass(const oc_config *c, const double *cp) {
  int nx = c->nx, ny = c->ny, nz = c->nz, ntr = c->num_tracers;
  size_t n3 = (size_t)nx * ny * nz;
  double dv = (c->xlen / nx) * (c->ylen / ny) * (c->zlen / nz);
  double s = 0;
  for (size_t id = 0; id < n3; id++) {
    double rho = cp[id];
    for (int tr = 0; tr < ntr; tr++) if (c->tracer_adds_mass[tr]) rho += cp[(5 + (size_t)tr) * n3 + id];
    s += rho;
  }
  return s * dv;
}

/* ----------------------------------------------------------------------------------------- */
/* LES closure (les_closure.h:30-536)                                                        */
/* ----------------------------------------------------------------------------------------- */

static int les_apply_impl(const oc_config *c, double *cp, const double *immersed_prop, double dtphys,
                          const oc_sgs_model *M, int bf16, float *tau_out, unsigned char *flags_out);

int oc_les_apply(const oc_config *c, double *cp, const double *immersed_prop, double dtphys) {
  return les_apply_impl(c, cp, immersed_prop, dtphys, NULL, 0, NULL, NULL);
}

int oc_les_apply_learned(const oc_config *c, double *cp, const double *immersed_prop, double dtphys,
                         const oc_sgs_model *M, int bf16, float *tau_out, unsigned char *flags_out) {
  if (!M) return -1;
  return les_apply_impl(c, cp, immersed_prop, dtphys, M, bf16, tau_out, flags_out);
}

/* The learned closure (no reference counterpart) replaces, on every face whose two cells both
 * hold the network's flag, the eddy-viscosity momentum fluxes by rho_face * avg(tau_i,dir) and
 * the theta flux by rho_face * avg(q_dir); TKE, tracers and all other faces keep
 * les_closure.h:92-229. */
static int les_apply_impl(const oc_config *c, double *cp, const double *immersed_prop, double dtphys,
                          const oc_sgs_model *M, int bf16, float *tau_out, unsigned char *flags_out) {
  set_threads(c);
  const int hs = 1;
  int nx = c->nx, ny = c->ny, nz = c->nz, ntr_all = c->num_tracers;
  double dx = c->xlen / nx, dy = c->ylen / ny, dz = c->zlen / nz;
  double grav = c->grav, nu = c->kinematic_viscosity, gamma = c->gamma_d, C0 = c->C0, R_d = c->R_d;
  int enable_gravity = c->enable_gravity, dns = c->dns;
  double delta = pow(dx * dy * dz, 1. / 3.);
  const double Pr = 0.7;
  double total_mult = c->les_total_mult, dissipation_mult = c->les_dissipation_mult, shear_prod_mult = c->les_shear_prod_mult;
  size_t n3 = (size_t)nx * ny * nz;
  size_t nh = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  if (c->idTKE < 0) return -1;
  int dtr[OC_MAX_TRACERS], ntr = 0;               /* diffused tracers, in tracer order */
  for (int tr = 0; tr < ntr_all; tr++) if (c->tracer_diffuse[tr]) dtr[ntr++] = tr;

  /* immersed_proportion_halos (sc_init.h:728-767): periodic x/y, z bottom halo 1, top 0 */
  double *imm = (double *)calloc(nh, sizeof(double));
  for (int k = 0; k < nz + 2; k++) for (int j = 0; j < ny + 2; j++) for (int i = 0; i < nx + 2; i++) {
    double v;
    if (k == 0) v = 1; else if (k == nz + 1) v = 0;
    else v = I3(immersed_prop, k - 1, wrap(j - 1, ny), wrap(i - 1, nx));
    H3(imm, k, j, i) = v;
  }
  double *state = (double *)calloc(5 * nh, sizeof(double));
  double *tracers = (double *)calloc((size_t)(ntr > 0 ? ntr : 1) * nh, sizeof(double));
  double *tke = (double *)calloc(nh, sizeof(double));
  /* convert_coupler_to_dynamics (:328-370) */
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    size_t id = ((size_t)k * ny + j) * nx + i;
    double rho_d = cp[id];
    H3(state + 0 * nh, hs + k, hs + j, hs + i) = rho_d;
    H3(state + 1 * nh, hs + k, hs + j, hs + i) = cp[n3 + id];
    H3(state + 2 * nh, hs + k, hs + j, hs + i) = cp[2 * n3 + id];
    H3(state + 3 * nh, hs + k, hs + j, hs + i) = cp[3 * n3 + id];
    H3(state + 4 * nh, hs + k, hs + j, hs + i) = pow(rho_d * R_d * cp[4 * n3 + id] / C0, 1. / gamma) / rho_d;
    H3(tke, hs + k, hs + j, hs + i) = fmax(0., cp[(5 + (size_t)c->idTKE) * n3 + id] / rho_d);
    for (int t = 0; t < ntr; t++) H3(tracers + t * nh, hs + k, hs + j, hs + i) = cp[(5 + (size_t)dtr[t]) * n3 + id] / rho_d;
  }
  /* halo_exchange (coupler.h:785-851): x over interior rows, then y over all i (corners filled) */
  int nfld = 5 + ntr + 1;
  for (int f = 0; f < nfld; f++) {
    double *a = f < 5 ? state + f * nh : f < 5 + ntr ? tracers + (f - 5) * nh : tke;
    for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) {
      H3(a, hs + k, hs + j, 0) = H3(a, hs + k, hs + j, nx);
      H3(a, hs + k, hs + j, nx + hs) = H3(a, hs + k, hs + j, hs);
    }
    for (int k = 0; k < nz; k++) for (int i = 0; i < nx + 2; i++) {
      H3(a, hs + k, 0, i) = H3(a, hs + k, ny, i);
      H3(a, hs + k, ny + hs, i) = H3(a, hs + k, hs, i);
    }
  }
  /* halo_bcs (:418-536) */
  double g = enable_gravity ? grav : 0;
  for (int f = 0; f < nfld; f++) {
    double *a = f < 5 ? state + f * nh : f < 5 + ntr ? tracers + (f - 5) * nh : tke;
    if (c->bc_x == OC_BC_PRECURSOR)
      for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) {
        H3(a, hs + k, hs + j, 0) = H3(a, hs + k, hs + j, hs);
        H3(a, hs + k, hs + j, hs + nx) = H3(a, hs + k, hs + j, hs + nx - 1);
      }
    if (c->bc_y == OC_BC_PRECURSOR)
      for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) {
        H3(a, hs + k, 0, hs + i) = H3(a, hs + k, hs, hs + i);
        H3(a, hs + k, hs + ny, hs + i) = H3(a, hs + k, hs + ny - 1, hs + i);
      }
  }
  if (c->bc_z == OC_BC_SOLID_WALL) {
    for (int j = 0; j < ny + 2; j++) for (int i = 0; i < nx + 2; i++) {
      int kk = 0;
      for (int f = 1; f < 5; f++) {
        double *a = state + f * nh;
        if (f == 3) { H3(a, kk, j, i) = 0; H3(a, hs + nz + kk, j, i) = 0; }
        else { H3(a, kk, j, i) = H3(a, hs, j, i); H3(a, hs + nz + kk, j, i) = H3(a, hs + nz - 1, j, i); }
      }
      H3(tke, kk, j, i) = H3(tke, hs, j, i); H3(tke, hs + nz + kk, j, i) = H3(tke, hs + nz - 1, j, i);
      for (int t = 0; t < ntr; t++) {
        double *a = tracers + t * nh;
        H3(a, kk, j, i) = H3(a, hs, j, i); H3(a, hs + nz + kk, j, i) = H3(a, hs + nz - 1, j, i);
      }
      {
        int k0 = hs, k = k0 - 1 - kk;
        double rho0 = H3(state, k0, j, i), theta0 = H3(state + 4 * nh, k0, j, i);
        double rho0_gm1 = pow(rho0, gamma - 1), theta0_g = pow(theta0, gamma);
        H3(state, k, j, i) = pow(rho0_gm1 + g * (gamma - 1) * dz * (kk + 1) / (gamma * C0 * theta0_g), 1. / (gamma - 1));
      }
      {
        int k0 = hs + nz - 1, k = k0 + 1 + kk;
        double rho0 = H3(state, k0, j, i), theta0 = H3(state + 4 * nh, k0, j, i);
        double rho0_gm1 = pow(rho0, gamma - 1), theta0_g = pow(theta0, gamma);
        H3(state, k, j, i) = pow(rho0_gm1 - g * (gamma - 1) * dz * (kk + 1) / (gamma * C0 * theta0_g), 1. / (gamma - 1));
      }
    }
  } else {
    for (int f = 0; f < nfld; f++) {
      double *a = f < 5 ? state + f * nh : f < 5 + ntr ? tracers + (f - 5) * nh : tke;
      for (int j = 0; j < ny + 2; j++) for (int i = 0; i < nx + 2; i++) {
        H3(a, 0, j, i) = H3(a, nz, j, i);
        H3(a, hs + nz, j, i) = H3(a, hs, j, i);
      }
    }
  }
#define S_(f, k, j, i) H3(state + (f) * nh, k, j, i)
#define K_(k, j, i) H3(tke, k, j, i)
#define Q_(t, k, j, i) H3(tracers + (size_t)(t) * nh, k, j, i)
#define IM(k, j, i) H3(imm, k, j, i)
  double visc_max_x = 0.25 * dx * dx / dtphys, visc_max_y = 0.25 * dy * dy / dtphys, visc_max_z = 0.25 * dz * dz / dtphys;
  size_t nfx = (size_t)nz * ny * (nx + 1), nfy = (size_t)nz * (ny + 1) * nx, nfz = (size_t)(nz + 1) * ny * nx;
  int nflx = 5 + ntr;     /* ru rv rw rt tke tracers */
  double *fx = (double *)calloc(nflx * nfx, sizeof(double));
  double *fy = (double *)calloc(nflx * nfy, sizeof(double));
  double *fz = (double *)calloc(nflx * nfz, sizeof(double));
#define FX(f, k, j, i) fx[(f) * nfx + ((size_t)(k) * ny + (j)) * (nx + 1) + (i)]
#define FY(f, k, j, i) fy[(f) * nfy + ((size_t)(k) * (ny + 1) + (j)) * nx + (i)]
#define FZ(f, k, j, i) fz[(f) * nfz + ((size_t)(k) * ny + (j)) * nx + (i)]
  /* face fluxes (:92-229) */
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz + 1; k++) for (int j = 0; j < ny + 1; j++) for (int i = 0; i < nx + 1; i++) {
    if (j < ny && k < nz) {
      if (IM(hs + k, hs + j, hs + i - 1) == 1 || IM(hs + k, hs + j, hs + i) == 1) {
        for (int f = 0; f < nflx; f++) FX(f, k, j, i) = 0;
      } else {
        double du_dz = 0.5 * ((S_(1, hs + k + 1, hs + j, hs + i - 1) - S_(1, hs + k - 1, hs + j, hs + i - 1)) / (2 * dz) +
                              (S_(1, hs + k + 1, hs + j, hs + i) - S_(1, hs + k - 1, hs + j, hs + i)) / (2 * dz));
        double dw_dz = 0.5 * ((S_(3, hs + k + 1, hs + j, hs + i - 1) - S_(3, hs + k - 1, hs + j, hs + i - 1)) / (2 * dz) +
                              (S_(3, hs + k + 1, hs + j, hs + i) - S_(3, hs + k - 1, hs + j, hs + i)) / (2 * dz));
        double dt_dz = 0.5 * ((S_(4, hs + k + 1, hs + j, hs + i - 1) - S_(4, hs + k - 1, hs + j, hs + i - 1)) / (2 * dz) +
                              (S_(4, hs + k + 1, hs + j, hs + i) - S_(4, hs + k - 1, hs + j, hs + i)) / (2 * dz));
        double du_dy = 0.5 * ((S_(1, hs + k, hs + j + 1, hs + i - 1) - S_(1, hs + k, hs + j - 1, hs + i - 1)) / (2 * dy) +
                              (S_(1, hs + k, hs + j + 1, hs + i) - S_(1, hs + k, hs + j - 1, hs + i)) / (2 * dy));
        double dv_dy = 0.5 * ((S_(2, hs + k, hs + j + 1, hs + i - 1) - S_(2, hs + k, hs + j - 1, hs + i - 1)) / (2 * dy) +
                              (S_(2, hs + k, hs + j + 1, hs + i) - S_(2, hs + k, hs + j - 1, hs + i)) / (2 * dy));
        double du_dx = (S_(1, hs + k, hs + j, hs + i) - S_(1, hs + k, hs + j, hs + i - 1)) / dx;
        double dv_dx = (S_(2, hs + k, hs + j, hs + i) - S_(2, hs + k, hs + j, hs + i - 1)) / dx;
        double dw_dx = (S_(3, hs + k, hs + j, hs + i) - S_(3, hs + k, hs + j, hs + i - 1)) / dx;
        double dt_dx = (S_(4, hs + k, hs + j, hs + i) - S_(4, hs + k, hs + j, hs + i - 1)) / dx;
        double dK_dx = (K_(hs + k, hs + j, hs + i) - K_(hs + k, hs + j, hs + i - 1)) / dx;
        double rho = 0.5 * (S_(0, hs + k, hs + j, hs + i - 1) + S_(0, hs + k, hs + j, hs + i));
        double K = 0.5 * (K_(hs + k, hs + j, hs + i - 1) + K_(hs + k, hs + j, hs + i));
        double tt = 0.5 * (S_(4, hs + k, hs + j, hs + i - 1) + S_(4, hs + k, hs + j, hs + i));
        double N = dt_dz >= 0 && enable_gravity ? sqrt(grav / tt * dt_dz) : 0;
        double ell = fmin(0.76 * sqrt(K) / (N + 1.e-20), delta);
        double km = total_mult * 0.1 * ell * sqrt(K);
        double Pr_t = delta / (1 + 2 * ell);
        double visc_tot = dns ? nu : fmin(km + nu, 0.5 * visc_max_x);
        double visc_tot_th = dns ? nu : fmin(km / Pr_t + nu / Pr, 0.5 * visc_max_x);
        FX(0, k, j, i) = -rho * visc_tot * (du_dx + du_dx - 2. / 3. * (du_dx + dv_dy + dw_dz));
        FX(1, k, j, i) = -rho * visc_tot * (dv_dx + du_dy);
        FX(2, k, j, i) = -rho * visc_tot * (dw_dx + du_dz);
        FX(3, k, j, i) = -rho * visc_tot_th * (dt_dx);
        FX(4, k, j, i) = -rho * visc_tot * 2 * (dK_dx);
        for (int t = 0; t < ntr; t++) {
          double dq = (Q_(t, hs + k, hs + j, hs + i) - Q_(t, hs + k, hs + j, hs + i - 1)) / dx;
          FX(5 + t, k, j, i) = -rho * visc_tot * dq;
        }
      }
    }
    if (i < nx && k < nz) {
      if (IM(hs + k, hs + j - 1, hs + i) == 1 || IM(hs + k, hs + j, hs + i) == 1) {
        for (int f = 0; f < nflx; f++) FY(f, k, j, i) = 0;
      } else {
        double dv_dz = 0.5 * ((S_(2, hs + k + 1, hs + j - 1, hs + i) - S_(2, hs + k - 1, hs + j - 1, hs + i)) / (2 * dz) +
                              (S_(2, hs + k + 1, hs + j, hs + i) - S_(2, hs + k - 1, hs + j, hs + i)) / (2 * dz));
        double dw_dz = 0.5 * ((S_(3, hs + k + 1, hs + j - 1, hs + i) - S_(3, hs + k - 1, hs + j - 1, hs + i)) / (2 * dz) +
                              (S_(3, hs + k + 1, hs + j, hs + i) - S_(3, hs + k - 1, hs + j, hs + i)) / (2 * dz));
        double dt_dz = 0.5 * ((S_(4, hs + k + 1, hs + j - 1, hs + i) - S_(4, hs + k - 1, hs + j - 1, hs + i)) / (2 * dz) +
                              (S_(4, hs + k + 1, hs + j, hs + i) - S_(4, hs + k - 1, hs + j, hs + i)) / (2 * dz));
        double du_dx = 0.5 * ((S_(1, hs + k, hs + j - 1, hs + i + 1) - S_(1, hs + k, hs + j - 1, hs + i - 1)) / (2 * dx) +
                              (S_(1, hs + k, hs + j, hs + i + 1) - S_(1, hs + k, hs + j, hs + i - 1)) / (2 * dx));
        double dv_dx = 0.5 * ((S_(2, hs + k, hs + j - 1, hs + i + 1) - S_(2, hs + k, hs + j - 1, hs + i - 1)) / (2 * dx) +
                              (S_(2, hs + k, hs + j, hs + i + 1) - S_(2, hs + k, hs + j, hs + i - 1)) / (2 * dx));
        double du_dy = (S_(1, hs + k, hs + j, hs + i) - S_(1, hs + k, hs + j - 1, hs + i)) / dy;
        double dv_dy = (S_(2, hs + k, hs + j, hs + i) - S_(2, hs + k, hs + j - 1, hs + i)) / dy;
        double dw_dy = (S_(3, hs + k, hs + j, hs + i) - S_(3, hs + k, hs + j - 1, hs + i)) / dy;
        double dt_dy = (S_(4, hs + k, hs + j, hs + i) - S_(4, hs + k, hs + j - 1, hs + i)) / dy;
        double dK_dy = (K_(hs + k, hs + j, hs + i) - K_(hs + k, hs + j - 1, hs + i)) / dy;
        double rho = 0.5 * (S_(0, hs + k, hs + j - 1, hs + i) + S_(0, hs + k, hs + j, hs + i));
        double K = 0.5 * (K_(hs + k, hs + j - 1, hs + i) + K_(hs + k, hs + j, hs + i));
        double tt = 0.5 * (S_(4, hs + k, hs + j - 1, hs + i) + S_(4, hs + k, hs + j, hs + i));
        double N = dt_dz >= 0 && enable_gravity ? sqrt(grav / tt * dt_dz) : 0;
        double ell = fmin(0.76 * sqrt(K) / (N + 1.e-20), delta);
        double km = total_mult * 0.1 * ell * sqrt(K);
        double Pr_t = delta / (1 + 2 * ell);
        double visc_tot = dns ? nu : fmin(km + nu, 0.5 * visc_max_y);
        double visc_tot_th = dns ? nu : fmin(km / Pr_t + nu / Pr, 0.5 * visc_max_y);
        FY(0, k, j, i) = -rho * visc_tot * (du_dy + dv_dx);
        FY(1, k, j, i) = -rho * visc_tot * (dv_dy + dv_dy - 2. / 3. * (du_dx + dv_dy + dw_dz));
        FY(2, k, j, i) = -rho * visc_tot * (dw_dy + dv_dz);
        FY(3, k, j, i) = -rho * visc_tot_th * (dt_dy);
        FY(4, k, j, i) = -rho * visc_tot * 2 * (dK_dy);
        for (int t = 0; t < ntr; t++) {
          double dq = (Q_(t, hs + k, hs + j, hs + i) - Q_(t, hs + k, hs + j - 1, hs + i)) / dy;
          FY(5 + t, k, j, i) = -rho * visc_tot * dq;
        }
      }
    }
    if (i < nx && j < ny) {
      if (IM(hs + k - 1, hs + j, hs + i) == 1 || IM(hs + k, hs + j, hs + i) == 1) {
        for (int f = 0; f < nflx; f++) FZ(f, k, j, i) = 0;
      } else {
        double du_dx = 0.5 * ((S_(1, hs + k - 1, hs + j, hs + i + 1) - S_(1, hs + k - 1, hs + j, hs + i - 1)) / (2 * dx) +
                              (S_(1, hs + k, hs + j, hs + i + 1) - S_(1, hs + k, hs + j, hs + i - 1)) / (2 * dx));
        double dw_dx = 0.5 * ((S_(3, hs + k - 1, hs + j, hs + i + 1) - S_(3, hs + k - 1, hs + j, hs + i - 1)) / (2 * dx) +
                              (S_(3, hs + k, hs + j, hs + i + 1) - S_(3, hs + k, hs + j, hs + i - 1)) / (2 * dx));
        double dv_dy = 0.5 * ((S_(2, hs + k - 1, hs + j + 1, hs + i) - S_(2, hs + k - 1, hs + j - 1, hs + i)) / (2 * dy) +
                              (S_(2, hs + k, hs + j + 1, hs + i) - S_(2, hs + k, hs + j - 1, hs + i)) / (2 * dy));
        double dw_dy = 0.5 * ((S_(3, hs + k - 1, hs + j + 1, hs + i) - S_(3, hs + k - 1, hs + j - 1, hs + i)) / (2 * dy) +
                              (S_(3, hs + k, hs + j + 1, hs + i) - S_(3, hs + k, hs + j - 1, hs + i)) / (2 * dy));
        double du_dz = (S_(1, hs + k, hs + j, hs + i) - S_(1, hs + k - 1, hs + j, hs + i)) / dz;
        double dv_dz = (S_(2, hs + k, hs + j, hs + i) - S_(2, hs + k - 1, hs + j, hs + i)) / dz;
        double dw_dz = (S_(3, hs + k, hs + j, hs + i) - S_(3, hs + k - 1, hs + j, hs + i)) / dz;
        double dt_dz = (S_(4, hs + k, hs + j, hs + i) - S_(4, hs + k - 1, hs + j, hs + i)) / dz;
        double dK_dz = (K_(hs + k, hs + j, hs + i) - K_(hs + k - 1, hs + j, hs + i)) / dz;
        double rho = 0.5 * (S_(0, hs + k - 1, hs + j, hs + i) + S_(0, hs + k, hs + j, hs + i));
        double K = 0.5 * (K_(hs + k - 1, hs + j, hs + i) + K_(hs + k, hs + j, hs + i));
        double tt = 0.5 * (S_(4, hs + k - 1, hs + j, hs + i) + S_(4, hs + k, hs + j, hs + i));
        double N = dt_dz >= 0 && enable_gravity ? sqrt(grav / tt * dt_dz) : 0;
        double ell = fmin(0.76 * sqrt(K) / (N + 1.e-20), delta);
        double km = total_mult * 0.1 * ell * sqrt(K);
        double Pr_t = delta / (1 + 2 * ell);
        double visc_tot = dns ? nu : fmin(km + nu, 0.5 * visc_max_z);
        double visc_tot_th = dns ? nu : fmin(km / Pr_t + nu / Pr, 0.5 * visc_max_z);
        FZ(0, k, j, i) = -rho * visc_tot * (du_dz + dw_dx);
        FZ(1, k, j, i) = -rho * visc_tot * (dv_dz + dw_dy);
        FZ(2, k, j, i) = -rho * visc_tot * (dw_dz + dw_dz - 2. / 3. * (du_dx + dv_dy + dw_dz));
        FZ(3, k, j, i) = -rho * visc_tot_th * (dt_dz);
        FZ(4, k, j, i) = -rho * visc_tot * 2 * (dK_dz);
        for (int t = 0; t < ntr; t++) {
          double dq = (Q_(t, hs + k, hs + j, hs + i) - Q_(t, hs + k - 1, hs + j, hs + i)) / dz;
          FZ(5 + t, k, j, i) = -rho * visc_tot * dq;
        }
      }
    }
  }
  float *tau = NULL;
  unsigned char *flg = NULL;
  if (M) {
    tau = (float *)malloc(sizeof(float) * 9 * n3);
    flg = (unsigned char *)malloc(n3);
    double nu_max = 0.5 * fmin(visc_max_x, fmin(visc_max_y, visc_max_z));
    oc_sgs_infer(M, state, imm, nx, ny, nz, dx, dy, dz, nu_max, bf16, tau, flg);
    /* tau component of momentum field f (0 u, 1 v, 2 w) on a face normal to direction d */
    static const int TIDX[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    int px = c->bc_x == OC_BC_PERIODIC, py = c->bc_y == OC_BC_PERIODIC, pz = c->bc_z == OC_BC_PERIODIC;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz + 1; k++) for (int j = 0; j < ny + 1; j++) for (int i = 0; i < nx + 1; i++) {
      for (int d = 0; d < 3; d++) {
        int ok = d == 0 ? (j < ny && k < nz) : d == 1 ? (i < nx && k < nz) : (i < nx && j < ny);
        if (!ok) continue;
        int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2), ic = i, jc = j, kc = k;
        int n_ = d == 0 ? nx : d == 1 ? ny : nz, per = d == 0 ? px : d == 1 ? py : pz;
        int lo = d == 0 ? im : d == 1 ? jm : km, hi = d == 0 ? ic : d == 1 ? jc : kc;
        if (lo < 0 || hi >= n_) {
          if (!per) continue;
          if (lo < 0) { if (d == 0) im = nx - 1; else if (d == 1) jm = ny - 1; else km = nz - 1; }
          if (hi >= n_) { if (d == 0) ic = 0; else if (d == 1) jc = 0; else kc = 0; }
        }
        size_t a = ((size_t)km * ny + jm) * nx + im, b = ((size_t)kc * ny + jc) * nx + ic;
        int both = flg[a] & flg[b];
        if (!both) continue;
        /* rho on the face, from the halo box as the eddy-viscosity flux uses it */
        double rho;
        if (d == 0) rho = 0.5 * (S_(0, hs + k, hs + j, hs + i - 1) + S_(0, hs + k, hs + j, hs + i));
        else if (d == 1) rho = 0.5 * (S_(0, hs + k, hs + j - 1, hs + i) + S_(0, hs + k, hs + j, hs + i));
        else rho = 0.5 * (S_(0, hs + k - 1, hs + j, hs + i) + S_(0, hs + k, hs + j, hs + i));
        for (int f = 0; f < 4; f++) {
          if (f < 3 && !(both & 1)) continue;
          if (f == 3 && !(both & 2)) continue;
          int t = f < 3 ? TIDX[f][d] : 6 + d;
          double v = rho * (0.5 * ((double)tau[t * n3 + a] + (double)tau[t * n3 + b]));
          if (d == 0) FX(f, k, j, i) = v; else if (d == 1) FY(f, k, j, i) = v; else FZ(f, k, j, i) = v;
        }
      }
    }
    if (tau_out) memcpy(tau_out, tau, sizeof(float) * 9 * n3);
    if (flags_out) memcpy(flags_out, flg, n3);
  }
  /* TKE source (:233-278) and forward-Euler update (:280-320), then convert back (:375-414) */
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    double rho = S_(0, hs + k, hs + j, hs + i);
    double K = K_(hs + k, hs + j, hs + i);
    double tt = S_(4, hs + k, hs + j, hs + i);
    double dt_dz = (S_(4, hs + k + 1, hs + j, hs + i) - S_(4, hs + k - 1, hs + j, hs + i)) / (2 * dz);
    double N = dt_dz >= 0 && enable_gravity ? sqrt(grav / tt * dt_dz) : 0;
    double ell = fmin(0.76 * sqrt(K) / (N + 1.e-20), delta);
    double km = 0.1 * ell * sqrt(K);
    double Pr_t = delta / (1 + 2 * ell);
    double src = 0;
    if (IM(hs + k, hs + j, hs + i) < 1) {
      if (enable_gravity) src += -(grav * rho * km) / (tt * Pr_t) * dt_dz;
      src -= dissipation_mult * rho * (0.19 + 0.51 * ell / delta) / delta * pow(K, 1.5);
      int im1 = IM(hs + k, hs + j, hs + i - 1) > 0 ? i : i - 1;
      int ip1 = IM(hs + k, hs + j, hs + i + 1) > 0 ? i : i + 1;
      int jm1 = IM(hs + k, hs + j - 1, hs + i) > 0 ? j : j - 1;
      int jp1 = IM(hs + k, hs + j + 1, hs + i) > 0 ? j : j + 1;
      int km1 = IM(hs + k - 1, hs + j, hs + i) > 0 ? k : k - 1;
      int kp1 = IM(hs + k + 1, hs + j, hs + i) > 0 ? k : k + 1;
      double du_dx = (S_(1, hs + k, hs + j, hs + i + 1) - S_(1, hs + k, hs + j, hs + i - 1)) / (2 * dx);
      double dv_dx = (S_(2, hs + k, hs + j, hs + ip1) - S_(2, hs + k, hs + j, hs + im1)) / (2 * dx);
      double dw_dx = (S_(3, hs + k, hs + j, hs + ip1) - S_(3, hs + k, hs + j, hs + im1)) / (2 * dx);
      double du_dy = (S_(1, hs + k, hs + jp1, hs + i) - S_(1, hs + k, hs + jm1, hs + i)) / (2 * dy);
      double dv_dy = (S_(2, hs + k, hs + j + 1, hs + i) - S_(2, hs + k, hs + j - 1, hs + i)) / (2 * dy);
      double dw_dy = (S_(3, hs + k, hs + jp1, hs + i) - S_(3, hs + k, hs + jm1, hs + i)) / (2 * dy);
      double du_dz = (S_(1, hs + kp1, hs + j, hs + i) - S_(1, hs + km1, hs + j, hs + i)) / (2 * dz);
      double dv_dz = (S_(2, hs + kp1, hs + j, hs + i) - S_(2, hs + km1, hs + j, hs + i)) / (2 * dz);
      double dw_dz = (S_(3, hs + k + 1, hs + j, hs + i) - S_(3, hs + k - 1, hs + j, hs + i)) / (2 * dz);
      double j1_i1 = (du_dx + du_dx) * du_dx, j1_i2 = (dv_dx + du_dy) * dv_dx, j1_i3 = (dw_dx + du_dz) * dw_dx;
      double j2_i1 = (du_dy + dv_dx) * du_dy, j2_i2 = (dv_dy + dv_dy) * dv_dy, j2_i3 = (dw_dy + dv_dz) * dw_dy;
      double j3_i1 = (du_dz + dw_dx) * du_dz, j3_i2 = (dv_dz + dw_dy) * dv_dz, j3_i3 = (dw_dz + dw_dz) * dw_dz;
      src += shear_prod_mult * rho * km * (j1_i1 + j1_i2 + j1_i3 + j2_i1 + j2_i2 + j2_i3 + j3_i1 + j3_i2 + j3_i3);
    }
    double tend[5];
    for (int f = 0; f < 5; f++)
      tend[f] = -(FX(f, k, j, i + 1) - FX(f, k, j, i)) / dx - (FY(f, k, j + 1, i) - FY(f, k, j, i)) / dy -
                (FZ(f, k + 1, j, i) - FZ(f, k, j, i)) / dz;
    tend[4] += src;
    size_t id = ((size_t)k * ny + j) * nx + i;
    double ru = S_(1, hs + k, hs + j, hs + i) * rho + dtphys * tend[0];
    double rv = S_(2, hs + k, hs + j, hs + i) * rho + dtphys * tend[1];
    double rw = S_(3, hs + k, hs + j, hs + i) * rho + dtphys * tend[2];
    double rt = S_(4, hs + k, hs + j, hs + i) * rho + dtphys * tend[3];
    double rk = fmax(0., K * rho + dtphys * tend[4]);
    cp[id] = rho;
    cp[n3 + id] = ru / rho;
    cp[2 * n3 + id] = rv / rho;
    cp[3 * n3 + id] = rw / rho;
    cp[4 * n3 + id] = C0 * pow(rt, gamma) / (rho * R_d);
    cp[(5 + (size_t)c->idTKE) * n3 + id] = rk;
    for (int t = 0; t < ntr; t++) {
      double tq = -(FX(5 + t, k, j, i + 1) - FX(5 + t, k, j, i)) / dx - (FY(5 + t, k, j + 1, i) - FY(5 + t, k, j, i)) / dy -
                  (FZ(5 + t, k + 1, j, i) - FZ(5 + t, k, j, i)) / dz;
      cp[(5 + (size_t)dtr[t]) * n3 + id] = Q_(t, hs + k, hs + j, hs + i) * rho + dtphys * tq;
    }
  }
#undef S_
#undef K_
#undef Q_
#undef IM
#undef FX
#undef FY
#undef FZ
  free(imm); free(state); free(tracers); free(tke); free(fx); free(fy); free(fz); free(tau); free(flg);
  return 0;
}
