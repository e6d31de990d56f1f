/* Exercises every entry point of the CPU oracle on small seeded states, for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py): out-of-bounds reads
 * and writes, use after free, signed overflow, misaligned or invalid shifts abort the run.
 * The cases cover every reconstruction order, the boundary kinds (periodic, precursor, solid wall,
 * periodic z, 2-D), buildings that reach the domain bottom, the LES closure with and without the
 * learned fluxes, surface fluxes, turbines and the column forcing.  Prints "OK" at the end. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.h"

static unsigned long long rng_state = 88172645463325252ull;
static double urand(void) {                     /* xorshift64, uniform [0, 1) */
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int fails = 0;
#define CHECK(c) do { if (!(c)) { printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); fails++; } } while (0)

static void base_config(oc_config *c, int nx, int ny, int nz, int ord, double dx) {
  memset(c, 0, sizeof(*c));
  c->nx = nx; c->ny = ny; c->nz = nz;
  c->xlen = nx * dx; c->ylen = ny * dx; c->zlen = nz * dx;
  c->ord = ord;
  c->num_tracers = 3;                    /* water_vapor, TKE, passive */
  c->idWV = 0; c->idTKE = 1;
  c->bc_x = OC_BC_PERIODIC; c->bc_y = OC_BC_PERIODIC; c->bc_z = OC_BC_SOLID_WALL;
  c->weno_all = 1; c->enable_gravity = 1; c->latitude = 0; c->cfl = 0.6; c->immersed_power = 5;
  c->R_d = 287.; c->R_v = 461.; c->cp_d = 1003.; c->grav = 9.81; c->p0 = 1.e5;
  c->gamma_d = c->cp_d / (c->cp_d - c->R_d);
  c->C0 = pow(c->R_d * pow(c->p0, -c->R_d / c->cp_d), c->gamma_d);
  c->tracer_positive[0] = 1; c->tracer_adds_mass[0] = 1; c->tracer_diffuse[0] = 1;
  c->tracer_positive[1] = 1;
  c->tracer_positive[2] = 1; c->tracer_diffuse[2] = 1;
  c->les_total_mult = 1; c->les_dissipation_mult = 1; c->les_shear_prod_mult = 1;
  c->nthreads = 1;
}

/* coupler state [5+ntr][nz][ny][nx]: isentropic background with noise, buildings on the bottom */
static void make_state(const oc_config *c, double *cp, double *imm, int buildings) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  size_t n3 = (size_t)nx * ny * nz;
  double dz = c->zlen / nz;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t id = ((size_t)k * ny + j) * nx + i;
        double z = (k + 0.5) * dz;
        double ex = 1. - c->grav * z / (c->cp_d * 300.);
        double p = c->p0 * pow(ex, c->cp_d / c->R_d);
        double rho = pow(p / c->C0, 1. / c->gamma_d) / 300. * (1 + 0.002 * (urand() - 0.5));
        double theta = 300. + (urand() - 0.5);
        cp[0 * n3 + id] = rho;
        cp[1 * n3 + id] = 3. + 2. * (urand() - 0.5);
        cp[2 * n3 + id] = 2. * (urand() - 0.5);
        cp[3 * n3 + id] = (urand() - 0.5);
        double rv = rho * 1e-3 * urand();
        cp[5 * n3 + id] = rv;
        cp[6 * n3 + id] = rho * (0.1 + 0.2 * urand());
        cp[7 * n3 + id] = rho * urand();
        double press = c->C0 * pow((rho + rv) * theta, c->gamma_d);
        cp[4 * n3 + id] = press / (rho * c->R_d + rv * c->R_v);
        imm[id] = 0;
      }
  for (int b = 0; b < buildings; b++) {
    int i0 = (int)(urand() * (nx - 2)), j0 = ny > 2 ? (int)(urand() * (ny - 2)) : 0, h = 1 + (int)(urand() * (nz / 3));
    for (int k = 0; k < h; k++)
      for (int j = j0; j < j0 + 2 && j < ny; j++)
        for (int i = i0; i < i0 + 2; i++) imm[((size_t)k * ny + j) * nx + i] = 1.;
  }
  imm[(size_t)(ny / 2) * nx + nx / 2] = 0.5;
}

static int all_finite(const double *a, size_t n) {
  for (size_t i = 0; i < n; i++) if (!isfinite(a[i])) return 0;
  return 1;
}

static void dycore_case(int nx, int ny, int nz, int ord, int bc_x, int bc_y, int bc_z, int weno_all, double lat,
                        int buildings) {
  oc_config c;
  base_config(&c, nx, ny, nz, ord, 40.);
  c.bc_x = bc_x; c.bc_y = bc_y; c.bc_z = bc_z; c.weno_all = weno_all; c.latitude = lat;
  size_t n3 = (size_t)nx * ny * nz, nv = 5 + c.num_tracers;
  double *cp = malloc(nv * n3 * sizeof(double)), *imm = malloc(n3 * sizeof(double));
  double *tend = malloc(nv * n3 * sizeof(double));
  make_state(&c, cp, imm, buildings);
  oc_dycore *d = oc_dycore_create(&c, cp, imm);
  CHECK(d != NULL);
  if (!d) { free(cp); free(imm); free(tend); return; }
  int hs = (ord - 1) / 2;
  double *hy = malloc((3 * (nz + 2 * hs) + 3 * (nz + 1)) * sizeof(double));
  CHECK(oc_dycore_hydrostatics(d, hy, hy + (nz + 2 * hs), hy + 2 * (nz + 2 * hs), hy + 3 * (nz + 2 * hs),
                               hy + 3 * (nz + 2 * hs) + (nz + 1), hy + 3 * (nz + 2 * hs) + 2 * (nz + 1)) == 0);
  unsigned char *any10 = malloc(n3);
  int *tau = malloc(n3 * sizeof(int));
  CHECK(oc_dycore_immersed_masks(d, any10, tau) == 0);
  CHECK(oc_dycore_tendencies(d, cp, tend) == 0);
  CHECK(all_finite(tend, nv * n3));
  double dt = oc_dycore_compute_time_step(&c);
  CHECK(oc_dycore_time_step(d, cp, 2.5 * dt) == 0);      /* three cycles */
  CHECK(all_finite(cp, nv * n3));
  double *pert = malloc(3 * n3 * sizeof(double));
  CHECK(oc_dycore_output_fields(d, cp, pert, pert + n3, pert + 2 * n3) == 0);
  CHECK(oc_total_mass(&c, cp) > 0);
  CHECK(oc_les_apply(&c, cp, imm, dt) == 0);
  CHECK(all_finite(cp, nv * n3));
  oc_dycore_destroy(d);
  free(cp); free(imm); free(tend); free(hy); free(any10); free(tau); free(pert);
}

static void limiter_cases(void) {
  for (int ord = 3; ord <= 11; ord += 2) {
    int S = (ord + 1) / 2, N = S;
    double *C = malloc((size_t)S * N * N * sizeof(double)), *fL = malloc((size_t)S * ord * sizeof(double));
    double *fR = malloc((size_t)S * ord * sizeof(double)), dL[8], dR[8], lL[16], lR[16];
    double *tv = malloc((size_t)N * N * sizeof(double));
    CHECK(oc_weno_tables(ord, C, fL, fR, dL, dR, lL, lR, tv) == 0);
    double s[16], qL, qR, wL[8], wR[8];
    for (int t = 0; t < 50; t++) {
      for (int i = 0; i < ord; i++) s[i] = t % 5 == 0 ? 1.0 : urand() * (t % 2 ? 1e-12 : 10.);
      for (int m = 0; m < 4; m++) {
        CHECK(oc_weno_limited_edges(ord, s, m & 1, (m >> 1) & 1, 0, &qL, &qR) == 0);
        CHECK(oc_weno_face_weights(ord, s, m & 1, 0, (m >> 1) & 1, wL, wR) == 0);
        CHECK(isfinite(qL) && isfinite(qR));
      }
    }
    free(C); free(fL); free(fR); free(tv);
  }
  CHECK(oc_weno_map_rs(0.3, 0.3) == 0.3);
}

static void physics_cases(void) {
  oc_config c;
  int nx = 24, ny = 16, nz = 20;
  base_config(&c, nx, ny, nz, 5, 5.);
  size_t n3 = (size_t)nx * ny * nz, nv = 5 + c.num_tracers;
  double *cp = malloc(nv * n3 * sizeof(double)), *imm = malloc(n3 * sizeof(double));
  make_state(&c, cp, imm, 3);
  /* surface fluxes */
  double *rough = malloc(n3 * sizeof(double)), *temp = malloc(n3 * sizeof(double));
  double *srough = malloc((size_t)nx * ny * sizeof(double)), *stemp = malloc((size_t)nx * ny * sizeof(double));
  for (size_t i = 0; i < n3; i++) { rough[i] = 0.1; temp[i] = urand() < 0.5 ? 0 : 290. + 10 * urand(); }
  for (int i = 0; i < nx * ny; i++) { srough[i] = 0.1; stemp[i] = 295.; }
  CHECK(oc_surface_fluxes(&c, cp, imm, rough, temp, srough, stemp, 0.1) == 0);
  /* column forcing */
  oc_pg_state pg = {0, 0, 0, 0, 0};
  CHECK(oc_sponge_layer(&c, cp, 0.1, 10., 0.2) == 0);
  CHECK(oc_geostrophic_forcing(&c, cp, 0.1, 40., 5., 1.) == 0);
  CHECK(oc_uniform_pg_forcing(&c, cp, 0.1, 50., 2, &pg) == 0);
  CHECK(oc_uniform_pg_forcing(&c, cp, 0.1, 50., 2, &pg) == 0);
  CHECK(oc_uniform_pg_forcing_height(&c, cp, 0.1, 30., 4., 1., 50.) == 0);
  /* turbines: two disks, one with its sampling box clipped by the domain */
  double vm[3] = {0., 8., 20.}, ct[3] = {0.8, 0.75, 0.5}, cpw[3] = {0.4, 0.45, 0.3}, pw[3] = {0, 1e-4, 2e-4};
  oc_turbine_ref ref = {3, vm, ct, cpw, pw, 50., 20., 0.01, -2., 0.6, 0.05};
  oc_turbine tu[2];
  memset(tu, 0, sizeof(tu));
  tu[0].base_x = 30.; tu[0].base_y = 40.; tu[0].apply_thrust = 1;
  tu[1].base_x = 90.; tu[1].base_y = 40.; tu[1].apply_thrust = 1; tu[1].yaw = 0.3;
  oc_turbine_opts topt = {0, 0, 0, 0.};
  double *pwt = malloc(n3 * sizeof(double)), *swt = malloc(n3 * sizeof(double));
  CHECK(oc_turbines_apply(&c, cp, &ref, tu, 2, &topt, 0.1, pwt, swt) == 0);
  CHECK(oc_turbines_apply(&c, cp, &ref, tu, 2, &topt, 0.1, NULL, NULL) == 0);
  CHECK(isfinite(oc_turbine_interp(vm, ct, 3, 25.)) && oc_turbine_blade_weight(5., 20.) > 0);
  CHECK(isfinite(oc_turbine_yaw_tend(&tu[0], 5., 1., 0.1, 0.01)));
  /* learned closure: a random network */
  float *P = malloc(OC_MLP_NPARAM * sizeof(float)), *G32 = malloc(OC_MLP_NPARAM * sizeof(float));
  float *m = calloc(OC_MLP_NPARAM, sizeof(float)), *v = calloc(OC_MLP_NPARAM, sizeof(float));
  double *G = malloc(OC_MLP_NPARAM * sizeof(double));
  for (int i = 0; i < OC_MLP_NPARAM; i++) P[i] = (float)(0.2 * (urand() - 0.5));
  long ns = 40;
  float *X = malloc(ns * OC_MLP_NIN * sizeof(float)), *T = malloc(ns * OC_MLP_NOUT * sizeof(float));
  float *Y = malloc(ns * OC_MLP_NOUT * sizeof(float));
  for (long i = 0; i < ns * OC_MLP_NIN; i++) X[i] = (float)(urand() - 0.5);
  for (long i = 0; i < ns * OC_MLP_NOUT; i++) T[i] = (float)(urand() - 0.5);
  oc_mlp_forward(P, X, ns, 1, Y);
  CHECK(isfinite(oc_mlp_loss_grad(P, X, T, ns, 1, G)));
  for (int i = 0; i < OC_MLP_NPARAM; i++) G32[i] = (float)G[i];
  oc_mlp_adam(P, G32, m, v, OC_MLP_NPARAM, 1, 1e-3f, 0.9f, 0.999f, 1e-8f);
  CHECK(oc_relu(0.3) == 0.3 && oc_relu_deriv(-0.3) == 0. && oc_bf16_round(1.0f) == 1.0f);
  int r = 2, cx = nx / r, cy = ny / r, cz = nz / r;
  size_t nb = (size_t)(cx + 2) * (cy + 2) * (cz + 2);
  double *les = malloc(6 * nb * sizeof(double)), *imm_c = malloc(nb * sizeof(double));
  float *tg = malloc(9 * (size_t)cx * cy * cz * sizeof(float));
  CHECK(oc_sgs_filter(&c, cp, imm, r, les, imm_c, tg) == 0);
  float fs[OC_MLP_NIN], fmul[OC_MLP_NIN], om[9], os[9];
  for (int i = 0; i < OC_MLP_NIN; i++) fs[i] = 1.f;
  for (int i = 0; i < 9; i++) { om[i] = 0.f; os[i] = 0.01f; }
  oc_sgs_fmul(fs, 10., fmul);
  long cells[3] = {0, 5, (long)cx * cy * cz - 1};
  float *Xc = malloc(3 * OC_MLP_NIN * sizeof(float));
  CHECK(oc_sgs_features(les, cx, cy, cz, fmul, cells, 3, Xc) == 0);
  oc_sgs_model M = {P, fs, 4.f, om, os};
  float *tau = malloc(9 * n3 * sizeof(float));
  unsigned char *fl = malloc(n3);
  double *les_f = malloc(6 * (size_t)(nx + 2) * (ny + 2) * (nz + 2) * sizeof(double));
  double *imm_f = malloc((size_t)(nx + 2) * (ny + 2) * (nz + 2) * sizeof(double));
  float *tg1 = malloc(9 * n3 * sizeof(float));
  CHECK(oc_sgs_filter(&c, cp, imm, 1, les_f, imm_f, tg1) == 0);
  oc_sgs_infer(&M, les_f, imm_f, nx, ny, nz, 5., 5., 5., 10., 1, tau, fl);
  CHECK(oc_les_apply_learned(&c, cp, imm, 0.05, &M, 1, tau, fl) == 0);
  CHECK(all_finite(cp, nv * n3));
  free(cp); free(imm); free(rough); free(temp); free(srough); free(stemp); free(pwt); free(swt);
  free(P); free(G32); free(m); free(v); free(G); free(X); free(T); free(Y); free(les); free(imm_c); free(tg);
  free(Xc); free(tau); free(fl); free(les_f); free(imm_f); free(tg1);
}

int main(void) {
  limiter_cases();
  const int ords[5] = {3, 5, 7, 9, 11};
  for (int o = 0; o < 5; o++) {
    int ord = ords[o], hs = (ord - 1) / 2;
    dycore_case(14, 12, 3 * hs + 4, ord, OC_BC_PERIODIC, OC_BC_PERIODIC, OC_BC_SOLID_WALL, 1, 0., 3);
  }
  dycore_case(16, 12, 12, 5, OC_BC_PRECURSOR, OC_BC_PERIODIC, OC_BC_SOLID_WALL, 1, 35., 2);
  dycore_case(16, 12, 14, 9, OC_BC_PERIODIC, OC_BC_PRECURSOR, OC_BC_PERIODIC, 0, 0., 2);
  dycore_case(20, 1, 12, 5, OC_BC_PERIODIC, OC_BC_PERIODIC, OC_BC_SOLID_WALL, 1, 20., 2);    /* 2-D */
  physics_cases();
  if (fails) { printf("%d failures\n", fails); return 1; }
  printf("OK\n");
  return 0;
}
