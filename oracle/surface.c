/*
 * surface.c -- CPU oracle of modules::apply_surface_fluxes (model/modules/surface_flux.h:9-126).
 * TEST INFRASTRUCTURE ONLY.
 *
 * Halo'd boxes as custom_modules::sc_init builds them (experiments/examples/custom_modules/
 * sc_init.h:728-767): immersed proportion / roughness / temperature exchanged periodically in x and
 * y (create_and_exchange_halos, corners included), the bottom halo = 1 / surface roughness /
 * surface temperature, the top halo = 0.  Per cell, for each of the six neighbours with immersed
 * proportion > 0: c = 0.4^2 / ln^2((delta/2 + z0)/z0) with delta the spacing across that face; the
 * two tangential velocities get -c u_t |u_t,pair| / delta, the temperature -c (T - T0) |.| / delta
 * when T0 != 0 (surface_flux.h:45-120); then u, v, w, T += dt * tendency (:122-127).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

static int wrp(int i, int n) { i %= n; return i < 0 ? i + n : i; }

int oc_surface_fluxes(const oc_config *c, double *cp, const double *prop, const double *rough, const double *temp,
                      const double *sfc_rough, const double *sfc_temp, double dt) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  size_t n3 = (size_t)nx * ny * nz, nxh = nx + 2, nyh = ny + 2, nh = nxh * nyh * (size_t)(nz + 2);
  double dx = c->xlen / nx, dy = c->ylen / ny, dz = c->zlen / nz;
  double *P = (double *)malloc(sizeof(double) * nh), *R = (double *)malloc(sizeof(double) * nh),
         *T = (double *)malloc(sizeof(double) * nh);
  for (size_t kh = 0; kh < (size_t)nz + 2; kh++) for (size_t jh = 0; jh < nyh; jh++) for (size_t ih = 0; ih < nxh; ih++) {
    int i = wrp((int)ih - 1, nx), j = wrp((int)jh - 1, ny), k = (int)kh - 1;
    size_t h = (kh * nyh + jh) * nxh + ih;
    if (k < 0) { P[h] = 1; R[h] = sfc_rough[(size_t)j * nx + i]; T[h] = sfc_temp[(size_t)j * nx + i]; }
    else if (k >= nz) { P[h] = 0; R[h] = 0; T[h] = 0; }
    else { size_t id = ((size_t)k * ny + j) * nx + i; P[h] = prop[id]; R[h] = rough[id]; T[h] = temp[id]; }
  }
  double *tend = (double *)calloc(4 * n3, sizeof(double));
  const double vk = 0.40;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
    size_t id = ((size_t)k * ny + j) * nx + i;
    double u = cp[n3 + id], v = cp[2 * n3 + id], w = cp[3 * n3 + id], t = cp[4 * n3 + id];
    double tu = 0, tv = 0, tw = 0, tt = 0;
    for (int f = 0; f < 6; f++) {
      int d = f / 2, s = (f % 2) ? 1 : -1;
      size_t nb = ((size_t)(k + 1 + (d == 2 ? s : 0)) * nyh + (j + 1 + (d == 1 ? s : 0))) * nxh + (i + 1 + (d == 0 ? s : 0));
      if (!(P[nb] > 0)) continue;
      double delta = d == 0 ? dx : (d == 1 ? dy : dz);
      double z0 = R[nb];
      double lg = log((delta / 2 + z0) / z0);
      double cd = vk * vk / (lg * lg);
      double T0 = T[nb];
      double a = d == 0 ? v : u, b = d == 2 ? v : w;         /* the two tangential velocities */
      double mag = sqrt(a * a + b * b);
      double ta = -cd * (a - 0) * mag / delta, tb = -cd * (b - 0) * mag / delta;
      if (d == 0) { tv += ta; tw += tb; }
      else if (d == 1) { tu += ta; tw += tb; }
      else { tu += ta; tv += tb; }
      if (T0 != 0) tt += -cd * (t - T0) * mag / delta;
    }
    tend[id] = tu; tend[n3 + id] = tv; tend[2 * n3 + id] = tw; tend[3 * n3 + id] = tt;
  }
  for (size_t id = 0; id < n3; id++) {
    cp[n3 + id] += dt * tend[id];
    cp[2 * n3 + id] += dt * tend[n3 + id];
    cp[3 * n3 + id] += dt * tend[2 * n3 + id];
    cp[4 * n3 + id] += dt * tend[3 * n3 + id];
  }
  free(P); free(R); free(T); free(tend);
  return 0;
}
