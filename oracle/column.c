/*
 * column.c -- CPU oracle of the column-mean forcing modules.  TEST INFRASTRUCTURE ONLY.
 *
 *   oc_sponge_layer            modules::sponge_layer            (model/modules/sponge_layer.h:7-60)
 *   oc_geostrophic_forcing     modules::geostrophic_wind_forcing
 *                                                   (model/modules/geostrophic_wind_forcing.h:7-36)
 *   oc_uniform_pg_forcing      modules::uniform_pg_wind_forcing (uniform_pg_wind_forcing.h:5-37)
 *   oc_uniform_pg_forcing_height
 *                              modules::uniform_pg_wind_forcing_height (:40-79)
 *
 * The coupler state is [5 + ntr][nz][ny][nx] (density_dry, uvel, vvel, wvel, temp, tracers).
 * Horizontal means follow the reference's accumulation: each term times 1/(nx ny), summed over j
 * then i in order (sponge_layer.h:37-46, geostrophic_wind_forcing.h:23-30); the domain mean of
 * uniform_pg_wind_forcing is the plain sum over all cells divided by nz ny nx (:18-23).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

static double level_mean(const double *f, int k, int nx, int ny) {
  double r = 1. / ((double)nx * ny), s = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) s += f[((size_t)k * ny + j) * nx + i] * r;
  return s;
}

int oc_sponge_layer(const oc_config *c, double *cp, double dt, double time_scale, double top_prop) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  size_t n3 = (size_t)nx * ny * nz, plane = (size_t)nx * ny;
  double *u = cp + n3, *v = cp + 2 * n3, *w = cp + 3 * n3, *T = cp + 4 * n3;
  double dz = c->zlen / nz, z1 = (1 - top_prop) * c->zlen, z2 = c->zlen, p = 3;
  int k1 = (int)floor(z1 / dz - 0.5);                /* sponge_layer.h:34 */
  if (k1 < 0) k1 = 0;                                /* the reference would index below the domain */
  for (int k = k1; k < nz; k++) {
    double cu = level_mean(u, k, nx, ny), cv = level_mean(v, k, nx, ny), cT = level_mean(T, k, nx, ny);
    double z = (k + 0.5) * dz;
    if (!(z > z1)) continue;
    double factor = pow((z - z1) / (z2 - z1), p) * dt / time_scale;
    for (size_t id = (size_t)k * plane; id < (size_t)(k + 1) * plane; id++) {
      u[id] = factor * cu + (1 - factor) * u[id];
      v[id] = factor * cv + (1 - factor) * v[id];
      w[id] = factor * 0 + (1 - factor) * w[id];
      T[id] = factor * cT + (1 - factor) * T[id];
    }
  }
  return 0;
}

int oc_geostrophic_forcing(const oc_config *c, double *cp, double dt, double lat_g, double u_g, double v_g) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  size_t n3 = (size_t)nx * ny * nz, plane = (size_t)nx * ny;
  double *u = cp + n3, *v = cp + 2 * n3;
  double fcor = 2 * 7.2921e-5 * sin(lat_g / 180 * M_PI);
  for (int k = 0; k < nz; k++) {
    double cu = level_mean(u, k, nx, ny), cv = level_mean(v, k, nx, ny);   /* before the update */
    for (size_t id = (size_t)k * plane; id < (size_t)(k + 1) * plane; id++) {
      u[id] += dt * (fcor * (cv - v_g));
      v[id] += dt * (-fcor * (cu - u_g));
    }
  }
  return 0;
}

int oc_uniform_pg_forcing(const oc_config *c, double *cp, double dt, double tau, unsigned long update_cycles,
                          oc_pg_state *st) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  size_t n3 = (size_t)nx * ny * nz;
  double *u = cp + n3, *v = cp + 2 * n3;
  if (st->counter % update_cycles == 0) {
    double su = 0, sv = 0;
    for (size_t id = 0; id < n3; id++) { su += u[id]; sv += v[id]; }
    st->u = su / ((double)nz * ny * nx);
    st->v = sv / ((double)nz * ny * nx);
    if (st->counter == 0) { st->u0 = st->u; st->v0 = st->v; }
  }
  double du = dt / tau * (st->u0 - st->u), dv = dt / tau * (st->v0 - st->v);
  for (size_t id = 0; id < n3; id++) { u[id] += du; v[id] += dv; }
  st->counter++;
  return 0;
}

int oc_uniform_pg_forcing_height(const oc_config *c, double *cp, double dt, double height, double u0, double v0,
                                 double tau) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  size_t n3 = (size_t)nx * ny * nz, plane = (size_t)nx * ny;
  double *u = cp + n3, *v = cp + 2 * n3;
  double dz = c->zlen / nz;
  int k1 = 0;
  for (int k = 0; k < nz; k++) {
    if ((k + 0.5) * dz > height) { k1 = k - 1; break; }
  }
  if (k1 < 0 || k1 + 1 >= nz) return 1;
  double s[2][2] = {{0, 0}, {0, 0}};
  for (int l = 0; l < 2; l++)
    for (size_t id = (size_t)(k1 + l) * plane; id < (size_t)(k1 + l + 1) * plane; id++) { s[l][0] += u[id]; s[l][1] += v[id]; }
  double a = (double)ny * nx;
  double u1 = s[0][0] / a, v1 = s[0][1] / a, u2 = s[1][0] / a, v2 = s[1][1] / a;
  double z1 = (k1 + 0.5) * dz, z2 = (k1 + 1.5) * dz;
  double w1 = (z2 - height) / dz, w2 = (height - z1) / dz;
  double um = w1 * u1 + w2 * u2, vm = w1 * v1 + w2 * v2;
  double du = dt / tau * (u0 - um), dv = dt / tau * (v0 - vm);
  for (size_t id = 0; id < n3; id++) { u[id] += du; v[id] += dv; }
  return 0;
}
