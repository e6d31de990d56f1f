/*
 * turbine.c -- CPU oracle of modules::WindmillActuators::apply, actuator-disk part
 * (model/modules/windmill_actuators_yaw.h:523-1043).  TEST INFRASTRUCTURE ONLY.
 *
 * Per turbine, in order: disk-box sums of u and v around the hub (:622-645) and the upstream
 * direction (:646-653); the disk projection as a point cloud in the rotor frame, shaped by
 * DefaultThrustShape2 (radius) times DefaultProjectionShape1D (axial) (:306-329,654-716), tilted,
 * yawed, moved to the hub and binned into cells by rounding (x/dx - 0.5), once at the disk
 * (projection) and once 4 radii upstream (sampling); normalisation of both weights (:834-881);
 * the sampled inflow (:882-897); table interpolation of C_T, C_P and power at the inertially
 * filtered inflow, with the induction-factor C_T unless turbine_orig_C_T (:898-913,1050-1063);
 * the 30 s inertial filter (:957-959); the thrust and TKE tendencies with the normal-wind limiter
 * (:966-1018); the yaw controller (:81-103,1020-1029).  After all turbines: u, v, TKE += dt * the
 * summed tendencies (:1035-1041).  Not restated (as in the reference's wind-farm drivers,
 * windfarm_ensemble.cpp:106-110): blades, floating motions, immersed hub/tower material.
 * Arithmetic in double (the reference computes these fields in float).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double thrust_shape2(double x, double x2, double x3, double x4, double a) {   /* :306-317 */
  if (x < x3) { double v = -((x * x) - 2 * x * x2) / (x2 * x2); return a == 1 ? v : pow(v, a); }
  if (x < x4) {
    double v3 = -((x3 * x3) - 2 * x3 * x2) / (x2 * x2);
    if (a != 1) v3 = pow(v3, a);
    return (v3 * x * x - 2 * v3 * x * x4 + v3 * x4 * x4) / (x3 * x3 - 2 * x3 * x4 + x4 * x4);
  }
  return 0;
}

static double proj_shape_1d(double x, double xr) {          /* :319-327, p = 2 */
  double term = 1 - (x / xr) * (x / xr);
  if (term <= 0) return 0;
  return term * term;
}

double oc_turbine_interp(const double *ref_umag, const double *ref_var, int n, double umag) {   /* :1050-1063 */
  int imax = n - 1;
  if (umag < ref_umag[0] || umag > ref_umag[imax]) return 0;
  int i = 0;
  while (umag > ref_umag[i]) i++;
  if (i > 0) i--;
  double fac = (ref_umag[i + 1] - umag) / (ref_umag[i + 1] - ref_umag[i]);
  return fac * ref_var[i] + (1 - fac) * ref_var[i + 1];
}

double oc_turbine_blade_weight(double dx, double rad) {     /* :828-834 */
  double dxloc = dx * (63 / rad);
  if (dxloc <= 2) return 1;
  double x = (log2(dxloc) - 1) / 3;
  return -2 * x * x * x + 3 * x * x;
}

double oc_turbine_yaw_tend(oc_turbine *t, double uvel, double vvel, double dt, double max_yaw_speed) {  /* :81-103 */
  const double tau = 60;
  t->yaw_uavg = (tau - dt) / tau * t->yaw_uavg + dt / tau * uvel;
  t->yaw_vavg = (tau - dt) / tau * t->yaw_vavg + dt / tau * vvel;
  double dir_upwind = atan2(t->yaw_vavg, t->yaw_uavg);
  double diff = dir_upwind - t->yaw;
  if (diff > M_PI) diff -= 2 * M_PI;
  if (diff < -M_PI) diff += 2 * M_PI;
  double tend = diff / dt;
  if (tend > 0) tend = fmin(max_yaw_speed, tend);
  else tend = fmax(-max_yaw_speed, tend);
  return t->yaw + dt * tend;
}

int oc_turbines_apply(const oc_config *c, double *cp, const oc_turbine_ref *ref, oc_turbine *turbs, int nturb,
                      const oc_turbine_opts *o, double dt, double *proj_weight_tot, double *samp_weight_tot) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  size_t n3 = (size_t)nx * ny * nz;
  if (c->idTKE < 0) return 1;
  double dx = c->xlen / nx, dy = c->ylen / ny, dz = c->zlen / nz;
  double *rho_d = cp, *uvel = cp + n3, *vvel = cp + 2 * n3, *tke = cp + (size_t)(5 + c->idTKE) * n3;
  double *tend_u = (double *)calloc(n3, sizeof(double)), *tend_v = (double *)calloc(n3, sizeof(double));
  double *tend_tke = (double *)calloc(n3, sizeof(double));
  double *proj = (double *)malloc(n3 * sizeof(double)), *samp = (double *)malloc(n3 * sizeof(double));
  if (proj_weight_tot) memset(proj_weight_tot, 0, n3 * sizeof(double));
  if (samp_weight_tot) memset(samp_weight_tot, 0, n3 * sizeof(double));
  double rad = ref->blade_radius, hub_height = ref->hub_height;
  double cos_tlt = cos(ref->shaft_tilt), sin_tlt = sin(ref->shaft_tilt);
  double decay = 3 * dx / rad;
  double blade_wt = oc_turbine_blade_weight(dx, rad);
  for (int it = 0; it < nturb; it++) {
    oc_turbine *t = &turbs[it];
    double cos_yaw = cos(t->yaw), sin_yaw = sin(t->yaw);
    double base_x = t->base_x, base_y = t->base_y;
    /* disk-box sums (:622-645) */
    double su = 0, sv = 0;
    for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
      double x = (i + 0.5) * dx, y = (j + 0.5) * dy, z = (k + 0.5) * dz;
      if (z >= hub_height - rad && z <= hub_height + rad && y >= base_y - rad && y <= base_y + rad &&
          x >= base_x - rad && x <= base_x + rad) {
        size_t id = ((size_t)k * ny + j) * nx + i;
        su += uvel[id]; sv += vvel[id];
      }
    }
    double upstream_dir = o->have_upstream_dir ? o->upstream_dir : atan2(sv, su);
    double up_x = -4 * rad * cos(upstream_dir), up_y = -4 * rad * sin(upstream_dir);
    /* projection (:654-716) */
    memset(proj, 0, n3 * sizeof(double));
    memset(samp, 0, n3 * sizeof(double));
    double decayloc = decay + 0.02;
    double xr = fmax(5., 5 * dx);
    int num_x = (int)ceil(20 / dx * xr * 2);
    int num_y = (int)ceil(20 / dx * rad * (1 + decayloc) * 2);
    int num_z = (int)ceil(20 / dx * rad * (1 + decayloc) * 2);
    double R = rad * (1 + decayloc);
    for (int k = 0; k < num_z; k++) for (int j = 0; j < num_y; j++) for (int i = 0; i < num_x; i++) {
      double x = -xr + (2 * xr * i) / (num_x - 1);
      double y = -R + (2 * R * j) / (num_y - 1);
      double z = -R + (2 * R * k) / (num_z - 1);
      double rloc = sqrt(y * y + z * z);
      if (!(rloc <= R)) continue;
      double shp = rloc <= ref->hub_radius ? 0 : thrust_shape2(rloc / rad, 0.7, 1, 1 + decayloc, 1) * proj_shape_1d(x, xr);
      double x1 = cos_tlt * (x + ref->overhang) + sin_tlt * z;
      double y1 = y;
      double z1 = -sin_tlt * (x + ref->overhang) + cos_tlt * z;
      double xp = base_x + cos_yaw * x1 - sin_yaw * y1;
      double yp = base_y + sin_yaw * x1 + cos_yaw * y1;
      double zp = hub_height + z1;
      int ti = (int)round(xp / dx - 0.5), tj = (int)round(yp / dy - 0.5), tk = (int)round(zp / dz - 0.5);
      if (ti >= 0 && ti < nx && tj >= 0 && tj < ny && tk >= 0 && tk < nz) proj[((size_t)tk * ny + tj) * nx + ti] += shp;
      xp += up_x;
      yp += up_y;
      ti = (int)round(xp / dx - 0.5); tj = (int)round(yp / dy - 0.5); tk = (int)round(zp / dz - 0.5);
      if (ti >= 0 && ti < nx && tj >= 0 && tj < ny && tk >= 0 && tk < nz) samp[((size_t)tk * ny + tj) * nx + ti] += shp;
    }
    double proj_tot = 0, samp_tot = 0;
    for (size_t id = 0; id < n3; id++) { proj_tot += proj[id]; samp_tot += samp[id]; }
    /* normalisation and sampled inflow (:870-897) */
    double s0 = 0, s1 = 0;
    for (size_t id = 0; id < n3; id++) {
      if (proj[id] > 0) {
        proj[id] /= proj_tot;
        if (proj_weight_tot) proj_weight_tot[id] += proj[id];
      }
      if (samp[id] > 0) {
        samp[id] /= samp_tot;
        if (samp_weight_tot) samp_weight_tot[id] += samp[id];
        s0 += samp[id] * uvel[id];
        s1 += samp[id] * vvel[id];
      }
    }
    double instant_u0 = s0 * cos_yaw, instant_v0 = s1 * sin_yaw;
    double instant_mag0 = fmax(0., instant_u0 + instant_v0);
    double inertial_mag0 = fmax(0., t->u_inertial + t->v_inertial);
    double C_T = oc_turbine_interp(ref->velmag, ref->thrust_coef, ref->ntab, inertial_mag0);
    double C_P = oc_turbine_interp(ref->velmag, ref->power_coef, ref->ntab, inertial_mag0);
    double pwr = oc_turbine_interp(ref->velmag, ref->power_mw, ref->ntab, inertial_mag0);
    if (inertial_mag0 > 1.e-10) {
      if (!o->orig_C_T) {
        double a = fmax(0., fmin(1., 1 - C_P / (C_T + 1.e-10)));
        C_T = 4 * a * (1 - a);
      }
      C_P = fmin(C_T, pwr * 1.e6 / (0.5 * 1.2 * M_PI * rad * rad * inertial_mag0 * inertial_mag0 * inertial_mag0));
    } else {
      C_T = 0;
      C_P = 0;
    }
    double inertial_tau = 30;
    t->u_inertial = instant_u0 * dt / inertial_tau + (inertial_tau - dt) / inertial_tau * t->u_inertial;
    t->v_inertial = instant_v0 * dt / inertial_tau + (inertial_tau - dt) / inertial_tau * t->v_inertial;
    t->power = pwr; t->ct = C_T; t->cp = C_P; t->u_samp = s0; t->v_samp = s1;
    t->upstream_u = su; t->upstream_v = sv;
    double turb_factor = M_PI * rad * rad / (dx * dy * dz);
    double C_TKE = 0.25 * (C_T - C_P);
    if (t->apply_thrust) {
      for (size_t id = 0; id < n3; id++) {
        if (!(proj[id] > 0)) continue;
        double wt = proj[id] * turb_factor;
        double un = uvel[id] * cos_yaw + vvel[id] * sin_yaw;
        if (!(un > 0)) continue;
        double t_u = -0.5 * C_T * instant_mag0 * instant_mag0 * cos_yaw * wt;
        double t_v = -0.5 * C_T * instant_mag0 * instant_mag0 * sin_yaw * wt;
        double t_tke = 0.5 * rho_d[id] * C_TKE * instant_mag0 * instant_mag0 * instant_mag0 * wt * (1 - blade_wt);
        double un_new = (uvel[id] + dt * t_u) * cos_yaw + (vvel[id] + dt * t_v) * sin_yaw;
        if (un_new < 0) {
          double t_old = fabs(t_u * cos_yaw + t_v * sin_yaw);
          double t_new = un;
          t_u *= (t_new / t_old);
          t_v *= (t_new / t_old);
          t_tke *= (t_new / t_old) * (t_new / t_old);
        }
        tend_u[id] += t_u;
        tend_v[id] += t_v;
        tend_tke[id] += t_tke;
      }
    }
    if (!o->fixed_yaw) t->yaw = oc_turbine_yaw_tend(t, su, sv, dt, ref->max_yaw_speed);
  }
  for (size_t id = 0; id < n3; id++) {
    uvel[id] += dt * tend_u[id];
    vvel[id] += dt * tend_v[id];
    tke[id] += dt * tend_tke[id];
  }
  free(tend_u); free(tend_v); free(tend_tke); free(proj); free(samp);
  return 0;
}
