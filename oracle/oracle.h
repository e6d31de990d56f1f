/*
 * oracle.h -- CPU oracle for the portUrb WENO finite-volume dycore step and the TKE LES closure.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the reference algorithm, used
 * as the parity checker by tests/, by __graft_entry__.smoke() and as bench.py's CPU baseline.
 * Product code never includes, links or calls it.
 *
 * It follows the reference's unfused kernel sequence:
 *   model/modules/dynamics_rk.h   (Dynamics_Euler_Stratified_WenoFV, :25-1352)
 *   model/modules/helpers/WenoLimiter.h (:10-96 helpers, :116-537 WenoLimiter<real,3..11>)
 *   model/modules/helpers/TransformMatrices.h (coefficient definitions; derived here numerically)
 *   model/modules/les_closure.h   (LES_Closure, :9-536)
 *   model/core/coupler.h          (halo exchange semantics, :726-1165)
 *   experiments/examples/custom_modules/sc_init.h:728-767 (immersed_proportion_halos)
 * with every float site of the reference promoted to double (the "FP64 oracle").
 *
 * Array conventions: every 3-D field is [nz][ny][nx], x fastest (the reference's styleC layout).
 * The coupler state is one contiguous block [5+num_tracers][nz][ny][nx] holding
 * density_dry, uvel, vvel, wvel, temp, then the tracer masses in tracer order.
 */
#ifndef PORTURB_ORACLE_H
#define PORTURB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OC_MAX_TRACERS 16

enum { OC_BC_PERIODIC = 0, OC_BC_PRECURSOR = 1, OC_BC_SOLID_WALL = 2 };

typedef struct {
  int    nx, ny, nz;
  double xlen, ylen, zlen;
  int    ord;                 /* 3, 5, 7, 9, 11 */
  int    num_tracers;
  int    idWV;                /* index of water_vapor among tracers (required) */
  int    idTKE;               /* index of TKE among tracers, -1 if none (closure only) */
  int    bc_x, bc_y, bc_z;
  int    weno_all;            /* option weno_all (default 1) */
  int    enable_gravity;      /* option enable_gravity (default 1) */
  double latitude;            /* option latitude (default 0) */
  double cfl;                 /* option cfl (default 0.6) */
  double immersed_power;      /* option immersed_power (default 5) */
  double R_d, R_v, cp_d, gamma_d, C0, grav, p0;
  unsigned char tracer_positive[OC_MAX_TRACERS];
  unsigned char tracer_adds_mass[OC_MAX_TRACERS];
  unsigned char tracer_diffuse[OC_MAX_TRACERS];
  double kinematic_viscosity; /* LES option kinematic_viscosity (default 0) */
  int    dns;                 /* LES option dns (default 0) */
  double les_total_mult, les_dissipation_mult, les_shear_prod_mult;
  int    nthreads;            /* OpenMP threads, <=0 means the runtime default */
} oc_config;

typedef struct oc_dycore oc_dycore;

/* Coefficient tables derived numerically (long double Gauss-Jordan) from their definitions.
 * C: [S][N][N] sub-stencil cell-average -> polynomial coefficient maps; faceL/faceR: [S][ord];
 * idealL/idealR: [S]; linL/linR: [ord]; tv: [N][N] upper-triangular quadratic form. */
int  oc_weno_tables(int ord, double *C, double *faceL, double *faceR, double *idealL,
                    double *idealR, double *linL, double *linR, double *tv);

/* WenoLimiter<real,ord>::compute_limited_edges (WenoLimiter.h) on one ord-point stencil. */
int  oc_weno_limited_edges(int ord, const double *stencil, int do_map, int imm_L, int imm_R,
                           double *qL, double *qR);

/* The final normalised weights of the two faces that compute_limited_edges combines the
 * sub-stencil values with (wL, wR: [(ord+1)/2]), the same code path as oc_weno_limited_edges. */
int  oc_weno_face_weights(int ord, const double *stencil, int do_map, int imm_L, int imm_R,
                          double *wL, double *wR);

/* The weight map of WenoLimiter.h:76-82 (map_rs, k = 6, m = 3, s = 2000) the limiter applies. */
double oc_weno_map_rs(double w, double d);

/* Dynamics_Euler_Stratified_WenoFV::compute_time_step (dynamics_rk.h:57-64). */
double oc_dycore_compute_time_step(const oc_config *cfg);

/* Dynamics_Euler_Stratified_WenoFV::init (dynamics_rk.h:917-1246): hydrostatic columns,
 * immersed halos, any_immersed10, hydrostatic edges. */
oc_dycore *oc_dycore_create(const oc_config *cfg, const double *coupler, const double *immersed_prop);
void       oc_dycore_destroy(oc_dycore *d);

/* Dynamics_Euler_Stratified_WenoFV::time_step (dynamics_rk.h:103-137), in place on the coupler. */
int  oc_dycore_time_step(oc_dycore *d, double *coupler, double dt_phys);

/* One call of compute_tendencies (dynamics_rk.h:498-732) on the dynamics state of a coupler
 * state (after convert_coupler_to_dynamics).  tend: [5+num_tracers][nz][ny][nx]. */
int  oc_dycore_tendencies(oc_dycore *d, const double *coupler, double *tend);

/* Hydrostatic profiles: cells [nz+2hs] x3, edges [nz+1] x3. */
int  oc_dycore_hydrostatics(const oc_dycore *d, double *dens_c, double *theta_c, double *pres_c,
                            double *dens_e, double *theta_e, double *pres_e);

/* dycore_any_immersed10 [nz][ny][nx] and dycore_immersed_tau [nz][ny][nx] (first relax call). */
int  oc_dycore_immersed_masks(oc_dycore *d, unsigned char *any10, int *tau);

/* density_pert, theta_pert, pressure_pert [nz][ny][nx] of the dycore's output hook. */
int  oc_dycore_output_fields(const oc_dycore *d, const double *coupler, double *density_pert, double *theta_pert,
                             double *pressure_pert);

/* LES_Closure::apply (les_closure.h:30-323), in place on the coupler state. */
int  oc_les_apply(const oc_config *cfg, double *coupler, const double *immersed_prop, double dt_phys);

/* ----------------------------------------------------------------------------------------- */
/* Learned subgrid closure (sgs.c; no reference counterpart)                                  */
/* ----------------------------------------------------------------------------------------- */
#define OC_MLP_NIN 108
#define OC_MLP_NH 64
#define OC_MLP_NOUT 9
/* parameters: W1[64][108], b1[64], W2[64][64], b2[64], W3[9][64], b3[9] (row = output unit) */
#define OC_MLP_NPARAM (OC_MLP_NH * OC_MLP_NIN + OC_MLP_NH + OC_MLP_NH * OC_MLP_NH + OC_MLP_NH + OC_MLP_NOUT * OC_MLP_NH + OC_MLP_NOUT)

typedef struct {
  const float *params;       /* OC_MLP_NPARAM */
  const float *feat_scale;   /* [108] 1/sigma of each raw input (difference / delta) */
  float xmax;                /* bound on |normalised input| */
  const float *out_mean;     /* [9] */
  const float *out_scale;    /* [9] tau = out_mean + out_scale * y */
} oc_sgs_model;

float  oc_bf16_round(float x);
double oc_relu(double x);
double oc_relu_deriv(double x);
void   oc_mlp_forward(const float *P, const float *X, long n, int bf16, float *Y);
double oc_mlp_loss_grad(const float *P, const float *X, const float *T, long n, int bf16, double *G);
void   oc_mlp_adam(float *P, const float *G, float *m, float *v, long np, int t, float lr, float b1, float b2,
                   float eps);
/* box filter by r of a fine coupler state: les [6][nz/r+2][ny/r+2][nx/r+2] (rho u v w theta k_sgs,
 * halo 1), imm_c [nz/r+2][ny/r+2][nx/r+2], targets [9][nz/r][ny/r][nx/r] */
int    oc_sgs_filter(const oc_config *c, const double *cp, const double *imm_fine, int r, double *les,
                     double *imm_c, float *targets);
void   oc_sgs_fmul(const float *feat_scale, double delta, float *fmul);
int    oc_sgs_features(const double *les, int nx, int ny, int nz, const float *fmul, const long *cells, long n,
                       float *X);
void   oc_sgs_infer(const oc_sgs_model *M, const double *les, const double *imm, int nx, int ny, int nz, double dx,
                    double dy, double dz, double nu_max, int bf16, float *tau, unsigned char *flags);
/* LES_Closure::apply with the learned momentum/heat fluxes where both cells of a face hold the
 * network's flag; tau_out [9][n3] and flags_out [n3] optional. */
int    oc_les_apply_learned(const oc_config *cfg, double *coupler, const double *immersed_prop, double dt_phys,
                            const oc_sgs_model *M, int bf16, float *tau_out, unsigned char *flags_out);

/* modules::apply_surface_fluxes (surface_flux.h:9-126), in place on the coupler state; prop,
 * rough, temp [nz][ny][nx], sfc_rough, sfc_temp [ny][nx]. */
int    oc_surface_fluxes(const oc_config *c, double *coupler, const double *prop, const double *rough,
                         const double *temp, const double *sfc_rough, const double *sfc_temp, double dt);

/* Column-mean forcing modules (column.c), in place on the coupler state. */
typedef struct {
  unsigned long counter;     /* option uniform_pg_wind_forcing_counter */
  double u0, v0, u, v;       /* options uniform_pg_wind_forcing_{u0,v0,u,v} */
} oc_pg_state;
int    oc_sponge_layer(const oc_config *c, double *coupler, double dt, double time_scale, double top_prop);
int    oc_geostrophic_forcing(const oc_config *c, double *coupler, double dt, double lat_g, double u_g, double v_g);
int    oc_uniform_pg_forcing(const oc_config *c, double *coupler, double dt, double tau, unsigned long update_cycles,
                             oc_pg_state *st);
int    oc_uniform_pg_forcing_height(const oc_config *c, double *coupler, double dt, double height, double u0,
                                    double v0, double tau);

/* modules::WindmillActuators::apply, actuator disks (turbine.c) */
typedef struct {
  int ntab;
  const double *velmag, *thrust_coef, *power_coef, *power_mw;
  double hub_height, blade_radius, max_yaw_speed, overhang, hub_radius, shaft_tilt;
} oc_turbine_ref;
typedef struct {
  double base_x, base_y, yaw, u_inertial, v_inertial, yaw_uavg, yaw_vavg;
  int apply_thrust;
  double power, ct, cp, u_samp, v_samp, upstream_u, upstream_v;   /* of the last apply */
} oc_turbine;
typedef struct {
  int fixed_yaw, orig_C_T, have_upstream_dir;
  double upstream_dir;
} oc_turbine_opts;
double oc_turbine_interp(const double *ref_umag, const double *ref_var, int n, double umag);
double oc_turbine_blade_weight(double dx, double rad);
double oc_turbine_yaw_tend(oc_turbine *t, double uvel, double vvel, double dt, double max_yaw_speed);
int    oc_turbines_apply(const oc_config *c, double *coupler, const oc_turbine_ref *ref, oc_turbine *turbs, int nturb,
                         const oc_turbine_opts *o, double dt, double *proj_weight_tot, double *samp_weight_tot);

/* Sum over cells of the dycore's full density rho (dry + mass-adding tracers), times cell volume. */
double oc_total_mass(const oc_config *cfg, const double *coupler);

#ifdef __cplusplus
}
#endif
#endif
