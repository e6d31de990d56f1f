/*
 * porturb.h -- C ABI of the MI355X-native portUrb dycore step and LES closure (libporturb.so).
 *
 * The reference has no C ABI: its dycore and LES closure are C++ structs called from drivers
 * through core::Coupler::run_module (model/core/coupler.h:344-380).  This header is the thin,
 * stream-ordered C layer under the drop-in C++ API (include/porturb/...), which keeps the
 * reference's class and method names.  Every entry point below names the reference interface it
 * replaces.  Plain pointers and sizes only; all field pointers are DEVICE pointers to
 * [nz][ny][nx] arrays of double (x fastest), as core::DataManager allocates them
 * (model/core/DataManager.h:54-56,127-209).
 *
 * Errors: functions return PU_OK (0) or a nonzero PU_ERR_*; pu_last_error() gives the message
 * (thread-local).  The reference aborts instead (endrun/Kokkos::abort, model/main_header.h:101-103);
 * the C++ layer maps nonzero returns to that behaviour.
 * Threading: one context per stream; calls on one context are not re-entrant.
 */
#ifndef PORTURB_H
#define PORTURB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PU_MAX_TRACERS 16

enum { PU_OK = 0, PU_ERR_ARG = 1, PU_ERR_HIP = 2, PU_ERR_STATE = 3, PU_ERR_UNSUPPORTED = 4 };

/* Boundary conditions: options bc_x / bc_y ("periodic" | "precursor") and bc_z
 * ("solid_wall" | "periodic"), read at dynamics_rk.h:746,752-797 and les_closure.h:431-534. */
enum { PU_BC_PERIODIC = 0, PU_BC_PRECURSOR = 1, PU_BC_SOLID_WALL = 2 };

/* Arithmetic of the dycore step.  PU_FP64: everything in double (the reference's float sites
 * promoted).  PU_FP32: FP32 state, reconstruction, fluxes and tendencies; FP64 hydrostatic
 * background, pressure perturbation p - p_hy, global reductions; compensated (two-float) storage of
 * rho and rho*theta across RK stages. */
enum { PU_FP64 = 0, PU_FP32 = 1 };

/* Model configuration: grid (Coupler::distribute_mpi_and_allocate_coupled_state / set_grid,
 * coupler.h:132-229,329-341), tracer attributes (Coupler::add_tracer, coupler.h:417-447) and the
 * Options keys the hot path reads (dynamics_rk.h:515-534, les_closure.h:36-51, sc_init.h:31-51). */
typedef struct {
  int    nx, ny, nz;
  double xlen, ylen, zlen;
  int    ord;                 /* reconstruction order: 3, 5, 7, 9 or 11 (PORTURB_ORD) */
  int    num_tracers;
  int    idWV;                /* option idWV: index of water_vapor (required) */
  int    idTKE;               /* index of the TKE tracer (LES closure), -1 if none */
  int    bc_x, bc_y, bc_z;
  int    weno_all;            /* option weno_all (default 1) */
  int    enable_gravity;      /* option enable_gravity (default 1) */
  double latitude;            /* option latitude (default 0) */
  double cfl;                 /* option cfl (default 0.6) */
  double immersed_power;      /* option immersed_power (default 5) */
  double R_d, R_v, cp_d, gamma_d, C0, grav, p0;
  unsigned char tracer_positive[PU_MAX_TRACERS];
  unsigned char tracer_adds_mass[PU_MAX_TRACERS];
  unsigned char tracer_diffuse[PU_MAX_TRACERS];
  double kinematic_viscosity; /* LES option kinematic_viscosity (default 0) */
  int    dns;                 /* LES option dns (default 0) */
  double les_total_mult, les_dissipation_mult, les_shear_prod_mult;
  int    precision;           /* PU_FP64 | PU_FP32 */
} pu_config;

/* The coupler state one module call reads and writes: DataManager entries density_dry, uvel,
 * vvel, wvel, temp, one entry per tracer (mass), and immersed_proportion
 * (dynamics_rk.h:1251-1352, les_closure.h:328-414). */
typedef struct {
  double       *density_dry, *uvel, *vvel, *wvel, *temp;
  double       *tracers[PU_MAX_TRACERS];
  const double *immersed_proportion;
} pu_coupler_state;

typedef struct pu_ctx pu_ctx;

/* Context: owns the device arena (dycore stage buffers, hydrostatics, immersed masks).
 * `stream` is a hipStream_t (NULL = the default stream).  Replaces the module objects
 * Dynamics_Euler_Stratified_WenoFV / LES_Closure and their DataManager-held helper entries. */
int  pu_create(const pu_config *cfg, void *stream, pu_ctx **out);
int  pu_destroy(pu_ctx *ctx);
const char *pu_last_error(void);

/* Dynamics_Euler_Stratified_WenoFV::init (dynamics_rk.h:917-1246): hydrostatic column means in
 * FP64, isentropic halo extrapolation, hydrostatic edges, immersed halos, any_immersed10, tau. */
int  pu_dycore_init(pu_ctx *ctx, const pu_coupler_state *cs);

/* Dynamics_Euler_Stratified_WenoFV::compute_time_step (dynamics_rk.h:57-64). */
double pu_dycore_compute_time_step(const pu_ctx *ctx);

/* Dynamics_Euler_Stratified_WenoFV::time_step (dynamics_rk.h:103-137): convert, ncycles SSPRK3
 * (time_step_rk_3_3 :279-367, enforce_immersed_boundaries :370-473, compute_tendencies :498-732),
 * convert back.  In place on the coupler state; stream-ordered (returns before completion). */
int  pu_dycore_time_step(pu_ctx *ctx, pu_coupler_state *cs, double dt_phys);

/* One compute_tendencies call (dynamics_rk.h:498-732) on the dynamics state converted from `cs`.
 * tend: device [5+num_tracers][nz][ny][nx] double.  For parity checking of a single stage. */
int  pu_dycore_tendencies(pu_ctx *ctx, const pu_coupler_state *cs, double *tend);

/* Reconstruction census of the first stage of a time step from `cs` (no counterpart in the
 * reference: a measurement aid for the work the sweeps execute).  For each direction x, y, z and
 * every interior cell and reconstructed field, the branch the WENO limiter takes on this state:
 * counts[3*dir + 0] the full limiter, + 1 the flat-stencil branch (the candidate stencils' total
 * variation <= 1e-20, where the limited edges equal the linear reconstruction to rounding), + 2
 * the linear reconstruction of a cell without WENO.  Host output, synchronous. */
int  pu_dycore_flat_census(pu_ctx *ctx, const pu_coupler_state *cs, long long *counts);

/* Hydrostatic profiles as the dycore holds them (host outputs): hy_{dens,theta,pressure}_cells
 * [nz+2hs] and hy_{dens,theta,pressure}_edges [nz+1] (dynamics_rk.h:961-1008,1145-1186). */
int  pu_dycore_hydrostatics(pu_ctx *ctx, double *dens_c, double *theta_c, double *pres_c,
                            double *dens_e, double *theta_e, double *pres_e);

/* dycore_any_immersed10 and dycore_immersed_tau (host outputs, [nz][ny][nx]) (:389-425,:1127-1143). */
int  pu_dycore_immersed_masks(pu_ctx *ctx, unsigned char *any10, int *tau);

/* The immersed_proportion_halos entry the LES closure reads (built by custom_modules::sc_init,
 * sc_init.h:728-767: halo 1, x/y periodic, bottom halo 1, top halo 0).  pu_dycore_init builds it
 * too; this entry point serves contexts that run the closure without the dycore. */
int  pu_les_closure_init(pu_ctx *ctx, const pu_coupler_state *cs);

/* A caller-held immersed_proportion_halos entry (device [nz+2][ny+2][nx+2]) replaces the one the
 * context built: the reference's closure and surface fluxes read that DataManager entry at every
 * apply (les_closure.h:45, surface_flux.h:28), and drivers may edit it (the turbines add their
 * projection weights, windmill_actuators_yaw.h:159).  Stream-ordered copy; the context keeps its
 * own copy until the next call. */
int  pu_les_closure_set_immersed_halos(pu_ctx *ctx, const double *halos);

/* LES_Closure::apply (les_closure.h:30-323).  In place on the coupler state; requires idTKE. */
int  pu_les_closure_apply(pu_ctx *ctx, pu_coupler_state *cs, double dt_phys);

/* modules::apply_surface_fluxes (model/modules/surface_flux.h:9-126): log-law drag (kappa = 0.4)
 * and temperature exchange against every immersed neighbour, forward Euler on uvel, vvel, wvel,
 * temp.  _init builds the halo'd roughness/temperature boxes the way custom_modules::sc_init builds
 * immersed_{roughness,temp}_halos (sc_init.h:728-767): immersed_roughness/immersed_temp
 * [nz][ny][nx] and surface_roughness/surface_temp [ny][nx] (device), x/y periodic, the bottom halo
 * from the surface fields, the top halo 0.  Needs pu_dycore_init or pu_les_closure_init first. */
int  pu_surface_fluxes_init(pu_ctx *ctx, const double *immersed_roughness, const double *immersed_temp,
                            const double *surface_roughness, const double *surface_temp);
int  pu_surface_fluxes_apply(pu_ctx *ctx, pu_coupler_state *cs, double dt_phys);

/* modules::sponge_layer (model/modules/sponge_layer.h:7-60): above z1 = (1 - top_prop) zlen, u, v
 * and temp relax towards their horizontal means at each level and w towards 0 with
 * factor = ((z - z1)/(zlen - z1))^3 dt / time_scale.  The means cover the levels from
 * k1 = floor(z1/dz - 0.5) up (k1 below 0 starts at level 0: no cell below the domain has z > z1).
 * Requires time_scale > 0 and 0 < top_prop <= 1.  Stream-ordered. */
int  pu_sponge_layer(pu_ctx *ctx, pu_coupler_state *cs, double dt, double time_scale, double top_prop);

/* modules::geostrophic_wind_forcing (model/modules/geostrophic_wind_forcing.h:7-36):
 * f = 2 * 7.2921e-5 * sin(lat_g pi/180); with <u>(k), <v>(k) the horizontal means per level,
 * uvel += dt f (<v> - v_g), vvel -= dt f (<u> - u_g).  Stream-ordered. */
int  pu_geostrophic_wind_forcing(pu_ctx *ctx, pu_coupler_state *cs, double dt, double lat_g, double u_g, double v_g);

/* Horizontal means of uvel and vvel on levels k_begin .. k_end-1, to host arrays of k_end - k_begin
 * values: the sums of modules::uniform_pg_wind_forcing (uniform_pg_wind_forcing.h:18-21, domain
 * mean = mean of the level means) and uniform_pg_wind_forcing_height (:56-60).  Synchronous. */
int  pu_level_mean_wind(pu_ctx *ctx, const pu_coupler_state *cs, int k_begin, int k_end, double *u_mean,
                        double *v_mean);

/* Allocation-free variant of pu_level_mean_wind.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight loops, harnesses
 * that forbid hipMalloc).  It calls no allocating routine: the partial sums live in a static device
 * array of PU_NOALLOC_MAX_LEVELS levels (256 KiB per GPU, held until the process ends), and it
 * rejects more levels than that with PU_ERR_ARG.  u_mean, v_mean are DEVICE arrays of
 * k_end - k_begin doubles; the result is bitwise that of pu_level_mean_wind.  Stream-ordered on
 * ctx's stream (no synchronisation).  Every call on one GPU shares the array: calls that overlap
 * on one GPU (other streams, other contexts, other threads) race. */
#define PU_NOALLOC_MAX_LEVELS 1024
int  pu_level_mean_wind_noalloc(pu_ctx *ctx, const pu_coupler_state *cs, int k_begin, int k_end, double *u_mean,
                                double *v_mean);

/* uvel += du, vvel += dv on every cell: the update of uniform_pg_wind_forcing(_height)
 * (uniform_pg_wind_forcing.h:31-34,75-78).  Stream-ordered. */
int  pu_add_uniform_wind(pu_ctx *ctx, pu_coupler_state *cs, double du, double dv);

/* ------------------------------------------------------------------------------------------
 * modules::WindmillActuators (model/modules/windmill_actuators_yaw.h:12-1065) as actuator disks:
 * the disk projection and upstream sampling, table-driven thrust / power coefficients at the
 * inertially filtered inflow, thrust and TKE tendencies with the normal-wind limiter, yaw
 * controller; no blades, floating motions or immersed hub/tower (the reference's wind-farm drivers
 * run without them, windfarm_ensemble.cpp:106-110).  The turbine table is the reference's turbine
 * file (RefTurbine::init, :34-79): velocity_magnitude, thrust_coef, power_coef, power_megawatts,
 * hub_height, blade_radius, max_yaw_speed (rad/s), overhang, hub_radius, shaft_tilt (rad). */
typedef struct {
  int    ntable;
  const double *velocity_magnitude, *thrust_coef, *power_coef, *power_megawatts;   /* host, ntable each */
  double hub_height, blade_radius, max_yaw_speed, overhang, hub_radius, shaft_tilt;
} pu_turbine_ref;
/* options turbine_initial_yaw (rad), turbine_fixed_yaw, turbine_orig_C_T, turbine_upstream_dir */
typedef struct {
  double initial_yaw;
  int    fixed_yaw, orig_C_T, have_upstream_dir;
  double upstream_dir;
} pu_turbine_opts;
/* one turbine after its last apply: yaw (rad), inertial inflow, table values, sampled inflow
 * sum(w u), sum(w v), disk-cube sums of u and v, instantaneous normal inflow, number of applies */
typedef struct {
  double yaw, u_samp_inertial, v_samp_inertial, power, ct, cp, u_samp, v_samp, upstream_u, upstream_v, instant_mag0;
  int    nsteps;
} pu_turbine_status;
/* time traces kept per apply (the reference's power/yaw/u_samp/v_samp/cp/ct traces, :111-125) */
enum { PU_TURB_POWER = 0, PU_TURB_YAW = 1, PU_TURB_U_SAMP = 2, PU_TURB_V_SAMP = 3, PU_TURB_CP = 4, PU_TURB_CT = 5,
       PU_TURB_NTRACE = 6 };
typedef struct pu_turbines pu_turbines;
/* WindmillActuators::init's turbine placement done by the caller: tower bases x[i], y[i] (m);
 * apply_thrust NULL = all (option turbine_apply_thrust).  Needs the TKE tracer (idTKE). */
int  pu_turbines_create(pu_ctx *ctx, const pu_turbine_ref *ref, int nturb, const double *x, const double *y,
                        const unsigned char *apply_thrust, const pu_turbine_opts *opts, pu_turbines **out);
/* WindmillActuators::apply (:523-1043): uvel, vvel and the TKE tracer in place; proj_weight_tot /
 * samp_weight_tot (device [nz][ny][nx], or NULL) receive windmill_proj_weight /
 * windmill_samp_weight.  Synchronises the stream once per turbine (table lookups on the host). */
int  pu_turbines_apply(pu_turbines *t, pu_coupler_state *cs, double dt, double *proj_weight_tot,
                       double *samp_weight_tot);
int  pu_turbines_status(pu_turbines *t, int iturb, pu_turbine_status *out);
/* trace `which` (PU_TURB_*) of turbine iturb: status.nsteps values (the yaw in radians) */
int  pu_turbines_trace(pu_turbines *t, int iturb, int which, double *out);
int  pu_turbines_destroy(pu_turbines *t);

/* ------------------------------------------------------------------------------------------
 * Output / restart files: NetCDF classic with fixed-size variables, the format the reference's
 * Coupler::write_output_file / overwrite_with_restart (model/core/coupler.h:533-723) write and
 * read through yakl::SimplePNetCDF (create_dim, create_var, enddef, redef, write, write_all,
 * read, read_all).  PU_NC_AUTO writes CDF-2 (64-bit offset) unless a variable needs CDF-5 (over
 * 4 GiB, or an unsigned-byte / 64-bit integer type); the reference always writes CDF-5
 * (NC_64BIT_DATA).  Reading accepts CDF-1, CDF-2 and CDF-5.  Device transfers (_device) take
 * double arrays and convert/byte-swap on the GPU, on ctx's stream (ctx NULL: the default
 * stream); they return after the data are in the file (put) or in device memory (get).  Host
 * transfers take double, float, int or unsigned char (PU_NC_UBYTE) arrays. */
enum { PU_NC_BYTE = 1, PU_NC_CHAR = 2, PU_NC_SHORT = 3, PU_NC_INT = 4, PU_NC_FLOAT = 5, PU_NC_DOUBLE = 6,
       PU_NC_UBYTE = 7, PU_NC_INT64 = 10 };
enum { PU_NC_AUTO = 0, PU_NC_CDF2 = 2, PU_NC_CDF5 = 5 };
typedef struct pu_ncfile pu_ncfile;
int  pu_nc_create(const char *path, int format, pu_ncfile **out);
int  pu_nc_open(const char *path, pu_ncfile **out);
int  pu_nc_def_dim(pu_ncfile *f, const char *name, long long len);
int  pu_nc_def_var(pu_ncfile *f, const char *name, int nc_type, int ndims, const char *const *dimnames);
int  pu_nc_redef(pu_ncfile *f);
int  pu_nc_enddef(pu_ncfile *f);
int  pu_nc_inq_var(pu_ncfile *f, const char *name, int *nc_type, int *ndims, long long *dims);
int  pu_nc_put_host(pu_ncfile *f, const char *name, const void *data, int host_type);
int  pu_nc_get_host(pu_ncfile *f, const char *name, void *data, int host_type);
int  pu_nc_put_device(pu_ncfile *f, const char *name, const double *d_data, pu_ctx *ctx);
int  pu_nc_get_device(pu_ncfile *f, const char *name, double *d_data, pu_ctx *ctx);
int  pu_nc_format(pu_ncfile *f, int *version);
int  pu_nc_close(pu_ncfile *f);

/* The dycore's output fields (the write hook registered at dynamics_rk.h:1192-1233): with the
 * coupler state converted as convert_coupler_to_dynamics does, density_pert = rho - hy_dens_cells,
 * theta_pert = rho theta / rho - hy_theta_cells, pressure_pert = C0 (rho theta)^gamma -
 * hy_pressure_cells (device outputs [nz][ny][nx]). */
int  pu_dycore_output_fields(pu_ctx *ctx, const pu_coupler_state *cs, double *density_pert, double *theta_pert,
                             double *pressure_pert);

/* The dycore's restart hook (dynamics_rk.h:1234-1241): hydrostatic cell profiles from a restart
 * file ([nz+2hs] each, host) replace the context's, the hydrostatic edges are recomputed
 * (compute_hydrostasis_edges) and the immersed halos rebuilt from the coupler's
 * immersed_proportion (create_immersed_proportion_halos). */
int  pu_dycore_restart(pu_ctx *ctx, const pu_coupler_state *cs, const double *dens_c, const double *theta_c,
                       const double *pres_c);

/* FP64 reductions on the device: total mass sum(rho)*dV (rho = dry + mass-adding tracers), max
 * wind speed and max |w| (the quantities of Coupler::inform_user, coupler.h:507-530). */
int  pu_diagnostics(pu_ctx *ctx, const pu_coupler_state *cs, double *mass, double *max_wind,
                    double *max_abs_w);

/* Device memory for DataManager entries (yakl::alloc_device / free_device, DataManager.h:54-56)
 * and host<->device copies (YAKL deep_copy_to).  Copies are synchronous. */
int  pu_alloc_device(size_t bytes, void **ptr);
int  pu_free_device(void *ptr);
int  pu_memcpy_h2d(void *dst, const void *src, size_t bytes);
int  pu_memcpy_d2h(void *dst, const void *src, size_t bytes);
int  pu_memcpy_d2d(void *dst, const void *src, size_t bytes);   /* device deep_copy_to; a plain copy kernel */
int  pu_memset_device(void *ptr, int value, size_t bytes);
int  pu_fill_device(double *ptr, double value, size_t count);   /* YAKL scalar assignment (arr = v) */
int  pu_synchronize(pu_ctx *ctx);            /* Kokkos::fence() on the context's stream */

/* Seeded synthetic initial state (the role of custom_modules::sc_init + sc_perturb for the
 * benchmark cases, sc_init.h:14-768, sc_perturb.h:13-136), generated on the device.
 * case_id: PU_CASE_BUBBLE (dry warm bubble), PU_CASE_ABL_URBAN (neutral ABL over an 8x8 array of
 * cuboid buildings), PU_CASE_WINDFARM (ABL with capping inversion).  immersed_out receives
 * immersed_proportion. */
enum { PU_CASE_BUBBLE = 1, PU_CASE_ABL_URBAN = 2, PU_CASE_WINDFARM = 3 };
int  pu_case_init(const pu_config *cfg, int case_id, unsigned long long seed,
                  pu_coupler_state *cs, double *immersed_out);

/* Timing on the context's stream with HIP events.  pu_timer_* brackets a region;
 * pu_kernel_timing turns on per-launch events around the named kernel family ("weno_x",
 * "weno_y", "weno_z", "prep", "combine", "les") and pu_kernel_time reads the accumulated
 * milliseconds and launch count. */
int  pu_timer_start(pu_ctx *ctx);
int  pu_timer_stop(pu_ctx *ctx, double *ms);
int  pu_kernel_timing(pu_ctx *ctx, int enable);
int  pu_kernel_time(pu_ctx *ctx, const char *name, double *total_ms, long long *count);

/* Benchmark driver on the drop-in C++ API (core::Coupler + modules), the analog of the reference's
 * experiments/examples/abl_neutral.cpp:65-179 main loop: constants as sc_init.h:31-51, tracers
 * water_vapor (+ TKE from LES_Closure::init when closure != 0, + passive_1/2 for the ABL cases),
 * pu_case_init, dycore init; pu_driver_step runs, with dt = compute_time_step * dyn_cycle, nsteps times:
 * run_module(geostrophic_wind_forcing) and run_module(uniform_pg_wind_forcing[_height]) when
 * enabled, run_module(dycore.time_step), run_module(sponge_layer) when enabled,
 * run_module(apply_surface_fluxes) for the ABL cases and run_module(les_closure.apply): the module
 * order of abl_neutral.cpp:98-103 and city.cpp:130-141.  sgs_model (NULL or a network file, with
 * closure != 0) sets option les_sgs_model: the learned subgrid closure. */
typedef struct pu_driver pu_driver;
int  pu_driver_create(int case_id, int nx, int ny, int nz, double dx, int ord, int precision,
                      unsigned long long seed, int closure, const char *sgs_model, pu_driver **out);
/* Column forcing of the driver step (off by default).  geostrophic: geostrophic_wind_forcing(c, dt,
 * lat_g, u_g, v_g) as abl_neutral.cpp:93; sponge: sponge_layer(c, dt, sponge_dt_multiple * dt,
 * sponge_top_prop) as abl_neutral.cpp:95 (100, 0.1); pg = 1: uniform_pg_wind_forcing(c, dt, pg_tau,
 * pg_update_cycles), pg = 2: uniform_pg_wind_forcing_height(c, dt, pg_height, pg_u0, pg_v0, pg_tau)
 * as city.cpp:131-136. */
typedef struct {
  int    geostrophic;
  double lat_g, u_g, v_g;
  int    sponge;
  double sponge_dt_multiple, sponge_top_prop;
  int    pg;
  double pg_height, pg_u0, pg_v0, pg_tau;
  unsigned long pg_update_cycles;
} pu_driver_forcing;
int  pu_driver_set_forcing(pu_driver *d, const pu_driver_forcing *f);
/* Coupler::write_output_file(prefix) (coupler.h:533-668): writes <prefix>_<counter>.nc (name
 * returned in path_out) and advances the file counter; double_precision 0 writes the state as
 * float like the reference (option output_precision).  pu_driver_restart sets option
 * restart_file and runs Coupler::overwrite_with_restart (coupler.h:673-723). */
int  pu_driver_write_output(pu_driver *d, const char *prefix, int double_precision, char *path_out, int path_len);
int  pu_driver_restart(pu_driver *d, const char *path);
/* Actuator-disk turbines in the driver step (after the surface fluxes, before the closure, as
 * windfarm_ensemble.cpp:226-229): modules::WindmillActuators with option turbine_file, a farm of
 * ncols_x x nrows_y turbines at x = (i+1) xlen/(ncols_x+1), y = (j+1/2) ylen/nrows_y, fixed yaw 0,
 * upstream direction 0, no blades; orig_C_T sets option turbine_orig_C_T.  Needs the closure. */
int  pu_driver_set_turbines(pu_driver *d, const char *turbine_file, int ncols_x, int nrows_y, int orig_C_T);
/* The physics step of pu_driver_step: dt = compute_time_step * dyn_cycle (default 1), as the
 * reference's ABL drivers set it (abl_neutral.cpp:27,86, city.cpp:52,123: dyn_cycle = 10).  The
 * dycore runs dyn_cycle SSPRK3 cycles per step (dynamics_rk.h:103-137), the forcing, surface
 * fluxes, turbines and closure once with the whole step. */
int  pu_driver_set_dyn_cycle(pu_driver *d, int dyn_cycle);
pu_turbines *pu_driver_turbines(pu_driver *d);
/* option elapsed_time of the driver's coupler (s) */
double pu_driver_elapsed_time(pu_driver *d);
int  pu_driver_step(pu_driver *d, int nsteps);
int  pu_driver_state(pu_driver *d, pu_coupler_state *out, int *num_tracers);
pu_ctx *pu_driver_context(pu_driver *d);
int  pu_driver_tracer_name(pu_driver *d, int index, char *buf, int buflen);
const char *pu_driver_last_error(void);
int  pu_driver_destroy(pu_driver *d);

/* ------------------------------------------------------------------------------------------
 * Learned subgrid closure.  No counterpart in the reference (its nearest prior art is the
 * explicit-vs-closure SGS comparison in build/postproc/compare_les.py:11-60 and the empty
 * external/ponni inference submodule, model/CMakeLists.txt:22).  A per-cell network
 * 108 -> 64 -> 64 -> 9 (ReLU) maps the differences of (u, v, w, theta) over a cell's 3x3x3
 * neighbourhood to the subgrid stresses tau_11 tau_12 tau_13 tau_22 tau_23 tau_33 and heat
 * fluxes q_1 q_2 q_3; attached to a context it replaces LES_Closure's eddy-viscosity momentum and
 * theta face fluxes (les_closure.h:92-229) on every face whose two cells hold the network's flag,
 * and keeps them elsewhere (the fallback: wall levels, immersed neighbourhoods, inputs outside the
 * training range, negative predicted dissipation, effective viscosity above the closure's cap).
 * Training runs on the device: BF16 matrix-core forward/backward with FP32 accumulation,
 * FP32 master weights, Adam.
 * ------------------------------------------------------------------------------------------ */
#define PU_SGS_NIN 108
#define PU_SGS_NHIDDEN 64
#define PU_SGS_NOUT 9
/* parameter vector: W1[64][108], b1[64], W2[64][64], b2[64], W3[9][64], b3[9] (row = output) */
#define PU_SGS_NPARAM 11721
typedef struct pu_mlp pu_mlp;

/* Network with zero parameters and unit normalisation. */
int  pu_mlp_create(pu_mlp **out);
int  pu_mlp_destroy(pu_mlp *m);
/* Host parameter vectors (PU_SGS_NPARAM floats).  set_params also resets the Adam state. */
int  pu_mlp_set_params(pu_mlp *m, const float *params);
int  pu_mlp_get_params(const pu_mlp *m, float *params);
/* Input/output normalisation: input f is (neighbour - centre) * feat_scale[f] / delta;
 * |input| > xmax marks a cell out of range; tau_o = out_mean[o] + out_scale[o] * y_o. */
int  pu_mlp_set_normalisation(pu_mlp *m, const float *feat_scale, float xmax, const float *out_mean,
                              const float *out_scale);
int  pu_mlp_get_normalisation(const pu_mlp *m, float *feat_scale, float *xmax, float *out_mean,
                              float *out_scale);
int  pu_mlp_save(const pu_mlp *m, const char *path);
int  pu_mlp_load(const char *path, pu_mlp **out);
/* Device inputs: X [N][112] bfloat16 bit patterns (inputs 108..111 zero), T [N][9] float
 * (normalised targets), idx [n] int sample rows of this batch (NULL: rows 0..n-1). */
int  pu_mlp_pack_inputs(const float *X, long long n, unsigned short *Xbf);   /* [n][108] f32 -> bf16 */
int  pu_mlp_forward(pu_mlp *m, const unsigned short *Xbf, const int *idx, int n, float *Y);   /* Y [n][9] */
/* Mean-squared loss over the batch and its gradient (device out, PU_SGS_NPARAM floats). */
int  pu_mlp_loss_grad(pu_mlp *m, const unsigned short *Xbf, const float *T, const int *idx, int n,
                      float *grad, float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the batch gradient; loss optional. */
int  pu_mlp_train_step(pu_mlp *m, const unsigned short *Xbf, const float *T, const int *idx, int n, float lr,
                       float *loss);
/* One Adam step with a given device gradient (PU_SGS_NPARAM floats). */
int  pu_mlp_adam_apply(pu_mlp *m, const float *grad, float lr);

/* Box filter of a fine run held by `fine` (its config and immersed_proportion) by r in each
 * direction: les [6][nz/r+2][ny/r+2][nx/r+2] double (rho u v w theta k_sgs, halo 1: x/y periodic,
 * z copies with w = 0), imm [nz/r+2][ny/r+2][nx/r+2] double (block max, z halos 1),
 * targets [9][nz/r][ny/r][nx/r] float (tau_ij, q_j).  Device pointers. */
int  pu_sgs_filter(pu_ctx *fine, const pu_coupler_state *cs, int r, double *les, double *imm, float *targets);
/* Inputs of the listed cells (interior ids of a halo-1 box les [>=5][nz+2][ny+2][nx+2]):
 * Xbf [n][112] bfloat16; fmul [108] host multipliers (feat_scale / delta). */
int  pu_sgs_features(const double *les, int nx, int ny, int nz, const float *fmul, const long long *cells,
                     long long n, unsigned short *Xbf);
/* Attach (m != NULL) or detach the network to a context's LES closure.  The context reads the
 * network's current parameters at every apply. */
int  pu_les_closure_set_model(pu_ctx *ctx, const pu_mlp *m);
/* The network's prediction for the coupler state: tau [9][nz][ny][nx] float and flags
 * [nz][ny][nx] (bit 0 momentum, bit 1 heat from the network), device outputs. */
int  pu_les_closure_sgs(pu_ctx *ctx, const pu_coupler_state *cs, double dt_phys, float *tau, unsigned char *flags);

/* Ceilings for the rooflines of this library's kernels, measured on the current device (no
 * counterpart in the reference).  pu_measure_fma_rate: the vector FMA rate (TFLOP/s, an FMA counted
 * as 2) of independent FMA chains in FP64 or non-packed FP32 (PU_FP64 / PU_FP32).
 * pu_measure_copy_rate: the HBM rate (GB/s, bytes read + written) of the plain streaming copy that
 * pu_memcpy_d2d runs, on two device buffers of `bytes` (a multiple of 16, >= 1 MiB) that it
 * allocates and frees.  Both synchronise the device. */
int  pu_measure_fma_rate(int precision, double *tflops);
int  pu_measure_copy_rate(size_t bytes, double *gbps);

/* Library build information (gfx target, precision variants compiled). */
const char *pu_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* PORTURB_H */
