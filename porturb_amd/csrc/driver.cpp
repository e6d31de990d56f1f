// driver.cpp -- benchmark driver on the drop-in C++ API (core::Coupler + modules).
//
// Mirrors the structure of the reference's ABL drivers (experiments/examples/abl_neutral.cpp:
// 65-179): set options and constants as custom_modules::sc_init does (sc_init.h:31-51), allocate
// the coupled state, add tracers, initialise the case, init the LES closure and the dycore, then
// per step, with dt = dycore.compute_time_step * dyn_cycle: the optional column forcing
// (geostrophic_wind_forcing, uniform_pg_wind_forcing[_height]), run_module(dycore.time_step), the
// optional sponge_layer, run_module(apply_surface_fluxes) for the rough-wall ABL cases, the
// optional actuator-disk turbines and run_module(les_closure.apply) -- the reference's module
// order (abl_neutral.cpp:98-103, city.cpp:130-141, windfarm_ensemble.cpp:226-229) without its
// time averager.  Exposed through pu_driver_* for bench.py and the
// tests.
#include <cmath>
#include <memory>
#include <string>

#include "porturb.h"
#include "porturb/core/coupler.h"
#include "porturb/modules/dynamics_rk.h"
#include "porturb/modules/geostrophic_wind_forcing.h"
#include "porturb/modules/les_closure.h"
#include "porturb/modules/sponge_layer.h"
#include "porturb/modules/surface_flux.h"
#include "porturb/modules/uniform_pg_wind_forcing.h"
#include "porturb/modules/windmill_actuators.h"

struct pu_driver {
  core::Coupler coupler;
  modules::Dynamics_Euler_Stratified_WenoFV dycore;
  modules::LES_Closure les;
  bool closure = true;
  bool surface = false;     // rough-wall surface fluxes (the ABL cases)
  pu_driver_forcing forcing{};
  modules::WindmillActuators windmills;
  bool turbines = false;
  int dyn_cycle = 1;        // dycore SSPRK3 cycles per physics step (abl_neutral.cpp:27,86)
  std::string error;
};

static thread_local std::string g_driver_error;

extern "C" {

int pu_driver_create(int case_id, int nx, int ny, int nz, double dx, int ord, int precision,
                     unsigned long long seed, int closure, const char *sgs_model, pu_driver **out) {
  if (!out) return PU_ERR_ARG;
  *out = nullptr;
  try {
    auto d = std::make_unique<pu_driver>();
    core::Coupler &c = d->coupler;
    d->closure = closure != 0;
    // constants (sc_init.h:31-51)
    double R_d = 287., cp_d = 1003., R_v = 461., cp_v = 1859., p0 = 1.e5, grav = 9.81;
    double cv_d = cp_d - R_d, gamma = cp_d / cv_d, kappa = R_d / cp_d;
    c.set_option<double>("R_d", R_d); c.set_option<double>("cp_d", cp_d); c.set_option<double>("R_v", R_v);
    c.set_option<double>("cp_v", cp_v); c.set_option<double>("p0", p0); c.set_option<double>("grav", grav);
    c.set_option<double>("cv_d", cv_d); c.set_option<double>("gamma_d", gamma); c.set_option<double>("kappa_d", kappa);
    c.set_option<double>("C0", std::pow(R_d * std::pow(p0, -kappa), gamma));
    c.set_option<std::string>("bc_x", case_id == PU_CASE_WINDFARM ? "precursor" : "periodic");
    c.set_option<std::string>("bc_y", "periodic");
    c.set_option<std::string>("bc_z", "solid_wall");
    c.set_option<int>("porturb_ord", ord);
    c.set_option<std::string>("dycore_precision", precision == PU_FP32 ? "fp32" : "fp64");
    c.set_option<bool>("function_timers", false);
    if (sgs_model && sgs_model[0] && closure) c.set_option<std::string>("les_sgs_model", sgs_model);
    c.distribute_mpi_and_allocate_coupled_state(nz, (size_t)ny, (size_t)nx);
    c.set_grid(nx * dx, ny * dx, nz * dx);
    auto &dm = c.get_data_manager_readwrite();
    for (const char *name : {"density_dry", "uvel", "vvel", "wvel", "temp", "immersed_proportion"})
      dm.register_and_allocate<double>(name, "", {nz, ny, nx}, {"z", "y", "x"});
    if (case_id != PU_CASE_BUBBLE) {
      // rough wall: sc_init's roughness entries (sc_init.h:52,63-88), roughness 0.1 m, no wall heating
      d->surface = true;
      c.set_option<double>("roughness", 0.1);
      dm.register_and_allocate<double>("immersed_roughness", "", {nz, ny, nx}, {"z", "y", "x"});
      dm.register_and_allocate<double>("immersed_temp", "", {nz, ny, nx}, {"z", "y", "x"});
      dm.register_and_allocate<double>("surface_roughness", "", {ny, nx}, {"y", "x"});
      dm.register_and_allocate<double>("surface_temp", "", {ny, nx}, {"y", "x"});
      auto fill = [&](const char *name, double v) {
        auto e = dm.get_collapsed<double>(name);
        if (pu_fill_device(e.data(), v, (size_t)e.size()) != PU_OK) core::endrun(pu_last_error());
      };
      fill("immersed_roughness", 0.1);
      fill("immersed_temp", 0.);
      fill("surface_roughness", 0.1);
      fill("surface_temp", 0.);
    }
    c.add_tracer("water_vapor", "water_vapor", true, true, true);
    c.set_option<int>("idWV", c.get_tracer_index("water_vapor"));
    if (d->closure) d->les.init(c);                  // adds TKE (= 0.1)
    if (case_id != PU_CASE_BUBBLE) {
      c.add_tracer("passive_1", "passive tracer", true, false, true);
      c.add_tracer("passive_2", "passive tracer", true, false, true);
    }
    // initial state on the device (the role of sc_init + sc_perturb)
    pu_config cfg = modules::detail::make_config(c, ord);
    pu_coupler_state cs = c.coupler_state();
    double *imm = dm.get<double, 3>("immersed_proportion").data();
    if (pu_case_init(&cfg, case_id, seed, &cs, imm) != PU_OK) core::endrun(pu_last_error());
    d->dycore.init(c);
    *out = d.release();
    return PU_OK;
  } catch (std::exception const &e) {
    g_driver_error = e.what();
    return PU_ERR_ARG;
  }
}

int pu_driver_set_forcing(pu_driver *d, const pu_driver_forcing *f) {
  if (!d || !f) return PU_ERR_ARG;
  if (f->pg < 0 || f->pg > 2 || (f->pg == 1 && f->pg_update_cycles == 0) || (f->pg && !(f->pg_tau > 0)) ||
      (f->sponge && (!(f->sponge_dt_multiple > 0) || !(f->sponge_top_prop > 0) || !(f->sponge_top_prop <= 1)))) {
    g_driver_error = "pu_driver_set_forcing: invalid forcing parameters";
    return PU_ERR_ARG;
  }
  d->forcing = *f;
  return PU_OK;
}

int pu_driver_set_turbines(pu_driver *d, const char *turbine_file, int ncols_x, int nrows_y, int orig_C_T) {
  if (!d || !turbine_file || ncols_x < 1 || nrows_y < 1) return PU_ERR_ARG;
  try {
    core::Coupler &c = d->coupler;
    if (!d->closure) core::endrun("ERROR: turbines need the LES closure (TKE)");
    std::vector<double> xs, ys;
    for (int j = 0; j < nrows_y; j++)
      for (int i = 0; i < ncols_x; i++) {
        xs.push_back((i + 1) * c.get_xlen() / (ncols_x + 1));
        ys.push_back((j + 0.5) * c.get_ylen() / nrows_y);
      }
    c.set_option<std::string>("turbine_file", turbine_file);
    c.set_option<std::vector<double>>("turbine_x_locs", xs);
    c.set_option<std::vector<double>>("turbine_y_locs", ys);
    c.set_option<bool>("turbine_do_blades", false);
    c.set_option<bool>("turbine_fixed_yaw", true);
    c.set_option<double>("turbine_initial_yaw", 0.);
    c.set_option<double>("turbine_upstream_dir", 0.);
    c.set_option<bool>("turbine_orig_C_T", orig_C_T != 0);
    d->windmills.init(c);
    d->turbines = true;
    return PU_OK;
  } catch (std::exception const &e) {
    g_driver_error = e.what();
    return PU_ERR_STATE;
  }
}

int pu_driver_set_dyn_cycle(pu_driver *d, int dyn_cycle) {
  if (!d || dyn_cycle < 1) {
    g_driver_error = "pu_driver_set_dyn_cycle: dyn_cycle must be >= 1";
    return PU_ERR_ARG;
  }
  d->dyn_cycle = dyn_cycle;
  return PU_OK;
}

pu_turbines *pu_driver_turbines(pu_driver *d) { return d && d->turbines ? d->windmills.handle() : nullptr; }

int pu_driver_step(pu_driver *d, int nsteps) {
  if (!d) return PU_ERR_ARG;
  try {
    core::Coupler &c = d->coupler;
    const pu_driver_forcing &f = d->forcing;
    for (int n = 0; n < nsteps; n++) {
      // the physics step: dyn_cycle dycore steps (abl_neutral.cpp:86); the dycore splits it into
      // dyn_cycle SSPRK3 cycles, every other module runs once with the whole step
      double dt = d->dycore.compute_time_step(c) * d->dyn_cycle;
      if (f.geostrophic)
        c.run_module([&](core::Coupler &cc) { modules::geostrophic_wind_forcing(cc, dt, f.lat_g, f.u_g, f.v_g); },
                     "geostrophic_forcing");
      if (f.pg == 1)
        c.run_module([&](core::Coupler &cc) { modules::uniform_pg_wind_forcing(cc, dt, f.pg_tau, f.pg_update_cycles); },
                     "pg_forcing");
      if (f.pg == 2)
        c.run_module([&](core::Coupler &cc) {
          modules::uniform_pg_wind_forcing_height(cc, dt, f.pg_height, f.pg_u0, f.pg_v0, f.pg_tau);
        }, "pg_forcing");
      c.run_module([&](core::Coupler &cc) { d->dycore.time_step(cc, dt); }, "dycore");
      if (f.sponge)
        c.run_module([&](core::Coupler &cc) { modules::sponge_layer(cc, dt, f.sponge_dt_multiple * dt, f.sponge_top_prop); },
                     "sponge");
      if (d->surface) c.run_module([&](core::Coupler &cc) { modules::apply_surface_fluxes(cc, dt); }, "surface_fluxes");
      if (d->turbines) c.run_module([&](core::Coupler &cc) { d->windmills.apply(cc, dt); }, "windmillactuators");
      if (d->closure) c.run_module([&](core::Coupler &cc) { d->les.apply(cc, dt); }, "les_closure");
      c.set_option<double>("elapsed_time", c.get_option<double>("elapsed_time") + dt);
    }
    return PU_OK;
  } catch (std::exception const &e) {
    g_driver_error = e.what();
    return PU_ERR_STATE;
  }
}

int pu_driver_write_output(pu_driver *d, const char *prefix, int double_precision, char *path_out, int path_len) {
  if (!d || !prefix) return PU_ERR_ARG;
  try {
    d->coupler.set_option<std::string>("output_precision", double_precision ? "double" : "float");
    std::string name = d->coupler.write_output_file(prefix, false);
    if (path_out && path_len > 0) snprintf(path_out, path_len, "%s", name.c_str());
    return PU_OK;
  } catch (std::exception const &e) {
    g_driver_error = e.what();
    return PU_ERR_STATE;
  }
}

int pu_driver_restart(pu_driver *d, const char *path) {
  if (!d || !path) return PU_ERR_ARG;
  try {
    d->coupler.set_option<std::string>("restart_file", path);
    d->coupler.overwrite_with_restart();
    return PU_OK;
  } catch (std::exception const &e) {
    g_driver_error = e.what();
    return PU_ERR_STATE;
  }
}

double pu_driver_elapsed_time(pu_driver *d) {
  return d ? d->coupler.get_option<double>("elapsed_time", 0.) : 0.;
}

int pu_driver_state(pu_driver *d, pu_coupler_state *out, int *num_tracers) {
  if (!d || !out) return PU_ERR_ARG;
  *out = d->coupler.coupler_state();
  if (num_tracers) *num_tracers = d->coupler.get_num_tracers();
  return PU_OK;
}

pu_ctx *pu_driver_context(pu_driver *d) {
  if (!d) return nullptr;
  auto &slot = d->coupler.module_state("porturb_context");
  return slot ? static_cast<modules::detail::Context *>(slot.get())->ctx : nullptr;
}

int pu_driver_tracer_name(pu_driver *d, int index, char *buf, int buflen) {
  if (!d || !buf || buflen < 1) return PU_ERR_ARG;
  auto names = d->coupler.get_tracer_names();
  if (index < 0 || index >= (int)names.size()) return PU_ERR_ARG;
  snprintf(buf, buflen, "%s", names[index].c_str());
  return PU_OK;
}

const char *pu_driver_last_error(void) { return g_driver_error.c_str(); }

int pu_driver_destroy(pu_driver *d) {
  delete d;
  return PU_OK;
}

}  // extern "C"
