// modules::WindmillActuators -- drop-in for model/modules/windmill_actuators_yaw.h:12-1065 with
// actuator disks (the configuration of the reference's wind-farm drivers, windfarm_ensemble.cpp:
// 106-110: turbine_do_blades false, no floating motions, no immersed hub/tower material).
//
// init reads the turbine file of option turbine_file (RefTurbine::init, :34-79: the numeric
// entries velocity_magnitude, thrust_coef, power_coef, power_megawatts, hub_height, blade_radius,
// max_yaw_speed [deg/s, default 0.5], overhang [-0.1 R], hub_radius [0.03 R], shaft_tilt_deg [0];
// option override_shaft_tilt_deg replaces the tilt as the reference does), places the turbines
// from options turbine_x_locs / turbine_y_locs or on the reference's regular grid (:366-393),
// reads turbine_apply_thrust, turbine_initial_yaw, turbine_fixed_yaw, turbine_orig_C_T,
// turbine_upstream_dir, registers windmill_proj_weight / windmill_samp_weight (output variables)
// and an output hook with the turbines' traces.  apply runs libporturb's kernels
// (pu_turbines_apply) on uvel, vvel and TKE.  A configuration that needs blades (a rotation_rpm
// table with turbine_do_blades and dx (63/R) < 16), floating motions or immersed material ends
// with endrun.
#pragma once

#include <cmath>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "porturb/core/coupler.h"
#include "porturb/modules/porturb_context.h"

namespace modules {

namespace detail {

// Numeric entries of a YAML turbine file: "key: number" and "key: [n, n, ...]" (flow sequences may
// span lines).  Entries holding nested lists or strings (the airfoil tables) are skipped.
inline std::map<std::string, std::vector<double>> read_turbine_yaml(std::string const &fname) {
  std::ifstream in(fname);
  if (!in) core::endrun("ERROR: Invalid turbine input file " + fname);
  std::map<std::string, std::vector<double>> out;
  std::string line, key, value;
  int depth = 0;
  auto finish = [&]() {
    if (key.empty()) return;
    std::string v = value;
    bool nested = false, list = false;
    int d = 0;
    for (char &ch : v) {
      if (ch == '[') { d++; if (d > 1) nested = true; list = true; ch = ' '; }
      else if (ch == ']') { d--; ch = ' '; }
      else if (ch == ',') ch = ' ';
      else if (ch == '"' || ch == '\'') nested = true;
    }
    if (!nested) {
      std::istringstream ss(v);
      std::vector<double> vals;
      std::string tok;
      bool ok = true;
      while (ss >> tok) {
        char *end = nullptr;
        double x = std::strtod(tok.c_str(), &end);
        if (!end || *end != '\0') { ok = false; break; }
        vals.push_back(x);
      }
      if (ok && !vals.empty() && (list || vals.size() == 1)) out[key] = vals;
    }
    key.clear();
    value.clear();
  };
  while (std::getline(in, line)) {
    auto hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    if (depth == 0) {
      auto colon = line.find(':');
      if (colon == std::string::npos) continue;
      finish();
      key = line.substr(0, colon);
      key.erase(0, key.find_first_not_of(" \t"));
      key.erase(key.find_last_not_of(" \t") + 1);
      line = line.substr(colon + 1);
    }
    value += " " + line;
    for (char ch : line) depth += ch == '[' ? 1 : (ch == ']' ? -1 : 0);
    if (depth == 0) finish();
  }
  finish();
  return out;
}

}  // namespace detail

struct WindmillActuators {
  using real = core::real;
  struct Handle {
    pu_turbines *t = nullptr;
    int nturb = 0;
    ~Handle() { if (t) pu_turbines_destroy(t); }
  };
  std::shared_ptr<Handle> h;

  void init(core::Coupler &coupler) {
#ifndef PORTURB_ORD
    constexpr int ord = 11;
#else
    constexpr int ord = PORTURB_ORD;
#endif
    auto &ctx = detail::get_context(coupler, ord);
    auto tab = detail::read_turbine_yaml(coupler.get_option<std::string>("turbine_file"));
    auto need = [&](const char *k) -> std::vector<double> & {
      if (!tab.count(k)) core::endrun(std::string("ERROR: turbine file lacks ") + k);
      return tab[k];
    };
    auto scalar = [&](const char *k, double dflt) { return tab.count(k) ? tab[k][0] : dflt; };
    std::vector<double> &vm = need("velocity_magnitude"), &ct = need("thrust_coef"), &cp = need("power_coef"),
                        &pw = need("power_megawatts");
    if (vm.size() != ct.size() || vm.size() != cp.size() || vm.size() != pw.size())
      core::endrun("ERROR: turbine arrays not all the same size");
    const double rad = need("blade_radius")[0];
    pu_turbine_ref ref{};
    ref.ntable = (int)vm.size();
    ref.velocity_magnitude = vm.data(); ref.thrust_coef = ct.data(); ref.power_coef = cp.data();
    ref.power_megawatts = pw.data();
    ref.hub_height = need("hub_height")[0];
    ref.blade_radius = rad;
    ref.max_yaw_speed = scalar("max_yaw_speed", 0.5) / 180. * M_PI;
    ref.overhang = scalar("overhang", -0.1 * rad);
    ref.hub_radius = scalar("hub_radius", 0.03 * rad);
    ref.shaft_tilt = scalar("shaft_tilt_deg", 0) / 180. * M_PI;
    if (coupler.option_exists("override_shaft_tilt_deg")) ref.shaft_tilt = coupler.get_option<real>("override_shaft_tilt_deg");
    const double dx = coupler.get_dx();
    if (tab.count("rotation_rpm") && coupler.get_option<bool>("turbine_do_blades", true) && dx * (63 / rad) < 16)
      core::endrun("ERROR: WindmillActuators: blade projection is not supported; set option turbine_do_blades to false");
    if (coupler.get_option<bool>("turbine_floating_motions", false))
      core::endrun("ERROR: WindmillActuators: floating motions are not supported");
    if (coupler.get_option<bool>("turbine_immerse_material", false))
      core::endrun("ERROR: WindmillActuators: immersed hub/tower material is not supported");
    // placement (:366-393)
    std::vector<double> xs, ys;
    if (coupler.option_exists("turbine_x_locs") && coupler.option_exists("turbine_y_locs")) {
      xs = coupler.get_option<std::vector<double>>("turbine_x_locs");
      ys = coupler.get_option<std::vector<double>>("turbine_y_locs");
      if (xs.size() != ys.size()) core::endrun("ERROR: turbine_x_locs and turbine_y_locs differ in length");
    } else {
      const double xlen = coupler.get_xlen(), ylen = coupler.get_ylen();
      const int num_x = (int)std::round(xlen / 10 / (2 * rad)), num_y = (int)std::round(ylen / 10 / (2 * rad));
      const double xinc = xlen / num_x, yinc = ylen / num_y;
      for (double y = yinc / 2; y < ylen; y += yinc)
        for (double x = xinc / 2; x < xlen; x += xinc) { xs.push_back(x); ys.push_back(y); }
    }
    std::vector<unsigned char> thrust(xs.size(), 1);
    if (coupler.option_exists("turbine_apply_thrust")) {
      auto at = coupler.get_option<std::vector<bool>>("turbine_apply_thrust");
      for (size_t i = 0; i < thrust.size() && i < at.size(); i++) thrust[i] = at[i];
    }
    pu_turbine_opts o{};
    o.initial_yaw = coupler.get_option<real>("turbine_initial_yaw", 0.);
    o.fixed_yaw = coupler.get_option<bool>("turbine_fixed_yaw", false);
    o.orig_C_T = coupler.get_option<bool>("turbine_orig_C_T", false);
    o.have_upstream_dir = coupler.option_exists("turbine_upstream_dir");
    o.upstream_dir = o.have_upstream_dir ? coupler.get_option<real>("turbine_upstream_dir") : 0.;
    h = std::make_shared<Handle>();
    h->nturb = (int)xs.size();
    detail::check(pu_turbines_create(ctx.ctx, &ref, h->nturb, xs.data(), ys.data(), thrust.data(), &o, &h->t));
    auto &dm = coupler.get_data_manager_readwrite();
    const int nx = coupler.get_nx(), ny = coupler.get_ny(), nz = coupler.get_nz();
    for (const char *n : {"windmill_proj_weight", "windmill_samp_weight"}) {
      if (!dm.entry_exists(n)) dm.register_and_allocate<real>(n, "", {nz, ny, nx}, {"z", "y", "x"});
      coupler.register_output_variable<real>(n, core::Coupler::DIMS_3D);
    }
    // traces in the output file (:401-517); the floating-motion traces are zero as in the
    // reference without floating motions, mag195 (an input of the floating model) is not written
    std::shared_ptr<Handle> hh = h;
    coupler.register_write_output_module([hh](core::Coupler &, core::NcFile &nc) {
      if (hh->nturb == 0) return;
      pu_turbine_status st;
      detail::check(pu_turbines_status(hh->t, 0, &st));
      if (st.nsteps == 0) return;
      nc.redef();
      nc.create_dim("num_time_steps", st.nsteps);
      const char *names[PU_TURB_NTRACE] = {"power_trace_turb_", "yaw_trace_turb_", "u_samp_trace_turb_",
                                           "v_samp_trace_turb_", "cp_trace_turb_", "ct_trace_turb_"};
      const char *zeros[6] = {"betti_trace_turb_", "surge_pos_trace_turb_", "surge_vel_trace_turb_",
                              "heave_pos_trace_turb_", "heave_vel_trace_turb_", "pitch_pos_trace_turb_"};
      for (int i = 0; i < hh->nturb; i++) {
        for (auto n : names) nc.create_var<real>(n + std::to_string(i), {"num_time_steps"});
        for (auto n : zeros) nc.create_var<real>(n + std::to_string(i), {"num_time_steps"});
        nc.create_var<real>("pitch_vel_trace_turb_" + std::to_string(i), {"num_time_steps"});
      }
      nc.enddef();
      std::vector<double> v(st.nsteps), z(st.nsteps, 0.);
      for (int i = 0; i < hh->nturb; i++) {
        for (int w = 0; w < PU_TURB_NTRACE; w++) {
          detail::check(pu_turbines_trace(hh->t, i, w, v.data()));
          if (w == PU_TURB_YAW) for (auto &y : v) y = y / M_PI * 180;
          nc.write(v, names[w] + std::to_string(i));
        }
        for (auto n : zeros) nc.write(z, n + std::to_string(i));
        nc.write(z, "pitch_vel_trace_turb_" + std::to_string(i));
      }
    });
  }

  void apply(core::Coupler &coupler, real dt) {
    if (!h) core::endrun("ERROR: WindmillActuators::init has not been called");
    auto &dm = coupler.get_data_manager_readwrite();
    pu_coupler_state cs = coupler.coupler_state();
    detail::check(pu_turbines_apply(h->t, &cs, dt, dm.get<real, 3>("windmill_proj_weight").data(),
                                    dm.get<real, 3>("windmill_samp_weight").data()));
  }

  pu_turbines *handle() const { return h ? h->t : nullptr; }
};

}  // namespace modules
