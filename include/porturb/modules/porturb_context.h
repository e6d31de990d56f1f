// Shared libporturb context of the dycore and LES closure modules on one coupler.
//
// The reference modules are stateless structs whose persistent data live in DataManager entries
// (dynamics_rk.h:917-1246); here the persistent device data (hydrostatics, immersed masks, stage
// buffers) live in one libporturb context per coupler, created from the coupler's options and
// tracer list by the first module that needs it.
#pragma once

#include <string>

#include "porturb.h"
#include "porturb/core/coupler.h"

namespace modules {
namespace detail {

struct Context {
  pu_ctx *ctx = nullptr;
  std::string signature;      // tracer list + order + precision the context was built for
  bool dycore_ready = false, les_ready = false;
  pu_mlp *sgs = nullptr;      // learned subgrid closure (option les_sgs_model), if any
  bool sgs_checked = false;
  bool sfc_ready = false;     // surface-flux roughness/temperature boxes built
  ~Context() {
    if (ctx) pu_destroy(ctx);
    if (sgs) pu_mlp_destroy(sgs);
  }
};

inline void check(int rc) {
  if (rc != PU_OK) core::endrun(std::string("ERROR: libporturb: ") + pu_last_error());
}

inline int bc_code(std::string const &s, bool z) {
  if (s == "periodic") return PU_BC_PERIODIC;
  if (!z && s == "precursor") return PU_BC_PRECURSOR;
  if (z && s == "solid_wall") return PU_BC_SOLID_WALL;
  core::endrun("ERROR: Specified invalid bc_" + std::string(z ? "z" : "x/y") + " in coupler options: " + s);
}

// pu_config from the coupler (the Options keys dynamics_rk.h:515-534 and les_closure.h:36-51 read)
inline pu_config make_config(core::Coupler const &c, int ord) {
  pu_config p{};
  p.nx = c.get_nx(); p.ny = c.get_ny(); p.nz = c.get_nz();
  p.xlen = c.get_xlen(); p.ylen = c.get_ylen(); p.zlen = c.get_zlen();
  p.ord = c.get_option<int>("porturb_ord", ord);
  auto names = c.get_tracer_names();
  p.num_tracers = (int)names.size();
  p.idWV = c.option_exists("idWV") ? c.get_option<int>("idWV") : c.get_tracer_index("water_vapor");
  p.idTKE = c.get_tracer_index("TKE");
  p.bc_x = bc_code(c.get_option<std::string>("bc_x", "periodic"), false);
  p.bc_y = bc_code(c.get_option<std::string>("bc_y", "periodic"), false);
  p.bc_z = bc_code(c.get_option<std::string>("bc_z", "solid_wall"), true);
  p.weno_all = c.get_option<bool>("weno_all", true);
  p.enable_gravity = c.get_option<bool>("enable_gravity", true);
  p.latitude = c.get_option<double>("latitude", 0.);
  p.cfl = c.get_option<double>("cfl", 0.6);
  p.immersed_power = c.get_option<double>("immersed_power", 5.);
  p.R_d = c.get_option<double>("R_d");
  p.R_v = c.get_option<double>("R_v");
  p.cp_d = c.get_option<double>("cp_d");
  p.gamma_d = c.get_option<double>("gamma_d");
  p.C0 = c.get_option<double>("C0");
  p.grav = c.get_option<double>("grav");
  p.p0 = c.get_option<double>("p0", 1.e5);
  for (size_t t = 0; t < names.size(); t++) {
    std::string desc;
    bool found, pos, mass, diff;
    c.get_tracer_info(names[t], desc, found, pos, mass, diff);
    p.tracer_positive[t] = pos; p.tracer_adds_mass[t] = mass; p.tracer_diffuse[t] = diff;
  }
  p.kinematic_viscosity = c.get_option<double>("kinematic_viscosity", 0.);
  p.dns = c.get_option<bool>("dns", false);
  p.les_total_mult = c.get_option<double>("les_total_mult", 1.);
  p.les_dissipation_mult = c.get_option<double>("les_dissipation_mult", 1.);
  p.les_shear_prod_mult = c.get_option<double>("les_shear_prod_mult", 1.);
  std::string prec = c.get_option<std::string>("dycore_precision", "fp64");
  if (prec != "fp64" && prec != "fp32") core::endrun("ERROR: option dycore_precision must be fp64 or fp32");
  p.precision = prec == "fp32" ? PU_FP32 : PU_FP64;
  return p;
}

inline std::string signature(pu_config const &p) {
  std::string s = std::to_string(p.ord) + "/" + std::to_string(p.precision) + "/" + std::to_string(p.num_tracers);
  for (int t = 0; t < p.num_tracers; t++)
    s += "/" + std::to_string(p.tracer_positive[t]) + std::to_string(p.tracer_adds_mass[t]) +
         std::to_string(p.tracer_diffuse[t]);
  return s;
}

// The context of this coupler, created on first use.  A context is sized by the tracer list, so
// tracers must all be added before the first module initialises (as the reference requires: the
// dycore's tracer_positive / tracer_adds_mass entries are sized at init, dynamics_rk.h:936-956).
inline Context &get_context(core::Coupler &c, int ord) {
  auto &slot = c.module_state("porturb_context");
  pu_config p = make_config(c, ord);
  std::string sig = signature(p);
  if (slot) {
    Context &ctx = *static_cast<Context *>(slot.get());
    if (ctx.signature != sig)
      core::endrun("ERROR: tracers, order or precision changed after the dycore/closure context was created");
    return ctx;
  }
  auto ctx = std::make_shared<Context>();
  check(pu_create(&p, nullptr, &ctx->ctx));
  ctx->signature = sig;
  slot = ctx;
  return *ctx;
}

}  // namespace detail
}  // namespace modules
