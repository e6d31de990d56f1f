// modules::LES_Closure -- drop-in for model/modules/les_closure.h:9-571.
//
// init (:19-26) adds the mass-weighted TKE tracer and sets it to 0.1; apply (:30-323) runs the
// 1.5-order TKE closure in libporturb's kernels on the coupler state.  Option les_sgs_model (a
// network file written by pu_mlp_save; no reference counterpart) switches on the learned subgrid
// closure for momentum and theta, with the eddy-viscosity fluxes as its fallback.
#pragma once

#include "porturb/core/coupler.h"
#include "porturb/modules/porturb_context.h"

namespace modules {

struct LES_Closure {
  static constexpr int hs = 1;
  static constexpr int num_state = 5;
  static constexpr int idR = 0, idU = 1, idV = 2, idW = 3, idT = 4;
  using real = core::real;

  void init(core::Coupler &coupler) const {
    coupler.add_tracer("TKE", "mass-weighted TKE", true, false, false);
    auto tke = coupler.get_data_manager_readwrite().get<real, 3>("TKE");
    detail::check(pu_fill_device(tke.data(), 0.1, (size_t)tke.size()));
  }

  void apply(core::Coupler &coupler, real dtphys) const {
#ifndef PORTURB_ORD
    constexpr int ord = 11;
#else
    constexpr int ord = PORTURB_ORD;
#endif
    auto &ctx = detail::get_context(coupler, ord);
    pu_coupler_state cs = coupler.coupler_state();
    if (!ctx.dycore_ready && !ctx.les_ready) {
      if (!coupler.get_data_manager_readonly().entry_exists("immersed_proportion"))
        core::endrun("ERROR: LES_Closure::apply needs the immersed_proportion entry");
      detail::check(pu_les_closure_init(ctx.ctx, &cs));
      ctx.les_ready = true;
    }
    if (!ctx.sgs_checked) {
      ctx.sgs_checked = true;
      std::string path = coupler.get_option<std::string>("les_sgs_model", "");
      if (!path.empty()) {
        detail::check(pu_mlp_load(path.c_str(), &ctx.sgs));
        detail::check(pu_les_closure_set_model(ctx.ctx, ctx.sgs));
      }
    }
    // a driver-held immersed_proportion_halos entry (les_closure.h:45 reads it at every apply; the
    // turbines add their projection weights to it) replaces the one the context built
    auto &dm = coupler.get_data_manager_readonly();
    if (dm.entry_exists("immersed_proportion_halos"))
      detail::check(pu_les_closure_set_immersed_halos(ctx.ctx, dm.get<real const, 3>("immersed_proportion_halos").data()));
    detail::check(pu_les_closure_apply(ctx.ctx, &cs, dtphys));
  }
};

}  // namespace modules
