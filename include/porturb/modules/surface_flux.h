// modules::apply_surface_fluxes -- drop-in for model/modules/surface_flux.h:9-126.
//
// Reads the DataManager entries immersed_roughness, immersed_temp ({nz,ny,nx}) and
// surface_roughness, surface_temp ({ny,nx}) that custom_modules::sc_init registers
// (sc_init.h:63-88); the halo'd boxes the reference reads (immersed_*_halos, sc_init.h:728-767)
// are built inside the shared libporturb context on the first call; an immersed_proportion_halos
// entry, when the driver holds one, is read at every call as the reference reads it.  The log-law
// drag and temperature exchange then run in libporturb's kernel on uvel, vvel, wvel, temp.
#pragma once

#include "porturb/core/coupler.h"
#include "porturb/modules/porturb_context.h"

namespace modules {

inline void apply_surface_fluxes(core::Coupler &coupler, core::real dt) {
#ifndef PORTURB_ORD
  constexpr int ord = 11;
#else
  constexpr int ord = PORTURB_ORD;
#endif
  auto &ctx = detail::get_context(coupler, ord);
  pu_coupler_state cs = coupler.coupler_state();
  auto &dm = coupler.get_data_manager_readwrite();
  if (!ctx.dycore_ready && !ctx.les_ready) {
    if (!dm.entry_exists("immersed_proportion"))
      core::endrun("ERROR: apply_surface_fluxes needs the immersed_proportion entry");
    detail::check(pu_les_closure_init(ctx.ctx, &cs));
    ctx.les_ready = true;
  }
  if (!ctx.sfc_ready) {
    for (const char *name : {"immersed_roughness", "immersed_temp", "surface_roughness", "surface_temp"})
      if (!dm.entry_exists(name)) core::endrun(std::string("ERROR: apply_surface_fluxes needs the entry ") + name);
    detail::check(pu_surface_fluxes_init(ctx.ctx, dm.get<core::real const, 3>("immersed_roughness").data(),
                                         dm.get<core::real const, 3>("immersed_temp").data(),
                                         dm.get<core::real const, 2>("surface_roughness").data(),
                                         dm.get<core::real const, 2>("surface_temp").data()));
    ctx.sfc_ready = true;
  }
  // a driver-held immersed_proportion_halos entry (surface_flux.h:28 reads it) replaces the context's
  if (dm.entry_exists("immersed_proportion_halos"))
    detail::check(pu_les_closure_set_immersed_halos(ctx.ctx, dm.get<core::real const, 3>("immersed_proportion_halos").data()));
  detail::check(pu_surface_fluxes_apply(ctx.ctx, &cs, dt));
}

}  // namespace modules
