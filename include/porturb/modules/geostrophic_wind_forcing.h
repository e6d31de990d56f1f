// modules::geostrophic_wind_forcing -- drop-in for model/modules/geostrophic_wind_forcing.h:7-36.
//
// Same signature: Coriolis turning of the horizontal-mean wind of each level against the
// geostrophic wind (u_g, v_g) at latitude lat_g (degrees), applied to every cell of the level.
// Level means and update run in libporturb's kernels (pu_geostrophic_wind_forcing).
#pragma once

#include "porturb/core/coupler.h"
#include "porturb/modules/porturb_context.h"

namespace modules {

inline void geostrophic_wind_forcing(core::Coupler &coupler, core::real dt, core::real lat_g, core::real u_g,
                                     core::real v_g) {
#ifndef PORTURB_ORD
  constexpr int ord = 11;
#else
  constexpr int ord = PORTURB_ORD;
#endif
  auto &ctx = detail::get_context(coupler, ord);
  pu_coupler_state cs = coupler.coupler_state();
  detail::check(pu_geostrophic_wind_forcing(ctx.ctx, &cs, dt, lat_g, u_g, v_g));
}

}  // namespace modules
