// modules::sponge_layer -- drop-in for model/modules/sponge_layer.h:7-60.
//
// Same signature: above z1 = (1 - top_prop) zlen, uvel, vvel and temp relax towards their
// horizontal means per level and wvel towards 0, with factor ((z - z1)/(zlen - z1))^3 dt /
// time_scale.  The level means and the relaxation run in libporturb's kernels (pu_sponge_layer)
// on the coupler's entries; the means stay on the device.
#pragma once

#include "porturb/core/coupler.h"
#include "porturb/modules/porturb_context.h"

namespace modules {

inline void sponge_layer(core::Coupler &coupler, core::real dt, core::real time_scale, core::real top_prop) {
#ifndef PORTURB_ORD
  constexpr int ord = 11;
#else
  constexpr int ord = PORTURB_ORD;
#endif
  auto &ctx = detail::get_context(coupler, ord);
  pu_coupler_state cs = coupler.coupler_state();
  detail::check(pu_sponge_layer(ctx.ctx, &cs, dt, time_scale, top_prop));
}

}  // namespace modules
