// modules::uniform_pg_wind_forcing and uniform_pg_wind_forcing_height -- drop-ins for
// model/modules/uniform_pg_wind_forcing.h:5-80.
//
// uniform_pg_wind_forcing: every update_cycles calls the domain-mean wind is measured and kept in
// the coupler options uniform_pg_wind_forcing_{u,v} (the first call also stores it as the target
// uniform_pg_wind_forcing_{u0,v0}); each call adds dt/tau (target - mean) to every cell and counts
// calls in uniform_pg_wind_forcing_counter (:16-36).
// uniform_pg_wind_forcing_height: the mean wind at `height`, linear between the two levels around
// it (the last level whose centre lies below `height` and the next; levels 0 and 1, extrapolated,
// when `height` is above every centre), is driven towards (u0, v0) (:41-79).  The level means come from libporturb's reductions (pu_level_mean_wind), the update
// from pu_add_uniform_wind.
#pragma once

#include <cstddef>
#include <vector>

#include "porturb/core/coupler.h"
#include "porturb/modules/porturb_context.h"

namespace modules {

namespace detail {
inline Context &forcing_context(core::Coupler &coupler) {
#ifndef PORTURB_ORD
  constexpr int ord = 11;
#else
  constexpr int ord = PORTURB_ORD;
#endif
  return get_context(coupler, ord);
}
}  // namespace detail

inline void uniform_pg_wind_forcing(core::Coupler &coupler, core::real dt, core::real tau = 10,
                                    size_t update_cycles = 1) {
  auto &ctx = detail::forcing_context(coupler);
  pu_coupler_state cs = coupler.coupler_state();
  const int nz = coupler.get_nz();
  auto counter = coupler.get_option<size_t>("uniform_pg_wind_forcing_counter", 0);
  if (counter % update_cycles == 0) {
    std::vector<double> um(nz), vm(nz);
    detail::check(pu_level_mean_wind(ctx.ctx, &cs, 0, nz, um.data(), vm.data()));
    double u = 0, v = 0;
    for (int k = 0; k < nz; k++) { u += um[k]; v += vm[k]; }
    u /= nz;
    v /= nz;
    coupler.set_option<core::real>("uniform_pg_wind_forcing_u", u);
    coupler.set_option<core::real>("uniform_pg_wind_forcing_v", v);
    if (counter == 0) {
      coupler.set_option<core::real>("uniform_pg_wind_forcing_u0", u);
      coupler.set_option<core::real>("uniform_pg_wind_forcing_v0", v);
    }
  }
  auto u0 = coupler.get_option<core::real>("uniform_pg_wind_forcing_u0");
  auto v0 = coupler.get_option<core::real>("uniform_pg_wind_forcing_v0");
  auto u = coupler.get_option<core::real>("uniform_pg_wind_forcing_u");
  auto v = coupler.get_option<core::real>("uniform_pg_wind_forcing_v");
  detail::check(pu_add_uniform_wind(ctx.ctx, &cs, dt / tau * (u0 - u), dt / tau * (v0 - v)));
  counter++;
  coupler.set_option<size_t>("uniform_pg_wind_forcing_counter", counter);
}

inline void uniform_pg_wind_forcing_height(core::Coupler &coupler, core::real dt, core::real height, core::real u0,
                                           core::real v0, core::real tau = 10) {
  auto &ctx = detail::forcing_context(coupler);
  pu_coupler_state cs = coupler.coupler_state();
  const auto dz = coupler.get_dz();
  const int nz = coupler.get_nz();
  int k1 = 0;
  for (int k = 0; k < nz; k++) {
    if ((k + 0.5) * dz > height) { k1 = k - 1; break; }
  }
  // a height below the first cell centre would read the level under the domain (k1 = -1)
  if (k1 < 0 || k1 + 1 >= nz)
    core::endrun("ERROR: uniform_pg_wind_forcing_height: height must lie above the first cell centre (and nz >= 2)");
  double um[2], vm[2];
  detail::check(pu_level_mean_wind(ctx.ctx, &cs, k1, k1 + 2, um, vm));
  const double z1 = (k1 + 0.5) * dz, z2 = (k1 + 1.5) * dz;
  const double w1 = (z2 - height) / dz, w2 = (height - z1) / dz;
  const double u = w1 * um[0] + w2 * um[1];
  const double v = w1 * vm[0] + w2 * vm[1];
  detail::check(pu_add_uniform_wind(ctx.ctx, &cs, dt / tau * (u0 - u), dt / tau * (v0 - v)));
}

}  // namespace modules
