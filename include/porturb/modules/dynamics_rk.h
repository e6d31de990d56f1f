// modules::Dynamics_Euler_Stratified_WenoFV -- drop-in for model/modules/dynamics_rk.h:25-1355.
//
// Same class name, constants and methods: compute_time_step (:57-64), time_step (:103-137),
// init (:917-1246).  The work runs in libporturb's gfx950 kernels through the C ABI; the
// coupler's DataManager keeps owning every state entry.  init registers the entries the reference
// registers for other modules to read: tracer_adds_mass, tracer_positive (as uint8) and the
// hydrostatic profiles hy_{dens,theta,pressure}_{cells,edges} (as double: this build keeps the
// hydrostatic background in FP64).
#pragma once

#include <vector>

#include "porturb/core/coupler.h"
#include "porturb/modules/porturb_context.h"

namespace modules {

struct Dynamics_Euler_Stratified_WenoFV {
#ifndef PORTURB_ORD
  static constexpr int ord = 11;
#else
  static constexpr int ord = PORTURB_ORD;
#endif
  static constexpr int hs = (ord - 1) / 2;
  static constexpr int num_state = 5;
  static constexpr int idR = 0, idU = 1, idV = 2, idW = 3, idT = 4, idP = 5;
  using real = core::real;

  real compute_time_step(core::Coupler const &coupler) const {
    real dx = coupler.get_dx(), dy = coupler.get_dy(), dz = coupler.get_dz();
    real constexpr maxwave = 350 + 100;
    real cfl = coupler.get_option<real>("cfl", 0.60);
    return cfl * std::fmin(std::fmin(dx, dy), dz) / maxwave;
  }

  void init(core::Coupler &coupler) const {
    auto &dm = coupler.get_data_manager_readwrite();
    if (!dm.entry_exists("immersed_proportion"))
      core::endrun("ERROR: Dynamics_Euler_Stratified_WenoFV::init needs the immersed_proportion entry");
    auto &ctx = detail::get_context(coupler, ord);
    pu_coupler_state cs = coupler.coupler_state();
    detail::check(pu_dycore_init(ctx.ctx, &cs));
    ctx.dycore_ready = true;
    // entries the reference registers at init (dynamics_rk.h:936-956, 959-984, 1145-1160)
    int ntr = coupler.get_num_tracers();
    auto names = coupler.get_tracer_names();
    std::vector<unsigned char> pos(ntr), mass(ntr);
    for (int t = 0; t < ntr; t++) {
      std::string desc;
      bool found, p, m, d;
      coupler.get_tracer_info(names[t], desc, found, p, m, d);
      pos[t] = p; mass[t] = m;
    }
    auto put_u8 = [&](const char *name, std::vector<unsigned char> const &v) {
      if (!dm.entry_exists(name)) dm.register_and_allocate<unsigned char>(name, "", {(long long)v.size()});
      detail::check(pu_memcpy_h2d(dm.get<unsigned char, 1>(name).data(), v.data(), v.size()));
    };
    put_u8("tracer_adds_mass", mass);
    put_u8("tracer_positive", pos);
    int nz = coupler.get_nz();
    int hsr = (coupler.get_option<int>("porturb_ord", ord) - 1) / 2;
    std::vector<double> c[3], e[3];
    for (int w = 0; w < 3; w++) { c[w].resize(nz + 2 * hsr); e[w].resize(nz + 1); }
    detail::check(pu_dycore_hydrostatics(ctx.ctx, c[0].data(), c[1].data(), c[2].data(), e[0].data(),
                                         e[1].data(), e[2].data()));
    const char *cn[3] = {"hy_dens_cells", "hy_theta_cells", "hy_pressure_cells"};
    const char *en[3] = {"hy_dens_edges", "hy_theta_edges", "hy_pressure_edges"};
    for (int w = 0; w < 3; w++) {
      if (!dm.entry_exists(cn[w])) dm.register_and_allocate<double>(cn[w], "", {(long long)c[w].size()});
      if (!dm.entry_exists(en[w])) dm.register_and_allocate<double>(en[w], "", {(long long)e[w].size()});
      detail::check(pu_memcpy_h2d(dm.get<double, 1>(cn[w]).data(), c[w].data(), c[w].size() * sizeof(double)));
      detail::check(pu_memcpy_h2d(dm.get<double, 1>(en[w]).data(), e[w].data(), e[w].size() * sizeof(double)));
    }
  }

  void time_step(core::Coupler &coupler, real dt_phys) const {
    auto &ctx = detail::get_context(coupler, ord);
    if (!ctx.dycore_ready) core::endrun("ERROR: Dynamics_Euler_Stratified_WenoFV::init has not been called");
    pu_coupler_state cs = coupler.coupler_state();
    detail::check(pu_dycore_time_step(ctx.ctx, &cs, dt_phys));
  }
};

}  // namespace modules
