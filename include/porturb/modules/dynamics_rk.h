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
    publish_hydrostatics(coupler, ctx.ctx);
    // output and restart (dynamics_rk.h:1190-1241)
    coupler.register_output_variable<real>("immersed_proportion", core::Coupler::DIMS_3D);
    coupler.register_write_output_module([](core::Coupler &c, core::NcFile &nc) { write_output(c, nc); });
    coupler.register_overwrite_with_restart_module([](core::Coupler &c, core::NcFile &nc) { read_restart(c, nc); });
  }

  static int halo_of(core::Coupler const &c) { return (c.get_option<int>("porturb_ord", ord) - 1) / 2; }

  // hy_{dens,theta,pressure}_{cells,edges} entries from the context's profiles
  static void publish_hydrostatics(core::Coupler &coupler, pu_ctx *ctx) {
    auto &dm = coupler.get_data_manager_readwrite();
    int nz = coupler.get_nz(), hsr = halo_of(coupler);
    std::vector<double> c[3], e[3];
    for (int w = 0; w < 3; w++) { c[w].resize(nz + 2 * hsr); e[w].resize(nz + 1); }
    detail::check(pu_dycore_hydrostatics(ctx, c[0].data(), c[1].data(), c[2].data(), e[0].data(), e[1].data(),
                                         e[2].data()));
    for (int w = 0; w < 3; w++) {
      if (!dm.entry_exists(cell_names[w])) dm.register_and_allocate<double>(cell_names[w], "", {(long long)c[w].size()});
      if (!dm.entry_exists(edge_names[w])) dm.register_and_allocate<double>(edge_names[w], "", {(long long)e[w].size()});
      detail::check(pu_memcpy_h2d(dm.get<double, 1>(cell_names[w]).data(), c[w].data(), c[w].size() * sizeof(double)));
      detail::check(pu_memcpy_h2d(dm.get<double, 1>(edge_names[w]).data(), e[w].data(), e[w].size() * sizeof(double)));
    }
  }

  // output hook (dynamics_rk.h:1192-1233): hy_*_cells on z_halo, density/theta/pressure_pert
  static void write_output(core::Coupler &c, core::NcFile &nc) {
    auto &ctx = detail::get_context(c, ord);
    const int nx = c.get_nx(), ny = c.get_ny(), nz = c.get_nz();
    nc.redef();
    nc.create_dim("z_halo", nz + 2 * halo_of(c));
    for (int w = 0; w < 3; w++) {
      if (c.output_double()) nc.create_var<double>(cell_names[w], {"z_halo"});
      else nc.create_var<float>(cell_names[w], {"z_halo"});
    }
    for (const char *n : {"theta_pert", "pressure_pert", "density_pert"}) nc.create_var<real>(n, {"z", "y", "x"});
    nc.enddef();
    auto &dm = c.get_data_manager_readonly();
    for (int w = 0; w < 3; w++) nc.write_all(dm.get<double const, 1>(cell_names[w]), cell_names[w]);
    const long long n3 = (long long)nx * ny * nz;
    void *buf = nullptr;
    detail::check(pu_alloc_device(3 * (size_t)n3 * sizeof(double), &buf));
    double *p = (double *)buf;
    pu_coupler_state cs = c.coupler_state();
    int rc = pu_dycore_output_fields(ctx.ctx, &cs, p, p + n3, p + 2 * n3);
    if (rc == PU_OK) {
      const char *names[3] = {"density_pert", "theta_pert", "pressure_pert"};
      for (int f = 0; f < 3; f++) {
        core::View<const double, 3> v;
        v.ptr = p + f * n3;
        v.dims = {nz, ny, nx};
        nc.write_all(v, names[f]);
      }
    }
    pu_free_device(buf);
    detail::check(rc);
  }

  // restart hook (dynamics_rk.h:1234-1241)
  static void read_restart(core::Coupler &c, core::NcFile &nc) {
    auto &ctx = detail::get_context(c, ord);
    const int nzh = c.get_nz() + 2 * halo_of(c);
    std::vector<double> h[3];
    for (int w = 0; w < 3; w++) {
      h[w].resize(nzh);
      nc.read(h[w], cell_names[w]);
    }
    pu_coupler_state cs = c.coupler_state();
    detail::check(pu_dycore_restart(ctx.ctx, &cs, h[0].data(), h[1].data(), h[2].data()));
    publish_hydrostatics(c, ctx.ctx);
  }

  static constexpr const char *cell_names[3] = {"hy_dens_cells", "hy_theta_cells", "hy_pressure_cells"};
  static constexpr const char *edge_names[3] = {"hy_dens_edges", "hy_theta_edges", "hy_pressure_edges"};

  void time_step(core::Coupler &coupler, real dt_phys) const {
    auto &ctx = detail::get_context(coupler, ord);
    if (!ctx.dycore_ready) core::endrun("ERROR: Dynamics_Euler_Stratified_WenoFV::init has not been called");
    pu_coupler_state cs = coupler.coupler_state();
    detail::check(pu_dycore_time_step(ctx.ctx, &cs, dt_phys));
  }
};

}  // namespace modules
