// core::Coupler -- the state container and module host (model/core/coupler.h:16-1171), one GPU.
//
// Keeps the reference's interface for what the dycore step, the LES closure and their drivers use:
// distribute_mpi_and_allocate_coupled_state (:132-229, here one rank: nproc_x = nproc_y = 1),
// set_grid (:329-341), the grid getters (:247-285), options (:288-326), tracers (add_tracer
// :417-447, get_tracer_names/info/index), run_module (:344-380) and the DataManager accessors.
// Output and restart files (write_output_file / overwrite_with_restart, :533-723) keep the
// reference's variables, names and hooks (register_output_variable, register_write_output_module,
// register_overwrite_with_restart_module) and write NetCDF classic through core::NcFile; option
// output_precision ("double", the default, or "float" as the reference writes) sets the type of
// the state variables.  Halo exchanges and MPI are not part of this build: on one GPU the periodic
// halos are index wraps inside the kernels.  Module contexts (the libporturb device arena of the
// dycore and closure) are kept per coupler through module_state().
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <sstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <typeinfo>
#include <vector>

#include "porturb.h"
#include "porturb/core/DataManager.h"
#include "porturb/core/NcFile.h"
#include "porturb/core/Options.h"

namespace core {

typedef double real;

class Coupler {
 public:
  static int constexpr DIMS_COLUMN = 1, DIMS_SURFACE = 2, DIMS_3D = 3;

  Coupler() = default;
  Coupler(Coupler &&) = default;
  Coupler &operator=(Coupler &&) = default;
  Coupler(Coupler const &) = delete;
  Coupler &operator=(Coupler const &) = delete;
  ~Coupler() {
    module_states.clear();      // contexts before the entries they computed on
    dm.finalize();
    options.finalize();
  }

  // one rank: the whole domain is local
  void distribute_mpi_and_allocate_coupled_state(int nz, size_t ny_glob, size_t nx_glob) {
    if (nz < 1 || ny_glob < 1 || nx_glob < 1) endrun("ERROR: grid sizes must be positive");
    this->nx_glob = nx_glob;
    this->ny_glob = ny_glob;
    dm.add_dimension("x", (long long)nx_glob);
    dm.add_dimension("y", (long long)ny_glob);
    dm.add_dimension("z", nz);
    set_option<real>("elapsed_time", 0.);
  }

  void set_grid(real xlen, real ylen, real zlen) { this->xlen = xlen; this->ylen = ylen; this->zlen = zlen; }

  real   get_xlen() const { return xlen; }
  real   get_ylen() const { return ylen; }
  real   get_zlen() const { return zlen; }
  size_t get_nx_glob() const { return nx_glob; }
  size_t get_ny_glob() const { return ny_glob; }
  int    get_nx() const { return dm.find_dimension("x") < 0 ? -1 : (int)dm.get_dimension_size("x"); }
  int    get_ny() const { return dm.find_dimension("y") < 0 ? -1 : (int)dm.get_dimension_size("y"); }
  int    get_nz() const { return dm.find_dimension("z") < 0 ? -1 : (int)dm.get_dimension_size("z"); }
  real   get_dx() const { return get_xlen() / nx_glob; }
  real   get_dy() const { return get_ylen() / ny_glob; }
  real   get_dz() const { return get_zlen() / get_nz(); }
  int    get_nranks() const { return 1; }
  int    get_myrank() const { return 0; }
  int    get_nproc_x() const { return 1; }
  int    get_nproc_y() const { return 1; }
  int    get_px() const { return 0; }
  int    get_py() const { return 0; }
  size_t get_i_beg() const { return 0; }
  size_t get_j_beg() const { return 0; }
  bool   is_sim2d() const { return ny_glob == 1; }
  bool   is_mainproc() const { return true; }
  DataManager const &get_data_manager_readonly() const { return dm; }
  DataManager &get_data_manager_readwrite() { return dm; }

  // options
  template <class T> void add_option(std::string key, T value) { options.add_option<T>(key, value); }
  template <class T> void add_option_if_empty(std::string key, T value) { if (!option_exists(key)) options.add_option<T>(key, value); }
  template <class T> void set_option(std::string key, T value) { options.set_option<T>(key, value); }
  template <class T> T get_option(std::string key) const { return options.get_option<T>(key); }
  template <class T> T get_option(std::string key, T val) const { return option_exists(key) ? options.get_option<T>(key) : val; }
  bool option_exists(std::string key) const { return options.option_exists(key); }
  void delete_option(std::string key) { options.delete_option(key); }

  // tracers (coupler.h:417-480)
  int add_tracer(std::string tracer_name, std::string tracer_desc = "", bool positive = true,
                 bool adds_mass = false, bool diffuse = true) {
    int ind = get_tracer_index(tracer_name);
    if (ind != -1) {
      Tracer const &t = tracers[ind];
      if (t.positive != positive || t.adds_mass != adds_mass || t.diffuse != diffuse)
        endrun("ERROR: adding tracer [" + tracer_name + "] that already exists with different attributes");
      return ind;
    }
    if ((int)tracers.size() >= PU_MAX_TRACERS) endrun("ERROR: too many tracers");
    dm.register_and_allocate<real>(tracer_name, tracer_desc, {get_nz(), get_ny(), get_nx()}, {"z", "y", "x"});
    tracers.push_back({tracer_name, tracer_desc, positive, adds_mass, diffuse});
    return (int)tracers.size() - 1;
  }
  std::vector<std::string> get_tracer_names() const {
    std::vector<std::string> out;
    for (auto &t : tracers) out.push_back(t.name);
    return out;
  }
  void get_tracer_info(std::string name, std::string &desc, bool &found, bool &positive, bool &adds_mass,
                       bool &diffuse) const {
    int i = get_tracer_index(name);
    found = i >= 0;
    if (found) {
      desc = tracers[i].desc; positive = tracers[i].positive; adds_mass = tracers[i].adds_mass;
      diffuse = tracers[i].diffuse;
    }
  }
  int get_tracer_index(std::string name) const {
    for (size_t i = 0; i < tracers.size(); i++) if (tracers[i].name == name) return (int)i;
    return -1;
  }
  bool tracer_exists(std::string name) const { return get_tracer_index(name) != -1; }
  int get_num_tracers() const { return (int)tracers.size(); }

  // run_module (coupler.h:344-380): wall time per module name is recorded (after a device sync)
  // when the option "function_timers" is true, as PORTURB_FUNCTION_TIMERS does with YAKL timers.
  template <class F> void run_module(F const &func, std::string name) {
    bool timers = get_option<bool>("function_timers", false);
    auto t0 = std::chrono::high_resolution_clock::now();
    func(*this);
    if (timers) {
      pu_synchronize(nullptr);
      auto t1 = std::chrono::high_resolution_clock::now();
      module_times[name] += std::chrono::duration<double>(t1 - t0).count();
    }
  }
  std::map<std::string, double> const &get_module_times() const { return module_times; }

  // Device views of the coupler state in the layout libporturb reads (DataManager entries
  // density_dry, uvel, vvel, wvel, temp, tracers, immersed_proportion).
  pu_coupler_state coupler_state() {
    pu_coupler_state cs{};
    cs.density_dry = dm.get<real, 3>("density_dry").data();
    cs.uvel = dm.get<real, 3>("uvel").data();
    cs.vvel = dm.get<real, 3>("vvel").data();
    cs.wvel = dm.get<real, 3>("wvel").data();
    cs.temp = dm.get<real, 3>("temp").data();
    auto names = get_tracer_names();
    for (size_t t = 0; t < names.size(); t++) cs.tracers[t] = dm.get<real, 3>(names[t]).data();
    if (dm.entry_exists("immersed_proportion"))
      cs.immersed_proportion = dm.get<real const, 3>("immersed_proportion").data();
    return cs;
  }

  // ---- output / restart (coupler.h:485-505, 533-723) ----
  template <class T> void register_output_variable(std::string name, int dims) {
    output_vars.push_back({name, dims, typeid(std::remove_cv_t<T>).hash_code()});
  }
  void register_write_output_module(std::function<void(Coupler &, NcFile &)> func) { out_write_funcs.push_back(func); }
  void register_overwrite_with_restart_module(std::function<void(Coupler &, NcFile &)> func) {
    restart_read_funcs.push_back(func);
  }
  int  get_file_counter() const { return file_counter; }
  void set_file_counter(int c) { file_counter = c; }
  // state variables are double unless option output_precision is "float" (the reference's type)
  bool output_double() const {
    std::string p = get_option<std::string>("output_precision", "double");
    if (p != "double" && p != "float") endrun("ERROR: option output_precision must be double or float");
    return p == "double";
  }

  // <prefix>_<file_counter, 8 digits>.nc with x, y, z, density_dry, uvel, vvel, wvel,
  // temperature, etime, file_counter, the tracers, the registered output variables and whatever
  // the write hooks add (coupler.h:533-668)
  std::string write_output_file(std::string prefix, bool verbose = true) {
    auto t0 = std::chrono::high_resolution_clock::now();
    const int nx = get_nx(), ny = get_ny(), nz = get_nz();
    const bool dbl = output_double();
    std::stringstream fname;
    fname << prefix << "_" << std::setw(8) << std::setfill('0') << file_counter << ".nc";
    NcFile nc;
    nc.create(fname.str());
    nc.create_dim("x", (long long)nx_glob);
    nc.create_dim("y", (long long)ny_glob);
    nc.create_dim("z", nz);
    nc.create_dim("t", 1);
    const std::vector<std::string> dcol = {"z"}, dsfc = {"y", "x"}, d3 = {"z", "y", "x"};
    auto state_var = [&](std::string const &name, std::vector<std::string> const &dims) {
      if (dbl) nc.create_var<double>(name, dims); else nc.create_var<float>(name, dims);
    };
    nc.create_var<float>("x", {"x"});
    nc.create_var<float>("y", {"y"});
    nc.create_var<float>("z", {"z"});
    for (const char *n : {"density_dry", "uvel", "vvel", "wvel", "temperature"}) state_var(n, d3);
    state_var("etime", {"t"});
    nc.create_var<int>("file_counter", {"t"});
    auto tracer_names = get_tracer_names();
    for (auto &t : tracer_names) state_var(t, d3);
    for (auto &ov : output_vars) {
      auto const &dims = ov.dims == DIMS_COLUMN ? dcol : (ov.dims == DIMS_SURFACE ? dsfc : d3);
      if (ov.type_hash == typeid(double).hash_code()) state_var(ov.name, dims);
      else if (ov.type_hash == typeid(float).hash_code()) nc.create_var<float>(ov.name, dims);
      else if (ov.type_hash == typeid(int).hash_code()) nc.create_var<int>(ov.name, dims);
      else if (ov.type_hash == typeid(unsigned char).hash_code()) nc.create_var<unsigned char>(ov.name, dims);
      else endrun("ERROR: output variable [" + ov.name + "] has an unsupported type");
    }
    nc.enddef();
    std::vector<float> xl(nx), yl(ny), zl(nz);
    for (int i = 0; i < nx; i++) xl[i] = (float)((i + 0.5) * get_dx());
    for (int j = 0; j < ny; j++) yl[j] = (float)((j + 0.5) * get_dy());
    for (int k = 0; k < nz; k++) zl[k] = (float)((k + 0.5) * get_dz());
    nc.write(xl, "x");
    nc.write(yl, "y");
    nc.write(zl, "z");
    nc.write(get_option<real>("elapsed_time", 0.), "etime");
    nc.write(file_counter, "file_counter");
    nc.write_all(dm.get<real const, 3>("density_dry"), "density_dry");
    nc.write_all(dm.get<real const, 3>("uvel"), "uvel");
    nc.write_all(dm.get<real const, 3>("vvel"), "vvel");
    nc.write_all(dm.get<real const, 3>("wvel"), "wvel");
    nc.write_all(dm.get<real const, 3>("temp"), "temperature");
    for (auto &t : tracer_names) nc.write_all(dm.get<real const, 3>(t), t);
    for (auto &ov : output_vars) with_entry(ov, [&](auto view) { nc.write_all(view, ov.name); });
    for (auto &f : out_write_funcs) f(*this, nc);
    nc.close();
    file_counter++;
    if (verbose) {
      std::chrono::duration<double> dt = std::chrono::high_resolution_clock::now() - t0;
      std::cout << "*** Output/restart file written ***  -->  Etime , Output time: " << std::scientific
                << get_option<real>("elapsed_time", 0.) << " , " << dt.count() << std::endl;
    }
    return fname.str();
  }

  // Reads option restart_file (coupler.h:673-723): elapsed_time, file_counter, the state, the
  // tracers, the registered output variables, then the restart hooks; any stored type is accepted.
  void overwrite_with_restart() {
    NcFile nc;
    nc.open(get_option<std::string>("restart_file"));
    double etime = 0;
    nc.read(etime, "etime");
    nc.read(file_counter, "file_counter");
    set_option<real>("elapsed_time", etime);
    nc.read_all(dm.get<real, 3>("density_dry"), "density_dry");
    nc.read_all(dm.get<real, 3>("uvel"), "uvel");
    nc.read_all(dm.get<real, 3>("vvel"), "vvel");
    nc.read_all(dm.get<real, 3>("wvel"), "wvel");
    nc.read_all(dm.get<real, 3>("temp"), "temperature");
    for (auto &t : get_tracer_names()) nc.read_all(dm.get<real, 3>(t), t);
    for (auto &ov : output_vars) with_entry_rw(ov, [&](auto view) { nc.read_all(view, ov.name); });
    for (auto &f : restart_read_funcs) f(*this, nc);
    nc.close();
    file_counter++;
  }

  // Module-held state that outlives one call (the libporturb context of the dycore/closure).
  std::shared_ptr<void> &module_state(std::string const &key) { return module_states[key]; }

  // Coupler::inform_user (coupler.h:507-530) through the device reductions of a module context.
  void inform_user(pu_ctx *ctx) {
    double mass = 0, wind = 0, w = 0;
    pu_coupler_state cs = coupler_state();
    if (pu_diagnostics(ctx, &cs, &mass, &wind, &w) != PU_OK) endrun(pu_last_error());
    auto t2 = std::chrono::high_resolution_clock::now();
    std::chrono::duration<double> dur = t2 - inform_timer;
    inform_timer = t2;
    std::printf("Etime [%10.4e s] , Walltime [%10.4e s] , max wind [%10.4e m/s] , max(abs(w)) [%10.4e m/s]\n",
                get_option<real>("elapsed_time", 0.), dur.count(), wind, w);
  }

 private:
  struct Tracer {
    std::string name, desc;
    bool positive, adds_mass, diffuse;
  };
  struct OutputVar {
    std::string name;
    int dims;
    size_t type_hash;
  };
  // f(view) for the registered entry with its element type and rank
  template <class T> struct TypeTag { using type = T; };
  template <bool RW, class DM, class F> static void visit_entry(DM &d, OutputVar const &ov, F &f) {
    auto go = [&](auto tag) {
      using T0 = typename decltype(tag)::type;
      using T = std::conditional_t<RW, T0, const T0>;
      if (ov.dims == DIMS_COLUMN) f(d.template get<T, 1>(ov.name));
      else if (ov.dims == DIMS_SURFACE) f(d.template get<T, 2>(ov.name));
      else f(d.template get<T, 3>(ov.name));
    };
    if (ov.type_hash == typeid(double).hash_code()) go(TypeTag<double>{});
    else if (ov.type_hash == typeid(float).hash_code()) go(TypeTag<float>{});
    else if (ov.type_hash == typeid(int).hash_code()) go(TypeTag<int>{});
    else go(TypeTag<unsigned char>{});
  }
  template <class F> void with_entry(OutputVar const &ov, F &&f) const { visit_entry<false>(dm, ov, f); }
  template <class F> void with_entry_rw(OutputVar const &ov, F &&f) { visit_entry<true>(dm, ov, f); }
  std::vector<OutputVar> output_vars;
  std::vector<std::function<void(Coupler &, NcFile &)>> out_write_funcs, restart_read_funcs;
  int file_counter = 0;
  Options options;
  real xlen = -1, ylen = -1, zlen = -1;
  size_t nx_glob = 0, ny_glob = 0;
  std::vector<Tracer> tracers;
  std::map<std::string, double> module_times;
  std::map<std::string, std::shared_ptr<void>> module_states;
  std::chrono::time_point<std::chrono::high_resolution_clock> inform_timer = std::chrono::high_resolution_clock::now();
  DataManager dm;   // declared last: destroyed first among members after module_states.clear()
};

}  // namespace core
