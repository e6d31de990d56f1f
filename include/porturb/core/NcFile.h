// core::NcFile -- the part of yakl::SimplePNetCDF that Coupler::write_output_file,
// Coupler::overwrite_with_restart and the module output/restart hooks use (model/core/coupler.h:
// 533-723, model/modules/dynamics_rk.h:1192-1241): create, open, create_dim, create_var<T>, redef,
// enddef, write / read (host values), write_all / read_all (device arrays of the DataManager),
// begin_indep_data / end_indep_data (no-ops on one rank), close.  Files are NetCDF classic
// (pu_nc_*, libporturb); device arrays of double are converted to the variable's type and
// byte-swapped on the GPU.
#pragma once

#include <string>
#include <type_traits>
#include <vector>

#include "porturb.h"
#include "porturb/core/DataManager.h"
#include "porturb/core/Options.h"

namespace core {

template <class T> constexpr int nc_type_of() {
  using U = std::remove_cv_t<T>;
  if constexpr (std::is_same<U, double>::value) return PU_NC_DOUBLE;
  else if constexpr (std::is_same<U, float>::value) return PU_NC_FLOAT;
  else if constexpr (std::is_same<U, int>::value) return PU_NC_INT;
  else if constexpr (std::is_same<U, unsigned char>::value) return PU_NC_UBYTE;
  else static_assert(sizeof(U) == 0, "NcFile: unsupported type");
}

class NcFile {
 public:
  NcFile() = default;
  NcFile(NcFile const &) = delete;
  NcFile &operator=(NcFile const &) = delete;
  ~NcFile() { if (f) pu_nc_close(f); }

  void create(std::string path, int format = PU_NC_AUTO) { close(); check(pu_nc_create(path.c_str(), format, &f)); }
  void open(std::string path) { close(); check(pu_nc_open(path.c_str(), &f)); }
  void close() {
    if (f) {
      pu_ncfile *g = f;
      f = nullptr;
      check(pu_nc_close(g));
    }
  }
  // the stream device transfers run on (a libporturb context; nullptr = the default stream)
  void set_context(pu_ctx *c) { ctx = c; }

  void create_dim(std::string name, long long len) { check(pu_nc_def_dim(f, name.c_str(), len)); }
  template <class T> void create_var(std::string name, std::vector<std::string> dims) {
    std::vector<const char *> d;
    for (auto &s : dims) d.push_back(s.c_str());
    check(pu_nc_def_var(f, name.c_str(), nc_type_of<T>(), (int)d.size(), d.data()));
  }
  void redef() { check(pu_nc_redef(f)); }
  void enddef() { check(pu_nc_enddef(f)); }
  void begin_indep_data() {}
  void end_indep_data() {}
  bool var_exists(std::string name) const { return pu_nc_inq_var(f, name.c_str(), nullptr, nullptr, nullptr) == PU_OK; }
  int var_type(std::string name) const {
    int t = 0;
    check(pu_nc_inq_var(f, name.c_str(), &t, nullptr, nullptr));
    return t;
  }

  // host values (scalars and vectors)
  template <class T> void write(T v, std::string name) { check(pu_nc_put_host(f, name.c_str(), &v, nc_type_of<T>())); }
  template <class T> void write(std::vector<T> const &v, std::string name) {
    check(pu_nc_put_host(f, name.c_str(), v.data(), nc_type_of<T>()));
  }
  template <class T> void read(T &v, std::string name) { check(pu_nc_get_host(f, name.c_str(), &v, nc_type_of<T>())); }
  template <class T> void read(std::vector<T> &v, std::string name) {
    check(pu_nc_get_host(f, name.c_str(), v.data(), nc_type_of<T>()));
  }

  // device arrays: DataManager views.  One rank holds the whole domain, so `start` is all zeros.
  template <class T, int N> void write_all(View<T, N> const &v, std::string name, std::vector<long long> = {}) {
    using U = std::remove_cv_t<T>;
    if constexpr (std::is_same<U, double>::value) {
      check(pu_nc_put_device(f, name.c_str(), v.data(), ctx));
    } else {
      std::vector<U> h((size_t)v.size());
      check(pu_memcpy_d2h(h.data(), v.data(), h.size() * sizeof(U)));
      check(pu_nc_put_host(f, name.c_str(), h.data(), nc_type_of<U>()));
    }
  }
  template <class T, int N> void read_all(View<T, N> const &v, std::string name, std::vector<long long> = {}) {
    static_assert(!std::is_const<T>::value, "NcFile::read_all needs a writable view");
    if constexpr (std::is_same<T, double>::value) {
      check(pu_nc_get_device(f, name.c_str(), v.data(), ctx));
    } else {
      std::vector<T> h((size_t)v.size());
      check(pu_nc_get_host(f, name.c_str(), h.data(), nc_type_of<T>()));
      check(pu_memcpy_h2d(v.data(), h.data(), h.size() * sizeof(T)));
    }
  }

 private:
  static void check(int rc) {
    if (rc != PU_OK) endrun(std::string("ERROR: NcFile: ") + pu_last_error());
  }
  pu_ncfile *f = nullptr;
  pu_ctx *ctx = nullptr;
};

}  // namespace core
