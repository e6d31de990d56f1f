// core::Options -- typed key/value store of run options (model/core/Options.h:11-206).
//
// Same interface as the reference: add_option / set_option / get_option<T> / option_exists /
// delete_option, with a fatal error (endrun) on a missing key or a type mismatch
// (Options.h:101-131).  Values are held in a std::variant instead of the reference's
// type-erased holder.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

namespace core {

// endrun (model/main_header.h:101-103): the reference aborts; here a C++ exception carries the
// message so a driver (or a test) can decide.  Uncaught, it terminates the program like the
// reference's Kokkos::abort.
struct PortUrbError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] inline void endrun(std::string const &msg = "") { throw PortUrbError(msg); }

class Options {
 public:
  using Value = std::variant<bool, int, long long, unsigned long, float, double, std::string, std::vector<double>,
                             std::vector<bool>>;

  template <class T> void add_option(std::string key, T value) {
    if (option_exists(key)) endrun("ERROR: Options::add_option: key [" + key + "] already exists");
    opts[key] = normalise(value);
  }
  template <class T> void set_option(std::string key, T value) { opts[key] = normalise(value); }
  template <class T> T get_option(std::string key) const {
    auto it = opts.find(key);
    if (it == opts.end()) endrun("ERROR: Options::get_option: key [" + key + "] does not exist");
    using U = stored_t<T>;
    if (!std::holds_alternative<U>(it->second))
      endrun("ERROR: Options::get_option: key [" + key + "] requested with the wrong type");
    return static_cast<T>(std::get<U>(it->second));
  }
  bool option_exists(std::string key) const { return opts.find(key) != opts.end(); }
  void delete_option(std::string key) { opts.erase(key); }
  void finalize() { opts.clear(); }
  void clone_into(Options &o) const { o.opts = opts; }

 private:
  // string literals are stored as std::string; all other option types as themselves
  template <class T> using stored_t =
      std::conditional_t<std::is_convertible<T, std::string>::value && !std::is_arithmetic<T>::value, std::string, T>;
  template <class T> static Value normalise(T const &v) { return Value(stored_t<T>(v)); }
  std::map<std::string, Value> opts;
};

}  // namespace core
