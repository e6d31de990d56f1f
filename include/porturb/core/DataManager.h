// core::DataManager -- named, typed, dimensioned device allocations (model/core/DataManager.h).
//
// Same interface as the reference for the calls the hot path and its drivers make:
// add_dimension / find_dimension / get_dimension_size, register_and_allocate<T>(name, desc, dims,
// dim_names) (:127-209), entry_exists, get<T,N>(name) returning a non-owning view (a non-const get
// marks the entry dirty, :265-299), get_dirty_entries / clean_all_entries (:222-251), finalize
// (:590-597).  Storage comes from libporturb's device allocator (pu_alloc_device); views carry a
// device pointer and extents, [N-1] fastest (styleC).
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <typeinfo>
#include <vector>

#include "porturb.h"
#include "porturb/core/Options.h"

namespace core {

// Non-owning device view (the role of yakl::Array<T,N,memDevice,styleC> returned by get<T,N>).
template <class T, int N> struct View {
  T *ptr = nullptr;
  std::array<long long, N> dims{};
  T *data() const { return ptr; }
  long long extent(int i) const { return dims[i]; }
  long long size() const {
    long long s = 1;
    for (int i = 0; i < N; i++) s *= dims[i];
    return s;
  }
  bool initialized() const { return ptr != nullptr; }
};

class DataManager {
 public:
  struct Entry {
    std::string name, desc;
    size_t type_hash = 0;
    size_t type_size = 0;
    std::vector<long long> dims;
    std::vector<std::string> dim_names;
    std::shared_ptr<void> ptr;      // device allocation, freed through pu_free_device
    bool dirty = false;
  };

  DataManager() = default;
  DataManager(DataManager &&) = default;
  DataManager &operator=(DataManager &&) = default;
  DataManager(DataManager const &) = delete;
  DataManager &operator=(DataManager const &) = delete;

  void add_dimension(std::string name, long long len) {
    int id = find_dimension(name);
    if (id >= 0) {
      if (dimensions[id].second != len) endrun("ERROR: dimension [" + name + "] exists with another size");
      return;
    }
    dimensions.push_back({name, len});
  }
  int find_dimension(std::string name) const {
    for (size_t i = 0; i < dimensions.size(); i++) if (dimensions[i].first == name) return (int)i;
    return -1;
  }
  long long get_dimension_size(std::string name) const {
    int id = find_dimension(name);
    if (id < 0) endrun("ERROR: dimension [" + name + "] does not exist");
    return dimensions[id].second;
  }

  template <class T>
  void register_and_allocate(std::string name, std::string desc, std::vector<long long> dims,
                             std::vector<std::string> dim_names = {}) {
    static std::mutex mtx;                                  // as DataManager.h:133
    std::lock_guard<std::mutex> lock(mtx);
    if (name.empty()) endrun("ERROR: register_and_allocate: empty name");
    if (entry_exists(name)) endrun("ERROR: register_and_allocate: entry [" + name + "] already exists");
    if (!dim_names.empty() && dim_names.size() != dims.size())
      endrun("ERROR: register_and_allocate: dims and dim_names differ in length for [" + name + "]");
    long long n = 1;
    for (auto d : dims) {
      if (d < 1) endrun("ERROR: register_and_allocate: non-positive dimension for [" + name + "]");
      n *= d;
    }
    for (size_t i = 0; i < dim_names.size(); i++) add_dimension(dim_names[i], dims[i]);
    void *p = nullptr;
    if (pu_alloc_device((size_t)n * sizeof(T), &p) != PU_OK) endrun(std::string("ERROR: ") + pu_last_error());
    pu_memset_device(p, 0, (size_t)n * sizeof(T));
    Entry e;
    e.name = name; e.desc = desc; e.type_hash = typeid(T).hash_code(); e.type_size = sizeof(T);
    e.dims = dims; e.dim_names = dim_names;
    e.ptr = std::shared_ptr<void>(p, [](void *q) { pu_free_device(q); });
    entries.push_back(std::move(e));
  }

  bool entry_exists(std::string name) const { return find_entry(name) >= 0; }

  // const access: no dirty mark; T may be const-qualified (get<real const,3>)
  template <class T, int N> View<T, N> get(std::string name) const {
    return make_view<T, N>(entry_or_die(name));
  }
  template <class T, int N> View<T, N> get(std::string name) {
    Entry &e = entries[index_or_die(name)];
    if (!std::is_const<T>::value) e.dirty = true;
    return make_view<T, N>(e);
  }
  // flattened view of any entry (get_collapsed, :360-393)
  template <class T> View<T, 1> get_collapsed(std::string name) const {
    Entry const &e = entry_or_die(name);
    check_type<T>(e);
    View<T, 1> v;
    v.ptr = static_cast<T *>(e.ptr.get());
    long long n = 1;
    for (auto d : e.dims) n *= d;
    v.dims[0] = n;
    return v;
  }

  std::vector<std::string> get_dirty_entries() const {
    std::vector<std::string> out;
    for (auto &e : entries) if (e.dirty) out.push_back(e.name);
    return out;
  }
  void clean_all_entries() { for (auto &e : entries) e.dirty = false; }
  void unregister_and_deallocate(std::string name) { entries.erase(entries.begin() + index_or_die(name)); }
  void finalize() { entries.clear(); dimensions.clear(); }
  std::vector<Entry> const &get_entries() const { return entries; }

 private:
  int find_entry(std::string const &name) const {
    for (size_t i = 0; i < entries.size(); i++) if (entries[i].name == name) return (int)i;
    return -1;
  }
  int index_or_die(std::string const &name) const {
    int id = find_entry(name);
    if (id < 0) endrun("ERROR: DataManager: entry [" + name + "] does not exist");
    return id;
  }
  Entry const &entry_or_die(std::string const &name) const { return entries[index_or_die(name)]; }
  template <class T> static void check_type(Entry const &e) {
    using U = std::remove_cv_t<T>;
    if (e.type_hash != typeid(U).hash_code())
      endrun("ERROR: DataManager: entry [" + e.name + "] requested with the wrong type");
  }
  template <class T, int N> static View<T, N> make_view(Entry const &e) {
    check_type<T>(e);
    if ((int)e.dims.size() != N) endrun("ERROR: DataManager: entry [" + e.name + "] requested with the wrong rank");
    View<T, N> v;
    v.ptr = static_cast<T *>(e.ptr.get());
    for (int i = 0; i < N; i++) v.dims[i] = e.dims[i];
    return v;
  }

  std::vector<Entry> entries;
  std::vector<std::pair<std::string, long long>> dimensions;
};

}  // namespace core
