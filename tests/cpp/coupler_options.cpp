// Host-only checks of the drop-in C++ API that need no device: core::Options typing and errors,
// and Coupler option defaults (coupler.h:288-326 semantics).  Built and run by tests/test_cpp_api.py.
#include <cstdio>
#include <string>

#include "porturb/core/coupler.h"
#include "porturb/modules/dynamics_rk.h"
#include "porturb/modules/les_closure.h"
#include "porturb/modules/geostrophic_wind_forcing.h"
#include "porturb/modules/sponge_layer.h"
#include "porturb/modules/surface_flux.h"
#include "porturb/modules/uniform_pg_wind_forcing.h"

static int fails = 0;
#define EXPECT(c) do { if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); fails++; } } while (0)

int main() {
  core::Options o;
  o.add_option<double>("cfl", 0.6);
  o.add_option<std::string>("bc_z", "solid_wall");
  o.add_option<bool>("weno_all", true);
  EXPECT(o.get_option<double>("cfl") == 0.6);
  EXPECT(o.get_option<std::string>("bc_z") == "solid_wall");
  bool threw = false;
  try { o.get_option<int>("cfl"); } catch (core::PortUrbError const &) { threw = true; }
  EXPECT(threw);                                          // wrong type -> endrun (Options.h:123-131)
  threw = false;
  try { o.get_option<double>("missing"); } catch (core::PortUrbError const &) { threw = true; }
  EXPECT(threw);
  threw = false;
  try { o.add_option<double>("cfl", 0.5); } catch (core::PortUrbError const &) { threw = true; }
  EXPECT(threw);                                          // add of an existing key
  o.set_option<size_t>("uniform_pg_wind_forcing_counter", 3);   // the counter type the forcing uses
  EXPECT(o.get_option<size_t>("uniform_pg_wind_forcing_counter") == 3);
  core::Coupler c;
  c.set_option<double>("cfl", 0.3);
  EXPECT(c.get_option<double>("cfl", 0.6) == 0.3);
  EXPECT(c.get_option<double>("latitude", 0.) == 0.);
  modules::Dynamics_Euler_Stratified_WenoFV dyn;
  EXPECT(dyn.hs == (dyn.ord - 1) / 2);
  c.set_grid(100., 50., 25.);
  EXPECT(c.get_xlen() == 100.);
  EXPECT(modules::detail::bc_code("periodic", false) == PU_BC_PERIODIC);
  threw = false;
  try { modules::detail::bc_code("open", true); } catch (core::PortUrbError const &) { threw = true; }
  EXPECT(threw);
  std::printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
