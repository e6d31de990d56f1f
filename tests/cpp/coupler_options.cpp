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
#include "porturb/modules/windmill_actuators.h"
#include <fstream>

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
  // turbine files: flat numeric entries kept, nested lists and strings (airfoil tables) skipped
  {
    const char *fn = "turbine_test.yaml";
    std::ofstream f(fn);
    f << "# comment line\nhub_height  : 90   # m\nblade_radius: 63\noverhang: -5.\n"
         "velocity_magnitude: [ 3 ,\n   4 ,  # inline comment\n   5.5 ]\n"
         "airfoil_names: [ \"Cylinder1\" ,\n \"DU40\" ]\n"
         "Cylinder1: [ [ -180.0 , 0.0 , 0.5 ] ,\n [ 180.0 , 0.0 , 0.5 ] ]\nthrust_coef: [1.1, 0.9, 0.7]\n";
    f.close();
    auto t = modules::detail::read_turbine_yaml(fn);
    EXPECT(t.count("hub_height") && t["hub_height"].size() == 1 && t["hub_height"][0] == 90);
    EXPECT(t["overhang"][0] == -5.);
    EXPECT(t["velocity_magnitude"] == (std::vector<double>{3, 4, 5.5}));
    EXPECT(t["thrust_coef"] == (std::vector<double>{1.1, 0.9, 0.7}));
    EXPECT(!t.count("airfoil_names") && !t.count("Cylinder1"));
    std::remove(fn);
  }
  o.set_option<std::vector<double>>("turbine_x_locs", {1., 2.});
  EXPECT(o.get_option<std::vector<double>>("turbine_x_locs").size() == 2);
  std::printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
