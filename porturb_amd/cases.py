"""The benchmark workloads as configurations (grid, order, tracers, boundaries, precision).

* BUBBLE_64        -- dry warm bubble, 64^3 cells over 10x10x10 km, WENO-5, FP64, 100 steps
* ABL_URBAN_FP64   -- neutral urban ABL, 1024x1024x128 cells at 2 m, 8x8 cuboid buildings,
                      water vapour + TKE + 2 passive tracers, WENO-9, TKE closure, FP64
* ABL_URBAN_FP32   -- the same in FP32 state/fluxes with FP64 hydrostatics and compensated
                      rho, rho*theta
* WINDFARM         -- 2048x1024x256 cells at 5 m, precursor x inflow, capping inversion, FP32

Tracer order: water_vapor (positive, adds mass, diffused; the reference drivers add it with
add_tracer("water_vapor", ..., true, true, true), e.g. abl_neutral.cpp:58), TKE (positive,
no mass, not diffused; LES_Closure::init, les_closure.h:23), then passive tracers
(positive, no mass, diffused).
"""
from . import (make_config, PU_BC_PERIODIC, PU_BC_PRECURSOR, PU_BC_SOLID_WALL, PU_FP64, PU_FP32,
               PU_CASE_BUBBLE, PU_CASE_ABL_URBAN, PU_CASE_WINDFARM)

WV = ("water_vapor", True, True, True)
TKE = ("TKE", True, False, False)
P1 = ("passive_1", True, False, True)
P2 = ("passive_2", True, False, True)


def bubble(n=64, ord=5, precision=PU_FP64):
    cfg = make_config(n, n, n, 10000., 10000., 10000., ord=ord, tracers=[WV], precision=precision)
    cfg["case_id"] = PU_CASE_BUBBLE
    cfg["name"] = "dry_warm_bubble_%dx%dx%d_weno%d_%s" % (n, n, n, ord, "fp64" if precision == PU_FP64 else "fp32")
    return cfg


def abl_urban(nx=1024, ny=1024, nz=128, dx=2., ord=9, precision=PU_FP64):
    cfg = make_config(nx, ny, nz, nx * dx, ny * dx, nz * dx, ord=ord, tracers=[WV, TKE, P1, P2],
                      precision=precision)
    cfg["case_id"] = PU_CASE_ABL_URBAN
    cfg["name"] = "neutral_urban_abl_%dx%dx%d_weno%d_%s" % (nx, ny, nz, ord,
                                                            "fp64" if precision == PU_FP64 else "fp32")
    return cfg


def windfarm(nx=2048, ny=1024, nz=256, dx=5., ord=9, precision=PU_FP32):
    cfg = make_config(nx, ny, nz, nx * dx, ny * dx, nz * dx, ord=ord, tracers=[WV, TKE, P1, P2],
                      bc_x=PU_BC_PRECURSOR, precision=precision)
    cfg["case_id"] = PU_CASE_WINDFARM
    cfg["name"] = "windfarm_abl_%dx%dx%d_weno%d" % (nx, ny, nz, ord)
    return cfg


BUBBLE_64 = dict(builder=bubble, kwargs=dict(), steps=100)
ABL_URBAN_FP64 = dict(builder=abl_urban, kwargs=dict(precision=PU_FP64), steps=200)
ABL_URBAN_FP32 = dict(builder=abl_urban, kwargs=dict(precision=PU_FP32), steps=200)
WINDFARM = dict(builder=windfarm, kwargs=dict(), steps=500)
