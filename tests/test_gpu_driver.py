"""The drop-in C++ path (core::Coupler + modules::Dynamics_Euler_Stratified_WenoFV +
modules::apply_surface_fluxes + modules::LES_Closure, via the benchmark driver) runs the same
kernels on the same inputs as the raw C ABI, so after several steps both states are bitwise
identical."""
import numpy as np
import pytest

import porturb_amd as pa
from porturb_amd import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32])
def test_cpp_driver_matches_raw_c_abi(precision):
    nx, ny, nz, dx = 64, 48, 32, 2.0
    drv = pa.Driver(pa.PU_CASE_ABL_URBAN, nx, ny, nz, dx, ord=9, precision=precision, seed=77)
    assert drv.tracer_names() == ["water_vapor", "TKE", "passive_1", "passive_2"]
    drv.step(3)
    a = drv.download()
    cfg = cases.abl_urban(nx, ny, nz, dx, ord=9, precision=precision)
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, pa.PU_CASE_ABL_URBAN, 77, cs)
    d = pa.Dycore(cfg)
    d.init(cs)
    # the driver's rough wall: roughness 0.1 m everywhere, no wall temperatures
    d.surface_init(np.full((nz, ny, nx), 0.1), np.zeros((nz, ny, nx)), np.full((ny, nx), 0.1), np.zeros((ny, nx)))
    dt = d.compute_time_step()
    for _ in range(3):
        d.time_step(cs, dt)
        d.surface_apply(cs, dt)
        d.les_apply(cs, dt)
    b = cs.download()
    names = list(pa.CouplerState.FIELDS) + [t[0] for t in cfg["tracers"]]
    for i, n in enumerate(names):
        assert np.array_equal(a[n], b[i]), n
    drv.destroy()


def test_cpp_driver_bubble_without_closure():
    drv = pa.Driver(pa.PU_CASE_BUBBLE, 32, 32, 32, 10000. / 32, ord=5, closure=False)
    assert drv.tracer_names() == ["water_vapor"]
    drv.step(2)
    st = drv.download()
    assert np.isfinite(st["temp"]).all()
    assert np.abs(st["wvel"]).max() > 0        # the bubble starts to rise
