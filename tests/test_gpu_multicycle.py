"""Physics steps longer than the dycore's step (Dynamics_Euler_Stratified_WenoFV::time_step,
dynamics_rk.h:103-137): ncycles = ceil(dt_phys / dt_dyn) SSPRK3 cycles of dt_phys / ncycles, with the
immersed relaxation that closes one cycle followed by the one that opens the next.  The GPU fuses
those two relaxations into the last stage of a cycle and hands the stage state from one cycle to the
next on the device; the reference's ABL drivers run this way (dyn_cycle = 10,
abl_neutral.cpp:27,86).

Bounds: FP64 per field 1e-10 relative (L1, L2, Linf) against the oracle; FP32 against the FP64
oracle as in the single-cycle FP32 test, widened for four times as many cycles.
"""
import numpy as np
import pytest

import porturb_amd as pa
from oracle import oracle_py
from porturb_amd import cases
from tests.helpers import random_state, small_config, rel_norms, WV, TKE, P1, P2

pytestmark = pytest.mark.gpu


def gpu_setup(cfg, st, imm):
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    d.init(cs)
    return d, cs


@pytest.mark.parametrize("ord_", (5, 9))
@pytest.mark.parametrize("mult", (10., 3.7))
def test_multi_cycle_steps_match_oracle(ord_, mult):
    cfg = small_config(nx=24, ny=20, nz=18, ord=ord_, tracers=[WV, TKE, P1, P2])
    st, imm = random_state(cfg, seed=41, buildings=3)
    print("multi-cycle ord", ord_, "dt_phys / dt_dyn", mult, flush=True)
    d, cs = gpu_setup(cfg, st, imm)
    od = oracle_py.Dycore(cfg, st, imm)
    dt = mult * d.compute_time_step()
    ref = st.copy()
    for _ in range(2):
        d.time_step(cs, dt)
        od.time_step(ref, dt)
    g = cs.download()
    d.destroy()
    nrm = rel_norms(g, ref)
    print("per-field [L1, L2, Linf]:\n", nrm, flush=True)
    assert nrm.max() < 1e-10, nrm


def test_multi_cycle_fp32_tracks_fp64():
    cfg64 = small_config(nx=24, ny=20, nz=18, ord=9, tracers=[WV, TKE, P1])
    cfg32 = dict(cfg64, precision=pa.PU_FP32)
    st, imm = random_state(cfg64, seed=42, buildings=2)
    print("multi-cycle fp32", flush=True)
    d, cs = gpu_setup(cfg32, st, imm)
    od = oracle_py.Dycore(cfg64, st, imm)
    dt = 10. * d.compute_time_step()
    ref = st.copy()
    for _ in range(2):
        d.time_step(cs, dt)
        od.time_step(ref, dt)
    nrm = rel_norms(cs.download(), ref)
    d.destroy()
    print("fp32 vs fp64 per-field [L1, L2, Linf]:\n", nrm, flush=True)
    assert nrm[[0, 4]].max() < 5e-6, nrm
    assert nrm.max() < 5e-4, nrm


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
def test_driver_dyn_cycle_matches_raw_c_abi(precision):
    """pu_driver_set_dyn_cycle(10): each driver step is one dycore time step of 10 dt_dyn, then the
    surface fluxes and the closure with the same step, bitwise as the raw C ABI calls."""
    nx, ny, nz, dx = 64, 48, 32, 2.0
    print("driver dyn_cycle 10 precision", precision, flush=True)
    drv = pa.Driver(pa.PU_CASE_ABL_URBAN, nx, ny, nz, dx, ord=9, precision=precision, seed=78)
    drv.set_dyn_cycle(10)
    drv.step(2)
    a = drv.download()
    t_drv = drv.elapsed_time()
    drv.destroy()
    cfg = cases.abl_urban(nx, ny, nz, dx, ord=9, precision=precision)
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, pa.PU_CASE_ABL_URBAN, 78, cs)
    d = pa.Dycore(cfg)
    d.init(cs)
    d.surface_init(np.full((nz, ny, nx), 0.1), np.zeros((nz, ny, nx)), np.full((ny, nx), 0.1), np.zeros((ny, nx)))
    dt = 10. * d.compute_time_step()
    for _ in range(2):
        d.time_step(cs, dt)
        d.surface_apply(cs, dt)
        d.les_apply(cs, dt)
    b = cs.download()
    d.destroy()
    assert t_drv == pytest.approx(2 * dt, rel=1e-15)
    names = list(pa.CouplerState.FIELDS) + [t[0] for t in cfg["tracers"]]
    for i, n in enumerate(names):
        assert np.array_equal(a[n], b[i]), n
