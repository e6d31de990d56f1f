"""The known-answer cases of tests/pin_cases.py on the GPU path (libporturb through the C ABI):
each case is checked against its closed form directly and against the oracle on the same input.

Bounds: FP64 tendencies 1e-11 of the largest tendency vs the oracle, time steps 1e-10 relative per
field (the warm bubble's velocities: within the oracle's own response to a rounding-level
perturbation, see tests.helpers.rounding_sensitivity); the closed forms as in the oracle's pins.
FP32 runs are held to the closed forms with the bounds stated at each test (float rounding of the
stencil and flux arithmetic).
"""
import numpy as np
import pytest

import porturb_amd as pa
from oracle import oracle_py
from tests import pin_cases as pc
from tests.helpers import rel_norms, rounding_sensitivity

pytestmark = pytest.mark.gpu


def gpu_setup(cfg, st, imm):
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    d.init(cs)
    return d, cs


def gpu_steps(cfg, st, imm, n, dt=None):
    d, cs = gpu_setup(cfg, st, imm)
    if dt is None:
        dt = d.compute_time_step()
    for _ in range(n):
        d.time_step(cs, dt)
    out = cs.download()
    d.destroy()
    return out, dt


def oracle_steps(cfg, st, imm, n, dt):
    od = oracle_py.Dycore(dict(cfg, precision=pa.PU_FP64), st, imm)
    s = st.copy()
    for _ in range(n):
        od.time_step(s, dt)
    return s


def tend_err(g, o):
    scale = np.maximum(np.abs(o).reshape(o.shape[0], -1).max(axis=1), 1e-300)
    return (np.abs(g - o).reshape(g.shape[0], -1).max(axis=1) / scale).max()


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("ord_", (5, 9))
@pytest.mark.parametrize("axis", (0, 1, 2))
@pytest.mark.parametrize("U", (6., -6.))
def test_step_tracer_is_upwinded(ord_, axis, U, precision):
    cfg, st, imm, expect = pc.step_advection(ord_, axis, U, precision=precision)
    print("step tracer ord", ord_, "axis", axis, "U", U, "precision", precision, flush=True)
    d, cs = gpu_setup(cfg, st, imm)
    g = d.tendencies(cs)
    d.destroy()
    pc.check_step_tendencies(cfg, g, expect, U, tol=1e-12 if precision == pa.PU_FP64 else 1e-6)
    if precision == pa.PU_FP64:
        o = oracle_py.Dycore(cfg, st, imm).tendencies(st)
        assert np.abs(g[5:] - o[5:]).max() <= 1e-12 * 1.1 * abs(U) * cfg["nx"] / cfg["xlen"]


@pytest.mark.parametrize("ord_", (5, 9))
def test_smooth_tracer_translates_at_flow_speed(ord_):
    cfg, st, imm, exact = pc.smooth_advection(ord_, 32)
    print("smooth tracer ord", ord_, flush=True)
    d, cs = gpu_setup(cfg, st, imm)
    dt = d.compute_time_step() * (1. if ord_ <= 5 else 0.25)
    n = int(round(10. / dt))
    d.destroy()
    g, _ = gpu_steps(cfg, st, imm, n, dt)
    q = g[5 + 1][0, 0] / g[0][0, 0]
    assert np.abs(q - exact(n * dt)).max() < (2e-5 if ord_ <= 5 else 2e-9)
    assert rel_norms(g, oracle_steps(cfg, st, imm, n, dt)).max() < 1e-10


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
def test_acoustic_pulse_splits_at_the_sound_speed(precision):
    cfg, st, imm, linear, cs_ = pc.acoustic_pulse(5, precision=precision)
    print("acoustic pulse precision", precision, flush=True)
    d, _ = gpu_setup(cfg, st, imm)
    dt = d.compute_time_step()
    d.destroy()
    n = int(round(0.9 / dt))
    g, _ = gpu_steps(cfg, st, imm, n, dt)
    pc.check_acoustics(g, linear, n * dt)
    if precision == pa.PU_FP64:
        assert rel_norms(g, oracle_steps(cfg, st, imm, n, dt)).max() < 1e-10


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("ord_", (5, 9))
def test_denser_cell_sinks(ord_, precision):
    cfg, st, imm, cell, drho = pc.buoyancy_cell(ord_, precision=precision)
    print("buoyancy ord", ord_, "precision", precision, flush=True)
    d, cs = gpu_setup(cfg, st, imm)
    g = d.tendencies(cs)
    d.destroy()
    # FP32 stores rho' and p' as float: the rest state's residual tendencies are float rounding of
    # the hydrostatic pressure gradient, ~1e-5 of g drho here
    pc.check_buoyancy_tendencies(cfg, st, g, cell, drho, tol=1e-9 if precision == pa.PU_FP64 else 1e-3)
    if precision == pa.PU_FP64:
        o = oracle_py.Dycore(cfg, st, imm).tendencies(st)
        # at rest the W tendency is the difference of hydrostatic edge pressures (~1e5 Pa, FP64
        # rounding 1e-11 Pa each) over dz, balanced against gravity: rounding of 1e-14 p / dz
        dz = cfg["zlen"] / cfg["nz"]
        assert np.abs(g - o).max() < 1e-14 * 1e5 / dz


def test_warm_bubble_rises():
    cfg, st, imm = pc.warm_bubble(n=24)
    print("warm bubble 24^3", flush=True)
    g, dt = gpu_steps(cfg, st, imm, 10)
    pc.check_bubble_rises(cfg, g)
    ref = oracle_steps(cfg, st, imm, 10, dt)
    nrm = rel_norms(g, ref)
    # The resting air around the bubble is ill-conditioned under the limiter's 1e-20 guard: the
    # oracle itself moves by ~5e-8 in u, v when its input is perturbed by 1e-15.  The GPU must stay
    # within that rounding response; the well-conditioned rho and theta within 1e-10.
    sens = rounding_sensitivity(lambda s: oracle_steps(cfg, s, imm, 10, dt), st, ref)
    print("GPU vs oracle:\n", nrm, "\noracle vs 1e-15-perturbed oracle:\n", sens, flush=True)
    assert (nrm <= 10 * sens + 1e-12).all()
    assert nrm[[0, 4]].max() < 1e-10


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("ord_", (5, 9))
def test_uniform_flow_rotates_inertially(ord_, precision):
    cfg, st, imm, f = pc.inertial(ord_, precision=precision)
    print("inertial ord", ord_, "precision", precision, flush=True)
    N = 1000
    g, dt = gpu_steps(cfg, st, imm, N)
    # FP32: the stage combination runs in double from float state, so each step rounds the
    # momentum to float (~6e-8 relative); 1000 steps accumulate at most ~1e-5
    pc.check_inertial(st, g, f, dt, N, tol=1e-11 if precision == pa.PU_FP64 else 2e-5)


LES_DT = 1.0


def gpu_les(cfg, st, imm):
    d, cs = gpu_setup(cfg, st, imm)
    d.les_apply(cs, LES_DT)
    out = cs.download()
    d.destroy()
    return out


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("gamma_theta", (0.01, -0.01))
def test_les_heat_flux_is_down_gradient_and_feeds_buoyancy(gamma_theta, precision):
    cfg, st, imm = pc.les_stratified(gamma_theta, precision=precision)
    cfg0, st0, _ = pc.les_stratified(gamma_theta, enable_gravity=0, precision=precision)
    print("les stratified", gamma_theta, "precision", precision, flush=True)
    out, out0 = gpu_les(cfg, st, imm), gpu_les(cfg0, st0, imm)
    pc.check_les_stratified(cfg, st, out, out0, gamma_theta, LES_DT,
                            tol=1e-9 if precision == pa.PU_FP64 else 2e-4)
    if precision == pa.PU_FP64:
        ref = st.copy()
        oracle_py.les_apply(cfg, ref, imm, LES_DT)
        assert rel_norms(out, ref).max() < 1e-12


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("S", (0.02, -0.03))
def test_les_momentum_flux_is_down_gradient_and_feeds_shear_production(S, precision):
    cfg, st, imm = pc.les_shear(S, precision=precision)
    cfg0, st0, _ = pc.les_shear(0., precision=precision)
    print("les shear", S, "precision", precision, flush=True)
    out, out0 = gpu_les(cfg, st, imm), gpu_les(cfg0, st0, imm)
    pc.check_les_shear(cfg, st, out, out0, S, LES_DT, tol=1e-9 if precision == pa.PU_FP64 else 2e-4)
    if precision == pa.PU_FP64:
        ref = st.copy()
        oracle_py.les_apply(cfg, ref, imm, LES_DT)
        assert rel_norms(out, ref).max() < 1e-12
