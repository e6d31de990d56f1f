"""Pins of the actuator-disk oracle (oracle/turbine.c, modules::WindmillActuators::apply,
windmill_actuators_yaw.h:523-1043) by closed forms: the table interpolation and blade weight by
hand, the momentum sink of actuator-disk theory (total thrust -1/2 C_T U^2 pi R^2 per unit
density in uniform inflow), the TKE source, normalised and symmetric projections, the induction
form of C_T, the inertial filter, the normal-wind limiter and the yaw controller."""
import math

import numpy as np
import pytest

from oracle import oracle_py as o
from tests.helpers import small_config, WV, TKE

DX = 4.0
CFG = dict(nx=64, ny=48, nz=40, tracers=[WV, TKE], dx=DX)      # 256 x 192 x 160 m
U0, RHO = 8.0, 1.2


def uniform_state(cfg, u=U0, v=0.0):
    n = (cfg["nz"], cfg["ny"], cfg["nx"])
    c = np.zeros((5 + cfg["num_tracers"],) + n)
    c[0], c[1], c[2], c[4] = RHO, u, v, 300.0
    c[6] = 0.1 * RHO
    return c


def turbines(n, **kw):
    t = (o.OcTurbine * n)()
    for i in range(n):
        t[i].apply_thrust = 1
        for k, v in kw.items():
            setattr(t[i], k, v[i] if isinstance(v, (list, tuple)) else v)
    return t


CONST = o.TurbineRef([0.0, 100.0], [0.75, 0.75], [0.0, 0.0], [0.0, 0.0], hub_height=80.0, blade_radius=20.0)


def test_interp_and_blade_weight_by_hand():
    v = np.array([3.0, 4.0, 6.0, 10.0])
    f = np.array([0.5, 1.0, 2.0, 6.0])
    L = o.lib()
    interp = lambda x: L.oc_turbine_interp(v.ctypes.data, f.ctypes.data, 4, x)
    assert interp(2.9) == 0 and interp(10.1) == 0             # idle outside the table
    assert interp(3.0) == 0.5 and interp(10.0) == 6.0
    assert interp(5.0) == pytest.approx(1.5) and interp(7.0) == pytest.approx(3.0)
    bw = L.oc_turbine_blade_weight
    assert bw(2.0, 63.0) == 1.0 and bw(16.0, 63.0) == pytest.approx(1.0)
    assert bw(4.0, 63.0) == pytest.approx(-2 / 27 + 3 / 9)    # x = 1/3: -2x^3 + 3x^2


def test_uniform_inflow_momentum_sink_and_tke_source():
    cfg = small_config(**CFG)
    c = uniform_state(cfg)
    t = turbines(1, base_x=150.0, base_y=96.0, u_inertial=U0)
    dt = 0.5
    pw, sw = o.turbines_apply(cfg, c, CONST, t, dt, fixed_yaw=True, orig_C_T=True, upstream_dir=0.0)
    V = DX ** 3
    A = math.pi * 20.0 ** 2
    assert pw.sum() == pytest.approx(1.0, rel=1e-13) and sw.sum() == pytest.approx(1.0, rel=1e-13)
    assert t[0].u_samp == pytest.approx(U0, rel=1e-13) and t[0].ct == 0.75 and t[0].cp == 0.0
    force = ((c[1] - U0) / dt).sum() * V
    assert force == pytest.approx(-0.5 * 0.75 * U0 ** 2 * A, rel=1e-12)
    assert np.array_equal(c[2], np.zeros_like(c[2]))                   # yaw 0: no lateral force
    bw = o.lib().oc_turbine_blade_weight(DX, 20.0)
    tke_src = ((c[6] - 0.1 * RHO) / dt).sum() * V
    assert tke_src == pytest.approx(0.5 * RHO * 0.25 * 0.75 * U0 ** 3 * A * (1 - bw), rel=1e-12)
    # the sampling image lies 4 R upstream: its weights are the disk's, shifted by 80 m = 20 cells
    assert np.allclose(np.roll(sw, 20, axis=2), pw, rtol=0, atol=1e-15)
    # inertial filter: tau = 30 s
    assert t[0].u_inertial == pytest.approx(U0 * dt / 30 + (30 - dt) / 30 * U0)


def test_projection_symmetry_and_no_thrust_before_the_filter_spins_up():
    cfg = small_config(**CFG)
    c = uniform_state(cfg)
    # hub on a cell face in y (96 m = 24 cells): the disk weights are mirror-symmetric in y
    t = turbines(1, base_x=150.0, base_y=96.0)
    pw, _ = o.turbines_apply(cfg, c, CONST, t, 0.5, fixed_yaw=True, orig_C_T=True, upstream_dir=0.0)
    assert np.allclose(pw, pw[:, ::-1, :], rtol=0, atol=1e-15)
    # zero inertial inflow: C_T = 0, nothing changes but the filter
    assert t[0].ct == 0.0 and np.array_equal(c[1], np.full_like(c[1], U0))
    assert t[0].u_inertial == pytest.approx(U0 * 0.5 / 30)


def test_induction_form_of_ct_and_power_cap():
    cfg = small_config(**CFG)
    ref = o.TurbineRef([0.0, 20.0], [0.8, 0.8], [0.4, 0.4], [5.0, 5.0], hub_height=80.0, blade_radius=20.0)
    c = uniform_state(cfg)
    t = turbines(1, base_x=150.0, base_y=96.0, u_inertial=U0)
    o.turbines_apply(cfg, c, ref, t, 0.5, fixed_yaw=True, orig_C_T=False, upstream_dir=0.0)
    a = 1 - 0.4 / (0.8 + 1e-10)
    ct = 4 * a * (1 - a)
    assert t[0].ct == pytest.approx(ct, rel=1e-12)
    cap = 5.0e6 / (0.5 * 1.2 * math.pi * 400.0 * U0 ** 3)
    assert t[0].cp == pytest.approx(min(ct, cap), rel=1e-12) and t[0].power == 5.0


def test_normal_wind_limiter():
    cfg = small_config(**CFG)
    c = uniform_state(cfg, u=0.5)
    ref = o.TurbineRef([0.0, 100.0], [50.0, 50.0], [0.0, 0.0], [0.0, 0.0], hub_height=80.0, blade_radius=20.0)
    t = turbines(1, base_x=150.0, base_y=96.0, u_inertial=0.5)
    dt = 50.0
    pw, _ = o.turbines_apply(cfg, c, ref, t, dt, fixed_yaw=True, orig_C_T=True, upstream_dir=0.0)
    # where the thrust would reverse the normal wind the tendency is scaled to -un:
    # u_new = u - dt * u (= -24.5 here); elsewhere u decreases but stays >= 0
    wt = pw * math.pi * 400.0 / DX ** 3
    would = 0.5 + dt * (-0.5 * 50.0 * 0.25 * wt)
    lim = (pw > 0) & (would < 0)
    assert lim.any()
    assert np.allclose(c[1][lim], 0.5 - dt * 0.5)
    assert np.allclose(c[1][(pw > 0) & ~lim], would[(pw > 0) & ~lim])


def test_yaw_controller_turns_at_the_maximum_rate():
    cfg = small_config(**CFG)
    th = 0.6                                                   # wind from 34 degrees
    c = uniform_state(cfg, u=U0 * math.cos(th), v=U0 * math.sin(th))
    ref = o.TurbineRef([0.0, 100.0], [0.75, 0.75], [0.0, 0.0], [0.0, 0.0], hub_height=80.0, blade_radius=20.0,
                       max_yaw_speed=0.01)
    t = turbines(1, base_x=150.0, base_y=96.0)
    o.turbines_apply(cfg, c, ref, t, 2.0, fixed_yaw=False, orig_C_T=True)
    assert math.atan2(t[0].upstream_v, t[0].upstream_u) == pytest.approx(th, rel=1e-12)
    assert t[0].yaw == pytest.approx(2.0 * 0.01)               # |diff|/dt = 0.3 rad/s > 0.01
    ref2 = o.TurbineRef([0.0, 100.0], [0.75, 0.75], [0.0, 0.0], [0.0, 0.0], hub_height=80.0, blade_radius=20.0,
                        max_yaw_speed=1.0)
    t2 = turbines(1, base_x=150.0, base_y=96.0)
    o.turbines_apply(cfg, uniform_state(cfg, u=U0 * math.cos(th), v=U0 * math.sin(th)), ref2, t2, 2.0)
    assert t2[0].yaw == pytest.approx(th, rel=1e-12)           # reached in one step
