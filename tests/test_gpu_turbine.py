"""GPU parity of the actuator-disk turbines (turbine_kernels.hip through the C ABI) with the oracle
(modules::WindmillActuators::apply, windmill_actuators_yaw.h:523-1043): projection and sampling
weights, sampled inflow, coefficients, traces and the updated u, v, TKE over several calls, with
yaw, shaft tilt and overhang, the yaw controller and the induction form of C_T."""
import math

import numpy as np
import pytest

import porturb_amd as pa
from oracle import oracle_py as o
from tests.helpers import small_config, WV, TKE

pytestmark = pytest.mark.gpu

DX = 4.0


def flow(cfg, seed, u0=8.0, v0=0.0):
    rng = np.random.default_rng(seed)
    n = (cfg["nz"], cfg["ny"], cfg["nx"])
    c = np.zeros((5 + cfg["num_tracers"],) + n)

    def smooth(a):
        f = rng.standard_normal(n)
        for ax in range(3):
            f = 0.25 * np.roll(f, 1, ax) + 0.5 * f + 0.25 * np.roll(f, -1, ax)
        return a * f
    c[0] = 1.2 + smooth(0.01)
    c[1] = u0 + smooth(1.0)
    c[2] = v0 + smooth(1.0)
    c[3] = smooth(0.3)
    c[4] = 300.0
    c[6] = 0.12 + np.abs(smooth(0.05))
    return c


def run_both(cfg, c, tab, kw, oracle_kw, nsteps, dt, xs, ys, **geo):
    cs = pa.CouplerState(cfg)
    cs.upload((c, np.zeros(c.shape[1:])))
    d = pa.Dycore(cfg)
    T = pa.Turbines(d, *tab, x=xs, y=ys, **geo, **kw)
    n3 = int(np.prod(c.shape[1:]))
    pb, sb = pa.DeviceBuffer(8 * n3), pa.DeviceBuffer(8 * n3)
    ref = o.TurbineRef(*tab, **geo)
    tu = (o.OcTurbine * len(xs))()
    for i in range(len(xs)):
        tu[i].base_x, tu[i].base_y, tu[i].apply_thrust = xs[i], ys[i], 1
        tu[i].yaw = kw.get("initial_yaw", 0.0)
    oc = c.copy()
    for _ in range(nsteps):
        T.apply(cs, dt, pb, sb)
        pw, sw = o.turbines_apply(cfg, oc, ref, tu, dt, **oracle_kw)
    g = cs.download()
    gp, gs = pb.download(np.float64, c.shape[1:]), sb.download(np.float64, c.shape[1:])
    return T, tu, g, oc, gp, gs, pw, sw


def close(a, b, tol):
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / s < tol


@pytest.mark.parametrize("yaw,tilt", [(0.0, 0.0), (0.35, 5 * math.pi / 180)])
def test_turbines_match_oracle(yaw, tilt):
    cfg = small_config(nx=64, ny=48, nz=40, tracers=[WV, TKE], dx=DX)
    c = flow(cfg, 11)
    tab = ([0.0, 6.0, 9.0, 12.0], [0.9, 0.8, 0.7, 0.5], [0.45, 0.45, 0.4, 0.3], [1e-5, 1e-5, 2e-5, 3e-5])
    kw = dict(initial_yaw=yaw, fixed_yaw=True, orig_C_T=False, upstream_dir=yaw)
    okw = dict(fixed_yaw=True, orig_C_T=False, upstream_dir=yaw)
    xs, ys = [120.0, 200.0], [70.0, 130.0]
    T, tu, g, oc, gp, gs, pw, sw = run_both(cfg, c, tab, kw, okw, 4, 0.8, xs, ys, hub_height=80.0, blade_radius=20.0,
                                            shaft_tilt=tilt)
    assert close(gp, pw, 1e-10) and close(gs, sw, 1e-10)
    assert np.array_equal(gp > 0, pw > 0) and np.array_equal(gs > 0, sw > 0)
    for f in (1, 2, 6):
        if f != 2 or yaw != 0:                               # yaw 0: the thrust has no y component
            assert np.abs(g[f] - c[f]).max() > 0
        assert close(g[f] - c[f], oc[f] - c[f], 1e-9) if yaw != 0 or f != 2 else np.array_equal(g[2], c[2]), f
    assert np.array_equal(g[[0, 3, 4, 5]], c[[0, 3, 4, 5]])
    for i in range(2):
        st = T.status(i)
        assert st.nsteps == 4
        assert st.ct == pytest.approx(tu[i].ct, rel=1e-10) and st.cp == pytest.approx(tu[i].cp, rel=1e-10)
        assert st.u_samp == pytest.approx(tu[i].u_samp, rel=1e-10)
        assert st.u_samp_inertial == pytest.approx(tu[i].u_inertial, rel=1e-10)
        assert st.ct > 0 and 0 < st.cp < st.ct             # induction C_T, power-capped C_P
    tr = T.trace(0, pa.PU_TURB_CT)
    assert tr.shape == (4,) and tr[0] == 0.0                 # the first call sees zero inertial inflow
    T.destroy()


def test_turbines_yaw_controller_and_computed_upstream_direction():
    cfg = small_config(nx=64, ny=48, nz=40, tracers=[WV, TKE], dx=DX)
    c = flow(cfg, 12, u0=7.0, v0=2.0)
    tab = ([0.0, 100.0], [0.75, 0.75], [0.0, 0.0], [0.0, 0.0])
    kw = dict(initial_yaw=0.0, fixed_yaw=False, orig_C_T=True)
    okw = dict(fixed_yaw=False, orig_C_T=True)
    T, tu, g, oc, gp, gs, pw, sw = run_both(cfg, c, tab, kw, okw, 3, 1.0, [150.0], [96.0], hub_height=80.0,
                                            blade_radius=20.0, max_yaw_speed=0.02)
    st = T.status(0)
    assert st.yaw == pytest.approx(tu[0].yaw, rel=1e-9) and st.yaw > 0
    assert close(gp, pw, 1e-9) and close(gs, sw, 1e-9)
    assert close(g[1] - c[1], oc[1] - c[1], 1e-8) and close(g[6] - c[6], oc[6] - c[6], 1e-8)
    yaws = T.trace(0, pa.PU_TURB_YAW)
    assert yaws[0] == 0.0 and np.all(np.diff(yaws) > 0)
    T.destroy()


def test_driver_windfarm_turbines_match_raw_calls():
    """WindmillActuators (turbine file, options, placement) in the driver step after the surface
    fluxes (windfarm_ensemble.cpp:226-229), bitwise against the same C ABI calls one by one."""
    from porturb_amd import cases
    import os
    nx, ny, nz, dx = 128, 64, 48, 5.0
    tfile = os.path.join(os.path.dirname(pa.__file__), "data", "turbine_disk_ct075_d120.yaml")
    drv = pa.Driver(pa.PU_CASE_WINDFARM, nx, ny, nz, dx, ord=9, precision=pa.PU_FP64, seed=90)
    drv.set_turbines(tfile, 1, 1, orig_C_T=True)
    drv.step(3)
    a = drv.download()
    st = drv.turbine_status(0)
    assert st.nsteps == 3 and st.ct == 0.75 and st.cp == 0.0
    cfg = cases.windfarm(nx, ny, nz, dx, ord=9, precision=pa.PU_FP64)
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, pa.PU_CASE_WINDFARM, 90, cs)
    d = pa.Dycore(cfg)
    d.init(cs)
    d.surface_init(np.full((nz, ny, nx), 0.1), np.zeros((nz, ny, nx)), np.full((ny, nx), 0.1), np.zeros((ny, nx)))
    T = pa.Turbines(d, [0.0, 100.0], [0.75, 0.75], [0.0, 0.0], [0.0, 0.0], hub_height=90.0, blade_radius=60.0,
                    x=[nx * dx / 2], y=[ny * dx / 2], fixed_yaw=True, orig_C_T=True, upstream_dir=0.0)
    n3 = nx * ny * nz
    pb, sb = pa.DeviceBuffer(8 * n3), pa.DeviceBuffer(8 * n3)
    dt = d.compute_time_step()
    for _ in range(3):
        d.time_step(cs, dt)
        d.surface_apply(cs, dt)
        T.apply(cs, dt, pb, sb)
        d.les_apply(cs, dt)
    b = cs.download()
    names = list(pa.CouplerState.FIELDS) + [t[0] for t in cfg["tracers"]]
    for i, n in enumerate(names):
        assert np.array_equal(a[n], b[i]), n
    assert T.status(0).ct == 0.75 and T.status(0).u_samp == st.u_samp
    T.destroy()
    drv.destroy()


def test_full_size_rotor_sums_do_not_overflow():
    """A 120 m rotor (R = 60 m) on 5 m cells, the wind-farm workload's turbine: its point cloud has
    200 x 610 x 610 = 7.4e7 points, whose fixed-point total must stay exact (at 2^40 per point it
    would pass 2^64).  The normalised projection and sampling weights sum to one, and the whole step
    matches the oracle."""
    cfg = small_config(nx=100, ny=40, nz=40, tracers=[WV, TKE], dx=5.0)
    c = flow(cfg, 13, u0=9.0)
    tab = ([0.0, 100.0], [0.75, 0.75], [0.0, 0.0], [0.0, 0.0])
    kw = dict(initial_yaw=0.0, fixed_yaw=True, orig_C_T=True, upstream_dir=0.0)
    okw = dict(fixed_yaw=True, orig_C_T=True, upstream_dir=0.0)
    print("rotor R = 60 m, dx = 5 m", flush=True)
    T, tu, g, oc, gp, gs, pw, sw = run_both(cfg, c, tab, kw, okw, 2, 0.5, [400.0], [100.0], hub_height=90.0,
                                            blade_radius=60.0)
    assert gp.sum() == pytest.approx(1.0, rel=1e-12) and gs.sum() == pytest.approx(1.0, rel=1e-12)
    assert pw.sum() == pytest.approx(1.0, rel=1e-12)
    assert close(gp, pw, 1e-10) and close(gs, sw, 1e-10)
    for f in (1, 6):
        assert close(g[f] - c[f], oc[f] - c[f], 1e-9), f
    assert T.status(0).u_samp == pytest.approx(tu[0].u_samp, rel=1e-10)
    T.destroy()
