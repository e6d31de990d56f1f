"""Pins of the column-mean forcing oracle (oracle/column.c): modules::sponge_layer
(sponge_layer.h:7-60), modules::geostrophic_wind_forcing (geostrophic_wind_forcing.h:7-36) and
modules::uniform_pg_wind_forcing(_height) (uniform_pg_wind_forcing.h:5-80), by invariants and
hand-computed values rather than by restating the formulas."""
import math

import numpy as np
import pytest

from oracle import oracle_py as o
from tests.helpers import small_config, WV


def state(cfg, seed=0):
    rng = np.random.default_rng(seed)
    n = (cfg["nz"], cfg["ny"], cfg["nx"])
    c = np.zeros((5 + cfg["num_tracers"],) + n)
    c[0] = 1.1
    c[1] = 5.0 + rng.standard_normal(n)
    c[2] = -2.0 + rng.standard_normal(n)
    c[3] = 0.3 * rng.standard_normal(n)
    c[4] = 300.0 + rng.standard_normal(n)
    return c


# nz = 10 levels of 100 m: with top_prop 0.2, z1 = 800 m; the centres 850 and 950 m lie above it
CFG = dict(nx=7, ny=5, nz=10, tracers=[WV], dx=100.)


def test_sponge_known_factors_and_untouched_levels():
    cfg = small_config(**CFG)
    c = state(cfg)
    out = c.copy()
    dt, ts = 2.0, 8.0
    o.sponge_layer(cfg, out, dt, ts, 0.2)
    assert np.array_equal(out[:, :8], c[:, :8])                  # z <= z1: unchanged
    assert np.array_equal(out[0], c[0])                          # density untouched
    # factor = ((z - z1)/(zlen - z1))^3 dt/ts: (50/200)^3/4 and (150/200)^3/4, by hand
    for k, f in ((8, 0.015625 / 4), (9, 0.421875 / 4)):
        for fld in (1, 2, 4):
            m = c[fld, k].mean()
            # deviations from the level mean shrink by (1 - f), the mean itself is kept
            assert np.allclose(out[fld, k] - m, (1 - f) * (c[fld, k] - m), rtol=0, atol=1e-12)
            assert abs(out[fld, k].mean() - m) < 1e-12 * abs(m)
        assert np.allclose(out[3, k], (1 - f) * c[3, k], rtol=1e-15, atol=0)   # w relaxes to 0


def test_sponge_whole_column_and_time_scale_equal_dt():
    cfg = small_config(**CFG)
    c = state(cfg, 1)
    out = c.copy()
    # top_prop 1: z1 = 0, every level is sponged (the first column index would be -1)
    o.sponge_layer(cfg, out, 1.0, 1.0, 1.0)
    # at the top cell centre factor = 0.95^3; the bottom one 0.05^3
    for k, f in ((0, 0.05 ** 3), (9, 0.95 ** 3)):
        m = c[1, k].mean()
        assert np.allclose(out[1, k] - m, (1 - f) * (c[1, k] - m), rtol=0, atol=1e-12)


def test_geostrophic_turning_closed_form():
    cfg = small_config(**CFG)
    c = state(cfg, 2)
    out = c.copy()
    dt, lat, ug, vg = 10.0, 45.0, 8.0, 1.0
    o.geostrophic_forcing(cfg, out, dt, lat, ug, vg)
    f = math.sqrt(2.0) * 7.2921e-5      # 2 Omega sin(45 deg) = sqrt(2) Omega
    for k in (0, 4, 9):
        mu, mv = c[1, k].mean(), c[2, k].mean()
        du, dv = out[1, k] - c[1, k], out[2, k] - c[2, k]
        # one uniform increment per level: the departure from the geostrophic wind turns
        # clockwise, du = f dt (<v> - v_g), dv = -f dt (<u> - u_g), and grows by sqrt(1 + (f dt)^2)
        assert np.ptp(du) < 1e-13 and np.ptp(dv) < 1e-13
        a0 = np.array([mu - ug, mv - vg])
        a1 = np.array([out[1, k].mean() - ug, out[2, k].mean() - vg])
        assert np.isclose(a1 @ a0, a0 @ a0, rtol=1e-12)           # increment orthogonal to a0
        assert np.isclose(np.linalg.norm(a1), np.linalg.norm(a0) * math.sqrt(1 + (f * dt) ** 2), rtol=1e-12)
        assert np.isclose(du[0, 0], f * dt * (mv - vg), rtol=1e-12)
    # at the geostrophic wind nothing changes; at the equator nothing changes
    g = c.copy()
    g[1] += ug - g[1].mean(axis=(1, 2), keepdims=True)
    g[2] += vg - g[2].mean(axis=(1, 2), keepdims=True)
    gg = g.copy()
    o.geostrophic_forcing(cfg, gg, dt, lat, ug, vg)
    assert np.abs(gg - g).max() < 1e-15 * 10
    e = c.copy()
    o.geostrophic_forcing(cfg, e, dt, 0.0, ug, vg)
    assert np.array_equal(e, c)


def test_uniform_pg_counter_and_stale_means():
    cfg = small_config(**CFG)
    c = state(cfg, 3)
    st = o.OcPgState()
    dt, tau = 0.5, 10.0
    u0, v0 = c[1].mean(), c[2].mean()
    a = c.copy()
    o.uniform_pg_forcing(cfg, a, dt, tau, 2, st)                  # first call: target = current mean
    assert np.allclose(a, c, rtol=0, atol=1e-15) and st.counter == 1
    assert math.isclose(st.u0, u0, rel_tol=1e-13) and math.isclose(st.v0, v0, rel_tol=1e-13)
    a[1] += 1.0                                                   # wind drifts by +1 m/s
    a[2] -= 2.0
    b = a.copy()
    o.uniform_pg_forcing(cfg, b, dt, tau, 2, st)                  # counter 1: means not re-measured
    assert np.allclose(b, a, rtol=0, atol=1e-15)
    o.uniform_pg_forcing(cfg, b, dt, tau, 2, st)                  # counter 2: new means, push back
    assert np.allclose(b[1] - a[1], -dt / tau * 1.0, rtol=1e-12)
    assert np.allclose(b[2] - a[2], dt / tau * 2.0, rtol=1e-12)
    assert np.array_equal(b[[0, 3, 4]], a[[0, 3, 4]]) and st.counter == 3


def test_uniform_pg_height_interpolates_linear_profiles():
    cfg = small_config(**CFG)
    c = state(cfg, 4)
    z = (np.arange(10) + 0.5) * 100.
    c[1] = 3.0 + 0.01 * z[:, None, None] + (c[1] - c[1].mean(axis=(1, 2), keepdims=True))
    c[2] = -1.0 + 0.002 * z[:, None, None]
    dt, tau, u0, v0 = 0.5, 10.0, 9.0, 0.0
    for h in (90.0, 333.0, 700.0):                                # between two centres: exact
        a = c.copy()
        assert o.uniform_pg_forcing_height(cfg, a, dt, h, u0, v0, tau)
        assert np.allclose(a[1] - c[1], dt / tau * (u0 - (3.0 + 0.01 * h)), rtol=1e-12)
        assert np.allclose(a[2] - c[2], dt / tau * (v0 - (-1.0 + 0.002 * h)), rtol=1e-12)
    # above every centre: levels 0 and 1, extrapolated along the same line
    a = c.copy()
    assert o.uniform_pg_forcing_height(cfg, a, dt, 1200.0, u0, v0, tau)
    assert np.allclose(a[1] - c[1], dt / tau * (u0 - (3.0 + 0.01 * 1200.0)), rtol=1e-12)
    # below the first centre the reference would read level -1: refused
    a = c.copy()
    assert not o.uniform_pg_forcing_height(cfg, a, dt, 20.0, u0, v0, tau)
    assert np.array_equal(a, c)


@pytest.mark.parametrize("top_prop", [0.05, 0.5])
def test_sponge_conserves_level_means_of_u_v_t(top_prop):
    cfg = small_config(nx=9, ny=4, nz=12, tracers=[WV], dx=25.)
    c = state(cfg, 5)
    out = c.copy()
    o.sponge_layer(cfg, out, 3.0, 7.0, top_prop)
    for fld in (1, 2, 4):
        assert np.allclose(out[fld].mean(axis=(1, 2)), c[fld].mean(axis=(1, 2)), rtol=1e-13)
