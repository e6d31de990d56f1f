"""Pins of the oracle's dycore, limiter map and LES closure by answers that a plausible mistake
fails (tests/pin_cases.py describes each case and what fixes its answer).

tests/test_oracle_mutants.py plants such mistakes in copies of the oracle's source and checks
that this file catches every one of them.
"""
import numpy as np
import pytest

from oracle import oracle_py
from tests import pin_cases as pc


# ------------------------------------------------------------------------------------------------
# map_rs (WenoLimiter.h:76-82): endpoints, fixed point, monotonicity, contact order and constant
# ------------------------------------------------------------------------------------------------
def test_map_rs_endpoints_fixed_point_and_monotonicity():
    for d in (0.1, 0.3, 0.6, 1. / 126.):
        assert oracle_py.weno_map_rs(0., d) == pytest.approx(0., abs=1e-15)
        assert oracle_py.weno_map_rs(1., d) == pytest.approx(1., abs=1e-15)
        assert oracle_py.weno_map_rs(d, d) == d
        w = np.linspace(0., 1., 2001)
        m = np.array([oracle_py.weno_map_rs(x, d) for x in w])
        assert (np.diff(m) >= -1e-15).all()


def _contact_constant(d, h):
    """Symmetric quotient (map(d+h) - d)/h^7, averaged over +-h (the O(h) term cancels), with one
    Richardson step between h and h/2."""
    def q(hh):
        return 0.5 * ((oracle_py.weno_map_rs(d + hh, d) - d) / hh ** 7 + (oracle_py.weno_map_rs(d - hh, d) - d) / (-hh) ** 7)
    return (4 * q(h / 2) - q(h)) / 3


@pytest.mark.parametrize("d", (0.3, 0.5, 0.6))
def test_map_rs_contact_of_order_seven_with_its_constant(d):
    """map(d + h) - d = h^7 / (h^6 + s (w (1-w))^m) -> h^7 / (s (d (1-d))^m) with k = 6 (contact of
    order k+1), m = 3, s = 2000.  k = 4 would diverge here, m = 2 or s = 1000 miss by 4x or 2x."""
    L = _contact_constant(d, 0.04)
    assert L * 2000. * (d * (1 - d)) ** 3 == pytest.approx(1., rel=3e-3)


def _cell_averages(coefs, offsets):
    out = []
    for o in offsets:
        lo, hi = o - 0.5, o + 0.5
        out.append(sum(c * (hi ** (n + 1) - lo ** (n + 1)) / (n + 1) for n, c in enumerate(coefs)))
    return np.array(out)


@pytest.mark.parametrize("ord_", (5, 9))
def test_limiter_applies_the_map_with_seventh_order_contact(ord_):
    """Through the limiter itself: as the unmapped weights w approach the ideal d (|w - d| = delta),
    the mapped weights approach d like delta^7 (slope of log|w' - d| against log delta), and the
    mapped weights are map_rs(w, d) renormalised."""
    hs = (ord_ - 1) // 2
    t = oracle_py.weno_tables(ord_)
    pts = []
    # a quadratic (all sub-stencils exact, ideal weights) plus eps x^(hs+1): as eps falls the
    # unmapped weights move towards the ideal ones along a fixed direction
    for eps in np.geomspace(0.2, 1e-3, 40):
        s = _cell_averages([0.3, 1.0, 0.5] + [0.] * (hs - 2) + [eps], range(-hs, hs + 1))
        w0L, _ = oracle_py.weno_face_weights(ord_, s, False)
        w1L, _ = oracle_py.weno_face_weights(ord_, s, True)
        mapped = np.array([oracle_py.weno_map_rs(w, d) for w, d in zip(w0L, t["idealL"])])
        np.testing.assert_allclose(w1L, mapped / mapped.sum(), rtol=1e-13, atol=1e-15)
        delta = np.abs(w0L - t["idealL"]).max()
        e = np.abs(w1L - t["idealL"]).max()
        if 0.03 < delta < 0.25 and e > 1e-13:
            pts.append((np.log(delta), np.log(e)))
    pts = np.array(pts)
    assert len(pts) >= 4
    slope = np.polyfit(pts[:, 0], pts[:, 1], 1)[0]
    assert 6.3 < slope < 7.7, slope


# ------------------------------------------------------------------------------------------------
# Dycore
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("ord_", (5, 9))
@pytest.mark.parametrize("axis", (0, 1, 2))
@pytest.mark.parametrize("U", (6., -6.))
def test_step_tracer_is_upwinded(ord_, axis, U):
    cfg, st, imm, expect = pc.step_advection(ord_, axis, U)
    tend = oracle_py.Dycore(cfg, st, imm).tendencies(st)
    pc.check_step_tendencies(cfg, tend, expect, U)


def _advect(cfg, st, imm, T, dt_factor=1.):
    d = oracle_py.Dycore(cfg, st, imm)
    dt = d.compute_time_step() * dt_factor
    n = int(round(T / dt))
    s = st.copy()
    for _ in range(n):
        d.time_step(s, dt)
    return s, n * dt


@pytest.mark.parametrize("ord_", (5, 9))
def test_smooth_tracer_translates_at_flow_speed(ord_):
    """After T = 10 s at U = 10 m/s the tracer equals q(x - U T) to truncation error, which falls at
    about the design order between 16 and 32 cells per wavelength; q(x + U T) (the wrong direction)
    is 0.35 away.  WENO-9 takes quarter steps so that the SSPRK3 error (third order in dt) stays
    below its spatial error."""
    errs = []
    for nx in (16, 32):
        cfg, st, imm, exact = pc.smooth_advection(ord_, nx)
        s, t = _advect(cfg, st, imm, 10., 1. if ord_ <= 5 else 0.25)
        q = s[5 + 1][0, 0] / s[0][0, 0]
        errs.append(np.abs(q - exact(t)).max())
        assert np.abs(q - exact(-t)).max() > 0.3
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < 2e-3 and rate > ord_ - 1.5, (errs, rate)


def test_acoustic_pulse_splits_at_the_sound_speed():
    cfg, st, imm, linear, cs = pc.acoustic_pulse(5)
    assert 340. < cs < 355.
    s, t = _advect(cfg, st, imm, 0.9)     # each half travels about 310 m = 31 cells
    pc.check_acoustics(s, linear, t)


@pytest.mark.parametrize("ord_", (5, 9))
def test_denser_cell_sinks(ord_):
    cfg, st, imm, cell, drho = pc.buoyancy_cell(ord_)
    tend = oracle_py.Dycore(cfg, st, imm).tendencies(st)
    pc.check_buoyancy_tendencies(cfg, st, tend, cell, drho)


def test_warm_bubble_rises():
    cfg, st, imm = pc.warm_bubble(n=24)
    d = oracle_py.Dycore(cfg, st, imm)
    s = st.copy()
    for _ in range(10):
        d.time_step(s, d.compute_time_step())
    pc.check_bubble_rises(cfg, s)


@pytest.mark.parametrize("ord_", (5, 9))
def test_uniform_flow_rotates_inertially(ord_):
    cfg, st, imm, f = pc.inertial(ord_)
    d = oracle_py.Dycore(cfg, st, imm)
    dt = d.compute_time_step()
    N = 1000
    s = st.copy()
    for _ in range(N):
        d.time_step(s, dt)
    assert f * dt * N > 0.4                     # a wrong third-order term would show at 1e-8
    pc.check_inertial(st, s, f, dt, N)


# ------------------------------------------------------------------------------------------------
# LES closure
# ------------------------------------------------------------------------------------------------
LES_DT = 1.0


def _les(cfg, st, imm):
    out = st.copy()
    oracle_py.les_apply(cfg, out, imm, LES_DT)
    return out


@pytest.mark.parametrize("gamma_theta", (0.01, -0.01))
def test_les_heat_flux_is_down_gradient_and_feeds_buoyancy(gamma_theta):
    cfg, st, imm = pc.les_stratified(gamma_theta)
    cfg0, st0, _ = pc.les_stratified(gamma_theta, enable_gravity=0)
    pc.check_les_stratified(cfg, st, _les(cfg, st, imm), _les(cfg0, st0, imm), gamma_theta, LES_DT)


@pytest.mark.parametrize("S", (0.02, -0.03))
def test_les_momentum_flux_is_down_gradient_and_feeds_shear_production(S):
    cfg, st, imm = pc.les_shear(S)
    cfg0, st0, _ = pc.les_shear(0.)
    pc.check_les_shear(cfg, st, _les(cfg, st, imm), _les(cfg0, st0, imm), S, LES_DT)


def test_2d_runs_have_no_v_tendency_and_no_coriolis():
    """ny == 1 (dynamics_rk.h:663,716,720-721): V's tendency is zero and the latitude changes
    nothing; the same state in 3-D (ny = 4 copies) has both."""
    from tests.helpers import small_config, random_state, WV, P1
    cfg = small_config(nx=30, ny=1, nz=12, ord=5, tracers=[WV, P1], latitude=35.)
    st, imm = random_state(cfg, seed=7)
    t_lat = oracle_py.Dycore(cfg, st, imm).tendencies(st)
    t_eq = oracle_py.Dycore(dict(cfg, latitude=0.), st, imm).tendencies(st)
    assert np.array_equal(t_lat[2], np.zeros_like(t_lat[2]))
    assert np.array_equal(t_lat, t_eq)
    cfg3 = small_config(nx=30, ny=4, nz=12, ord=5, tracers=[WV, P1], latitude=35.)
    st3 = np.repeat(st, 4, axis=2)
    t3 = oracle_py.Dycore(cfg3, st3, np.zeros(st3.shape[1:])).tendencies(st3)
    assert np.abs(t3[2]).max() > 0
    # x and z fluxes are the same in 2-D and in the y-uniform 3-D state: only V and the Coriolis
    # term of U differ
    t3_eq = oracle_py.Dycore(dict(cfg3, latitude=0.), st3, np.zeros(st3.shape[1:])).tendencies(st3)
    np.testing.assert_allclose(t3_eq[[0, 3, 4, 5, 6], :, :1], t_eq[[0, 3, 4, 5, 6]], rtol=1e-9,
                               atol=1e-12 * np.abs(t_eq).max())
