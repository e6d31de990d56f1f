"""Closed-form tests of the oracle's LES closure (LES_Closure::apply, les_closure.h:30-323).

* uniform flow: every gradient vanishes, so only TKE dissipation acts:
  rho*K -> rho*(K - dt*(0.19+0.51)*K^1.5/Delta) with mixing length ell = Delta (N = 0).
* uniform shear u = S z (no gravity): momentum fluxes are uniform (no divergence in the interior)
  and shear production adds rho*km*S^2 with km = 0.1*Delta*sqrt(K).
* diffused tracer mass is conserved (zero flux through the immersed bottom halo, zero-gradient top).
"""
import numpy as np
import pytest

from oracle import oracle_py
from porturb_amd import physical_constants
from tests.helpers import small_config, random_state, WV, TKE, P1


def uniform(cfg, u=3., v=-1., K=0.4, rho=1.1, theta=300., shear=0.):
    c = physical_constants()
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    st = np.zeros((5 + cfg["num_tracers"], nz, ny, nx))
    dz = cfg["zlen"] / nz
    z = (np.arange(nz) + 0.5) * dz
    st[0] = rho
    st[1] = u + shear * z[:, None, None]
    st[2] = v
    st[4] = c["C0"] * (rho * theta) ** c["gamma_d"] / (rho * c["R_d"])
    st[5 + cfg["idTKE"]] = rho * K
    return st


def test_uniform_flow_only_dissipates_tke():
    cfg = small_config(tracers=[WV, TKE, P1])
    st = uniform(cfg)
    imm = np.zeros(st.shape[1:])
    out = st.copy()
    dt = 0.5
    oracle_py.les_apply(cfg, out, imm, dt)
    delta = (cfg["xlen"] / cfg["nx"] * cfg["ylen"] / cfg["ny"] * cfg["zlen"] / cfg["nz"]) ** (1 / 3)
    K = 0.4
    expect = 1.1 * (K - dt * (0.19 + 0.51) / delta * K ** 1.5)
    np.testing.assert_allclose(out[5 + cfg["idTKE"]], expect, rtol=1e-13)
    for f in (0, 1, 2, 3):
        np.testing.assert_allclose(out[f], st[f], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(out[4], st[4], rtol=1e-12)


def test_uniform_shear_production():
    cfg = small_config(tracers=[WV, TKE], enable_gravity=0)
    S = 0.05
    st = uniform(cfg, shear=S, v=0.)
    imm = np.zeros(st.shape[1:])
    out = st.copy()
    dt = 0.5
    oracle_py.les_apply(cfg, out, imm, dt)
    delta = (cfg["xlen"] / cfg["nx"] * cfg["ylen"] / cfg["ny"] * cfg["zlen"] / cfg["nz"]) ** (1 / 3)
    K = 0.4
    km = 0.1 * delta * np.sqrt(K)
    expect = 1.1 * K + dt * (1.1 * km * S * S - 1.1 * (0.19 + 0.51) / delta * K ** 1.5)
    inner = slice(2, cfg["nz"] - 2)
    np.testing.assert_allclose(out[5 + cfg["idTKE"]][inner], expect, rtol=1e-12)
    np.testing.assert_allclose(out[1][inner], st[1][inner], rtol=1e-12)


def test_tracer_mass_is_conserved():
    cfg = small_config(tracers=[WV, TKE, P1])
    st, imm = random_state(cfg, seed=4, buildings=2)
    out = st.copy()
    oracle_py.les_apply(cfg, out, imm, 0.5)
    for t in (0, 2):
        m0, m1 = st[5 + t].sum(), out[5 + t].sum()
        assert abs(m1 - m0) / abs(m0) < 1e-12
