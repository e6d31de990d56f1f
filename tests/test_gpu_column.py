"""GPU parity of the column-mean forcing modules (column_kernels.hip through the C ABI) with the
oracle: modules::sponge_layer, modules::geostrophic_wind_forcing and the reductions/updates of
modules::uniform_pg_wind_forcing(_height); and the C++ drop-ins of all four inside the driver step,
bitwise against the same C ABI calls made one by one."""
import numpy as np
import pytest

import porturb_amd as pa
from porturb_amd import cases
from oracle import oracle_py
from tests.helpers import small_config, random_state, rel_norms, WV, TKE, P1

pytestmark = pytest.mark.gpu


def setup(cfg, seed):
    st, imm = random_state(cfg, seed=seed)
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    return st, d, cs


@pytest.mark.parametrize("top_prop", [0.3, 1.0, 0.01])
def test_sponge_matches_oracle(top_prop):
    # ragged sizes: the chunked level sums end mid-row
    cfg = small_config(nx=37, ny=23, nz=19, tracers=[WV, TKE, P1], dx=7.)
    st, d, cs = setup(cfg, 41)
    d.sponge_layer(cs, 0.9, 4.0, top_prop)
    ref = st.copy()
    oracle_py.sponge_layer(cfg, ref, 0.9, 4.0, top_prop)
    g = cs.download()
    assert np.array_equal(g[0], st[0]) and np.array_equal(g[5:], st[5:])
    if top_prop > 0.1:
        assert np.abs(g[1:5] - st[1:5]).max() > 0
    nrm = rel_norms(g, ref)
    assert nrm.max() < 1e-14, nrm


def test_geostrophic_matches_oracle():
    cfg = small_config(nx=37, ny=23, nz=19, tracers=[WV, TKE, P1], dx=7.)
    st, d, cs = setup(cfg, 42)
    d.geostrophic_forcing(cs, 2.5, 41.0, 10.0, -3.0)
    ref = st.copy()
    oracle_py.geostrophic_forcing(cfg, ref, 2.5, 41.0, 10.0, -3.0)
    g = cs.download()
    assert np.abs(g[1:3] - st[1:3]).max() > 0
    nrm = rel_norms(g, ref)
    assert nrm.max() < 1e-14, nrm


def test_level_means_and_uniform_wind_update():
    cfg = small_config(nx=37, ny=23, nz=19, tracers=[WV, TKE, P1], dx=7.)
    st, d, cs = setup(cfg, 43)
    u, v = d.level_mean_wind(cs, 0, cfg["nz"])
    assert np.allclose(u, st[1].mean(axis=(1, 2)), rtol=1e-14, atol=1e-15)
    assert np.allclose(v, st[2].mean(axis=(1, 2)), rtol=1e-14, atol=1e-15)
    u2, v2 = d.level_mean_wind(cs, 5, 7)
    assert np.array_equal(u2, u[5:7]) and np.array_equal(v2, v[5:7])   # same partition, same sums
    with pytest.raises(RuntimeError):
        d.level_mean_wind(cs, 3, cfg["nz"] + 1)
    d.add_uniform_wind(cs, 0.25, -0.5)
    g = cs.download()
    assert np.array_equal(g[1], st[1] + 0.25) and np.array_equal(g[2], st[2] - 0.5)
    assert np.array_equal(g[3:], st[3:])


def _raw_case(nx, ny, nz, dx, seed):
    cfg = cases.abl_urban(nx, ny, nz, dx, ord=9, precision=pa.PU_FP64)
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, pa.PU_CASE_ABL_URBAN, seed, cs)
    d = pa.Dycore(cfg)
    d.init(cs)
    d.surface_init(np.full((nz, ny, nx), 0.1), np.zeros((nz, ny, nx)), np.full((ny, nx), 0.1), np.zeros((ny, nx)))
    return cfg, cs, d


def _compare(drv, cfg, cs):
    a = drv.download()
    b = cs.download()
    names = list(pa.CouplerState.FIELDS) + [t[0] for t in cfg["tracers"]]
    for i, n in enumerate(names):
        assert np.array_equal(a[n], b[i]), n


def test_driver_abl_forcing_matches_raw_calls():
    """abl_neutral.cpp's module order: geostrophic forcing, dycore, sponge, surface fluxes, closure."""
    nx, ny, nz, dx = 48, 32, 32, 2.0
    drv = pa.Driver(pa.PU_CASE_ABL_URBAN, nx, ny, nz, dx, ord=9, seed=78)
    drv.set_forcing(geostrophic=1, lat_g=40.0, u_g=8.0, v_g=0.5, sponge=1, sponge_dt_multiple=100.0,
                    sponge_top_prop=0.1)
    drv.step(3)
    cfg, cs, d = _raw_case(nx, ny, nz, dx, 78)
    dt = d.compute_time_step()
    for _ in range(3):
        d.geostrophic_forcing(cs, dt, 40.0, 8.0, 0.5)
        d.time_step(cs, dt)
        d.sponge_layer(cs, dt, 100.0 * dt, 0.1)
        d.surface_apply(cs, dt)
        d.les_apply(cs, dt)
    _compare(drv, cfg, cs)
    drv.destroy()


@pytest.mark.parametrize("pg", [1, 2])
def test_driver_pg_forcing_matches_raw_calls(pg):
    """uniform_pg_wind_forcing (counter, stale means every 2nd call) and _height (city.cpp)."""
    nx, ny, nz, dx = 48, 32, 32, 2.0
    drv = pa.Driver(pa.PU_CASE_ABL_URBAN, nx, ny, nz, dx, ord=9, seed=79)
    drv.set_forcing(pg=pg, pg_height=21.0, pg_u0=6.0, pg_v0=1.0, pg_tau=0.05, pg_update_cycles=2)
    drv.step(3)
    cfg, cs, d = _raw_case(nx, ny, nz, dx, 79)
    dt = d.compute_time_step()
    u0 = v0 = um = vm = None
    for n in range(3):
        if pg == 1:
            if n % 2 == 0:
                u, v = d.level_mean_wind(cs, 0, nz)
                um = vm = 0.0
                for k in range(nz):               # the C++ module's summation order
                    um += u[k]
                    vm += v[k]
                um, vm = um / nz, vm / nz
                if n == 0:
                    u0, v0 = um, vm
            d.add_uniform_wind(cs, dt / 0.05 * (u0 - um), dt / 0.05 * (v0 - vm))
        else:
            k1 = int(np.nonzero((np.arange(nz) + 0.5) * dx > 21.0)[0][0]) - 1
            u, v = d.level_mean_wind(cs, k1, k1 + 2)
            z1, z2 = (k1 + 0.5) * dx, (k1 + 1.5) * dx
            w1, w2 = (z2 - 21.0) / dx, (21.0 - z1) / dx
            d.add_uniform_wind(cs, dt / 0.05 * (6.0 - (w1 * u[0] + w2 * u[1])), dt / 0.05 * (1.0 - (w1 * v[0] + w2 * v[1])))
        d.time_step(cs, dt)
        d.surface_apply(cs, dt)
        d.les_apply(cs, dt)
    _compare(drv, cfg, cs)
    drv.destroy()
