"""Pins of the surface-flux oracle (modules::apply_surface_fluxes, surface_flux.h:9-126) by closed
forms: the ground below the first level and a building face act as log-law drag, nothing else
changes, and the temperature exchange needs a nonzero wall temperature."""
import numpy as np

from oracle import oracle_py as o
from tests.helpers import small_config, WV

KAPPA = 0.4


def state(cfg, u=3.0, v=-1.0, w=0.5, T=290.):
    n = (cfg["nz"], cfg["ny"], cfg["nx"])
    c = np.zeros((5 + cfg["num_tracers"],) + n)
    c[0], c[1], c[2], c[3], c[4] = 1.2, u, v, w, T
    return c


def cd(delta, z0):
    return KAPPA ** 2 / np.log((delta / 2 + z0) / z0) ** 2


def run(cfg, c, prop=None, rough=0.1, temp=0., sfc_rough=0.1, sfc_temp=0., dt=0.5):
    n = (cfg["nz"], cfg["ny"], cfg["nx"])
    prop = np.zeros(n) if prop is None else prop
    arr = lambda x, shape: np.broadcast_to(np.asarray(x, float), shape).copy()
    out = c.copy()
    o.surface_fluxes(cfg, out, prop, arr(rough, n), arr(temp, n), arr(sfc_rough, n[1:]), arr(sfc_temp, n[1:]), dt)
    return out


def test_ground_drag_closed_form_and_interior_untouched():
    cfg = small_config(nx=6, ny=5, nz=4, tracers=[WV], dx=2.)
    c = state(cfg)
    out = run(cfg, c, dt=0.5)
    k = cd(2., 0.1)
    mag = np.hypot(3.0, -1.0)
    assert np.allclose(out[1, 0], 3.0 - 0.5 * k * 3.0 * mag / 2.)
    assert np.allclose(out[2, 0], -1.0 - 0.5 * k * (-1.0) * mag / 2.)
    assert np.array_equal(out[3], c[3]) and np.array_equal(out[4], c[4])      # w has no z-face drag; T0 = 0
    assert np.array_equal(out[1:, 1:], c[1:, 1:])                               # no immersed neighbour above k=0


def test_building_face_and_wall_temperature():
    cfg = small_config(nx=6, ny=5, nz=4, tracers=[WV], dx=2.)
    c = state(cfg)
    prop = np.zeros((4, 5, 6))
    prop[2, 2, 3] = 1.0                        # a building cell: its x-neighbour (2,2,2) sees it
    out = run(cfg, c, prop=prop, temp=300., sfc_temp=295., dt=0.5)
    k = cd(2., 0.1)
    mag = np.hypot(-1.0, 0.5)                  # tangential (v, w) of an x face
    assert np.allclose(out[2, 2, 2, 2], -1.0 - 0.5 * k * (-1.0) * mag / 2.)
    assert np.allclose(out[3, 2, 2, 2], 0.5 - 0.5 * k * 0.5 * mag / 2.)
    assert np.allclose(out[4, 2, 2, 2], 290. - 0.5 * k * (290. - 300.) * mag / 2.)
    assert np.array_equal(out[1, 2, 2, 2], c[1, 2, 2, 2])      # u is normal to that face
    # ground temperature exchange at the first level
    magz = np.hypot(3.0, -1.0)
    assert np.allclose(out[4, 0, 0, 0], 290. - 0.5 * k * (290. - 295.) * magz / 2.)


def test_periodic_neighbour_across_x_boundary():
    cfg = small_config(nx=6, ny=5, nz=4, tracers=[WV], dx=2.)
    c = state(cfg)
    prop = np.zeros((4, 5, 6))
    prop[2, 1, 5] = 0.5                        # last column: seen by (2,1,0) through the periodic wrap
    out = run(cfg, c, prop=prop)
    assert not np.array_equal(out[2, 2, 1, 0], c[2, 2, 1, 0])
    assert not np.array_equal(out[2, 2, 1, 4], c[2, 2, 1, 4])
    assert np.array_equal(out[2, 2, 1, 2], c[2, 2, 1, 2])
