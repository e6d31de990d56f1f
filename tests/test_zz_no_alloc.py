"""pu_level_mean_wind_noalloc (the allocation-free entry point) against pu_level_mean_wind: bitwise
equal level means on two calls in a row, the second with a smaller grid and fewer levels, and a
clear error beyond PU_NOALLOC_MAX_LEVELS."""
import numpy as np
import pytest

import porturb_amd as pa
from tests.helpers import random_state, small_config, WV, TKE

pytestmark = pytest.mark.gpu


def _context(nx, ny, nz, seed, ord=5):
    cfg = small_config(nx=nx, ny=ny, nz=nz, ord=ord, tracers=[WV, TKE], dx=7.)
    st, imm = random_state(cfg, seed=seed)
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    return cfg, st, cs, pa.Dycore(cfg)


def _noalloc(d, cs, k0, k1):
    n = k1 - k0
    U, V = pa.DeviceBuffer(8 * n), pa.DeviceBuffer(8 * n)
    d.level_mean_wind_noalloc(cs, k0, k1, U.ptr, V.ptr)
    d.synchronize()
    u, v = U.download(np.float64, (n,)), V.download(np.float64, (n,))
    U.free()
    V.free()
    return u, v


def test_noalloc_level_means_equal_the_allocating_entry_point():
    print("noalloc: call 1, 64x48x40 FP64, levels [0, 40)", flush=True)
    cfg, st, cs, d = _context(64, 48, 40, seed=71)
    u, v = _noalloc(d, cs, 0, 40)
    uh, vh = d.level_mean_wind(cs, 0, 40)
    assert np.array_equal(u, uh) and np.array_equal(v, vh)
    assert np.allclose(u, st[1].mean(axis=(1, 2)), rtol=1e-14, atol=1e-15)
    assert np.allclose(v, st[2].mean(axis=(1, 2)), rtol=1e-14, atol=1e-15)
    # second call: a smaller grid and fewer levels on the same GPU (the scratch keeps the first
    # call's larger contents beyond what this call writes)
    print("noalloc: call 2, 24x20x12 FP64, levels [2, 9)", flush=True)
    cfg2, st2, cs2, d2 = _context(24, 20, 12, seed=72)
    u2, v2 = _noalloc(d2, cs2, 2, 9)
    uh2, vh2 = d2.level_mean_wind(cs2, 2, 9)
    assert np.array_equal(u2, uh2) and np.array_equal(v2, vh2)
    assert np.allclose(u2, st2[1, 2:9].mean(axis=(1, 2)), rtol=1e-14, atol=1e-15)
    d.destroy()
    d2.destroy()


def test_noalloc_level_means_reject_more_levels_than_the_scratch_holds():
    print("noalloc: 4x4x1100 ord 3, levels beyond the maximum", flush=True)
    cfg, st, cs, d = _context(4, 4, 1100, seed=73, ord=3)
    U, V = pa.DeviceBuffer(8 * 1100), pa.DeviceBuffer(8 * 1100)
    with pytest.raises(RuntimeError, match="maximum"):
        d.level_mean_wind_noalloc(cs, 0, 1100, U.ptr, V.ptr)
    with pytest.raises(RuntimeError):
        d.level_mean_wind_noalloc(cs, 5, 1101, U.ptr, V.ptr)
    U.free()
    V.free()
    u, v = _noalloc(d, cs, 76, 1100)          # exactly the maximum of 1024 levels
    uh, vh = d.level_mean_wind(cs, 76, 1100)
    assert np.array_equal(u, uh) and np.array_equal(v, vh)
    d.destroy()
