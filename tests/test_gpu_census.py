"""The reconstruction census (pu_dycore_flat_census) against known answers: which branch of the
limiter the sweeps take for every interior cell, field and direction.

* smooth random fields, WENO everywhere: every reconstruction takes the full limiter, except a
  passive tracer of mixing ratio ~1e-13, whose every stencil has a total variation far below the
  1e-20 guard (all of its reconstructions take the flat-stencil branch);
* weno_all = 0 without buildings: every reconstruction is linear;
* weno_all = 0 with buildings: the cells with an immersed cell within the 10-cell box
  (dycore_any_immersed10) take WENO, the others the linear reconstruction.
"""
import numpy as np
import pytest

import porturb_amd as pa
from tests.helpers import random_state, small_config, WV, TKE, P1

pytestmark = pytest.mark.gpu


def census_of(cfg, st, imm):
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    d.init(cs)
    c = d.flat_census(cs)
    any10, _ = d.immersed_masks()
    d.destroy()
    return c, any10


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32])
def test_census_limiter_and_flat_tracer(precision):
    cfg = small_config(nx=24, ny=20, nz=16, ord=9, tracers=[WV, TKE, P1], precision=precision)
    st, imm = random_state(cfg, seed=3)
    st[5 + 2] = st[0] * 1e-13 * (1.0 + 0.5 * np.sin(np.arange(st[0].size).reshape(st[0].shape)))
    c, _ = census_of(cfg, st, imm)
    n3 = cfg["nx"] * cfg["ny"] * cfg["nz"]
    nf = 6 + cfg["num_tracers"]
    for d in range(3):
        assert c[d].sum() == n3 * nf, (d, c[d])
        assert c[d][1] == n3, (d, c[d])            # the tiny tracer, nothing else
        assert c[d][0] == n3 * (nf - 1), (d, c[d])
        assert c[d][2] == 0


def test_census_linear_without_weno():
    cfg = small_config(nx=24, ny=20, nz=16, ord=7, weno_all=0)
    st, imm = random_state(cfg, seed=4)
    c, any10 = census_of(cfg, st, imm)
    n3 = cfg["nx"] * cfg["ny"] * cfg["nz"]
    nf = 6 + cfg["num_tracers"]
    assert any10.sum() == 0
    assert (c[:, 2] == n3 * nf).all() and (c[:, :2] == 0).all(), c


def test_census_weno_only_near_buildings():
    cfg = small_config(nx=32, ny=24, nz=16, ord=5, weno_all=0)
    st, imm = random_state(cfg, seed=5, buildings=3)
    c, any10 = census_of(cfg, st, imm)
    n3 = cfg["nx"] * cfg["ny"] * cfg["nz"]
    nf = 6 + cfg["num_tracers"]
    na = int((any10 != 0).sum())
    assert 0 < na < n3
    for d in range(3):
        assert c[d][2] == (n3 - na) * nf, (d, c[d], na)
        assert c[d][0] + c[d][1] == na * nf, (d, c[d], na)
