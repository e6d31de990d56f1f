"""The whole-line sweeps rearrange which thread computes what, not the operations, and are checked
against the segment kernels (forced by the environment variable PORTURB_SEGMENT_SWEEPS, narrowed to
some directions by PORTURB_SEGMENT_DIRS, when the context is created):
  * k_sweep_zwall (z between solid walls) agrees bitwise -- tendencies and whole time steps, FP64
    and FP32, columns shorter than one chunk of 32 cells, exact multiples and ragged last chunks;
  * k_sweep_pline (periodic y lines) agrees to rounding (the compiler fuses some multiply-adds
    differently): tendencies within 1e-13 of each field's largest, states after two steps within
    1e-12 relative in FP64, both within 1e-5 in FP32 -- lines shorter than one chunk, exact
    multiples, a last chunk holding only the periodic end cell and two-cell lines.  A wrong carry, wrap or ownership
    would differ at the size of the tendencies themselves."""
import os

import numpy as np
import pytest

import porturb_amd as pa
from tests.helpers import random_state, small_config, WV, TKE, P1

pytestmark = pytest.mark.gpu


def run(cfg, st, imm, segment_dirs):
    if segment_dirs:
        os.environ["PORTURB_SEGMENT_SWEEPS"] = "1"
        os.environ["PORTURB_SEGMENT_DIRS"] = segment_dirs
    try:
        cs = pa.CouplerState(cfg)
        cs.upload((st, imm))
        d = pa.Dycore(cfg)
    finally:
        os.environ.pop("PORTURB_SEGMENT_SWEEPS", None)
        os.environ.pop("PORTURB_SEGMENT_DIRS", None)
    d.init(cs)
    tend = d.tendencies(cs)
    dt = d.compute_time_step()
    for _ in range(2):
        d.time_step(cs, dt)
    return tend, cs.download()


def setup(case):
    kw = dict(case)
    nx, ny, nz = kw.pop("nx"), kw.pop("ny"), kw.pop("nz")
    cfg = small_config(nx=nx, ny=ny, nz=nz, tracers=[WV, TKE, P1], **kw)
    st, imm = random_state(cfg, seed=5, buildings=3 if min(nx, ny) > 4 else 0)
    print("case", case, flush=True)
    return cfg, st, imm


ZCASES = [
    dict(ord=9, nx=40, ny=12, nz=70),
    dict(ord=9, nx=37, ny=10, nz=33, precision=pa.PU_FP32),
    dict(ord=5, nx=24, ny=16, nz=64),
    dict(ord=11, nx=20, ny=12, nz=14),
    dict(ord=3, nx=16, ny=8, nz=129, precision=pa.PU_FP32),
]


@pytest.mark.parametrize("case", ZCASES, ids=[str(c) for c in ZCASES])
def test_zwall_sweep_matches_segment_sweep_bitwise(case):
    cfg, st, imm = setup(case)
    t1, s1 = run(cfg, st, imm, None)
    t2, s2 = run(cfg, st, imm, "z")
    assert np.abs(t1).max() > 0
    for f in range(t1.shape[0]):
        assert np.array_equal(t1[f], t2[f]), ("tendency", f, np.abs(t1[f] - t2[f]).max())
    for f in range(s1.shape[0]):
        assert np.array_equal(s1[f], s2[f]), ("state", f, np.abs(s1[f] - s2[f]).max())


YCASES = [
    dict(ord=9, nx=20, ny=70, nz=12),
    dict(ord=9, nx=20, ny=70, nz=12, precision=pa.PU_FP32),
    dict(ord=3, nx=16, ny=2, nz=10, precision=pa.PU_FP32),
    dict(ord=5, nx=24, ny=64, nz=10),
    dict(ord=7, nx=16, ny=65, nz=10, bc_z=pa.PU_BC_PERIODIC),
    dict(ord=9, nx=18, ny=20, nz=12),
    dict(ord=3, nx=22, ny=2, nz=10),
]


@pytest.mark.parametrize("case", YCASES, ids=[str(c) for c in YCASES])
def test_periodic_line_sweep_matches_segment_sweep(case):
    cfg, st, imm = setup(case)
    t1, s1 = run(cfg, st, imm, None)
    t2, s2 = run(cfg, st, imm, "y")
    fp32 = case.get("precision") == pa.PU_FP32
    scale = np.maximum(np.abs(t2).reshape(t2.shape[0], -1).max(axis=1), 1e-300)
    err = np.abs(t1 - t2).reshape(t1.shape[0], -1).max(axis=1) / scale
    assert err.max() < (1e-5 if fp32 else 1e-13), err
    for f in range(s1.shape[0]):
        d = np.abs(s1[f] - s2[f]).max() / max(np.abs(s2[f]).max(), 1e-300)
        assert d < (1e-5 if fp32 else 1e-12), (f, d)
