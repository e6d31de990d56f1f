"""Size-independent properties at the benchmark's full size (neutral urban ABL, 1024 x 1024 x 128,
WENO-9, the bench workload), where the oracle would take hours:

  * translation equivariance: x and y are periodic and every kernel is the same at every cell, so
    two dycore steps from a state shifted by (37, 13) cells equal the shifted result of two steps
    from the original state.  The shift moves data across the chunks of the whole-column and
    whole-line sweeps and across the periodic ends of their lines.  Agreement is to 1e-10 of each
    field's largest value in FP64 (the bound of the time-step parity tests; measured at most 1.2e-11,
    for w) and 1e-5 in FP32 (measured at most 3.6e-7), not bitwise: the hydrostatic background
    is a horizontal mean whose sum order changes with the shift, and the limiter carries those
    roundings on; a wrong carry, wrap or chunk ownership would differ at the size of the flow's own
    variations;
  * global mass (sum of rho dV, FP64 reduction on the device) is conserved to 1e-13 relative
    when the buildings are taken out (the immersed relaxation pulls the density inside buildings
    towards the hydrostatic background, dynamics_rk.h:370-473, which moves mass);
  * every value stays finite.
FP64 and FP32 modes."""
import gc

import numpy as np
import pytest

import porturb_amd as pa
from porturb_amd import cases

pytestmark = pytest.mark.gpu

SHIFT = (13, 37)            # cells in y, x


def two_steps(cfg, st, imm, dt=None):
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    d.init(cs)
    m0 = d.diagnostics(cs)[0]
    if dt is None:
        dt = d.compute_time_step()
    for _ in range(2):
        d.time_step(cs, dt)
    m1 = d.diagnostics(cs)[0]
    out = cs.download()
    d.destroy()
    del cs
    gc.collect()
    return out, m0, m1, dt


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32], ids=["fp64", "fp32"])
def test_full_size_translation_equivariance_and_mass(precision):
    cfg = cases.abl_urban(precision=precision)
    print("workload", cfg["name"], flush=True)
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, cfg["case_id"], 11, cs)
    st, imm = cs.download(), cs.download_immersed()
    del cs
    gc.collect()
    assert imm.max() > 0                                    # the city's buildings are there
    a, _, _, dt = two_steps(cfg, st, imm)
    assert np.isfinite(a).all()
    c, m0, m1, _ = two_steps(cfg, st, np.zeros_like(imm), dt)
    assert np.isfinite(c).all()
    assert abs(m1 - m0) / m0 < 1e-13, (m0, m1)
    del c
    st = np.roll(st, SHIFT, axis=(2, 3))
    imm = np.roll(imm, SHIFT, axis=(1, 2))
    b, _, _, _ = two_steps(cfg, st, imm, dt)
    del st, imm
    for f in range(a.shape[0]):
        ref = np.roll(a[f], SHIFT, axis=(1, 2))
        err = np.abs(b[f] - ref).max() / max(np.abs(ref).max(), 1e-300)
        print("field", f, "shifted-run difference", err, flush=True)
        assert err < (1e-10 if precision == pa.PU_FP64 else 1e-5), (f, err)
