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
  * every value stays finite;
  * the bench driver's whole physics step (dycore + rough-wall surface fluxes + closure; FP64 with
    the TKE closure, FP32 with the learned closure) is translation-equivariant in the same way.
FP64 and FP32 modes."""
import gc
import os

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


SGS_MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "porturb_amd", "data",
                         "sgs_abl_urban_dx2.bin")


def two_driver_steps(cfg, st, imm, learned, dt=None):
    """Two steps of the bench driver's physics step through the C ABI: dycore, rough-wall surface
    fluxes (roughness 0.1 m, no wall temperatures, as the driver sets them) and the closure."""
    nz, ny, nx = cfg["nz"], cfg["ny"], cfg["nx"]
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    d.init(cs)
    d.surface_init(np.full((nz, ny, nx), 0.1), np.zeros((nz, ny, nx)), np.full((ny, nx), 0.1), np.zeros((ny, nx)))
    mlp = None
    if learned:
        d.les_init(cs)
        mlp = pa.Mlp.load(SGS_MODEL)
        d.set_sgs_model(mlp)
    if dt is None:
        dt = d.compute_time_step()
    for _ in range(2):
        d.time_step(cs, dt)
        d.surface_apply(cs, dt)
        d.les_apply(cs, dt)
    out = cs.download()
    d.destroy()
    del cs, mlp
    gc.collect()
    return out, dt


@pytest.mark.parametrize("precision,learned", [(pa.PU_FP64, False), (pa.PU_FP32, True)],
                         ids=["fp64_tke", "fp32_learned"])
def test_full_size_driver_step_translation_equivariance(precision, learned):
    """The bench driver's whole physics step (dycore + surface fluxes + closure) at 1024x1024x128
    with the city's buildings: two steps from the state and buildings shifted by (37, 13) cells
    equal the shifted result.  FP64 with the TKE closure: 1e-10 of each field's largest value, as
    for the dycore alone (measured at most 2.1e-13, for w).  FP32 with the learned closure: the
    velocities on the flow's speed scale, the other fields on their own, 1e-5 at the worst cell, as
    for the FP32 dycore alone, and 1e-8 in RMS (measured 6.4e-7 and 6.2e-10); a rounding-level
    difference that flipped the network's fallback flag of a cell would change that cell's fluxes by
    the network-vs-eddy-viscosity difference and show here."""
    cfg = cases.abl_urban(precision=precision)
    print("workload", cfg["name"], "learned" if learned else "tke", flush=True)
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, cfg["case_id"], 11, cs)
    st, imm = cs.download(), cs.download_immersed()
    del cs
    gc.collect()
    a, dt = two_driver_steps(cfg, st, imm, learned)
    assert np.isfinite(a).all()
    st = np.roll(st, SHIFT, axis=(2, 3))
    imm = np.roll(imm, SHIFT, axis=(1, 2))
    b, _ = two_driver_steps(cfg, st, imm, learned, dt)
    del st, imm
    assert np.isfinite(b).all()
    vscale = np.abs(a[1]).max()
    for f in range(a.shape[0]):
        ref = np.roll(a[f], SHIFT, axis=(1, 2))
        diff = np.abs(b[f] - ref)
        scale = vscale if (learned and f in (1, 2, 3)) else max(np.abs(ref).max(), 1e-300)
        err, rms = diff.max() / scale, np.sqrt((diff ** 2).mean()) / scale
        print("field", f, "shifted-run difference max %.3e rms %.3e" % (err, rms), flush=True)
        if precision == pa.PU_FP64:
            assert err < 1e-10, (f, err)
        else:
            assert err < 1e-5 and rms < 1e-8, (f, err, rms)
        del ref, diff
