"""FP32 (compensated) dycore against the FP64 dycore on the same seeded state: per-field errors and
the mean-wind / resolved-TKE profile differences after a number of steps.

Both runs start from one seeded neutral urban ABL state (``pu_case_init``), take the same dt each
step (the FP64 run's ``compute_time_step``) and apply the dycore step plus the TKE closure.  After
each checkpoint the coupler entries of both runs are compared field by field:

* ``l2``  = ||a32 - a64||_2 / ||a64||_2
* ``linf`` = max|a32 - a64| / max|a64|
* for the velocities also the error relative to the FP64 field's spread (max - min), which is the
  meaningful scale for a field whose mean may be near zero;

and the horizontal-mean profiles the reference's analysis plots (build/postproc/
abl_neutral_analysis.py:122-131): <u>(z), <|U|>(z) and the resolved TKE 1/2<u'^2+v'^2+w'^2>(z),
primes taken against the horizontal mean of each level, all cells included as there.

With --sgs-model the FP32 run also uses the learned subgrid closure (a network file written by
pu_mlp_save), and the comparison is FP32 + learned closure against the FP64 plain closure.

    python -m porturb_amd.precision_study --nx 1024 --ny 1024 --nz 128 --steps 200 --out res.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import porturb_amd as pa  # noqa: E402
from porturb_amd import cases  # noqa: E402

FIELD_NAMES = ["density_dry", "uvel", "vvel", "wvel", "temp", "water_vapor", "TKE", "passive_1", "passive_2"]


def profiles(u, v, w):
    um, vm, wm = u.mean(axis=(1, 2)), v.mean(axis=(1, 2)), w.mean(axis=(1, 2))
    mag = np.sqrt(u * u + v * v + w * w).mean(axis=(1, 2))
    up, vp, wp = u - um[:, None, None], v - vm[:, None, None], w - wm[:, None, None]
    tke = 0.5 * (up * up + vp * vp + wp * wp).mean(axis=(1, 2))
    return {"u_mean": um, "wind_speed": mag, "tke_resolved": tke}


def compare(cs64, cs32, names):
    res = {}
    for i, name in enumerate(names):
        a = cs64.field(i)
        b = cs32.field(i)
        d = b - a
        r = {"l2": float(np.linalg.norm(d) / max(np.linalg.norm(a), 1e-300)),
             "linf": float(np.abs(d).max() / max(np.abs(a).max(), 1e-300)),
             "abs_linf": float(np.abs(d).max())}
        if name in ("uvel", "vvel", "wvel"):
            r["linf_vs_range"] = float(np.abs(d).max() / max(a.max() - a.min(), 1e-300))
        res[name] = r
    u64, v64, w64 = (cs64.field(i) for i in (1, 2, 3))
    p64 = profiles(u64, v64, w64)
    del u64, v64, w64
    u32, v32, w32 = (cs32.field(i) for i in (1, 2, 3))
    p32 = profiles(u32, v32, w32)
    prof = {}
    for k in p64:
        scale = max(np.abs(p64[k]).max(), 1e-300)
        prof[k] = {"max_rel_diff": float(np.abs(p32[k] - p64[k]).max() / scale),
                   "fp64_max": float(np.abs(p64[k]).max())}
    return res, prof


def learned_fraction(d, cs, dt):
    """Fractions of cells whose momentum / heat fluxes come from the network for the current
    state (the attached network's flags)."""
    n3 = int(np.prod(cs.shape))
    fl = pa.DeviceBuffer(n3)
    pa.check(pa.lib().pu_les_closure_sgs(d.ctx, pa.C.byref(cs.c), dt, None, fl.ptr))
    f = fl.download(np.uint8, (n3,))
    fl.free()
    return {"momentum": float(((f & 1) != 0).mean()), "heat": float(((f & 2) != 0).mean())}


class State:
    def __init__(self, cfg):
        self.cfg = cfg
        self.cs = pa.CouplerState(cfg)
        self.shape = self.cs.shape

    def field(self, i):
        buf = self.cs.bufs[pa.CouplerState.FIELDS[i]] if i < 5 else self.cs.tracers[i - 5]
        return buf.download(np.float64, self.shape)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--steps", type=int, nargs="+", default=[200])
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--out", default=None)
    ap.add_argument("--sgs-model", default=None)
    args = ap.parse_args()
    cfg64 = cases.abl_urban(nx=args.nx, ny=args.ny, nz=args.nz, precision=pa.PU_FP64)
    cfg32 = cases.abl_urban(nx=args.nx, ny=args.ny, nz=args.nz, precision=pa.PU_FP32)
    s64, s32 = State(cfg64), State(cfg32)
    pa.case_init(cfg64, cfg64["case_id"], args.seed, s64.cs)
    pa.case_init(cfg32, cfg32["case_id"], args.seed, s32.cs)
    d64, d32 = pa.Dycore(cfg64), pa.Dycore(cfg32)
    d64.init(s64.cs)
    d32.init(s32.cs)
    mlp = None
    if args.sgs_model:
        mlp = pa.Mlp.load(args.sgs_model)
        d32.set_sgs_model(mlp)
    names = FIELD_NAMES[:5 + cfg64["num_tracers"]]
    out = {"workload": cfg64["name"].replace("_fp64", ""), "seed": args.seed, "checkpoints": [],
           "fp32_closure": "learned (%s)" % os.path.basename(args.sgs_model) if args.sgs_model else "tke"}
    step, t0 = 0, time.time()
    for target in sorted(args.steps):
        while step < target:
            dt = d64.compute_time_step()
            d64.time_step(s64.cs, dt)
            d64.les_apply(s64.cs, dt)
            d32.time_step(s32.cs, dt)
            d32.les_apply(s32.cs, dt)
            step += 1
            if step % 50 == 0:
                d64.synchronize()
                d32.synchronize()
                print("precision_study: step %d (%.0f s)" % (step, time.time() - t0), file=sys.stderr, flush=True)
        d64.synchronize()
        d32.synchronize()
        fields, prof = compare(s64, s32, names)
        m64 = d64.diagnostics(s64.cs)[0]
        m32 = d32.diagnostics(s32.cs)[0]
        ck = {"steps": step, "fields": fields, "profiles": prof, "mass_rel_diff": abs(m32 - m64) / abs(m64)}
        if mlp is not None:
            ck["learned_fraction"] = learned_fraction(d32, s32.cs, d64.compute_time_step())
        out["checkpoints"].append(ck)
        print("precision_study: checkpoint %d done" % step, file=sys.stderr, flush=True)
    d64.destroy()
    d32.destroy()
    js = json.dumps(out)
    print(js)
    if args.out:
        with open(args.out, "w") as f:
            f.write(js + "\n")


if __name__ == "__main__":
    main()
