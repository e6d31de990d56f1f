"""A-posteriori judge of the learned subgrid closure: does the network move a coarse run's
horizontal-mean wind and resolved-TKE profiles towards a finer run's more than the eddy-viscosity
closure does?

Runs through the drop-in driver (dycore time_step + apply_surface_fluxes + LES_Closure::apply per
step, the module sequence of abl_neutral.cpp:98-103) on the neutral urban ABL case (8x8 cuboid
buildings, rough wall, seeded theta noise):

* truth   : 2nx x 2ny x 2nz cells at dx/2, FP64, TKE closure, 2 N steps (the time step halves
            with the cell size: compute_time_step = cfl dx / 450, dynamics_rk.h:57-64, so both
            grids reach the same simulated time);
* plain   : nx x ny x nz at dx, FP64, TKE closure, N steps;
* fp32    : the same in FP32 (compensated rho, rho theta), TKE closure;
* learned : the same in FP32 with the network attached (option les_sgs_model), the TKE closure
            as its fallback.

At each checkpoint the truth's u, v, w are box-filtered by 2 in each direction onto the coarse
grid and the profiles of build/postproc/abl_neutral_analysis.py:122-131 are taken from every run:
<u>(z), <|U|>(z) and the resolved TKE 1/2 <u'^2 + v'^2 + w'^2>(z) (primes against the level
mean, all cells included).  Reported per run: max |profile - truth| / max |truth| for each
profile, and the fraction of cells whose momentum / heat fluxes come from the network.

    python -m porturb_amd.sgs_aposteriori --nx 192 --ny 192 --nz 64 --dx 2 --steps 500 1000 2000 3000 \
        --out profiles/sgs_aposteriori_abl_urban_192x192x64.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import porturb_amd as pa  # noqa: E402
from porturb_amd.precision_study import profiles  # noqa: E402

MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "sgs_abl_urban_dx2.bin")


def box_filter(a, r):
    nz, ny, nx = a.shape
    return a.reshape(nz // r, r, ny // r, r, nx // r, r).mean(axis=(1, 3, 5))


def learned_fraction(drv, dt):
    """Fractions of cells whose momentum / heat fluxes come from the attached network."""
    cs = pa.PuCouplerState()
    n = pa.C.c_int()
    pa.check(pa.lib().pu_driver_state(drv.d, pa.C.byref(cs), pa.C.byref(n)))
    ctx = pa.lib().pu_driver_context(drv.d)
    n3 = int(np.prod(drv.shape))
    fl = pa.DeviceBuffer(n3)
    pa.check(pa.lib().pu_les_closure_sgs(ctx, pa.C.byref(cs), dt, None, fl.ptr))
    f = fl.download(np.uint8, (n3,))
    fl.free()
    return {"momentum": float(((f & 1) != 0).mean()), "heat": float(((f & 2) != 0).mean())}


def run_profiles(drv, r=1):
    u, v, w = (drv.download_field(n) for n in ("uvel", "vvel", "wvel"))
    if r > 1:
        u, v, w = box_filter(u, r), box_filter(v, r), box_filter(w, r)
    p = profiles(u, v, w)
    return {k: np.asarray(x, np.float64) for k, x in p.items()}


def errors(p, truth):
    return {k: float(np.abs(p[k] - truth[k]).max() / max(np.abs(truth[k]).max(), 1e-300)) for k in truth}


def run(nx, ny, nz, dx, steps, seed=1234, model=MODEL, log=print):
    """The four runs to the coarse step counts `steps`; returns the report dict."""
    steps = sorted(steps)
    cfgs = {"truth": (2 * nx, 2 * ny, 2 * nz, dx / 2, pa.PU_FP64, None, 2),
            "plain": (nx, ny, nz, dx, pa.PU_FP64, None, 1),
            "fp32": (nx, ny, nz, dx, pa.PU_FP32, None, 1),
            "learned": (nx, ny, nz, dx, pa.PU_FP32, model, 1)}
    prof = {name: [] for name in cfgs}
    frac = []
    walls = {}
    for name, (mx, my, mz, h, prec, sgs, mult) in cfgs.items():
        t0 = time.time()
        drv = pa.Driver(pa.PU_CASE_ABL_URBAN, mx, my, mz, h, ord=9, precision=prec, seed=seed, sgs_model=sgs)
        done = 0
        for s in steps:
            drv.step(mult * s - done)
            done = mult * s
            drv.synchronize()
            prof[name].append(run_profiles(drv, r=2 if name == "truth" else 1))
            if name == "learned":
                frac.append(learned_fraction(drv, drv.elapsed_time() / done))
            log("sgs_aposteriori: %s step %d (%.0f s)" % (name, done, time.time() - t0))
        walls[name] = time.time() - t0
        drv.destroy()
    rep = {"case": "neutral urban ABL (8x8 buildings, rough wall), WENO-9",
           "coarse": [nx, ny, nz, dx], "truth": [2 * nx, 2 * ny, 2 * nz, dx / 2, "FP64 TKE closure, box-filtered by 2"],
           "model": os.path.basename(model), "seed": seed, "checkpoints": []}
    for i, s in enumerate(steps):
        ck = {"coarse_steps": s, "truth": {k: v.tolist() for k, v in prof["truth"][i].items()}}
        for name in ("plain", "fp32", "learned"):
            ck[name] = errors(prof[name][i], prof["truth"][i])
            ck[name + "_profiles"] = {k: v.tolist() for k, v in prof[name][i].items()}
        ck["learned_vs_plain"] = errors(prof["learned"][i], prof["plain"][i])
        ck["learned_fraction"] = frac[i]
        rep["checkpoints"].append(ck)
    rep["wall_s"] = walls
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--dx", type=float, default=2.0)
    ap.add_argument("--steps", type=int, nargs="+", default=[500, 1000, 2000, 3000])
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--model", default=MODEL)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    rep = run(a.nx, a.ny, a.nz, a.dx, a.steps, a.seed, a.model,
              log=lambda m: print(m, file=sys.stderr, flush=True))
    for ck in rep["checkpoints"]:
        print("steps %d: max |profile - truth| / max truth" % ck["coarse_steps"])
        for name in ("plain", "fp32", "learned"):
            print("  %-8s u_mean %.4f  wind_speed %.4f  tke_resolved %.4f" %
                  (name, ck[name]["u_mean"], ck[name]["wind_speed"], ck[name]["tke_resolved"]))
        print("  learned vs plain: u_mean %.2e tke_resolved %.2e; network fraction momentum %.3f heat %.3f" %
              (ck["learned_vs_plain"]["u_mean"], ck["learned_vs_plain"]["tke_resolved"],
               ck["learned_fraction"]["momentum"], ck["learned_fraction"]["heat"]))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
