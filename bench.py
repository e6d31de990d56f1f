#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the dycore step (SSPRK3 WENO-FV time_step + LES closure apply) on one
MI355X, with the dominant kernel's roofline and the CPU oracle timed on the host as a baseline.

Workload (default): neutral urban ABL, 1024x1024x128 cells at 2 m, 8x8 cuboid buildings, rough
wall (log-law surface fluxes, z0 = 0.1 m), water vapour + TKE + 2 passive tracers, WENO-9, TKE
closure, FP64.  One step = one Dynamics_Euler_Stratified_WenoFV::time_step (3 stages) + one
apply_surface_fluxes + one LES_Closure::apply, inputs resident in HBM.  Timed with HIP events on
the library's stream.

    python bench.py [--gpus 1] [--steps 10] [--warmup 200]
                    [--config abl_fp64|abl_fp32|abl_fp32_learned|bubble|windfarm] [--no-modes]

The timed steps start from a spun-up state: the seeded initialisation advanced by --warmup steps.
The dominant sweep's roofline counts the FLOPs the sweep executes on that state: a census of every
reconstruction's branch (the full limiter, the flat-stencil branch that takes the linear
reconstruction where the stencil's total variation is <= 1e-20, or the linear reconstruction of a
cell without WENO) at the start and the end of the timed steps; the full-limiter figure is kept
beside it, and the chip's own ceilings measured in the same run (pu_measure_fma_rate,
pu_measure_copy_rate: roofline.measured_ceiling).  With the default workload (abl_fp64) the same
spun-up state is then advanced by the FP32 dycore with the TKE closure and with the learned closure: their ms/step, speed-up over FP64
and per-field error against the FP64 state after the same steps are in "modes".

abl_fp32_learned: the FP32 dycore with the learned subgrid closure (network file
porturb_amd/data/sgs_abl_urban_dx2.bin, trained by porturb_amd.sgs_training) in place of the
eddy-viscosity momentum/heat fluxes, the TKE closure as its fallback.
windfarm: 2048x1024x256 cells at 5 m, precursor x inflow, capping inversion, FP32 + learned
closure, 3x4 actuator-disk turbines (D = 120 m, hub 90 m, C_T = 0.75,
porturb_amd/data/turbine_disk_ct075_d120.yaml) after the surface fluxes.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/sec per SSPRK3 step, 1 MI355X; % HBM roofline; err vs FP64"
SGS_MODEL = os.path.join(ROOT, "porturb_amd", "data", "sgs_abl_urban_dx2.bin")
TURBINE_FILE = os.path.join(ROOT, "porturb_amd", "data", "turbine_disk_ct075_d120.yaml")


def make_workload(name):
    from porturb_amd import cases, PU_FP64, PU_FP32
    if name == "abl_fp64":
        return cases.abl_urban(precision=PU_FP64), True
    if name in ("abl_fp32", "abl_fp32_learned"):
        return cases.abl_urban(precision=PU_FP32), True
    if name == "bubble":
        return cases.bubble(n=64), False
    if name == "windfarm":
        return cases.windfarm(precision=PU_FP32), True
    raise SystemExit("unknown config " + name)


def cpu_baseline(cfg_full, les, seconds_budget=25.0):
    """The CPU oracle (FP64 restatement of the reference, OpenMP over all host cores) on a bounded
    sample of the same workload: the same case, order, tracers and closure on a 256x256xnz box
    (bubble: the full 64^3 case), one or two steps from the seeded initialisation."""
    import numpy as np
    import porturb_amd as pa
    from oracle import oracle_py
    cores = len(os.sched_getaffinity(0))
    cfg = dict(cfg_full)
    if cfg["nx"] > 256:
        dx = cfg["xlen"] / cfg["nx"]
        cfg.update(nx=256, ny=256, xlen=256 * dx, ylen=256 * dx)
    cfg["precision"] = pa.PU_FP64
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, cfg["case_id"], 1234, cs)
    st = cs.download()
    imm = cs.download_immersed()
    od = oracle_py.Dycore(cfg, st, imm, nthreads=cores)
    dt = od.compute_time_step()
    n3 = cfg["nx"] * cfg["ny"] * cfg["nz"]
    shape = (cfg["nz"], cfg["ny"], cfg["nx"])
    rough, zero3 = np.full(shape, 0.1), np.zeros(shape)
    srough, zero2 = np.full(shape[1:], 0.1), np.zeros(shape[1:])
    steps, t_tot = 0, 0.0
    while True:
        t0 = time.perf_counter()
        od.time_step(st, dt)
        if les:
            oracle_py.surface_fluxes(cfg, st, imm, rough, zero3, srough, zero2, dt)
            oracle_py.les_apply(cfg, st, imm, dt, nthreads=cores)
        t_tot += time.perf_counter() - t0
        steps += 1
        if t_tot > seconds_budget / 2 or steps >= 2:
            break
    return {"value": n3 * steps / t_tot, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": "%s on %dx%dx%d cells, %d step(s), %.1f s, FP64 OpenMP oracle" %
                      (cfg["name"].split("_1024")[0], cfg["nx"], cfg["ny"], cfg["nz"], steps, t_tot)}


def sweep_kernel_name(cfg, fam):
    """The kernel that runs sweep family `fam` (the launch choice of dycore_kernels.hip sweep_one):
    z between solid walls -> k_sweep_zwall, periodic y lines -> k_sweep_pline, otherwise the segment
    kernel k_sweep<ORD, T, DIR, MODE> (x finishes the stage: MODE 2)."""
    import porturb_amd as pa
    t = "double" if cfg["precision"] == pa.PU_FP64 else "float"
    if fam == "weno_z" and cfg["bc_z"] == pa.PU_BC_SOLID_WALL:
        return "k_sweep_zwall<%d, %s>" % (cfg["ord"], t)
    if fam == "weno_y" and cfg["bc_y"] == pa.PU_BC_PERIODIC and cfg["ny"] >= 2:
        return "k_sweep_pline<%d, %s, 1, 1>" % (cfg["ord"], t)
    d = {"weno_x": 0, "weno_y": 1, "weno_z": 2}[fam]
    return "k_sweep<%d, %s, %d, %d>" % (cfg["ord"], t, d, {0: 2, 1: 1, 2: 0}[d])


def snapshot(d):
    """Device copies of every coupler field of driver `d`: {name: DeviceBuffer}."""
    import porturb_amd as pa
    nb = int(d.shape[0] * d.shape[1] * d.shape[2]) * 8
    out = {}
    for name, ptr in d.field_pointers().items():
        b = pa.DeviceBuffer(nb)
        pa.check(pa.lib().pu_memcpy_d2d(pa.C.c_void_p(b.ptr), pa.C.c_void_p(ptr), nb))
        out[name] = b
    d.synchronize()
    return out


def restore(d, snap):
    import porturb_amd as pa
    nb = int(d.shape[0] * d.shape[1] * d.shape[2]) * 8
    for name, ptr in d.field_pointers().items():
        pa.check(pa.lib().pu_memcpy_d2d(pa.C.c_void_p(ptr), pa.C.c_void_p(snap[name].ptr), nb))
    d.synchronize()


def compare_with(d, ref, shape):
    """Per-field relative L2 / Linf of driver `d`'s state against the snapshot `ref`, and the
    largest relative differences of the horizontal-mean wind and resolved-TKE profiles (the
    profiles of build/postproc/abl_neutral_analysis.py:122-131)."""
    import numpy as np
    from porturb_amd.precision_study import profiles
    fields, vel = {}, {}
    for name, b in ref.items():
        a = b.download(np.float64, shape)
        x = d.download_field(name)
        dd = x - a
        fields[name] = {"l2": float(np.linalg.norm(dd) / max(np.linalg.norm(a), 1e-300)),
                        "linf": float(np.abs(dd).max() / max(np.abs(a).max(), 1e-300))}
        if name in ("uvel", "vvel", "wvel"):
            vel[name] = (a, x)
        del a, x, dd
    p_ref = profiles(*(vel[n][0] for n in ("uvel", "vvel", "wvel")))
    p_got = profiles(*(vel[n][1] for n in ("uvel", "vvel", "wvel")))
    prof = {k: float(np.abs(p_got[k] - p_ref[k]).max() / max(np.abs(p_ref[k]).max(), 1e-300)) for k in p_ref}
    return fields, prof


def time_steps(d, steps):
    d.synchronize()
    d.timer_start()
    for _ in range(steps):
        d.step(1)
    ms = d.timer_stop()
    d.synchronize()
    return ms / steps


FAMS = ["weno_x", "weno_y", "weno_z", "prep", "surface", "les", "sgs", "turbines"]


def kernel_times(d, steps=2):
    """Per-launch HIP events around every kernel family for `steps` steps: {family: (ms, launches)}."""
    d.kernel_timing(True)
    for _ in range(steps):
        d.step(1)
    d.synchronize()
    times = {f: d.kernel_time(f) for f in FAMS}
    d.kernel_timing(False)
    return times


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=200,
                    help="steps from the seeded initialisation before the timed steps (spin-up)")
    ap.add_argument("--config", default="abl_fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-modes", action="store_true",
                    help="abl_fp64: skip the FP32 and FP32 + learned closure runs from the same state")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each MI355X box here has "
              "one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import porturb_amd as pa
    from porturb_amd import roofline as rl

    cfg, les = make_workload(args.config)
    n3 = cfg["nx"] * cfg["ny"] * cfg["nz"]
    print("bench: %s, %d cells, warmup %d, steps %d" % (cfg["name"], n3, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    # the drop-in C++ path: core::Coupler + Dynamics_Euler_Stratified_WenoFV (+ LES_Closure)
    learned = args.config in ("abl_fp32_learned", "windfarm")
    turbines = args.config == "windfarm"
    modes = args.config == "abl_fp64" and not args.no_modes
    if (learned or modes) and not os.path.exists(SGS_MODEL):
        raise SystemExit("bench: %s missing (train it with python -m porturb_amd.sgs_training)" % SGS_MODEL)
    d = pa.Driver(cfg["case_id"], cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlen"] / cfg["nx"], ord=cfg["ord"],
                  precision=cfg["precision"], seed=1234, closure=les, sgs_model=SGS_MODEL if learned else None)
    if turbines:
        d.set_turbines(TURBINE_FILE, 3, 4, orig_C_T=True)

    t0 = time.time()
    for w in range(args.warmup):
        d.step(1)
        if (w + 1) % 25 == 0:
            d.synchronize()
            print("bench: warm-up step %d/%d (%.0f s)" % (w + 1, args.warmup, time.time() - t0),
                  file=sys.stderr, flush=True)
    d.synchronize()
    sim_time = d.elapsed_time()
    # reconstruction census at the start and the end of the timed steps: the work they execute
    census0 = d.flat_census()
    snap0 = snapshot(d) if modes else None
    ms_per_step = time_steps(d, args.steps)
    value = n3 / (ms_per_step * 1e-3)
    census1 = d.flat_census()
    snap1 = snapshot(d) if modes else None
    print("bench: %.3f ms/step, %.4e cell-updates/s" % (ms_per_step, value), file=sys.stderr, flush=True)

    times = kernel_times(d)
    dom = max(["weno_x", "weno_y", "weno_z"], key=lambda f: times[f][0] / max(times[f][1], 1))
    tot_ms, cnt = times[dom]
    avg_ms = tot_ms / max(cnt, 1)
    nf = 6 + cfg["num_tracers"]
    fp64 = cfg["precision"] == pa.PU_FP64
    drow = {"weno_x": 0, "weno_y": 1, "weno_z": 2}[dom]
    census = (census0 + census1) / 2.0
    flops_full = n3 * rl.sweep_flops_per_cell(cfg["ord"], nf)
    flops_exec = rl.executed_sweep_flops(cfg["ord"], nf, census[drow], n3)
    achieved = flops_exec / (avg_ms * 1e-3) / 1e12
    achieved_full = flops_full / (avg_ms * 1e-3) / 1e12
    peak = rl.MI355X["fp64_TFLOPs"] if fp64 else rl.MI355X["fp32_TFLOPs"]
    step_bytes = rl.step_bytes_per_cell(5 + cfg["num_tracers"], fp64)
    hbm_frac = value * step_bytes / (rl.MI355X["hbm_GBps"] * 1e9)
    # HBM bytes per launch of the dominant kernel, from the calibrated FETCH_SIZE/WRITE_SIZE runs
    # recorded under profiles/ for this workload (null for workloads not profiled)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s.json" % cfg["name"].replace("neutral_urban_abl", "abl_urban")
                         .replace("_weno9_", "_"))
    if os.path.exists(tfile) and not learned:
        fam = json.load(open(tfile)).get("families", {}).get(dom)
        if fam:
            traffic = fam["read"] + fam["write"]
    recons = census.sum(axis=1)
    roof = {"bound": "valu", "kernel": sweep_kernel_name(cfg, dom),
            "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic,
            "work": "executed: limiter, flat-stencil and linear reconstructions counted on the timed state",
            "flops_per_launch": flops_exec, "avg_launch_ms": round(avg_ms, 4),
            "achieved_full_limiter": round(achieved_full, 3), "frac_full_limiter": round(achieved_full / peak, 4),
            "flops_per_launch_full_limiter": flops_full,
            "flat_fraction": {k: round(float(census[i][1] / max(recons[i], 1)), 4) for i, k in enumerate("xyz")},
            "census": {"x": census[0].tolist(), "y": census[1].tolist(), "z": census[2].tolist(),
                       "columns": ["limiter", "flat", "linear"]},
            "step_hbm_frac": round(hbm_frac, 4), "step_bytes_per_cell": step_bytes,
            "kernel_ms_per_step": {f: round(times[f][0] / 2, 3) for f in FAMS}}
    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if fp64 else "f32", "data": "synthetic",
           "config": {"workload": cfg["name"] + ("_learned" if learned else "") + ("_turbines3x4" if turbines else ""),
                      "cells": n3, "ord": cfg["ord"],
                      "path": "C++ Coupler + modules",
                      "closure": ("learned (bf16 matrix cores) + tke fallback" if learned else "tke") if les else "none",
                      "tracers": cfg["num_tracers"],
                      "state": "case initialisation + %d warm-up steps (%.3f s of simulated flow)" % (args.warmup, sim_time),
                      "step": ("dycore.time_step + apply_surface_fluxes + " +
                               ("windmills.apply + " if turbines else "") + "les_closure.apply") if les else "dycore.time_step"},
           "roofline": roof}
    d.destroy()
    # the same chip's own ceilings beside the published peak: independent FMA chains (FP64 and
    # non-packed FP32, the forms the kernels are compiled to) and the plain streaming copy
    ceil = pa.measure_ceilings()
    fma = ceil["fma_fp64_TFLOPs"] if fp64 else ceil["fma_fp32_TFLOPs"]
    roof["measured_ceiling"] = dict(ceil, frac_of_fma=round(achieved / fma, 4),
                                    step_frac_of_copy=round(value * step_bytes / (ceil["copy_GBps"] * 1e9), 4))
    print("bench: measured FMA %.1f (FP64) / %.1f (FP32) TFLOP/s, copy %.0f GB/s; %s at %.3f of the measured FMA rate"
          % (ceil["fma_fp64_TFLOPs"], ceil["fma_fp32_TFLOPs"], ceil["copy_GBps"], dom, achieved / fma),
          file=sys.stderr, flush=True)

    if modes:
        # the same spun-up state advanced by the FP32 dycore (compensated rho, rho*theta) with the
        # TKE closure and with the learned closure: speed against the FP64 run above and the
        # per-field error after the same steps
        shape = d.shape
        out["modes"] = {}
        for name, model in (("fp32", None), ("fp32_learned", SGS_MODEL)):
            d2 = pa.Driver(cfg["case_id"], cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlen"] / cfg["nx"], ord=cfg["ord"],
                           precision=pa.PU_FP32, seed=1234, closure=les, sgs_model=model)
            d2.step(2)                       # first launches of every kernel, outside the timing
            restore(d2, snap0)
            ms2 = time_steps(d2, args.steps)
            fields, prof = compare_with(d2, snap1, shape)
            t2 = kernel_times(d2)
            out["modes"][name] = {
                "ms_per_step": ms2, "value": n3 / (ms2 * 1e-3), "speedup_vs_fp64": ms_per_step / ms2,
                "closure": "learned (bf16 matrix cores) + tke fallback" if model else "tke",
                "err_vs_fp64_after_steps": args.steps, "err_vs_fp64": fields, "profile_rel_diff_vs_fp64": prof,
                "kernel_ms_per_step": {f: round(t2[f][0] / 2, 3) for f in FAMS}}
            print("bench: %s %.3f ms/step (%.3fx FP64)" % (name, ms2, ms_per_step / ms2), file=sys.stderr, flush=True)
            d2.destroy()
        for b in list(snap0.values()) + list(snap1.values()):
            b.free()

    if not args.no_cpu_baseline:
        print("bench: CPU oracle baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, les, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
