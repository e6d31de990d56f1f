#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the dycore step (SSPRK3 WENO-FV time_step + LES closure apply) on one
MI355X, with the dominant kernel's roofline and the CPU oracle timed on the host as a baseline.

Workload (default): neutral urban ABL, 1024x1024x128 cells at 2 m, 8x8 cuboid buildings, rough
wall (log-law surface fluxes, z0 = 0.1 m), water vapour + TKE + 2 passive tracers, WENO-9, TKE
closure, FP64.  One step = one Dynamics_Euler_Stratified_WenoFV::time_step (3 stages) + one
apply_surface_fluxes + one LES_Closure::apply, inputs resident in HBM.  Timed with HIP events on
the library's stream.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
                    [--config abl_fp64|abl_fp32|abl_fp32_learned|bubble|windfarm]

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
    sample of the same workload: the same case, order, tracers and closure on a 128x128xnz box
    (bubble: the full 64^3 case), one step."""
    import numpy as np
    import porturb_amd as pa
    from oracle import oracle_py
    cores = len(os.sched_getaffinity(0))
    cfg = dict(cfg_full)
    if cfg["nx"] > 128:
        dx = cfg["xlen"] / cfg["nx"]
        cfg.update(nx=128, ny=128, xlen=128 * dx, ylen=128 * dx)
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
        if t_tot > seconds_budget / 3 or steps >= 3:
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="abl_fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    if learned and not os.path.exists(SGS_MODEL):
        raise SystemExit("bench: %s missing (train it with python -m porturb_amd.sgs_training)" % SGS_MODEL)
    d = pa.Driver(cfg["case_id"], cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlen"] / cfg["nx"], ord=cfg["ord"],
                  precision=cfg["precision"], seed=1234, closure=les, sgs_model=SGS_MODEL if learned else None)
    if turbines:
        d.set_turbines(TURBINE_FILE, 3, 4, orig_C_T=True)

    def step():
        d.step(1)

    for _ in range(args.warmup):
        step()
    d.synchronize()
    d.timer_start()
    for _ in range(args.steps):
        step()
    ms = d.timer_stop()
    d.synchronize()
    ms_per_step = ms / args.steps
    value = n3 / (ms_per_step * 1e-3)
    print("bench: %.3f ms/step, %.4e cell-updates/s" % (ms_per_step, value), file=sys.stderr, flush=True)

    # per-kernel timing pass (HIP events around each launch on the library stream)
    d.kernel_timing(True)
    for _ in range(2):
        step()
    d.synchronize()
    fams = ["weno_x", "weno_y", "weno_z", "prep", "surface", "les", "sgs", "turbines"]
    times = {f: d.kernel_time(f) for f in fams}
    d.kernel_timing(False)
    dom = max(["weno_x", "weno_y", "weno_z"], key=lambda f: times[f][0])
    tot_ms, cnt = times[dom]
    avg_ms = tot_ms / max(cnt, 1)
    nf = 6 + cfg["num_tracers"]
    fp64 = cfg["precision"] == pa.PU_FP64
    flops_launch = n3 * rl.sweep_flops_per_cell(cfg["ord"], nf)
    achieved = flops_launch / (avg_ms * 1e-3) / 1e12
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
    roof = {"bound": "valu", "kernel": sweep_kernel_name(cfg, dom),
            "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic,
            "flops_per_launch": flops_launch, "avg_launch_ms": round(avg_ms, 4),
            "step_hbm_frac": round(hbm_frac, 4), "step_bytes_per_cell": step_bytes,
            "kernel_ms_per_step": {f: round(times[f][0] / 2, 3) for f in fams}}
    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if fp64 else "f32", "data": "synthetic",
           "config": {"workload": cfg["name"] + ("_learned" if learned else "") + ("_turbines3x4" if turbines else ""),
                      "cells": n3, "ord": cfg["ord"],
                      "path": "C++ Coupler + modules",
                      "closure": ("learned (bf16 matrix cores) + tke fallback" if learned else "tke") if les else "none",
                      "tracers": cfg["num_tracers"],
                      # the timed steps follow the warm-up steps from the seeded case initialisation:
                      # the air above the boundary layer is still uniform, where the limiters take
                      # their flat-stencil branch (the linear reconstruction, equal to rounding)
                      "state": "case initialisation + %d warm-up steps" % args.warmup,
                      "step": ("dycore.time_step + apply_surface_fluxes + " +
                               ("windmills.apply + " if turbines else "") + "les_closure.apply") if les else "dycore.time_step"},
           "roofline": roof}
    d.destroy()
    if not args.no_cpu_baseline:
        print("bench: CPU oracle baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, les, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
