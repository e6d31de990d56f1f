"""Workload for kernel traces of the modules around the dycore step: the neutral urban ABL through
the C++ driver with the column forcing of experiments/examples/abl_neutral.cpp:93-95
(geostrophic_wind_forcing, sponge_layer) for a few steps, then one output file
(Coupler::write_output_file) and a restart from it (Coupler::overwrite_with_restart).

    rocprofv3 --kernel-trace --stats -d OUT -o k -- python3 porturb_amd/profile_modules.py \\
        --nx 1024 --ny 1024 --nz 128 --steps 3 --out /tmp/abl

Prints one JSON line: ms per step and per module kernel (HIP events), and the output/restart
wall times with the file size.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import porturb_amd as pa  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--out", default=None, help="output prefix (no file without it)")
    ap.add_argument("--float-output", action="store_true")
    args = ap.parse_args()
    d = pa.Driver(pa.PU_CASE_ABL_URBAN, args.nx, args.ny, args.nz, 2.0, ord=9, seed=1234)
    d.set_forcing(geostrophic=1, lat_g=40.0, u_g=8.0, v_g=0.0, sponge=1, sponge_dt_multiple=100.0,
                  sponge_top_prop=0.1)
    d.step(1)
    d.synchronize()
    d.kernel_timing(True)
    d.timer_start()
    d.step(args.steps)
    ms = d.timer_stop() / args.steps
    names = ["weno_x", "weno_y", "weno_z", "prep", "surface", "les", "geostrophic", "sponge"]
    res = {"workload": "abl_urban_%dx%dx%d" % (args.nx, args.ny, args.nz), "ms_per_step": ms,
           "kernel_ms_per_step": {n: d.kernel_time(n)[0] / args.steps for n in names}}
    if args.out:
        t0 = time.time()
        path = d.write_output(args.out, double_precision=not args.float_output)
        t1 = time.time()
        d.restart(path)
        d.synchronize()
        t2 = time.time()
        res.update({"file_bytes": os.path.getsize(path), "write_s": t1 - t0, "restart_s": t2 - t1})
        os.remove(path)
    print(json.dumps(res), flush=True)
    d.destroy()


if __name__ == "__main__":
    main()
