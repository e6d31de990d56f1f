"""Workload for the HBM-traffic counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE):

1. a plain 2 GiB device-to-device copy (k_copy) -- the calibration: it must read and write
   exactly 2 GiB, so counter/2 GiB is the scale factor applied to the other kernels;
2. one step of a benchmark workload through the C++ driver, after --warmup steps from the seeded
   initialisation (the counters of the last step's launches describe the spun-up flow).

    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex "k_copy|k_sweep" -d OUT -o fetch -- \\
        python3 -m porturb_amd.profile_traffic --config abl_fp64
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import porturb_amd as pa  # noqa: E402

CALIB_BYTES = 2 << 30


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="abl_fp64")
    ap.add_argument("--warmup", type=int, default=0, help="steps before the last (profiled) step")
    args = ap.parse_args()
    a, b = pa.DeviceBuffer(CALIB_BYTES), pa.DeviceBuffer(CALIB_BYTES)
    pa.check(pa.lib().pu_memset_device(a.ptr, 1, CALIB_BYTES))
    pa.check(pa.lib().pu_memcpy_d2d(b.ptr, a.ptr, CALIB_BYTES))
    a.free(); b.free()
    from bench import make_workload, SGS_MODEL
    cfg, les = make_workload(args.config)
    learned = args.config in ("abl_fp32_learned", "windfarm")
    d = pa.Driver(cfg["case_id"], cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlen"] / cfg["nx"], ord=cfg["ord"],
                  precision=cfg["precision"], seed=1234, closure=les, sgs_model=SGS_MODEL if learned else None)
    for w in range(args.warmup):
        d.step(1)
        if (w + 1) % 50 == 0:
            d.synchronize()
            print("profile_traffic: warm-up step %d" % (w + 1), flush=True)
    d.step(1)
    d.synchronize()
    print("profile_traffic: copy %d bytes + %d + 1 %s steps done" % (CALIB_BYTES, args.warmup, cfg["name"]), flush=True)
    d.destroy()


if __name__ == "__main__":
    main()
