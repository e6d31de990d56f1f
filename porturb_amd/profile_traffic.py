"""Workload for the HBM-traffic counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE):

1. a plain 2 GiB device-to-device copy (k_copy) -- the calibration: it must read and write
   exactly 2 GiB, so counter/2 GiB is the scale factor applied to the other kernels;
2. one step of a benchmark workload through the C++ driver.

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
    args = ap.parse_args()
    a, b = pa.DeviceBuffer(CALIB_BYTES), pa.DeviceBuffer(CALIB_BYTES)
    pa.check(pa.lib().pu_memset_device(a.ptr, 1, CALIB_BYTES))
    pa.check(pa.lib().pu_memcpy_d2d(b.ptr, a.ptr, CALIB_BYTES))
    a.free(); b.free()
    from bench import make_workload
    cfg, les = make_workload(args.config)
    d = pa.Driver(cfg["case_id"], cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlen"] / cfg["nx"], ord=cfg["ord"],
                  precision=cfg["precision"], seed=1234, closure=les)
    d.step(1)
    d.synchronize()
    print("profile_traffic: copy %d bytes + one %s step done" % (CALIB_BYTES, cfg["name"]), flush=True)
    d.destroy()


if __name__ == "__main__":
    main()
