"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code, run on the CPU:

* the oracle (oracle/*.c) with a driver that calls every entry point on small seeded states
  (tests/cpp/oracle_sanitize.c: all reconstruction orders, periodic / precursor / wall / 2-D
  boundaries, buildings, the LES closure with and without the learned fluxes, surface fluxes,
  turbines, column forcing), leak checking on;
* the drop-in C++ API's host-only checks (tests/cpp/coupler_options.cpp) against libporturb.

Any out-of-bounds access, use after free, leak (oracle) or undefined behaviour aborts the run."""
import os
import subprocess
import tempfile

import porturb_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    src = [os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c")] + \
          [os.path.join(ROOT, "oracle", f) for f in ("oracle.c", "sgs.c", "surface.c", "column.c", "turbine.c")]
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "oracle_sanitize")
        subprocess.run(["gcc", "-std=gnu11", "-fopenmp", "-ffp-contract=off"] + SAN + src + ["-lm", "-o", exe], check=True)
        env = dict(os.environ, OMP_NUM_THREADS="1", ASAN_OPTIONS="detect_leaks=1")
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-2000:] + r.stderr[-4000:]


def test_cpp_api_under_address_and_undefined_behaviour_sanitizers():
    assert os.path.exists(porturb_amd.LIB_PATH)
    libdir = os.path.dirname(porturb_amd.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "coupler_options")
        subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include")] + SAN +
                       [os.path.join(ROOT, "tests", "cpp", "coupler_options.cpp"), "-o", exe,
                        "-L" + libdir, "-lporturb", "-Wl,-rpath," + libdir], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"),
                           timeout=300)
        assert r.returncode == 0 and "OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
