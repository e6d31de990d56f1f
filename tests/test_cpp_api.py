"""The drop-in C++ API (include/porturb/...) compiles against libporturb and its host-only parts
(Options typing and error behaviour, option defaults, BC parsing) behave as the reference's."""
import os
import subprocess
import tempfile

import porturb_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_api_compiles_and_host_checks_pass():
    assert os.path.exists(porturb_amd.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "coupler_options")
        subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "coupler_options.cpp"), "-o", exe,
                        "-L" + os.path.dirname(porturb_amd.LIB_PATH), "-lporturb",
                        "-Wl,-rpath," + os.path.dirname(porturb_amd.LIB_PATH)], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "OK" in r.stdout
