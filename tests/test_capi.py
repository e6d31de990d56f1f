"""The C ABI library loads and exports every entry point include/porturb.h declares (no device
calls: this runs without a GPU)."""
import os
import re
import subprocess

import pytest

import porturb_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "porturb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(porturb_amd.LIB_PATH), "build the library first (__graft_entry__.build())"
    syms = declared_symbols()
    assert len(syms) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", porturb_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s+(pu_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_library_loads_and_reports():
    L = porturb_amd.lib()
    info = L.pu_build_info().decode()
    assert "gfx950" in info


def test_config_validation_without_device():
    """pu_create validates its config before touching the device."""
    import ctypes as C
    cfg = porturb_amd.make_config(16, 16, 16, 100., 100., 100., ord=4)
    pc = porturb_amd.to_pu_config(cfg)
    p = C.c_void_p()
    rc = porturb_amd.lib().pu_create(C.byref(pc), None, C.byref(p))
    assert rc == porturb_amd.PU_ERR_UNSUPPORTED
    assert "ord" in porturb_amd.lib().pu_last_error().decode()
    cfg = porturb_amd.make_config(16, 16, 16, 100., 100., 100., ord=5, bc_z=7)
    rc = porturb_amd.lib().pu_create(C.byref(porturb_amd.to_pu_config(cfg)), None, C.byref(p))
    assert rc == porturb_amd.PU_ERR_ARG
    assert "bc_z" in porturb_amd.lib().pu_last_error().decode()


def test_struct_layout_matches_header():
    """sizeof(pu_config) from a C compile of the header equals the ctypes mirror."""
    import ctypes as C
    import tempfile
    src = '#include "porturb.h"\n#include <stdio.h>\nint main(){printf("%zu %zu\\n", sizeof(pu_config), sizeof(pu_coupler_state));return 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "t.c"), "w").write(src)
        subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), os.path.join(d, "t.c"), "-o", os.path.join(d, "t")], check=True)
        a, b = subprocess.run([os.path.join(d, "t")], capture_output=True, text=True).stdout.split()
    assert int(a) == C.sizeof(porturb_amd.PuConfig)
    assert int(b) == C.sizeof(porturb_amd.PuCouplerState)


def test_ceiling_measurements_check_their_arguments_without_device():
    """pu_measure_fma_rate / pu_measure_copy_rate reject bad arguments before any device call."""
    import ctypes as C
    L = porturb_amd.lib()
    out = C.c_double()
    assert L.pu_measure_fma_rate(porturb_amd.PU_FP64, None) == porturb_amd.PU_ERR_ARG
    assert L.pu_measure_fma_rate(7, C.byref(out)) == porturb_amd.PU_ERR_ARG
    assert L.pu_measure_copy_rate(1 << 10, C.byref(out)) == porturb_amd.PU_ERR_ARG
    assert L.pu_measure_copy_rate((1 << 20) + 8, C.byref(out)) == porturb_amd.PU_ERR_ARG
    assert "multiple of 16" in L.pu_last_error().decode()
