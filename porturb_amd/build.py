"""Build libporturb.so (HIP kernels for gfx950 + C ABI + C++ host layer) in-tree.

    python -m porturb_amd.build          # or porturb_amd.build.build()

Every translation unit is compiled with hipcc for --offload-arch=gfx950 and linked into
porturb_amd/libporturb.so.  Object files go to porturb_amd/_build/ (git-ignored).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libporturb.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["capi.hip", "dycore_kernels.hip", "sweep_zwall_f32.hip", "sweep_pline_f32.hip", "sweep_seg_x_f32.hip", "init_kernels.hip", "les_kernels.hip", "sgs_kernels.hip", "surface_kernels.hip", "column_kernels.hip", "turbine_kernels.hip", "ncio.hip", "ceiling_kernels.hip", "driver.cpp"]

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-result",
]


# per-source extra flags: the FP32 whole-column z and whole-line y sweeps keep the limiter's
# constants out of registers when machine-code loop-invariant motion is off
EXTRA = {"sweep_zwall_f32.hip": ["-mllvm", "-disable-machine-licm"],
         "sweep_pline_f32.hip": ["-mllvm", "-disable-machine-licm"],
         "sweep_seg_x_f32.hip": ["-mllvm", "-disable-machine-licm"]}


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    """Regenerate the WENO tables, compile every source and link the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_weno_tables.py")], check=True,
                   capture_output=not verbose)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    errors = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for src, obj, rc, out in ex.map(_compile, srcs):
            if rc != 0:
                errors.append("== %s ==\n%s" % (src, out))
            elif verbose and out.strip():
                print(out)
            objs.append(obj)
    if errors:
        raise RuntimeError("HIP compilation failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
