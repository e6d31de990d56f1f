"""Planted mistakes: each entry below changes one line of a copy of the oracle's C source into a
plausible error (a swapped upwind side, a flipped sign, a wrong exponent or coefficient), builds
the copy, and runs tests/test_oracle_pins.py against it.  Every mutant must fail at least one pin:
a mistake the pins let through would leave that part of the oracle unpinned.

The copies are built in a temporary directory; the committed oracle is untouched.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SOURCES = ["oracle.c", "sgs.c", "surface.c", "column.c", "turbine.c"]

# (name, file, original text, mutated text): the original must occur in the file
MUTANTS = [
    ("upwind side swapped (x)", "oracle.c",
     "int ind = ru_upw > 0 ? 0 : 1;", "int ind = ru_upw > 0 ? 1 : 0;"),
    ("upwind side swapped (y)", "oracle.c",
     "int ind = rv_upw > 0 ? 0 : 1;", "int ind = rv_upw > 0 ? 1 : 0;"),
    ("upwind side swapped (z)", "oracle.c",
     "int ind = rw_upw > 0 ? 0 : 1;", "int ind = rw_upw > 0 ? 1 : 0;"),
    ("acoustic dissipation sign in p_upw (x)", "oracle.c",
     "double p_upw = 0.5 * (p_L + p_R - cs * (ru_R - ru_L));", "double p_upw = 0.5 * (p_L + p_R + cs * (ru_R - ru_L));"),
    ("pressure term sign in ru_upw (x)", "oracle.c",
     "double ru_upw = 0.5 * (ru_L + ru_R - (p_R - p_L) / cs);", "double ru_upw = 0.5 * (ru_L + ru_R + (p_R - p_L) / cs);"),
    ("pressure sign in the momentum flux (x)", "oracle.c",
     "LX(0, idU, k, j, i) = ru_upw * LX(ind, idU, k, j, i) + p_upw;", "LX(0, idU, k, j, i) = ru_upw * LX(ind, idU, k, j, i) - p_upw;"),
    ("hydrostatic density not added back (x)", "oracle.c",
     "LX(0, idR, k, j, i) += hyr[hs + k]; LX(1, idR, k, j, i) += hyr[hs + k];", ";"),
    ("tracer flux takes the neighbouring field slot (x)", "oracle.c",
     "LX(0, NUM_STATE + tr, k, j, i) = ru_upw * LX(ind, NUM_STATE + 1 + tr, k, j, i);",
     "LX(0, NUM_STATE + tr, k, j, i) = ru_upw * LX(ind, NUM_STATE + tr, k, j, i);"),
    ("y divergence sign", "oracle.c",
     "- ydiv_s * (LY(0, NUM_STATE + l, k, j + 1, i) - LY(0, NUM_STATE + l, k, j, i)) * r_dy",
     "+ ydiv_s * (LY(0, NUM_STATE + l, k, j + 1, i) - LY(0, NUM_STATE + l, k, j, i)) * r_dy"),
    ("gravity sign", "oracle.c",
     "tnd += -grav * (H3(state + idR * nh, hs + k, hs + j, hs + i) - hyr[hs + k]);",
     "tnd += grav * (H3(state + idR * nh, hs + k, hs + j, hs + i) - hyr[hs + k]);"),
    ("gravity on rho instead of rho - hydrostatic rho", "oracle.c",
     "tnd += -grav * (H3(state + idR * nh, hs + k, hs + j, hs + i) - hyr[hs + k]);",
     "tnd += -grav * (H3(state + idR * nh, hs + k, hs + j, hs + i) - hyr[hs + k] * 0.999);"),
    ("Coriolis sign", "oracle.c",
     "if (c->latitude != 0 && !sim2d && l == idU) tnd += fcor", "if (c->latitude != 0 && !sim2d && l == idU) tnd -= fcor"),
    ("SSPRK3 second stage 1/2, 1/2", "oracle.c",
     "*pt = (3. / 4.) * *pn + (1. / 4.) * *pt + (1. / 4.) * dt * tn;",
     "*pt = (1. / 2.) * *pn + (1. / 2.) * *pt + (1. / 2.) * dt * tn;"),
    ("SSPRK3 third-stage tendency weight", "oracle.c",
     "*pn = (1. / 3.) * *pn + (2. / 3.) * *pt + (2. / 3.) * dt * tn;",
     "*pn = (1. / 3.) * *pn + (2. / 3.) * *pt + (1. / 2.) * dt * tn;"),
    ("map_rs exponent k = 4", "oracle.c",
     "for (int i = 1; i < 6; i++) term1 *= w - d;", "for (int i = 1; i < 4; i++) term1 *= w - d;"),
    ("map_rs exponent m = 2", "oracle.c",
     "for (int i = 1; i < 3; i++) term2 *= w * (1 - w);", "for (int i = 1; i < 2; i++) term2 *= w * (1 - w);"),
    ("map_rs constant s = 1000", "oracle.c", "2000. * term2", "1000. * term2"),
    ("LES buoyancy sign", "oracle.c",
     "src += -(grav * rho * km) / (tt * Pr_t) * dt_dz;", "src += (grav * rho * km) / (tt * Pr_t) * dt_dz;"),
    ("LES vertical heat flux sign", "oracle.c",
     "FZ(3, k, j, i) = -rho * visc_tot_th * (dt_dz);", "FZ(3, k, j, i) = rho * visc_tot_th * (dt_dz);"),
    ("LES heat diffusivity without Pr_t", "oracle.c",
     "FZ(3, k, j, i) = -rho * visc_tot_th * (dt_dz);", "FZ(3, k, j, i) = -rho * visc_tot * (dt_dz);"),
    ("LES vertical momentum flux sign", "oracle.c",
     "FZ(0, k, j, i) = -rho * visc_tot * (du_dz + dw_dx);", "FZ(0, k, j, i) = rho * visc_tot * (du_dz + dw_dx);"),
]


def _run_mutant(m):
    name, fname, old, new = m
    tmp = tempfile.mkdtemp(prefix="oracle_mutant_")
    try:
        for f in SOURCES + ["oracle.h"]:
            shutil.copy(os.path.join(ORACLE, f), tmp)
        path = os.path.join(tmp, fname)
        src = open(path).read()
        if old not in src:
            return name, "pattern not found"
        open(path, "w").write(src.replace(old, new, 1))
        lib = os.path.join(tmp, "liboracle.so")
        r = subprocess.run(["gcc", "-O1", "-fopenmp", "-fPIC", "-ffp-contract=off", "-std=gnu11", "-shared", "-o", lib]
                           + [os.path.join(tmp, f) for f in SOURCES] + ["-lm"], capture_output=True, text=True)
        if r.returncode:
            return name, "build failed: " + r.stderr[-400:]
        env = dict(os.environ, PORTURB_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
        r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                            os.path.join(ROOT, "tests", "test_oracle_pins.py")],
                           cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
        return name, "caught" if r.returncode == 1 else ("NOT CAUGHT" if r.returncode == 0 else "error %d: %s" % (r.returncode, r.stdout[-400:]))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_every_planted_mistake_fails_a_pin():
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_run_mutant, MUTANTS))
    for name, res in results:
        print("%-55s %s" % (name, res))
    bad = [(n, r) for n, r in results if r != "caught"]
    assert not bad, bad
