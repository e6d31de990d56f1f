#!/usr/bin/env python3
"""Extract the WENO numerical constants that portUrb's headers hold, as golden vectors.

The reference ships its reconstruction constants as 38-digit decimal literals inside two
generated C++ headers.  This script reads those headers *as text* and records the numbers
(nothing else) in ``weno_reference_constants.json``:

* ``coefs{N}_shift{m}`` -- the N x N matrix mapping sub-stencil cell averages to polynomial
  coefficients about the centre cell
  (``model/modules/helpers/TransformMatrices.h:19776-`` e.g. ``coefs3_shift1`` :19795,
  ``coefs5_shift1`` :19841);
* ``coefs_to_tv`` for N points -- the total-variation quadratic form in the coefficients
  (``TransformMatrices.h:39,105,271,429,836,...``);
* ``WenoLimiter<real,ord>::gll_1 / gll_2`` -- ideal weights and the face-value weights of each
  sub-stencil at the left (x=-1/2) and right (x=+1/2) interface
  (``model/modules/helpers/WenoLimiter.h:116-180`` ord 3, ``:186-260`` ord 5, ``:264-345`` ord 7,
  ``:349-437`` ord 9, ``:441-537`` ord 11);
* ``sten_to_coefs`` for N = ord and ``coefs_to_gll_lower`` with 2 points -- the full-stencil
  linear branch (``TransformMatrices.h:138,525,1385,3028`` and ``:45-..``).

Run in the development container only (it reads /root/reference):

    python tests/golden/extract_reference_constants.py

The output JSON is committed; tests compare the product's generated tables and the oracle's
numerically derived ones against it.
"""
import json
import os
import re
import sys

REF = "/root/reference/model/modules/helpers"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weno_reference_constants.json")

NUM = r"([0-9]+\.[0-9]*(?:[eE][+-]?[0-9]+)?)"


def _func_body(text, signature_regex):
    m = re.search(signature_regex, text)
    if not m:
        raise RuntimeError("not found: " + signature_regex)
    start = text.index("{", m.end() - 1)
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
    raise RuntimeError("unbalanced braces")


def _linear_terms(expr, var):
    """Parse '+/- static_cast<T>(c)*var(i)' or '*var i' style terms into {index: value}."""
    out = {}
    pat = re.compile(r"([+-]?)\s*static_cast<\w+>\(\s*" + NUM + r"\s*\)\s*\*\s*" + var)
    for sign, val, idx in pat.findall(expr):
        v = float(val) * (-1.0 if sign == "-" else 1.0)
        out[int(idx)] = out.get(int(idx), 0.0) + v
    return out


def extract_coefs_shift(tm, N):
    mats = {}
    for m in range(1, N + 1):
        body = _func_body(tm, r"void coefs%d_shift%d\(SArray<FP,1,%d> &coefs" % (N, m, N))
        mat = [[0.0] * N for _ in range(N)]
        for line in body.split(";"):
            mm = re.search(r"coefs\((\d+)\)\s*=(.*)", line, re.S)
            if not mm:
                continue
            row = int(mm.group(1))
            terms = _linear_terms(mm.group(2), r"s(\d+)")
            for idx, v in terms.items():
                mat[row][idx] = v
        mats["coefs%d_shift%d" % (N, m)] = mat
    return mats


def extract_tv(tm, N):
    body = _func_body(tm, r"float coefs_to_tv\(SArray<FP,1,%d> &a\)" % N)
    expr = body.split("rslt =", 1)[1].split(";")[0]
    pat = re.compile(r"([+-]?)\s*static_cast<FP>\(\s*" + NUM +
                     r"\s*\)\s*\*\s*\(?\s*a\((\d+)\)\s*\*\s*a\((\d+)\)\s*\)?")
    q = []
    for sign, val, i, j in pat.findall(expr):
        v = float(val) * (-1.0 if sign == "-" else 1.0)
        i, j = sorted((int(i), int(j)))
        q.append([i, j, v])
    return q


def extract_matrix(tm, name, rows, cols):
    body = _func_body(tm, r"void %s\(SArray<FP,2,%d,%d> &rslt\)" % (name, rows, cols))
    mat = [[0.0] * cols for _ in range(rows)]
    pat = re.compile(r"rslt\((\d+),(\d+)\)\s*=\s*([+-]?)\s*(?:static_cast<FP>\(\s*" + NUM +
                     r"\s*\)|([0-9]+))\s*;")
    for r, c, sign, v1, v2 in pat.findall(body):
        v = float(v1 if v1 else v2) * (-1.0 if sign == "-" else 1.0)
        mat[int(r)][int(c)] = v
    return mat


def extract_limiter(wl, ord_):
    S = (ord_ + 1) // 2
    start = wl.index("struct WenoLimiter<real,%d>" % ord_)
    nxt = wl.find("template <class real> struct WenoLimiter<real,", start + 10)
    block = wl[start:nxt if nxt > 0 else len(wl)]
    out = {}
    for face, fname in (("left", "gll_1"), ("right", "gll_2")):
        fb = _func_body(block, r"static real %s\(" % fname)
        # ideal weights, in declaration order
        ideal = [float(v) * (-1.0 if s == "-" else 1.0) for s, v in
                 re.findall(r"real\s+i_\w+\s*=\s*([+-]?)\s*static_cast<real>\(\s*" + NUM + r"\s*\)", fb)]
        ret = fb.split("return", 1)[1].split(";")[0]
        # groups "( ... )*w_X"
        groups = re.findall(r"\(([^()]*(?:\([^()]*\)[^()]*)*)\)\s*\*\s*w_(\w+)", ret)
        weights = {}
        for g, wname in groups:
            weights[wname] = _linear_terms(g, r"s\((\d+)\)")
        names = [n for n in re.findall(r"real\s+w_(\w+)\s*=", fb)]
        rows = []
        for n in names:
            row = [0.0] * ord_
            for idx, v in weights[n].items():
                row[idx] = v
            rows.append(row)
        uses_map_rs = "map_rs(" in fb
        out[face] = {"ideal": ideal, "face_weights": rows, "map": "map_rs" if uses_map_rs else "map"}
        assert len(ideal) == S and len(rows) == S, (ord_, face, len(ideal), len(rows))
    return out


def main():
    if not os.path.isdir(REF):
        print("reference not present; nothing to do", file=sys.stderr)
        return 1
    tm = open(os.path.join(REF, "TransformMatrices.h")).read()
    wl = open(os.path.join(REF, "WenoLimiter.h")).read()
    data = {"source": "portUrb model/modules/helpers/{TransformMatrices.h,WenoLimiter.h} (numbers only)",
            "coefs_shift": {}, "tv": {}, "limiter": {}, "linear": {}}
    for N in (2, 3, 4, 5, 6):
        data["coefs_shift"].update(extract_coefs_shift(tm, N))
        data["tv"][str(N)] = extract_tv(tm, N)
    for ord_ in (3, 5, 7, 9, 11):
        data["limiter"][str(ord_)] = extract_limiter(wl, ord_)
        data["linear"][str(ord_)] = {
            "sten_to_coefs": extract_matrix(tm, "sten_to_coefs", ord_, ord_),
            "coefs_to_gll_lower": extract_matrix(tm, "coefs_to_gll_lower", ord_, 2),
        }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)
    return 0


if __name__ == "__main__":
    sys.exit(main())
