#!/usr/bin/env python3
"""Generate ``weno_tables.h``: exact-rational WENO tables for reconstruction orders 3..11.

Everything is derived from first principles with ``fractions.Fraction`` (no floating point until
the final rounding to C literals):

* sub-stencil m (0-based, m = 0..S-1, S = hs+1, N = hs+1 points) covers stencil cells
  m .. m+N-1; the centre cell is stencil index hs.  Its polynomial p_m(x) = sum_n a_n x^n
  (x in cell widths about the centre cell) matches the N cell averages, i.e.
  a = A^-1 s with A[r][n] = integral of x^n over [o_r-1/2, o_r+1/2], o_r = m+r-hs.
  Reference: the same definition in ``helpers/sage/transformation_matrices.sage:19-27``,
  whose outputs are ``coefs{N}_shift{m+1}`` (``TransformMatrices.h:19776-``).
* total variation TV(p) = sum_{l=1}^{N-1} integral_{-1/2}^{1/2} trunc_N[(d^l p/dx^l)^2] dx,
  where trunc_N keeps powers x^0..x^{N-1} of the squared derivative (the reference integrates
  with ``integrate_poly(N, ...)``, which drops higher powers; ``poly_utils.sage:23-26`` and
  ``transformation_matrices.sage:140-146``).  Emitted as a quadratic form in a_1..a_{N-1}.
* face weights: p_m(-1/2) and p_m(+1/2) as linear weights over the full ord-point stencil
  (``WenoLimiter.h`` ``gll_1``/``gll_2`` return expressions).
* ideal weights d: the unique convex combination with sum_m d_m p_m(face) == P(face) for every
  stencil, P being the degree ord-1 polynomial over all ord cells; solved by forward
  substitution because sub-stencil m is the first one to cover cell m+N-1... (see _ideal()).
* the full-stencil linear branch ``s2e = coefs_to_gll_lower * sten_to_coefs``
  (``dynamics_rk.h:539-545``): P(-1/2), P(+1/2) as weights over the ord cells.

Run:  python porturb_amd/csrc/gen_weno_tables.py   (writes weno_tables.h beside this file)
"""
from fractions import Fraction as F
import os
import sys

ORDERS = (3, 5, 7, 9, 11)


def cell_avg_row(o, ncoef):
    """Row of integrals of x^n over [o-1/2, o+1/2] (unit cells) for n = 0..ncoef-1."""
    lo, hi = F(2 * o - 1, 2), F(2 * o + 1, 2)
    return [(hi ** (n + 1) - lo ** (n + 1)) / (n + 1) for n in range(ncoef)]


def inverse(M):
    n = len(M)
    A = [list(r) + [F(int(i == j)) for j in range(n)] for i, r in enumerate(M)]
    for c in range(n):
        p = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[p] = A[p], A[c]
        piv = A[c][c]
        A[c] = [v / piv for v in A[c]]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c]
                A[r] = [a - f * b for a, b in zip(A[r], A[c])]
    return [r[n:] for r in A]


def poly_from_averages(offsets):
    """Matrix C with a_n = sum_i C[n][i] s_i for a polynomial matching averages at offsets."""
    N = len(offsets)
    return inverse([cell_avg_row(o, N) for o in offsets])


def eval_row(C, x):
    """Weights over the stencil points giving p(x) = sum_n a_n x^n."""
    N = len(C)
    return [sum(C[n][i] * x ** n for n in range(N)) for i in range(len(C[0]))]


def int_monomial(q):
    """Integral of x^q over [-1/2, 1/2]."""
    if q % 2:
        return F(0)
    return F(2, 1) * F(1, 2) ** (q + 1) / (q + 1)


def tv_form(N):
    """Quadratic form {(n1,n2): coef} with n1 <= n2 for TV in terms of the coefficients a_n."""
    from math import factorial
    Q = {}
    for l in range(1, N):
        for n1 in range(l, N):
            for n2 in range(l, N):
                q = (n1 - l) + (n2 - l)
                if q >= N:           # truncation to degree < N (integrate_poly)
                    continue
                c1 = F(factorial(n1), factorial(n1 - l))
                c2 = F(factorial(n2), factorial(n2 - l))
                v = c1 * c2 * int_monomial(q)
                if v == 0:
                    continue
                key = (min(n1, n2), max(n1, n2))
                Q[key] = Q.get(key, F(0)) + v
    return {k: v for k, v in Q.items() if v != 0}


def tables(ord_):
    hs = (ord_ - 1) // 2
    N = hs + 1
    S = hs + 1
    subC = []      # per sub-stencil: N x N
    face = {}
    for m in range(S):
        offs = [m + r - hs for r in range(N)]
        subC.append(poly_from_averages(offs))
    for side, x in (("L", F(-1, 2)), ("R", F(1, 2))):
        rows = []
        for m in range(S):
            r = eval_row(subC[m], x)
            full = [F(0)] * ord_
            for i in range(N):
                full[m + i] = r[i]
            rows.append(full)
        face[side] = rows
    Cfull = poly_from_averages([i - hs for i in range(ord_)])
    lin = {"L": eval_row(Cfull, F(-1, 2)), "R": eval_row(Cfull, F(1, 2))}
    ideal = {}
    for side in ("L", "R"):
        d = _ideal(face[side], lin[side], S, N)
        ideal[side] = d
    return {"hs": hs, "N": N, "S": S, "C": subC, "face": face, "lin": lin, "ideal": ideal,
            "tv": tv_form(N)}


def _ideal(rows, target, S, N):
    """Solve sum_m d_m rows[m] == target.  Cell index c < S is covered by sub-stencils 0..c only,
    so d_c follows from cell c once d_0..d_{c-1} are known (forward substitution); the remaining
    cells are checked for consistency (exact for odd ord)."""
    d = []
    for c in range(S):
        acc = target[c] - sum(d[m] * rows[m][c] for m in range(c))
        d.append(acc / rows[c][c])
    for c in range(len(target)):
        assert sum(d[m] * rows[m][c] for m in range(S)) == target[c], "inconsistent ideal weights"
    assert sum(d) == 1
    return d


def lit(v):
    return "%.17g" % float(v)


def emit(path):
    out = []
    out.append("// Generated by gen_weno_tables.py -- exact rational WENO tables rounded to double.")
    out.append("// Do not edit by hand.  Definitions: see the generator's docstring.")
    out.append("#pragma once")
    out.append("namespace pu_weno {")
    out.append("template <int ORD> struct Tab;")
    for ord_ in ORDERS:
        t = tables(ord_)
        hs, N, S = t["hs"], t["N"], t["S"]
        out.append("template <> struct Tab<%d> {" % ord_)
        out.append("  static constexpr int ord = %d, hs = %d, N = %d, S = %d;" % (ord_, hs, N, S))
        out.append("  static constexpr bool use_map_rs = %s;" % ("true" if ord_ > 3 else "false"))
        # C[m][n][i] for n >= 1 only (a_0 is not needed by the limiter)
        vals = []
        for m in range(S):
            for n in range(N):
                for i in range(N):
                    vals.append(lit(t["C"][m][n][i]))
        out.append("  static constexpr double C[%d][%d][%d] = {%s};" % (S, N, N, ",".join(vals)))
        for side in ("L", "R"):
            rows = []
            for m in range(S):
                rows.append(",".join(lit(t["face"][side][m][m + i]) for i in range(N)))
            out.append("  static constexpr double face%s[%d][%d] = {%s};" % (side, S, N, ",".join(rows)))
            out.append("  static constexpr double ideal%s[%d] = {%s};" % (side, S, ",".join(lit(v) for v in t["ideal"][side])))
            out.append("  static constexpr double lin%s[%d] = {%s};" % (side, ord_, ",".join(lit(v) for v in t["lin"][side])))
        tv = sorted(t["tv"].items())
        out.append("  static constexpr int ntv = %d;" % len(tv))
        out.append("  static constexpr int tv_i[%d] = {%s};" % (len(tv), ",".join(str(k[0]) for k, _ in tv)))
        out.append("  static constexpr int tv_j[%d] = {%s};" % (len(tv), ",".join(str(k[1]) for k, _ in tv)))
        out.append("  static constexpr double tv_c[%d] = {%s};" % (len(tv), ",".join(lit(v) for _, v in tv)))
        out.append("};")
    out.append("}  // namespace pu_weno")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def tv_ldl(ord_):
    """TV of each sub-stencil as a weighted sum of squares of first differences.

    TV is translation invariant, so with d_j = s_{j+1} - s_j the coefficients a_n (n >= 1) of
    sub-stencil m are a = M d (M[n][j] = sum_{i>j} C[n][i]) and TV = d^T G d with G = M^T Q M
    (Q the quadratic form of coefs_to_tv).  The exact LDL^T factorisation G = L D L^T gives
    TV = sum_k D_k (d_k + sum_{j>k} L[j][k] d_j)^2.  Returns per sub-stencil (D, L) as Fractions;
    the differences are indexed relative to the sub-stencil's first point."""
    t = tables(ord_)
    N, S = t["N"], t["S"]
    Q = [[F(0)] * N for _ in range(N)]
    for (i, j), c in t["tv"].items():
        if i == j:
            Q[i][i] += c
        else:
            Q[i][j] += c / 2
            Q[j][i] += c / 2
    out = []
    for m in range(S):
        C = t["C"][m]
        M = [[sum(C[n][i] for i in range(j + 1, N)) for j in range(N - 1)] for n in range(N)]
        G = [[sum(M[a][p] * Q[a][b] * M[b][q] for a in range(1, N) for b in range(1, N))
              for q in range(N - 1)] for p in range(N - 1)]
        n = N - 1
        L = [[F(int(i == j)) for j in range(n)] for i in range(n)]
        D = [F(0)] * n
        for j in range(n):
            D[j] = G[j][j] - sum(L[j][k] ** 2 * D[k] for k in range(j))
            for i in range(j + 1, n):
                L[i][j] = (G[i][j] - sum(L[i][k] * L[j][k] * D[k] for k in range(j))) / D[j]
        out.append((D, L))
    return out


def _lin_expr(coefs, var):
    """Straight-line sum of the nonzero terms c*var[i] (left to right)."""
    terms = ["T(%s)*%s[%d]" % (lit(c), var, i) for i, c in enumerate(coefs) if c != 0]
    return " + ".join(terms) if terms else "T(0)"


def emit_ops(path):
    """Device code: one struct per order with the TV, face and linear expressions written out."""
    out = []
    out.append("// Generated by gen_weno_tables.py -- WENO reconstruction expressions (device code).")
    out.append("// Do not edit by hand.  Only nonzero coefficients appear.")
    out.append("#pragma once")
    out.append("#include <hip/hip_runtime.h>")
    out.append("namespace pu_weno {")
    out.append("template <int ORD> struct Ops;")
    for ord_ in ORDERS:
        t = tables(ord_)
        hs, N, S = t["hs"], t["N"], t["S"]
        out.append("template <> struct Ops<%d> {" % ord_)
        out.append("  static constexpr int ord = %d, hs = %d, N = %d, S = %d;" % (ord_, hs, N, S))
        out.append("  static constexpr bool use_map_rs = %s;" % ("true" if ord_ > 3 else "false"))
        # TV of each sub-stencil: sum of D_k * (first-difference combination)^2 (see tv_ldl)
        out.append("  template <class T> __device__ static __forceinline__ void tv(const T *s, T *TV) {")
        out.append("    T d[%d];" % (ord_ - 1))
        out.append("    #pragma unroll")
        out.append("    for (int i = 0; i < %d; i++) d[i] = s[i + 1] - s[i];" % (ord_ - 1))
        for m, (Dk, L) in enumerate(tv_ldl(ord_)):
            n = N - 1
            terms = []
            for k in range(n):
                comb = "d[%d]" % (m + k)
                for j in range(k + 1, n):
                    if L[j][k] != 0:
                        comb += " + T(%s)*d[%d]" % (lit(L[j][k]), m + j)
                out.append("    const T c%d_%d = %s;" % (m, k, comb))
                terms.append("T(%s)*(c%d_%d*c%d_%d)" % (lit(Dk[k]), m, k, m, k))
            out.append("    TV[%d] = %s;" % (m, " + ".join(terms)))
        out.append("  }")
        for side in ("L", "R"):
            out.append("  template <class T> __device__ static __forceinline__ T comb%s(const T *s, const T *w) {" % side)
            parts = []
            for m in range(S):
                parts.append("(%s)*w[%d]" % (_lin_expr(t["face"][side][m], "s"), m))
            out.append("    return %s;" % " + ".join(parts))
            out.append("  }")
            out.append("  template <class T> __device__ static __forceinline__ T lin%s(const T *s) {" % side)
            out.append("    return %s;" % _lin_expr(t["lin"][side], "s"))
            out.append("  }")
            out.append("  __device__ static __forceinline__ double ideal%s(int m) {" % side)
            out.append("    constexpr double d[%d] = {%s};" % (S, ",".join(lit(v) for v in t["ideal"][side])))
            out.append("    return d[m];")
            out.append("  }")
        out.append("};")
    out.append("}  // namespace pu_weno")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def as_float_dict():
    """Tables as plain floats, for tests."""
    res = {}
    for ord_ in ORDERS:
        t = tables(ord_)
        res[ord_] = {
            "C": [[[float(v) for v in row] for row in m] for m in t["C"]],
            "faceL": [[float(v) for v in r] for r in t["face"]["L"]],
            "faceR": [[float(v) for v in r] for r in t["face"]["R"]],
            "idealL": [float(v) for v in t["ideal"]["L"]],
            "idealR": [float(v) for v in t["ideal"]["R"]],
            "linL": [float(v) for v in t["lin"]["L"]],
            "linR": [float(v) for v in t["lin"]["R"]],
            "tv": {k: float(v) for k, v in t["tv"].items()},
        }
    return res


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    emit(os.path.join(here, "weno_tables.h"))
    emit_ops(os.path.join(here, "weno_ops.h"))
    print("wrote", os.path.join(here, "weno_tables.h"), os.path.join(here, "weno_ops.h"))
    sys.exit(0)
