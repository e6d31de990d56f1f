"""WENO tables pinned against the reference's own constants (golden vectors).

tests/golden/weno_reference_constants.json holds the decimal literals of
model/modules/helpers/TransformMatrices.h (coefs{N}_shift{m}, coefs_to_tv, sten_to_coefs,
coefs_to_gll_lower) and WenoLimiter.h (ideal weights and gll_1/gll_2 face weights), extracted as
numbers by tests/golden/extract_reference_constants.py.  Two independent derivations are checked
against them: the product's exact-rational generator and the oracle's long-double solver.
"""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "porturb_amd", "csrc"))
import gen_weno_tables as gen  # noqa: E402
from oracle import oracle_py  # noqa: E402

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "weno_reference_constants.json")))
ORDERS = (3, 5, 7, 9, 11)


def golden_for(ord_):
    N = (ord_ - 1) // 2 + 1
    C = np.array([GOLD["coefs_shift"]["coefs%d_shift%d" % (N, m + 1)] for m in range(N)])
    tv = np.zeros((N, N))
    for i, j, v in GOLD["tv"][str(N)]:
        tv[i, j] += v
    L = GOLD["limiter"][str(ord_)]
    lin = GOLD["linear"][str(ord_)]
    s2e = np.array(lin["sten_to_coefs"]) @ np.array(lin["coefs_to_gll_lower"])
    return dict(C=C, tv=tv, idealL=np.array(L["left"]["ideal"]), idealR=np.array(L["right"]["ideal"]),
                faceL=np.array(L["left"]["face_weights"]), faceR=np.array(L["right"]["face_weights"]),
                linL=s2e[:, 0], linR=s2e[:, 1], map=L["left"]["map"])


@pytest.mark.parametrize("ord_", ORDERS)
def test_generator_matches_reference_constants(ord_):
    g = golden_for(ord_)
    t = gen.as_float_dict()[ord_]
    np.testing.assert_allclose(np.array(t["C"]), g["C"], rtol=1e-15, atol=1e-15)
    tv = np.zeros_like(g["tv"])
    for (i, j), v in t["tv"].items():
        tv[i, j] = v
    np.testing.assert_allclose(tv, g["tv"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(t["idealL"], g["idealL"], rtol=1e-15)
    np.testing.assert_allclose(t["idealR"], g["idealR"], rtol=1e-15)
    np.testing.assert_allclose(np.array(t["faceL"]), g["faceL"], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(np.array(t["faceR"]), g["faceR"], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(t["linL"], g["linL"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(t["linR"], g["linR"], rtol=1e-13, atol=1e-14)
    assert (g["map"] == "map_rs") == (ord_ > 3)


@pytest.mark.parametrize("ord_", ORDERS)
def test_oracle_tables_match_reference_constants(ord_):
    g = golden_for(ord_)
    t = oracle_py.weno_tables(ord_)
    np.testing.assert_allclose(t["C"], g["C"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(t["tv"], g["tv"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(t["idealL"], g["idealL"], rtol=1e-13)
    np.testing.assert_allclose(t["idealR"], g["idealR"], rtol=1e-13)
    np.testing.assert_allclose(t["faceL"], g["faceL"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(t["faceR"], g["faceR"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(t["linL"], g["linL"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(t["linR"], g["linR"], rtol=1e-12, atol=1e-12)


def test_tv_truncation_is_what_the_reference_does():
    """The reference's TV drops powers >= N of each squared derivative (integrate_poly(N,...)):
    for N=5 the a4^2 weight is 625.8, not the untruncated 625.8 + 16/448."""
    g = golden_for(9)
    assert g["tv"][4, 4] == pytest.approx(625.8, rel=1e-15)
    assert abs(g["tv"][4, 4] - (625.8 + 16. / 448.)) > 1e-2


def test_ideal_weights_are_textbook():
    """WENO-5 0.3/0.6/0.1 and WENO-9 (1, 20, 60, 40, 5)/126 (Balsara & Shu 2000)."""
    np.testing.assert_allclose(golden_for(5)["idealL"], [0.3, 0.6, 0.1], rtol=1e-15)
    np.testing.assert_allclose(golden_for(9)["idealR"], np.array([1, 20, 60, 40, 5]) / 126., rtol=1e-15)


@pytest.mark.parametrize("ord_", ORDERS)
def test_difference_form_of_tv_is_exact(ord_):
    """The kernels evaluate TV as sum_k D_k (d_k + sum_j L_jk d_j)^2 over first differences d; in
    exact rational arithmetic that equals the reference's coefs_to_tv(coefs{N}_shift{m}(s)) for
    every sub-stencil (reference constants from the golden file, stencils random rationals)."""
    from fractions import Fraction as Fr
    import random
    g = golden_for(ord_)
    N = (ord_ - 1) // 2 + 1
    t = gen.tables(ord_)
    ldl = gen.tv_ldl(ord_)
    rnd = random.Random(ord_)
    for trial in range(5):
        s = [Fr(rnd.randint(-1000, 1000), rnd.randint(1, 50)) for _ in range(ord_)]
        d = [s[i + 1] - s[i] for i in range(ord_ - 1)]
        for m in range(N):
            a = [sum(t["C"][m][n][i] * s[m + i] for i in range(N)) for n in range(N)]
            tv_ref = sum(c * a[i] * a[j] for (i, j), c in t["tv"].items())
            D, L = ldl[m]
            tv_ldl = Fr(0)
            for k in range(N - 1):
                comb = d[m + k] + sum(L[j][k] * d[m + j] for j in range(k + 1, N - 1))
                tv_ldl += D[k] * comb * comb
            assert tv_ldl == tv_ref
            # and the rational table agrees with the reference's decimal constants
            a_f = [float(x) for x in a]
            tv_gold = sum(g["tv"][i, j] * a_f[i] * a_f[j] for i in range(N) for j in range(i, N))
            assert float(tv_ref) == pytest.approx(tv_gold, rel=1e-12, abs=1e-9)
