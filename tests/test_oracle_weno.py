"""Known-answer tests of the oracle's limiter (WenoLimiter<real,ord>::compute_limited_edges).

Each test is chosen so a plausible mistake fails it: constant data (normalisation guard, eps),
polynomial exactness (sub-stencil matrices, face weights, ideal weights), mirror symmetry (left/right
face tables swapped), the map fixed point, and imm_L/imm_R (which TV is replaced by the maximum).
"""
import numpy as np
import pytest

from oracle import oracle_py

ORDERS = (3, 5, 7, 9, 11)


def cell_averages(coefs, offsets):
    """Exact cell averages over [o-1/2, o+1/2] of the polynomial sum c_n x^n."""
    out = []
    for o in offsets:
        lo, hi = o - 0.5, o + 0.5
        out.append(sum(c * (hi ** (n + 1) - lo ** (n + 1)) / (n + 1) for n, c in enumerate(coefs)))
    return np.array(out)


def poly_at(coefs, x):
    return sum(c * x ** n for n, c in enumerate(coefs))


@pytest.mark.parametrize("ord_", ORDERS)
@pytest.mark.parametrize("do_map", [False, True])
def test_constant_stencil_is_exact(ord_, do_map):
    qL, qR = oracle_py.weno_edges(ord_, np.full(ord_, 3.25), do_map)
    assert qL == pytest.approx(3.25, rel=1e-15) and qR == pytest.approx(3.25, rel=1e-15)


@pytest.mark.parametrize("ord_", (5, 7, 9, 11))
@pytest.mark.parametrize("do_map", [False, True])
def test_low_degree_polynomial_is_reproduced(ord_, do_map):
    """For a polynomial of degree <= hs every sub-stencil recovers it exactly, so every weighting
    gives the exact interface values p(-1/2), p(+1/2)."""
    hs = (ord_ - 1) // 2
    rng = np.random.default_rng(ord_)
    for trial in range(5):
        coefs = rng.standard_normal(hs + 1)
        s = cell_averages(coefs, range(-hs, hs + 1))
        qL, qR = oracle_py.weno_edges(ord_, s, do_map)
        assert qL == pytest.approx(poly_at(coefs, -0.5), abs=1e-11 * (1 + np.abs(coefs).sum()))
        assert qR == pytest.approx(poly_at(coefs, 0.5), abs=1e-11 * (1 + np.abs(coefs).sum()))


@pytest.mark.parametrize("ord_", ORDERS)
def test_smooth_data_recovers_full_order(ord_):
    """With ideal weights (smooth data, small TV differences) the reconstruction error of a smooth
    profile falls at the design order ord under refinement."""
    errs = []
    for h in (0.4, 0.2):
        offs = np.arange(-(ord_ // 2), ord_ // 2 + 1)
        # averages of sin(x) over cells of width h centred at x0 + o*h
        x0 = 0.3
        s = np.array([(np.cos(x0 + (o - .5) * h) - np.cos(x0 + (o + .5) * h)) / h for o in offs])
        qL, qR = oracle_py.weno_edges(ord_, s, True)
        errs.append(max(abs(qL - np.sin(x0 - h / 2)), abs(qR - np.sin(x0 + h / 2))))
    rate = np.log2(errs[0] / errs[1])
    assert rate > ord_ - 1.5, (errs, rate)


@pytest.mark.parametrize("ord_", ORDERS)
@pytest.mark.parametrize("do_map", [False, True])
def test_mirror_symmetry(ord_, do_map):
    rng = np.random.default_rng(7)
    for _ in range(10):
        s = rng.standard_normal(ord_) * np.array([1, 10, 1, 100, 1, 1, 1, 1, 1, 1, 1])[:ord_]
        qL, qR = oracle_py.weno_edges(ord_, s, do_map)
        rL, rR = oracle_py.weno_edges(ord_, s[::-1].copy(), do_map)
        assert qR == pytest.approx(rL, rel=1e-12, abs=1e-12)
        assert qL == pytest.approx(rR, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("ord_", (5, 9))
def test_discontinuity_is_not_oscillatory(ord_):
    """A step: the interface values stay inside the data range (non-oscillatory weights)."""
    hs = (ord_ - 1) // 2
    for jump_at in range(1, ord_):
        s = np.where(np.arange(ord_) < jump_at, 0.0, 1.0)
        qL, qR = oracle_py.weno_edges(ord_, s, True)
        assert -0.02 < qL < 1.02 and -0.02 < qR < 1.02, (jump_at, qL, qR)


@pytest.mark.parametrize("ord_", (5, 9))
def test_immersed_flags_suppress_the_far_stencil(ord_):
    """imm_L replaces the TV of the last sub-stencil (the one not containing the left interface)
    by the maximum; with a step on the far right, the left-face value must then come from the
    smooth sub-stencils only -- and imm flags must change the result relative to no flags."""
    hs = (ord_ - 1) // 2
    s = np.linspace(0, 1, ord_) ** 2
    s[-1] += 5.0
    base = oracle_py.weno_edges(ord_, s, False)
    withL = oracle_py.weno_edges(ord_, s, False, imm_L=True)
    withR = oracle_py.weno_edges(ord_, s, False, imm_R=True)
    assert withL != base or withR != base
    # symmetry of the two flags
    r = oracle_py.weno_edges(ord_, s[::-1].copy(), False, imm_R=True)
    assert withL[0] == pytest.approx(r[1], rel=1e-12, abs=1e-12)


def test_map_rs_fixed_point_via_ideal_weights():
    """When all TVs are equal the normalised weights are the ideal ones; map_rs(d, d) = d keeps them,
    so mapped and unmapped results agree (a wrong map exponent or constant would break this only
    away from the fixed point, which test_mirror_symmetry/test_low_degree cover)."""
    for ord_ in (5, 7, 9, 11):
        hs = (ord_ - 1) // 2
        s = cell_averages([0.3, -1.2], range(-hs, hs + 1))   # linear: all TVs equal
        a = oracle_py.weno_edges(ord_, s, False)
        b = oracle_py.weno_edges(ord_, s, True)
        assert a[0] == pytest.approx(b[0], rel=1e-14) and a[1] == pytest.approx(b[1], rel=1e-14)
