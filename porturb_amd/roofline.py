"""Algorithmic work of the dycore step, counted from the generated WENO expressions.

FLOPs are counted as the arithmetic the algorithm performs (one per add, multiply, compare-free
reciprocal or divide; a fused multiply-add counts two), independent of how the compiler schedules
it.  Bytes are the compulsory HBM traffic of a fully fused step: each state value read and written
once per stage.
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc"))
import gen_weno_tables as _gen  # noqa: E402

MI355X = {"hbm_GBps": 8000.0, "fp64_TFLOPs": 78.6, "fp32_TFLOPs": 157.3}


def recon_flops(ord_, do_map=True):
    """Both interface values of one field in one cell (weno_edges, WENO branch)."""
    t = _gen.tables(ord_)
    N, S = t["N"], t["S"]
    f = 0
    for m in range(S):
        for n in range(1, N):
            nnz = sum(1 for i in range(N) if t["C"][m][n][i] != 0)
            f += 2 * nnz - 1
        for (i, j), c in t["tv"].items():
            f += 1 if c == 1 else 2
        f += len(t["tv"]) - 1
    conv = (S - 1) + 1 + S
    f += conv                       # normalise TV
    f += 3 * S                      # 1/(TV^2 + eps)
    f += 2 * S                      # ideal * 1/(...)
    f += 2 * conv
    if do_map:
        f += 2 * S * (14 if ord_ > 3 else 9)
        f += 2 * conv
    for side in ("L", "R"):
        for m in range(S):
            nnz = sum(1 for v in t["face"][side][m] if v != 0)
            f += 2 * nnz - 1
        f += S + (S - 1)
    return f


def _tv_flops(ord_):
    """The candidate stencils' total variations and their sum (evaluated by every branch)."""
    t = _gen.tables(ord_)
    N, S = t["N"], t["S"]
    f = 0
    for m in range(S):
        for n in range(1, N):
            nnz = sum(1 for i in range(N) if t["C"][m][n][i] != 0)
            f += 2 * nnz - 1
        for (i, j), c in t["tv"].items():
            f += 1 if c == 1 else 2
        f += len(t["tv"]) - 1
    return f + (S - 1)


def linear_flops(ord_):
    """Both interface values from the full-stencil polynomial (the linear reconstruction)."""
    t = _gen.tables(ord_)
    return sum(2 * sum(1 for v in t["lin"][side] if v != 0) - 1 for side in ("L", "R"))


def recon_flops_flat(ord_):
    """Both interface values of a flat stencil (total variation <= 1e-20): the total variations
    that decide the branch, then the linear reconstruction."""
    return _tv_flops(ord_) + linear_flops(ord_)


def sweep_overhead_flops(nf):
    """Per cell and direction, besides the reconstructions: the acoustic Riemann solve at one
    face, the upwinded fluxes and the flux divergence of nf-1 conserved variables."""
    riemann = 16
    advected = (nf - 3) * 1 + 2      # ru_upw * q_upw, hydrostatic theta add-back
    divergence = (nf - 1) * 3
    return riemann + advected + divergence


def executed_sweep_flops(ord_, nf, census_row, cells):
    """FLOPs one sweep launch executes over `cells` cells, from the reconstruction census of the
    state (counts of the limiter, flat-stencil and linear branches)."""
    n_lim, n_flat, n_lin = (int(v) for v in census_row)
    return (n_lim * recon_flops(ord_) + n_flat * recon_flops_flat(ord_) + n_lin * linear_flops(ord_)
            + cells * sweep_overhead_flops(nf))


def sweep_flops_per_cell(ord_, nf):
    """One direction: nf reconstructions, the acoustic Riemann solve at one face, upwinded fluxes
    and the flux divergence of nf-1 conserved variables."""
    return nf * recon_flops(ord_) + sweep_overhead_flops(nf)


def step_flops_per_cell(ord_, nf):
    """One SSPRK3 step (three stages of three sweeps) plus prep and combine."""
    prep = 30 + 2 * (nf - 6)          # pow (counted as ~20), reciprocal, perturbations
    combine = 6 * (nf - 1) + 10
    return 3 * (3 * sweep_flops_per_cell(ord_, nf) + prep + combine)


def step_bytes_per_cell(nv, fp64=True):
    """Compulsory traffic of a fused step: per stage read the stage state and the step-start
    state and write the new state (8*nv*s per stage on average over the three stages), plus the
    immersed mask; FP32 adds the compensation words of rho and rho*theta."""
    s = 8 if fp64 else 4
    b = 8 * nv * s + 12
    if not fp64:
        b += 64
    return b
