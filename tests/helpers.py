"""Seeded host-side inputs and norms shared by the tests."""
import numpy as np

from porturb_amd import make_config, physical_constants, PU_BC_PERIODIC, PU_BC_SOLID_WALL, PU_FP64

WV = ("water_vapor", True, True, True)
TKE = ("TKE", True, False, False)
P1 = ("passive_1", True, False, True)
P2 = ("passive_2", True, False, True)


def isentropic(z, theta0=300.):
    """Point values of the constant-theta hydrostatic state (profiles.h:12-21)."""
    c = physical_constants()
    exner = 1. - c["grav"] * z / (c["cp_d"] * theta0)
    p = c["p0"] * exner ** (c["cp_d"] / c["R_d"])
    rt = (p / c["C0"]) ** (1. / c["gamma_d"])
    return rt / theta0


def random_state(cfg, seed=0, amp_u=3.0, amp_t=1.0, buildings=0, vapor=True, smooth=True):
    """Coupler state [5+ntr][nz][ny][nx] (density_dry, u, v, w, temp, tracer masses) and
    immersed_proportion, built from an isentropic background plus seeded perturbations.
    `buildings` cuboids of fully immersed cells (value 1) and one partially immersed cell."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    dx, dy, dz = cfg["xlen"] / nx, cfg["ylen"] / ny, cfg["zlen"] / nz
    c = physical_constants()
    z = (np.arange(nz) + 0.5) * dz
    rho = np.broadcast_to(isentropic(z)[:, None, None], (nz, ny, nx)).copy()

    def field(amp):
        f = rng.standard_normal((nz, ny, nx))
        if smooth:
            for ax in range(3):
                f = 0.25 * np.roll(f, 1, ax) + 0.5 * f + 0.25 * np.roll(f, -1, ax)
        return amp * f

    u, v, w = field(amp_u) + 2.0, field(amp_u), field(amp_u) * 0.5
    theta = 300. + field(amp_t)
    rho = rho * (1. + field(0.002))
    ntr = cfg["num_tracers"]
    st = np.zeros((5 + ntr, nz, ny, nx))
    st[0], st[1], st[2], st[3] = rho, u, v, w
    for t, (name, pos, mass, diff) in enumerate(cfg["tracers"]):
        if name == "water_vapor":
            st[5 + t] = rho * (np.abs(field(1e-3)) if vapor else 0.)
        elif name == "TKE":
            st[5 + t] = rho * (0.1 + np.abs(field(0.2)))
        else:
            q = np.clip(field(1.0) + 0.5, 0, None)
            st[5 + t] = rho * q
    rho_v = st[5 + cfg["idWV"]]
    # temperature such that p = C0 (rho_full theta)^gamma with rho_full = rho_d + vapour
    rho_full = rho + (rho_v if cfg["tracers"][cfg["idWV"]][2] else 0)
    p = c["C0"] * (rho_full * theta) ** c["gamma_d"]
    st[4] = p / (rho * c["R_d"] + rho_v * c["R_v"])
    imm = np.zeros((nz, ny, nx))
    for b in range(buildings):
        bx, by = rng.integers(0, nx - 3), rng.integers(0, ny - 3)
        wx, wy = rng.integers(2, max(3, nx // 4)), rng.integers(2, max(3, ny // 4))
        h = rng.integers(1, max(2, nz // 3))
        imm[:h, by:by + wy, bx:bx + wx] = 1.
    if buildings:
        imm[0, ny // 2, nx // 2] = 0.5
    st[1:4, imm == 1] = 0.
    return st, imm


def rel_norms(a, b, scale=None):
    """Per-field relative L1, L2, Linf of a-b, normalised by the field's own norms of b."""
    out = []
    for f in range(a.shape[0]):
        d = (a[f] - b[f]).ravel()
        ref = b[f].ravel()
        s1 = np.abs(ref).sum() or 1.
        s2 = np.sqrt((ref ** 2).sum()) or 1.
        si = np.abs(ref).max() or 1.
        if scale is not None:
            s1, s2, si = scale[f] * ref.size, scale[f] * np.sqrt(ref.size), scale[f]
        out.append((np.abs(d).sum() / s1, np.sqrt((d ** 2).sum()) / s2, np.abs(d).max() / si))
    return np.array(out)


def rounding_sensitivity(run, st, ref, rel=1e-15, seed=0):
    """Per-field relative [L1, L2, Linf] by which `run` (a function of an initial state, e.g. N oracle
    steps) moves when the initial temperature is perturbed by `rel` relative (random signs), against
    `ref` = run(st): how far a rounding-level difference carries in this case.  Resting air is
    ill-conditioned under the WENO limiter: a stencil whose candidates' total variation crosses the
    1e-20 guard switches from the ideal to the nonlinear weights (WenoLimiter.h convexify), so
    velocities of ~1e-10 m/s at the edge of the resting region change by O(1) of themselves."""
    s = st.copy()
    s[4] = s[4] * (1 + rel * np.random.default_rng(seed).standard_normal(s[4].shape))
    return rel_norms(run(s), ref)


def small_config(nx=20, ny=18, nz=16, ord=5, tracers=None, precision=PU_FP64, dx=50., **kw):
    if tracers is None:
        tracers = [WV, TKE, P1]
    return make_config(nx, ny, nz, nx * dx, ny * dx, nz * dx, ord=ord, tracers=tracers,
                       precision=precision, **kw)
