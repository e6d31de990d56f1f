"""Pins of the learned-closure oracle (oracle/sgs.c) against independent implementations and
closed forms: torch's bfloat16 cast, ReLU and autograd, torch.optim.Adam, finite differences,
and known answers of the box filter, the inputs and the fallback rule."""
import numpy as np
import pytest
import torch

from oracle import oracle_py as o
from porturb_amd import cases
import porturb_amd as pa

NIN, NH, NOUT = o.NIN, o.NH, o.NOUT


def rand_params(seed=0, scale=0.15):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(o.NPARAM) * scale).astype(np.float32)


def unpack(P):
    P = np.asarray(P, np.float64)
    i = 0
    out = []
    for shape in [(NH, NIN), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)]:
        n = int(np.prod(shape))
        out.append(P[i:i + n].reshape(shape))
        i += n
    assert i == o.NPARAM
    return out


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.standard_normal(2000).astype(np.float32) * 10 ** rng.uniform(-8, 8, 2000).astype(np.float32),
                         np.float32([1 + 2 ** -8, 1 + 3 * 2 ** -9, -(1 + 2 ** -8), 1 + 2 ** -8 + 2 ** -20, 0.0, -0.0])])
    ref = torch.from_numpy(xs).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([o.bf16_round(x) for x in xs], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert o.bf16_round(1 + 2 ** -8) == 1.0                 # tie goes to even
    assert o.bf16_round(1 + 3 * 2 ** -9) == 1 + 2 ** -7


def test_relu_and_derivative():
    x = np.concatenate([np.linspace(-6, 6, 241), [0.0, -0.0, 1e-300, -1e-300]])
    xt = torch.from_numpy(x).requires_grad_(True)
    y = torch.relu(xt)
    y.sum().backward()
    got = np.array([o.lib().oc_relu(v) for v in x])
    d = np.array([o.lib().oc_relu_deriv(v) for v in x])
    assert np.array_equal(got, y.detach().numpy())
    assert np.array_equal(d, xt.grad.numpy())           # torch takes the derivative at 0 as 0


def torch_forward(P, X, bf16):
    W1, b1, W2, b2, W3, b3 = [torch.from_numpy(a) for a in unpack(P)]
    Xt = torch.from_numpy(np.asarray(X, np.float64))
    if bf16:
        r = lambda t: t.float().bfloat16().double()
        W1, W2, W3, Xt = r(W1), r(W2), r(W3), r(Xt)
    else:
        r = lambda t: t
    g = torch.relu
    a1 = r(g(Xt @ W1.T + b1))
    a2 = r(g(a1 @ W2.T + b2))
    return (a2 @ W3.T + b3).numpy()


@pytest.mark.parametrize("bf16", [False, True])
def test_forward_matches_matrix_form(bf16):
    P = rand_params(2)
    X = np.random.default_rng(3).standard_normal((64, NIN)).astype(np.float32)
    got = o.mlp_forward(P, X, bf16)
    ref = torch_forward(P, X, bf16)
    tol = 1e-5 if not bf16 else 2e-2       # bf16: one activation may round to the other side
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    if bf16:
        assert np.median(np.abs(got - ref)) <= 1e-5 * np.abs(ref).max()


def test_gradient_matches_autograd():
    P = rand_params(4)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((32, NIN)).astype(np.float32)
    T = rng.standard_normal((32, NOUT)).astype(np.float32)
    loss, G = o.mlp_loss_grad(P, X, T, bf16=False)
    Pt = torch.from_numpy(P.astype(np.float64)).requires_grad_(True)
    parts, i = [], 0
    for shape in [(NH, NIN), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)]:
        n = int(np.prod(shape))
        parts.append(Pt[i:i + n].reshape(shape))
        i += n
    W1, b1, W2, b2, W3, b3 = parts
    g = torch.relu
    Y = g(g(torch.from_numpy(X.astype(np.float64)) @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    L = ((Y - torch.from_numpy(T.astype(np.float64))) ** 2).mean()
    L.backward()
    ref = Pt.grad.numpy()
    assert abs(loss - L.item()) <= 1e-6 * L.item()
    assert np.abs(G - ref).max() <= 1e-5 * np.abs(ref).max()


def preactivation_signs(P, X):
    """Signs of every hidden pre-activation of the network on X (float64)."""
    W1, b1, W2, b2, _, _ = unpack(P)
    h1 = np.asarray(X, np.float64) @ W1.T + b1
    h2 = np.maximum(h1, 0.) @ W2.T + b2
    return np.concatenate([np.sign(h1).ravel(), np.sign(h2).ravel()])


def test_gradient_finite_differences():
    """Central differences of the oracle's loss against its gradient.  Within a region where no
    hidden unit switches on or off the loss is a quadratic in any one parameter, so the central
    difference is exact up to rounding; each hidden-layer entry is checked with a step that keeps
    every unit's state (the output layer is linear in its parameters at any step)."""
    P = rand_params(6)
    rng = np.random.default_rng(7)
    X = rng.standard_normal((8, NIN)).astype(np.float32)
    T = rng.standard_normal((8, NOUT)).astype(np.float32)
    _, G = o.mlp_loss_grad(P, X, T, bf16=False)
    # one entry from each block
    for idx in [5, NH * NIN + 3, NH * NIN + NH + 100, o.NPARAM - NOUT - 10, o.NPARAM - 1]:
        h = 1e-2
        while True:
            Pp, Pm = P.copy(), P.copy()
            Pp[idx] += h
            Pm[idx] -= h
            if np.array_equal(preactivation_signs(Pp, X), preactivation_signs(Pm, X)):
                break
            h /= 2
        assert h >= 1e-3, (idx, h)
        fd = (o.mlp_loss_grad(Pp, X, T, False)[0] - o.mlp_loss_grad(Pm, X, T, False)[0]) / (2 * h)
        assert abs(fd - G[idx]) <= 2e-3 * max(abs(G[idx]), 1e-3), idx


def test_adam_matches_torch():
    rng = np.random.default_rng(8)
    P = rand_params(9)
    m = np.zeros_like(P)
    v = np.zeros_like(P)
    Pt = torch.from_numpy(P.copy()).requires_grad_(True)
    opt = torch.optim.Adam([Pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = (rng.standard_normal(P.size) * 10 ** rng.uniform(-4, 0)).astype(np.float32)
        o.mlp_adam(P, g, m, v, t)
        Pt.grad = torch.from_numpy(g.copy())
        opt.step()
    assert np.abs(P - Pt.detach().numpy()).max() <= 1e-6


def test_adam_first_step_is_signed_lr():
    P = np.zeros(100, np.float32)
    m = np.zeros_like(P)
    v = np.zeros_like(P)
    g = np.linspace(-1, 1, 100).astype(np.float32) + 0.005
    o.mlp_adam(P, g, m, v, 1, lr=0.01)
    assert np.allclose(P, -0.01 * np.sign(g), rtol=1e-5)


def fine_state(nx, ny, nz, dx, fields, imm=None):
    cfg = pa.make_config(nx, ny, nz, nx * dx, ny * dx, nz * dx, ord=5, tracers=[("water_vapor", True, True, False)])
    n = (nz, ny, nx)
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) * dx, (np.arange(ny) + 0.5) * dx, (np.arange(nx) + 0.5) * dx,
                          indexing="ij")
    rho, u, v, w, th = fields(x, y, z)
    c = np.zeros((6,) + n)
    c[0], c[1], c[2], c[3] = rho, u, v, w
    c[4] = cfg["C0"] * (rho * th) ** cfg["gamma_d"] / (rho * cfg["R_d"])
    return cfg, c, np.zeros(n) if imm is None else imm


def test_filter_constant_field_has_no_subgrid_stress():
    cfg, c, imm = fine_state(8, 8, 8, 1.0, lambda x, y, z: (1.1 + 0 * x, 3 + 0 * x, -2 + 0 * x, 0.5 + 0 * x, 300 + 0 * x))
    les, immc, tg = o.sgs_filter(cfg, c, imm, 2)
    assert np.abs(tg).max() < 1e-9
    assert np.allclose(les[1, 1:-1, 1:-1, 1:-1], 3) and np.allclose(les[4, 1:-1, 1:-1, 1:-1], 300, rtol=1e-12)


def test_filter_linear_field_closed_form():
    a, b, s, dx, r = 0.3, -0.7, 0.05, 1.0, 4
    cfg, c, imm = fine_state(16, 16, 16, dx, lambda x, y, z: (1.0 + 0 * x, a * x, b * y, 0 * x, 300 + s * z))
    les, immc, tg = o.sgs_filter(cfg, c, imm, r)
    var = dx * dx * (r * r - 1) / 12.            # variance of r equally spaced points
    assert np.allclose(tg[0], a * a * var, rtol=1e-5)      # tau11
    assert np.allclose(tg[3], b * b * var, rtol=1e-5)      # tau22
    assert np.abs(tg[1]).max() < 1e-6                      # tau12: x and y independent in a box
    assert np.abs(tg[5]).max() < 1e-9                      # w = 0
    assert np.abs(tg[6]).max() < 1e-5 and np.abs(tg[8]).max() < 1e-9   # q1: u(x) vs theta(z); q3: w = 0
    # filtered means at the coarse cell centres
    xc = (np.arange(4) * r + (r - 1) / 2 + 0.5) * dx
    assert np.allclose(les[1, 1, 1, 1:-1], a * xc)
    # halos: periodic in x/y, w = 0 and immersed = 1 in z
    assert np.allclose(les[1, 1:-1, 1:-1, 0], les[1, 1:-1, 1:-1, -2])
    assert np.all(les[3, 0] == 0) and np.all(immc[0] == 1) and np.all(immc[-1] == 1)


def test_filter_immersed_is_block_max():
    imm = np.zeros((8, 8, 8))
    imm[0, 0, 0] = 1.0
    imm[5, 3, 6] = 0.25
    cfg, c, imm = fine_state(8, 8, 8, 1.0, lambda x, y, z: (1 + 0 * x, 0 * x, 0 * x, 0 * x, 300 + 0 * x), imm)
    _, immc, _ = o.sgs_filter(cfg, c, imm, 2)
    assert immc[1, 1, 1] == 1.0 and immc[3, 2, 4] == 0.25 and immc[1:-1, 1:-1, 1:-1].sum() == 1.25


def halo_box(fields, nz, ny, nx):
    """A halo-1 closure box [6][nz+2][ny+2][nx+2] from functions of the halo index."""
    k, j, i = np.meshgrid(np.arange(nz + 2) - 1, np.arange(ny + 2) - 1, np.arange(nx + 2) - 1, indexing="ij")
    les = np.zeros((6, nz + 2, ny + 2, nx + 2))
    for f, fn in enumerate(fields):
        les[f] = fn(k, j, i)
    return les


def test_features_linear_field_and_galilean_invariance():
    nz = ny = nx = 6
    g = (0.1, -0.2, 0.05, 0.3)     # per-index gradients of u, v, w, theta along x, y, z combined
    les = halo_box([lambda k, j, i: 1 + 0 * i,
                    lambda k, j, i: g[0] * i, lambda k, j, i: g[1] * j,
                    lambda k, j, i: g[2] * k, lambda k, j, i: 300 + g[3] * (i + j + k),
                    lambda k, j, i: 0 * i], nz, ny, nx)
    fmul = np.linspace(0.5, 2.0, NIN).astype(np.float32)
    cells = np.array([((2 * ny) + 3) * nx + 2, 0, nz * ny * nx - 1])
    X = o.sgs_features(les, fmul, cells)
    expect = np.zeros(NIN)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                n = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                expect[4 * n:4 * n + 4] = [g[0] * dx, g[1] * dy, g[2] * dz, g[3] * (dx + dy + dz)]
    for s in range(3):
        assert np.allclose(X[s], (expect * fmul).astype(np.float32), rtol=1e-5, atol=1e-6)
    assert np.all(X[:, 4 * 13:4 * 13 + 4] == 0)          # the centre's own differences
    les2 = les.copy()
    les2[1] += 7.0
    les2[2] -= 3.0
    X2 = o.sgs_features(les2, fmul, cells)
    assert np.abs(X2 - X).max() < 1e-5


def const_model(out_mean, xmax=1e30):
    P = np.zeros(o.NPARAM, np.float32)
    return o.SgsModel(P, np.ones(NIN), xmax, np.asarray(out_mean, np.float32), np.ones(9))


def test_fallback_rule_known_answers():
    nz = ny = nx = 6
    dudz = 0.4
    les = halo_box([lambda k, j, i: 1 + 0 * i, lambda k, j, i: dudz * k, lambda k, j, i: 0 * i,
                    lambda k, j, i: 0 * i, lambda k, j, i: 300 + 0.01 * k, lambda k, j, i: 0 * i], nz, ny, nx)
    imm = np.zeros((nz + 2, ny + 2, nx + 2))
    imm[3, 3, 3] = 0.5            # halo coords: cell (2,2,2)
    nu_max = 1e9
    # tau13 < 0 with du/dz > 0: dissipation -2 tau13 S13 > 0 -> network kept; q3 < 0 with dtheta/dz > 0 -> kept
    tau, fl = o.sgs_infer(const_model([0, 0, -0.01, 0, 0, 0, 0, 0, -0.001]), les, imm, 1., 1., 1., nu_max)
    assert np.allclose(tau[2], -0.01) and np.allclose(tau[8], -0.001)
    assert np.all(fl[0] == 0) and np.all(fl[-1] == 0)                   # wall levels
    near = np.zeros_like(fl, bool)
    near[1:4, 1:4, 1:4] = True                                          # 3x3x3 around the immersed cell
    inner = np.zeros_like(fl, bool)
    inner[1:-1] = True
    assert np.all(fl[inner & near] == 0) and np.all(fl[inner & ~near] == 3)
    # the opposite signs are backscatter -> fallback
    _, fl2 = o.sgs_infer(const_model([0, 0, 0.01, 0, 0, 0, 0, 0, 0.001]), les, imm, 1., 1., 1., nu_max)
    assert np.all(fl2 == 0)
    # effective viscosity above the cap -> fallback: eps = -2 tau13 S13 = 2*0.01*0.2 = 0.004 >
    # nu_max * 2 S:S = 0.02 * 2*2*0.2^2 = 0.0032; heat chi = 1e-5*0.01 = 1e-7 <= 0.02 * 0.01^2 kept
    _, fl3 = o.sgs_infer(const_model([0, 0, -0.01, 0, 0, 0, 0, 0, -1e-5]), les, imm, 1., 1., 1., 0.02)
    assert np.all(fl3[inner & ~near] == 2)
    # out of range inputs -> fallback for both
    _, fl4 = o.sgs_infer(const_model([0, 0, -0.01, 0, 0, 0, 0, 0, -0.001], xmax=0.1), les, imm, 1., 1., 1., nu_max)
    assert np.all(fl4 == 0)
    # a zero theta gradient bounds no heat flux -> heat fallback; a zero strain likewise for momentum
    flat = halo_box([lambda k, j, i: 1 + 0 * i, lambda k, j, i: dudz * k, lambda k, j, i: 0 * i,
                     lambda k, j, i: 0 * i, lambda k, j, i: 300 + 0 * i, lambda k, j, i: 0 * i], nz, ny, nx)
    _, fl5 = o.sgs_infer(const_model([0, 0, -0.01, 0, 0, 0, 0, 0, -0.001]), flat, imm, 1., 1., 1., nu_max)
    assert np.all(fl5[inner & ~near] == 1)
    still = halo_box([lambda k, j, i: 1 + 0 * i, lambda k, j, i: 0 * i, lambda k, j, i: 0 * i,
                      lambda k, j, i: 0 * i, lambda k, j, i: 300 + 0.01 * k, lambda k, j, i: 0 * i], nz, ny, nx)
    _, fl6 = o.sgs_infer(const_model([0, 0, 0, 0, 0, 0, 0, 0, -0.001]), still, imm, 1., 1., 1., nu_max)
    assert np.all(fl6[inner & ~near] == 2)
    # theta constant up to rounding (one-ulp noise on 300 K) and velocities with rounding-level
    # differences count as zero gradient / strain: fallback, whatever the noise's signs
    rng = np.random.default_rng(3)
    ulp = np.spacing(300.)
    noisy = flat.copy()
    noisy[4] = 300. + ulp * rng.integers(-2, 3, noisy[4].shape)
    _, fl7 = o.sgs_infer(const_model([0, 0, -0.01, 0, 0, 0, 0, 0, -0.001]), noisy, imm, 1., 1., 1., nu_max)
    assert np.all(fl7[inner & ~near] == 1)
    calm = still.copy()
    calm[1] = 5. + np.spacing(5.) * rng.integers(-2, 3, calm[1].shape)
    _, fl8 = o.sgs_infer(const_model([0, 0, -0.01, 0, 0, 0, 0, 0, -0.001]), calm, imm, 1., 1., 1., nu_max)
    assert np.all(fl8[inner & ~near] == 2)
    # a gradient just above the threshold (2e-6 theta per cell width) is kept
    weak = halo_box([lambda k, j, i: 1 + 0 * i, lambda k, j, i: dudz * k, lambda k, j, i: 0 * i,
                     lambda k, j, i: 0 * i, lambda k, j, i: 300 * (1 + 4e-6 * k), lambda k, j, i: 0 * i], nz, ny, nx)
    _, fl9 = o.sgs_infer(const_model([0, 0, -0.01, 0, 0, 0, 0, 0, -0.001]), weak, imm, 1., 1., 1., nu_max)
    assert np.all(fl9[inner & ~near] == 3)


def shear_case(n=8, nz=10):
    cfg = cases.abl_urban(nx=n, ny=n, nz=nz, dx=2.)
    rng = np.random.default_rng(11)
    shape = (nz, n, n)
    z = (np.arange(nz) + 0.5) * 2.
    c = np.zeros((5 + cfg["num_tracers"],) + shape)
    c[0] = 1.16
    c[1] = (0.02 * z * z)[:, None, None] + 0 * rng.standard_normal(shape)
    c[4] = 290.
    c[5 + cfg["idTKE"]] = 0.2 * c[0]
    return cfg, c, np.zeros(shape)


def test_learned_apply_all_fallback_equals_plain():
    cfg, c, imm = shear_case()
    a, b = c.copy(), c.copy()
    o.les_apply(cfg, a, imm, 0.01)
    _, fl = o.les_apply_learned(cfg, b, imm, 0.01, const_model(np.zeros(9), xmax=-1.))
    assert np.all(fl == 0) and np.array_equal(a, b)


def test_learned_apply_zero_stress_removes_interior_momentum_flux():
    cfg, c, imm = shear_case()
    a, b = c.copy(), c.copy()
    o.les_apply(cfg, a, imm, 0.01)
    _, fl = o.les_apply_learned(cfg, b, imm, 0.01, const_model(np.zeros(9)))
    nz = cfg["nz"]
    # theta is uniform here: no gradient bounds a heat flux, so only the momentum bit is set
    assert np.all(fl[1:nz - 1] == 1) and np.all(fl[0] == 0) and np.all(fl[nz - 1] == 0)
    # the plain closure diffuses the curved profile on every level; learned faces (both cells
    # flagged) carry zero flux, so u is unchanged on levels 2..nz-3, and level 0 keeps the plain update
    assert np.abs(a[1] - c[1])[1].min() > 1e-6 and np.abs(a[1] - c[1])[3].min() > 1e-6
    assert np.abs(b[1, 2:nz - 2] - c[1, 2:nz - 2]).max() < 1e-13
    assert np.allclose(b[1, 0], a[1, 0], rtol=0, atol=1e-13)
