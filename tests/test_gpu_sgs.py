"""Learned subgrid closure on the GPU (libporturb through the C ABI) vs the CPU oracle.

Tolerances.  The network runs BF16 operands with FP32 accumulation on the matrix cores; the
oracle rounds the same operands to BF16 and accumulates in double.  Sums therefore differ by FP32
rounding (~1e-6 relative), and an activation whose FP32 value lies within that distance of a
BF16 rounding boundary rounds to the neighbouring BF16 value (2^-8 relative) in one of the two:
outputs agree to 1e-4 of their scale in the median and 2e-2 at worst.  Integer/byte work (the
BF16 packing of inputs) is bit-exact; the FP64 closure with a network whose outputs are exact
(zero weights) agrees to 1e-12 like the plain closure.
"""
import numpy as np
import pytest

import porturb_amd as pa
from oracle import oracle_py as o
from porturb_amd import cases, sgs_training as st_
from tests.helpers import random_state, small_config, rel_norms, WV, TKE, P1, P2

pytestmark = pytest.mark.gpu

NIN, NOUT = o.NIN, o.NOUT


def rand_params(seed, scale=0.15):
    return (np.random.default_rng(seed).standard_normal(o.NPARAM) * scale).astype(np.float32)


def test_pack_inputs_bit_exact():
    X = np.random.default_rng(1).standard_normal((300, NIN)).astype(np.float32) * 3
    Xb = pa.pack_inputs(X)
    got = pa.bf16_to_float(Xb.download(np.uint16, (300, pa.SGS_NINP)))
    ref = np.vectorize(o.bf16_round, otypes=[np.float32])(X)
    assert np.array_equal(got[:, :NIN].view(np.uint32), ref.view(np.uint32))
    assert np.all(got[:, NIN:] == 0)


def test_forward_matches_oracle():
    P = rand_params(2)
    X = np.random.default_rng(3).standard_normal((1000, NIN)).astype(np.float32)   # not a multiple of 64
    m = pa.Mlp(P)
    Y = m.forward(pa.pack_inputs(X), X.shape[0])
    ref = o.mlp_forward(P, X, bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(Y - ref)
    assert np.median(err) <= 1e-4 * scale, np.median(err) / scale
    assert err.max() <= 2e-2 * scale, err.max() / scale


def test_forward_with_index_list():
    P = rand_params(4)
    X = np.random.default_rng(5).standard_normal((256, NIN)).astype(np.float32)
    m = pa.Mlp(P)
    Xb = pa.pack_inputs(X)
    idx = np.random.default_rng(6).permutation(256)[:100].astype(np.int32)
    di = pa.DeviceBuffer(idx.nbytes)
    di.upload(idx)
    Y = m.forward(Xb, 100, di)
    Yall = m.forward(Xb, 256)
    assert np.array_equal(Y, Yall[idx])


def test_gradient_matches_oracle():
    P = rand_params(7)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((777, NIN)).astype(np.float32)
    T = rng.standard_normal((777, NOUT)).astype(np.float32)
    m = pa.Mlp(P)
    dT = pa.DeviceBuffer(T.nbytes)
    dT.upload(T)
    loss, G = m.loss_grad(pa.pack_inputs(X), dT, 777)
    lref, Gref = o.mlp_loss_grad(P, X, T, bf16=True)
    assert abs(loss - lref) <= 1e-4 * lref
    assert np.linalg.norm(G - Gref) <= 5e-3 * np.linalg.norm(Gref)
    # each parameter block
    off = 0
    for n in [64 * 108, 64, 64 * 64, 64, 9 * 64, 9]:
        a, b = G[off:off + n], Gref[off:off + n]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), off
        off += n
    # and the BF16 training gradient stays close to the exact one
    _, Gx = o.mlp_loss_grad(P, X, T, bf16=False)
    assert np.linalg.norm(G - Gx) <= 5e-2 * np.linalg.norm(Gx)


def test_adam_matches_oracle():
    P = rand_params(9)
    m = pa.Mlp(P)
    Pr, mr, vr = P.copy(), np.zeros_like(P), np.zeros_like(P)
    rng = np.random.default_rng(10)
    for t in range(1, 4):
        g = (rng.standard_normal(P.size) * 10 ** rng.uniform(-3, 0)).astype(np.float32)
        m.adam_apply(g, 1e-3)
        o.mlp_adam(Pr, g, mr, vr, t, lr=1e-3)
    assert np.abs(m.get_params() - Pr).max() <= 1e-7


def test_training_fits_a_teacher_network():
    rng = np.random.default_rng(11)
    teacher = rand_params(12, scale=0.1)
    X = rng.standard_normal((8192, NIN)).astype(np.float32)
    T = o.mlp_forward(teacher, X, bf16=False)
    T = (T - T.mean(0)) / T.std(0)
    m = pa.Mlp(st_.init_params(13))
    Xb = pa.pack_inputs(X)
    dT = pa.DeviceBuffer(T.astype(np.float32).nbytes)
    dT.upload(T.astype(np.float32))
    l0, _ = m.loss_grad(Xb, dT, 8192)
    idx = pa.DeviceBuffer(4 * 8192)
    for it in range(300):
        if it % 2 == 0:
            idx.upload(rng.permutation(8192).astype(np.int32))
        m.train_step(Xb, dT, st_._Ptr(idx.ptr + 4 * 4096 * (it % 2)), 4096, 3e-3)
    l1, _ = m.loss_grad(Xb, dT, 8192)
    assert l1 < 0.2 * l0, (l0, l1)


def test_model_file_roundtrip(tmp_path):
    P = rand_params(14)
    m = pa.Mlp(P)
    fs = np.linspace(0.5, 2, NIN).astype(np.float32)
    om, os_ = np.arange(9, dtype=np.float32), np.linspace(1, 2, 9).astype(np.float32)
    m.set_normalisation(fs, 7.5, om, os_)
    path = str(tmp_path / "net.bin")
    m.save(path)
    m2 = pa.Mlp.load(path)
    assert np.array_equal(m2.get_params(), P)
    fs2, xm, om2, os2 = m2.get_normalisation()
    assert np.array_equal(fs2, fs) and xm == 7.5 and np.array_equal(om2, om) and np.array_equal(os2, os_)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"not a network")
    with pytest.raises(RuntimeError):
        pa.Mlp.load(str(bad))


def fine_state(nx=16, ny=16, nz=16, dx=1.0, seed=15):
    cfg = cases.abl_urban(nx=nx, ny=ny, nz=nz, dx=dx)
    st, imm = random_state(cfg, seed=seed, buildings=2)
    return cfg, st, imm


def test_filter_matches_oracle():
    cfg, st, imm = fine_state()
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    les, immc, tg = pa.sgs_filter(d.ctx, cs.c, cfg, 2)
    n = 8
    g_les = les.download(np.float64, (6, n + 2, n + 2, n + 2))
    g_imm = immc.download(np.float64, (n + 2, n + 2, n + 2))
    g_tg = tg.download(np.float32, (9, n, n, n))
    r_les, r_imm, r_tg = o.sgs_filter(cfg, st, imm, 2)
    assert np.array_equal(g_imm, r_imm)
    for f in range(6):
        s = max(np.abs(r_les[f]).max(), 1e-30)
        assert np.abs(g_les[f] - r_les[f]).max() <= 1e-12 * s, f
    for f in range(9):
        s = max(np.abs(r_tg[f]).max(), 1e-30)
        assert np.abs(g_tg[f] - r_tg[f]).max() <= 1e-6 * s, f


def test_features_bit_exact():
    cfg, st, imm = fine_state(seed=16)
    r_les, _, _ = o.sgs_filter(cfg, st, imm, 2)
    les = pa.DeviceBuffer(r_les.nbytes)
    les.upload(r_les)
    fmul = np.random.default_rng(17).uniform(0.2, 3., NIN).astype(np.float32)
    cells = np.arange(8 * 8 * 8, dtype=np.int64)[::3]
    Xb = pa.sgs_features(les, 8, 8, 8, fmul, cells)
    got = pa.bf16_to_float(Xb.download(np.uint16, (cells.size, pa.SGS_NINP)))
    ref = np.vectorize(o.bf16_round, otypes=[np.float32])(o.sgs_features(r_les, fmul, cells))
    assert np.array_equal(got[:, :NIN].view(np.uint32), ref.view(np.uint32))
    assert np.all(got[:, NIN:] == 0)


def les_case(seed, nx=24, ny=20, nz=14):
    cfg = small_config(nx=nx, ny=ny, nz=nz, ord=5, tracers=[WV, TKE, P1, P2], dx=2.)
    st, imm = random_state(cfg, seed=seed, buildings=2, amp_u=1.0)
    return cfg, st, imm


def model_for(P, out_mean, out_scale, xmax=1e30, fs=None):
    fs = np.full(NIN, 0.5, np.float32) if fs is None else fs
    m = pa.Mlp(P)
    m.set_normalisation(fs, xmax, out_mean, out_scale)
    return m, o.SgsModel(P, fs, xmax, out_mean, out_scale)


def run_learned(cfg, st, imm, gm, dt):
    cs = pa.CouplerState(cfg)
    cs.upload((st, imm))
    d = pa.Dycore(cfg)
    d.les_init(cs)
    d.set_sgs_model(gm)
    tau, fl = d.sgs(cs, dt)
    d.les_apply(cs, dt)
    return cs.download(), tau, fl


@pytest.mark.parametrize("shape,learned_min", [((24, 20, 14), 0.2), ((150, 6, 10), 0.1)], ids=["24x20x14", "150x6x10"])
def test_inference_matches_oracle(shape, learned_min):
    """A random network on a closure box with buildings; 150 cells in x span two 128-cell groups of
    the inference kernel, the second one ending inside a wave's second 16-cell tile (its narrow y
    and z put more cells next to walls and buildings, hence the lower share of learned cells)."""
    nx, ny, nz = shape
    cfg, st, imm = les_case(18, nx, ny, nz)
    P = rand_params(19, scale=0.1)
    om = np.array([0.02, 0, -0.01, 0.02, 0, 0.02, 0, 0, -1e-3], np.float32)
    os_ = np.array([0.01, 0.005, 0.005, 0.01, 0.005, 0.01, 1e-3, 1e-3, 1e-3], np.float32)
    gm, om_ = model_for(P, om, os_)
    dt = 0.02
    _, tau, fl = run_learned(cfg, st, imm, gm, dt)
    ref = st.copy()
    rtau, rfl = o.les_apply_learned(cfg, ref, imm, dt, om_)
    assert (fl == rfl).mean() >= 0.99
    assert (rfl == 3).sum() > learned_min * rfl.size
    for f in range(9):
        s = max(np.abs(rtau[f]).max(), 1e-30)
        err = np.abs(tau[f] - rtau[f])
        assert np.median(err) <= 1e-4 * s and err.max() <= 2e-2 * s, f


def test_inference_fp32_mode_matches_oracle():
    """FP32 dycore mode: the closure box is float, so the network's inputs are differences of
    float-rounded values (the FP64 oracle takes them from doubles).  Differences of nearby theta
    values of ~300 K lose ~3e-5 K each, which can move an input across a BF16 rounding boundary.
    The bounds are the FP64 mode's: flags equal on 99 % of the cells, outputs within 1e-4 (median)
    and 2e-2 (max) of each output's scale (measured: every flag equal, median 0, max 3.7e-3)."""
    cfg, st, imm = les_case(18)
    cfg = dict(cfg, precision=pa.PU_FP32)
    P = rand_params(19, scale=0.1)
    om = np.array([0.02, 0, -0.01, 0.02, 0, 0.02, 0, 0, -1e-3], np.float32)
    os_ = np.array([0.01, 0.005, 0.005, 0.01, 0.005, 0.01, 1e-3, 1e-3, 1e-3], np.float32)
    gm, om_ = model_for(P, om, os_)
    dt = 0.02
    _, tau, fl = run_learned(cfg, st, imm, gm, dt)
    ref = st.copy()
    rtau, rfl = o.les_apply_learned(dict(cfg, precision=pa.PU_FP64), ref, imm, dt, om_)
    print("fp32 inference: flags agree on %.4f of the cells" % (fl == rfl).mean(), flush=True)
    assert (fl == rfl).mean() >= 0.99
    assert (rfl == 3).sum() > 0.2 * rfl.size
    both = (fl == rfl)
    for f in range(9):
        s = max(np.abs(rtau[f]).max(), 1e-30)
        err = np.abs(tau[f] - rtau[f])[both]
        print("output", f, "median %.2e max %.2e of scale" % (np.median(err) / s, err.max() / s), flush=True)
        assert np.median(err) <= 1e-4 * s and err.max() <= 2e-2 * s, f


def neighbourhood_agree(fa, fb):
    """Cells whose own flags and their six face neighbours' flags agree (x/y periodic)."""
    same = fa == fb
    ok = same.copy()
    for ax in (1, 2):
        ok &= np.roll(same, 1, ax) & np.roll(same, -1, ax)
    up = np.ones_like(same)
    up[:-1] = same[1:]
    dn = np.ones_like(same)
    dn[1:] = same[:-1]
    return ok & up & dn


@pytest.mark.parametrize("kind", ["zero", "constant"])
def test_learned_apply_exact_outputs_fp64(kind):
    """Networks whose outputs are exact (zero weights): the learned closure agrees with the
    oracle to FP64 round-off."""
    cfg, st, imm = les_case(20)
    P = np.zeros(o.NPARAM, np.float32)
    om = np.zeros(9, np.float32) if kind == "zero" else np.array([0, 0, -2e-3, 0, 0, 0, 0, 0, -1e-4], np.float32)
    gm, om_ = model_for(P, om, np.ones(9, np.float32))
    dt = 0.02
    g, tau, fl = run_learned(cfg, st, imm, gm, dt)
    ref = st.copy()
    rtau, rfl = o.les_apply_learned(cfg, ref, imm, dt, om_)
    assert np.array_equal(tau, rtau)
    agree = neighbourhood_agree(fl, rfl)
    assert agree.mean() >= 0.99 and (rfl > 0).sum() > 0
    plain = st.copy()
    o.les_apply(cfg, plain, imm, dt)
    assert np.abs(plain[1] - ref[1]).max() > 0           # the network changed the update
    for f in range(g.shape[0]):
        s = max(np.abs(ref[f]).max(), 1e-300)
        assert np.abs(g[f] - ref[f])[agree].max() <= 1e-12 * s, f


def test_learned_apply_random_network():
    cfg, st, imm = les_case(21)
    P = rand_params(22, scale=0.1)
    om = np.array([0.02, 0, -0.01, 0.02, 0, 0.02, 0, 0, -1e-3], np.float32)
    os_ = np.array([0.01, 0.005, 0.005, 0.01, 0.005, 0.01, 1e-3, 1e-3, 1e-3], np.float32)
    gm, om_ = model_for(P, om, os_)
    dt = 0.02
    g, tau, fl = run_learned(cfg, st, imm, gm, dt)
    ref = st.copy()
    o.les_apply_learned(cfg, ref, imm, dt, om_)
    _, rfl = o.les_apply_learned(cfg, st.copy(), imm, dt, om_)
    agree = neighbourhood_agree(fl, rfl)
    assert agree.mean() >= 0.97
    for f in (1, 2, 3, 4):                                   # u, v, w, temp
        inc_g, inc_r = g[f] - st[f], ref[f] - st[f]
        s = max(np.abs(inc_r).max(), 1e-300)
        assert np.abs(inc_g - inc_r)[agree].max() <= 2e-2 * s, f


def test_a_priori_gain_over_eddy_viscosity():
    """Train on a small fine run with the library's kernels; on held-out cells the network's
    subgrid stresses beat the eddy-viscosity closure's, even rescaled to its best coefficient."""
    mlp, res = st_.run(96, 96, 48, 1.0, 2, [40], iters=800, batch=1024, lr=3e-3, seed=5)
    print("a priori NMSE network %.3f, eddy viscosity %.3f, best-scaled %.3f" %
          (res["nmse_network_mean"], res["nmse_eddy_viscosity_mean"], res["nmse_eddy_viscosity_best_scaled_mean"]))
    assert res["nmse_network_mean"] < 0.5 * res["nmse_eddy_viscosity_best_scaled_mean"]
    assert res["nmse_network_mean"] < 0.6
    assert res["loss_history"][-1][1] < 0.5 * res["loss_history"][0][1]


def test_driver_with_learned_closure(tmp_path):
    """The C++ driver loads a network file (option les_sgs_model) and steps; profiles stay close
    to the plain closure's over a short run."""
    mlp, _ = st_.run(64, 64, 32, 1.0, 2, [20], iters=300, batch=1024, lr=3e-3, seed=6)
    path = str(tmp_path / "sgs.bin")
    mlp.save(path)
    a = pa.Driver(pa.PU_CASE_ABL_URBAN, 32, 32, 16, 2.0, ord=9, precision=pa.PU_FP32, seed=7, sgs_model=path)
    b = pa.Driver(pa.PU_CASE_ABL_URBAN, 32, 32, 16, 2.0, ord=9, precision=pa.PU_FP64, seed=7)
    a.step(20)
    b.step(20)
    sa, sb = a.download(), b.download()
    for f in ("uvel", "vvel", "wvel", "temp", "TKE"):
        assert np.all(np.isfinite(sa[f])), f
    ua, ub = sa["uvel"].mean(axis=(1, 2)), sb["uvel"].mean(axis=(1, 2))
    assert np.abs(ua - ub).max() <= 0.02 * np.abs(ub).max()


def test_trained_network_flags_and_outputs_match_oracle():
    """The shipped network (porturb_amd/data/sgs_abl_urban_dx2.bin) on a 96x96x32 piece of the
    case it was trained for (neutral urban ABL at 2 m, seeded state with buildings): outputs to the
    BF16 bounds above, and each fallback bit (momentum, heat) set on the same cells as the oracle
    for >= 99 % of the cells, with the same fraction of learned cells to 0.5 %."""
    import os
    path = os.path.join(os.path.dirname(pa.__file__), "data", "sgs_abl_urban_dx2.bin")
    gm = pa.Mlp.load(path)
    fs, xmax, om, os_ = gm.get_normalisation()
    model = o.SgsModel(gm.get_params(), fs, xmax, om, os_)
    cfg = cases.abl_urban(96, 96, 32, 2.0, ord=9)
    cs = pa.CouplerState(cfg)
    pa.case_init(cfg, cfg["case_id"], 5, cs)
    st, imm = cs.download(), cs.download_immersed()
    dt = 0.02
    _, tau, fl = run_learned(cfg, st, imm, gm, dt)
    ref = st.copy()
    rtau, rfl = o.les_apply_learned(cfg, ref, imm, dt, model)
    for bit in (1, 2):
        g, r = (fl & bit) != 0, (rfl & bit) != 0
        assert (g == r).mean() >= 0.99, (bit, (g == r).mean())
        assert abs(g.mean() - r.mean()) < 5e-3, (bit, g.mean(), r.mean())
        assert r.mean() > 0.05, (bit, r.mean())
    for f in range(9):
        s = max(np.abs(rtau[f]).max(), 1e-30)
        err = np.abs(tau[f] - rtau[f])
        assert np.median(err) <= 1e-4 * s and err.max() <= 2e-2 * s, f
    gm.destroy()


def test_rounding_level_theta_gradient_falls_back_like_the_oracle():
    """Levels whose theta is 300 K up to the rounding of the closure's theta = (rho R_d T / C0)^(1/gamma)
    / rho: the heat flag is cleared on every one of their interior cells, on the GPU and in the
    oracle alike (the rounding-level gradient is below the fallback rule's noise threshold), while
    the momentum flags follow the resolved strain there."""
    cfg = small_config(nx=24, ny=20, nz=14, ord=5, tracers=[WV, TKE, P1, P2], dx=2.)
    st, imm = random_state(cfg, seed=23, amp_u=1.0, vapor=False)
    c = pa.physical_constants()
    rho = st[0]
    upper = np.arange(cfg["nz"]) >= 7
    st[4, upper] = c["C0"] * (rho[upper] * 300.) ** c["gamma_d"] / (rho[upper] * c["R_d"])
    P = np.zeros(o.NPARAM, np.float32)
    om = np.array([0, 0, -2e-3, 0, 0, 0, 0, 0, -1e-4], np.float32)
    gm, om_ = model_for(P, om, np.ones(9, np.float32))
    dt = 0.02
    _, tau, fl = run_learned(cfg, st, imm, gm, dt)
    _, rfl = o.les_apply_learned(cfg, st.copy(), imm, dt, om_)
    # levels 8 .. nz-2 have all 27 neighbours at theta = 300 up to rounding
    band = fl[8:-1]
    assert np.all((band & 2) == 0) and np.all((rfl[8:-1] & 2) == 0)
    assert (band & 1).mean() > 0.2                       # the momentum flags are not affected
    assert (fl == rfl).mean() >= 0.99
    gm.destroy()


def test_a_posteriori_learned_closure_tracks_the_plain_closure():
    """A posteriori on a small neutral urban ABL (48x48x32 at 2 m, 600 steps, against a 96x96x64 FP64
    truth at 1 m filtered onto the same grid; porturb_amd.sgs_aposteriori): the FP32 run with the
    shipped network keeps the mean-wind and resolved-TKE profiles within 2 % of the FP64 plain
    closure's, and its distance to the finer run is no more than 1.25x the plain closure's (the
    network is active on a sizeable share of the cells, but does not degrade the flow)."""
    from porturb_amd import sgs_aposteriori as ap
    rep = ap.run(48, 48, 32, 2.0, [300, 600], seed=1234)
    for ck in rep["checkpoints"]:
        print(ck["coarse_steps"], {k: ck[k] for k in ("plain", "learned", "learned_vs_plain", "learned_fraction")},
              flush=True)
        assert ck["learned_fraction"]["momentum"] > 0.05
        for p in ("u_mean", "tke_resolved"):
            assert ck["learned_vs_plain"][p] < 0.02, p
            assert ck["learned"][p] <= 1.25 * ck["plain"][p] + 0.005, p
