"""Training of the learned subgrid closure on one GPU, with the library's own kernels.

Pipeline (no counterpart in the reference; its nearest prior art is the explicit-vs-closure SGS
comparison of build/postproc/compare_les.py:11-60):

1. a fine run of the neutral urban ABL case (dx_fine = dx / r) through the C++ driver;
2. ``pu_sgs_filter``: box filter by r onto the deployment grid -> filtered (rho, u, v, w, theta,
   k_sgs) with halo, the block-max immersed proportion and the subgrid stresses/heat fluxes
   tau_ij = <u_i u_j> - <u_i><u_j>, q_j = <u_j theta> - <u_j><theta>;
3. training cells = interior levels whose 3x3x3 neighbourhood holds no immersed cell (the cells the
   closure would hand to the network); input scales 1/sigma of each raw input (difference / delta)
   and target mean/scale from those cells; ``pu_sgs_features`` builds the BF16 inputs on the device;
4. ``pu_mlp_train_step``: Adam on mini-batches (BF16 matrix-core forward/backward, FP32 master
   weights), learning rate cosine-decayed;
5. a priori evaluation on held-out cells against the eddy-viscosity closure's own stress model
   tau_ij = 2/3 k delta_ij - 2 nu_t S_ij, nu_t = 0.1 delta sqrt(k), q_j = -(nu_t/Pr_t) dtheta/dx_j,
   Pr_t = delta / (1 + 2 delta) (les_closure.h:123-127 with ell = delta), given the true filtered
   subgrid energy k -- and a least-squares rescaled variant, its best case.

    python -m porturb_amd.sgs_training --nx 1024 --ny 1024 --nz 256 --dx 1.0 --r 2 \\
        --snapshots 1500 2500 --iters 20000 --batch 8192 --out porturb_amd/data/sgs_abl_urban_dx2.bin
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import porturb_amd as pa  # noqa: E402

NIN, NOUT = pa.SGS_NIN, pa.SGS_NOUT
OFFSETS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def log(msg):
    print("sgs_training: " + msg, file=sys.stderr, flush=True)


class Snapshot:
    """The filtered coarse data of one fine-run state (device + host copies)."""

    def __init__(self, driver, nx, ny, nz, dx_fine, r):
        cs = pa.PuCouplerState()
        n = pa.C.c_int()
        pa.check(pa.lib().pu_driver_state(driver.d, pa.C.byref(cs), pa.C.byref(n)))
        cfg = dict(nx=nx, ny=ny, nz=nz)
        self.les, self.imm_d, self.tg = pa.sgs_filter(driver.ctx, cs, cfg, r)
        self.nx, self.ny, self.nz = nx // r, ny // r, nz // r
        self.delta = dx_fine * r
        nh = (self.nx + 2) * (self.ny + 2) * (self.nz + 2)
        self.les_h = self.les.download(np.float64, (6, self.nz + 2, self.ny + 2, self.nx + 2))
        self.imm_h = self.imm_d.download(np.float64, (self.nz + 2, self.ny + 2, self.nx + 2))
        self.tg_h = self.tg.download(np.float32, (9, self.nz, self.ny, self.nx))
        assert self.les_h.size == 6 * nh

    def valid_cells(self):
        """Interior levels 1..nz-2 with no immersed cell in the 3x3x3 neighbourhood."""
        im = self.imm_h > 0
        nz, ny, nx = self.nz, self.ny, self.nx
        bad = np.zeros((nz, ny, nx), bool)
        for dz, dy, dx in OFFSETS:
            bad |= im[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        bad[0] = True
        bad[-1] = True
        return np.flatnonzero(~bad.ravel()).astype(np.int64)

    def raw_inputs(self, cells, q, off):
        """(neighbour - centre) of quantity q at one offset, for the given cells (double)."""
        k, rem = np.divmod(cells, self.nx * self.ny)
        j, i = np.divmod(rem, self.nx)
        F = self.les_h[1 + q]
        dz, dy, dx = off
        return F[k + 1 + dz, j + 1 + dy, i + 1 + dx] - F[k + 1, j + 1, i + 1]

    def targets(self, cells):
        return self.tg_h.reshape(9, -1)[:, cells].T.copy()      # [n][9]

    def strain_and_gradients(self, cells):
        """Centred-difference resolved strain S_ij (6) and dtheta/dx_j (3) at the cells."""
        h = 2. * self.delta
        d = {}
        for q, name in enumerate("uvwt"):
            for a, off in enumerate([(0, 0, 1), (0, 1, 0), (1, 0, 0)]):
                m = (-off[0], -off[1], -off[2])
                d[name + "xyz"[a]] = (self.raw_inputs(cells, q, off) - self.raw_inputs(cells, q, m)) / h
        S = np.stack([d["ux"], 0.5 * (d["uy"] + d["vx"]), 0.5 * (d["uz"] + d["wx"]), d["vy"],
                      0.5 * (d["vz"] + d["wy"]), d["wz"]], axis=1)
        G = np.stack([d["tx"], d["ty"], d["tz"]], axis=1)
        return S, G

    def ksgs(self, cells):
        k, rem = np.divmod(cells, self.nx * self.ny)
        j, i = np.divmod(rem, self.nx)
        return np.maximum(self.les_h[5][k + 1, j + 1, i + 1], 0.)

    def free(self):
        for b in (self.les, self.imm_d, self.tg):
            b.free()


def input_scales(snaps, cells_list, max_cells=400000, seed=0):
    """feat_scale = 1/sigma of each raw input (difference / delta) over the training cells."""
    rng = np.random.default_rng(seed)
    fs = np.zeros(NIN, np.float32)
    for n, off in enumerate(OFFSETS):
        for q in range(4):
            vals = []
            for s, cells in zip(snaps, cells_list):
                sub = cells if cells.size <= max_cells else rng.choice(cells, max_cells, replace=False)
                vals.append(s.raw_inputs(sub, q, off) / s.delta)
            sd = float(np.concatenate(vals).std())
            fs[4 * n + q] = 1. / sd if sd > 0 else 0.
    return fs


def fmul_of(feat_scale, delta):
    """The per-input multipliers the kernels use: float32(feat_scale / delta), rounded once."""
    return (np.asarray(feat_scale, np.float32).astype(np.float64) / delta).astype(np.float32)


def max_abs_input(snap, cells, fmul, max_cells=2000000, seed=0):
    """Largest |normalised input| over the cells (a random subset of at most max_cells)."""
    if cells.size > max_cells:
        cells = np.random.default_rng(seed).choice(cells, max_cells, replace=False)
    m = 0.
    for n, off in enumerate(OFFSETS):
        for q in range(4):
            x = snap.raw_inputs(cells, q, off).astype(np.float32) * fmul[4 * n + q]
            if x.size:
                m = max(m, float(np.abs(x).max()))
    return m


def init_params(seed=0):
    """He-scaled normal weights for the ReLU layers, Xavier for the output layer, zero biases."""
    rng = np.random.default_rng(seed)
    H = pa.SGS_NH
    parts = [rng.standard_normal((H, NIN)) * math.sqrt(2. / NIN), np.zeros(H),
             rng.standard_normal((H, H)) * math.sqrt(2. / H), np.zeros(H),
             rng.standard_normal((NOUT, H)) * math.sqrt(1. / H), np.zeros(NOUT)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def eddy_viscosity_tau(snap, cells):
    """The TKE closure's stress model at the cells, given the true filtered subgrid energy."""
    S, G = snap.strain_and_gradients(cells)
    k = snap.ksgs(cells)
    d = snap.delta
    nut = 0.1 * d * np.sqrt(k)
    prt = d / (1 + 2 * d)
    iso = np.array([1, 0, 0, 1, 0, 1], np.float64)
    tau = (2. / 3.) * k[:, None] * iso[None, :] - 2. * nut[:, None] * S
    q = -(nut / prt)[:, None] * G
    return np.concatenate([tau, q], axis=1), S, G, k


def nmse(pred, truth):
    """Per-output sum (pred - truth)^2 / sum (truth - mean)^2."""
    var = ((truth - truth.mean(axis=0)) ** 2).sum(axis=0)
    return ((pred - truth) ** 2).sum(axis=0) / np.maximum(var, 1e-300)


def best_scaled_ev(snap, cells, truth):
    """Eddy viscosity with the least-squares best coefficient for the deviatoric stress and the
    heat flux (its best case on this data)."""
    ev, S, G, k = eddy_viscosity_tau(snap, cells)
    d = snap.delta
    base = -2. * d * np.sqrt(k)[:, None] * S            # deviatoric model direction
    iso = np.array([1, 0, 0, 1, 0, 1], np.float64)
    tr = (truth[:, 0] + truth[:, 3] + truth[:, 5]) / 3.
    dev = truth[:, :6] - tr[:, None] * iso[None, :]
    c = float((base * dev).sum() / max((base * base).sum(), 1e-300))
    hb = -d * np.sqrt(k)[:, None] * G
    ch = float((hb * truth[:, 6:]).sum() / max((hb * hb).sum(), 1e-300))
    pred = np.concatenate([c * base + (2. / 3.) * k[:, None] * iso[None, :], ch * hb], axis=1)
    return pred, c, ch


class Trainer:
    """Training inputs on the device and the Adam loop."""

    def __init__(self, mlp, snaps, cells_list, feat_scale, out_mean, out_scale):
        self.mlp = mlp
        self.n = int(sum(c.size for c in cells_list))
        Xs, Ts = [], []
        for s, cells in zip(snaps, cells_list):
            Xs.append(pa.sgs_features(s.les, s.nx, s.ny, s.nz, fmul_of(feat_scale, s.delta), cells))
            Ts.append(((s.targets(cells) - out_mean) / out_scale).astype(np.float32))
        # one contiguous input/target buffer
        self.X = pa.DeviceBuffer(2 * pa.SGS_NINP * self.n)
        off = 0
        for X in Xs:
            pa.check(pa.lib().pu_memcpy_d2d(self.X.ptr + off, X.ptr, X.nbytes))
            off += X.nbytes
            X.free()
        self.T = pa.DeviceBuffer(4 * NOUT * self.n)
        self.T.upload(np.concatenate(Ts))

    def train(self, iters, batch=4096, lr=2e-3, lr_min=1e-4, seed=1, log_every=500):
        """Adam over shuffled mini-batches (epochs of random permutations), cosine learning rate."""
        rng = np.random.default_rng(seed)
        chunk = max(1, min(iters, (8 << 20) // (4 * batch)))    # iterations per index upload
        idx = pa.DeviceBuffer(4 * batch * chunk)
        pending = np.empty(0, np.int64)
        hist = []
        t0 = time.time()
        for it in range(iters):
            if it % chunk == 0:
                need = batch * chunk
                while pending.size < need:
                    pending = np.concatenate([pending, rng.permutation(self.n)])
                idx.upload(pending[:need].astype(np.int32))
                pending = pending[need:]
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / max(iters, 1)))
            want = (it % log_every == 0) or it == iters - 1
            ptr = _Ptr(idx.ptr + 4 * batch * (it % chunk))
            loss = self.mlp.train_step(self.X, self.T, ptr, batch, lr_t, want_loss=want)
            if want:
                hist.append((it, loss))
                log("iter %d loss %.5f lr %.2e (%.1f s)" % (it, loss, lr_t, time.time() - t0))
        idx.free()
        return hist

    def free(self):
        self.X.free()
        self.T.free()


class _Ptr:
    def __init__(self, ptr):
        self.ptr = ptr


def predict(mlp, snap, cells, feat_scale, out_mean, out_scale):
    X = pa.sgs_features(snap.les, snap.nx, snap.ny, snap.nz, fmul_of(feat_scale, snap.delta), cells)
    Y = mlp.forward(X, cells.size)
    X.free()
    return Y.astype(np.float64) * out_scale + out_mean


def fine_driver(nx, ny, nz, dx, seed, precision):
    return pa.Driver(pa.PU_CASE_ABL_URBAN, nx, ny, nz, dx, ord=9, precision=precision, seed=seed, closure=True)


def split_cells(snap, cells, frac_train=0.8):
    """Train on y < frac*ny, evaluate on the rest (spatially disjoint)."""
    j = (cells // snap.nx) % snap.ny
    cut = int(frac_train * snap.ny)
    return cells[j < cut], cells[j >= cut + 1]


def run(nx, ny, nz, dx, r, snapshots, iters, batch, lr, seed=1234, precision=pa.PU_FP32, out=None, report=None):
    drv = fine_driver(nx, ny, nz, dx, seed, precision)
    snaps, step = [], 0
    for target in sorted(snapshots):
        while step < target:
            n = min(50, target - step)
            drv.step(n)
            step += n
            drv.synchronize()
            log("fine run step %d" % step)
        snaps.append(Snapshot(drv, nx, ny, nz, dx, r))
        log("snapshot at step %d filtered" % step)
    drv.destroy()
    train_cells, test_cells = [], []
    for s in snaps:
        a, b = split_cells(s, s.valid_cells())
        train_cells.append(a)
        test_cells.append(b)
    feat_scale = input_scales(snaps, train_cells)
    T_all = np.concatenate([s.targets(c) for s, c in zip(snaps, train_cells)]).astype(np.float64)
    out_mean = T_all.mean(axis=0).astype(np.float32)
    out_scale = np.maximum(T_all.std(axis=0), 1e-30).astype(np.float32)
    # inputs beyond the training range (with a 25 % margin over the sampled maximum) fall back
    xmax = 1.25 * max(max_abs_input(s, c, fmul_of(feat_scale, s.delta)) for s, c in zip(snaps, train_cells))
    log("training cells %d, held-out %d, xmax %.2f" % (sum(c.size for c in train_cells),
                                                       sum(c.size for c in test_cells), xmax))
    mlp = pa.Mlp(init_params(seed))
    mlp.set_normalisation(feat_scale, xmax, out_mean, out_scale)
    tr = Trainer(mlp, snaps, train_cells, feat_scale, out_mean, out_scale)
    hist = tr.train(iters, batch=batch, lr=lr, seed=seed)
    tr.free()
    # a priori evaluation on the held-out cells
    preds, truths, evs, evbs = [], [], [], []
    for s, c in zip(snaps, test_cells):
        truth = s.targets(c).astype(np.float64)
        preds.append(predict(mlp, s, c, feat_scale, out_mean, out_scale))
        truths.append(truth)
        evs.append(eddy_viscosity_tau(s, c)[0])
        evbs.append(best_scaled_ev(s, c, truth)[0])
    truth = np.concatenate(truths)
    res = {"fine_grid": [nx, ny, nz], "dx_fine": dx, "r": r, "snapshots": sorted(snapshots), "iters": iters,
           "batch": batch, "train_cells": int(sum(c.size for c in train_cells)), "test_cells": int(truth.shape[0]),
           "loss_history": hist,
           "nmse_network": nmse(np.concatenate(preds), truth).tolist(),
           "nmse_eddy_viscosity": nmse(np.concatenate(evs), truth).tolist(),
           "nmse_eddy_viscosity_best_scaled": nmse(np.concatenate(evbs), truth).tolist()}
    for k in ("nmse_network", "nmse_eddy_viscosity", "nmse_eddy_viscosity_best_scaled"):
        res[k + "_mean"] = float(np.mean(res[k]))
    log("held-out NMSE: network %.3f, eddy viscosity %.3f, best-scaled eddy viscosity %.3f" %
        (res["nmse_network_mean"], res["nmse_eddy_viscosity_mean"], res["nmse_eddy_viscosity_best_scaled_mean"]))
    if out:
        mlp.save(out)
        log("saved " + out)
    if report:
        with open(report, "w") as f:
            json.dump(res, f, indent=1)
    for s in snaps:
        s.free()
    return mlp, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--dx", type=float, default=1.0)
    ap.add_argument("--r", type=int, default=2)
    ap.add_argument("--snapshots", type=int, nargs="+", default=[200, 400])
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    _, res = run(args.nx, args.ny, args.nz, args.dx, args.r, args.snapshots, args.iters, args.batch, args.lr,
                 seed=args.seed, out=args.out, report=args.report)
    print(json.dumps({k: v for k, v in res.items() if k != "loss_history"}))


if __name__ == "__main__":
    main()
