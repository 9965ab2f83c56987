"""Learned simplex proposal: training data, training and evaluation.

The network (librl2hip.so, rl2_proposal_*) reads a galaxy's binned facet (16 radial x 16
azimuthal uv bins of re, im and weight, plus log flux and log scalelength) and predicts the
likelihood peak (e1, e2, log r) with a 3x3 Cholesky factor of its precision.  The fit uses
the peak as the simplex start and twice the predicted e width (clamped to [0.01, 0.1]) as the
first step, and keeps the reference's start (0, 0), step 0.1 whenever that start has the
higher likelihood (both are evaluated in the first launch).  The exact likelihood makes
every other decision.

Training data come from the method itself: observations simulated on the device with the
seeded catalog generator and noise, processed by the FP64 path (rl2_process), which writes
each galaxy's network inputs and its fitted (e1, e2) with the scalelength that maximises
L_r + log prior there.  Training runs the library's BF16 MFMA forward / backward kernels and
Adam on the GPU.

    python -m radiolensfit2_amd.proposal train --obs 24 --ngal 1000 --out weights.npy
    python -m radiolensfit2_amd.proposal evaluate --weights weights.npy --ngal 1000
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _native as N
from . import pipeline as P
from . import synth

N_IN, N_OUT = 770, 9


class Proposal:
    """Owns an rl2_proposal network on the device."""

    def __init__(self, seed=0):
        H = N.hip()
        N.check(H.rl2_init(0), "rl2_init")
        self.h = C.c_void_p()
        N.check(H.rl2_proposal_create(C.byref(self.h), int(seed)), "rl2_proposal_create")
        self.n_params = H.rl2_proposal_num_params()

    def close(self):
        if self.h:
            N.hip().rl2_proposal_destroy(self.h)
            self.h = None

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        N.check(N.hip().rl2_proposal_get_params(self.h, N.fptr(p)), "rl2_proposal_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        N.check(N.hip().rl2_proposal_set_params(self.h, N.fptr(p)), "rl2_proposal_set_params")

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((len(x), N_OUT), np.float32)
        N.check(N.hip().rl2_proposal_forward(self.h, len(x), N.fptr(x), N.fptr(out)),
                "rl2_proposal_forward")
        return out

    def train_step(self, x, y, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_float()
        N.check(N.hip().rl2_proposal_train_step(self.h, len(x), N.fptr(x), N.fptr(y),
                                                C.c_float(lr), C.byref(loss)),
                "rl2_proposal_train_step")
        return loss.value


def simulate(obs, cat, res_prec=N.RL2_PREC_FP64, noise_seed=12345):
    """Observation in HBM: sheared models + noise, round sky model subtracted."""
    H = N.hip()
    N.check(H.rl2_init(0), "rl2_init")
    h = C.c_void_p()
    N.check(H.rl2_obs_create(C.byref(h), obs.nch, obs.nrow, N.dptr(obs.u), N.dptr(obs.v),
                             N.dptr(obs.w), N.dptr(obs.wn), N.dptr(obs.spec),
                             N.fptr(obs.sigma2), None, obs.len, res_prec), "rl2_obs_create")
    dev = P.DeviceObservation.__new__(P.DeviceObservation)
    dev.obs, dev.precision, dev.h = obs, res_prec, h
    dev.add_sources(cat, 1.0)
    dev.add_noise(float(obs.sigma2[0]), noise_seed)
    dev.add_sources(cat, -1.0, round_models=True)
    return dev


def geometry(cfg):
    u, v, w, wn, spec, sigma2, flag, len_lambda = synth.make_geometry(cfg)
    return synth.Observation(cfg, u, v, w, wn, spec, sigma2, flag, len_lambda,
                             np.zeros(0, np.complex128))


def collect(cfg, n_obs, ngal, seed0=100, batch=64, log=print):
    """Network inputs and labels of n_obs simulated observations of ngal galaxies each."""
    obs = geometry(cfg)
    X, Y = [], []
    for s in range(n_obs):
        t0 = time.time()
        cat = synth.make_catalog(ngal, seed=seed0 + s, g=(0.0, 0.0))
        dev = simulate(obs, cat, noise_seed=1000 + seed0 + s)
        feats = np.zeros((ngal, N_IN), np.float32)
        labels = np.full((ngal, 3), np.nan)
        dev.process(cat, batch=batch, features_out=feats, labels_out=labels)
        dev.close()
        keep = np.isfinite(labels).all(1)
        X.append(feats[keep])
        Y.append(labels[keep].astype(np.float32))
        log(f"[proposal] observation {s + 1}/{n_obs}: {keep.sum()} labelled galaxies, "
            f"{time.time() - t0:.1f}s")
    return np.concatenate(X), np.concatenate(Y)


def train(X, Y, epochs=40, batch=256, lr=3e-4, seed=0, val_frac=0.1, patience=5, log=print):
    """Adam on the GPU with early stopping on a held-out tenth; returns (network with the
    best validation parameters, history of (epoch, train loss, val loss))."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(X))
    nval = max(1, int(len(X) * val_frac))
    va, tr = idx[:nval], idx[nval:]
    net = Proposal(seed)
    hist = []
    best, best_params, since = np.inf, net.params(), 0
    for ep in range(epochs):
        order = rng.permutation(tr)
        lr_ep = lr * (0.5 ** (ep // 10))
        losses = []
        for b in range(0, len(order) - batch + 1, batch):
            j = order[b:b + batch]
            losses.append(net.train_step(X[j], Y[j], lr_ep))
        vloss = validation_loss(net, X[va], Y[va])
        hist.append((ep, float(np.mean(losses)), vloss))
        log(f"[proposal] epoch {ep}: train {np.mean(losses):.4f} val {vloss:.4f}")
        if vloss < best:  # early stopping on the held-out galaxies
            best, best_params, since = vloss, net.params(), 0
        else:
            since += 1
            if since >= patience:
                break
    net.set_params(best_params)
    return net, hist


def validation_loss(net, X, Y):
    out = net.forward(X).astype(np.float64)
    d = Y - out[:, :3]
    l00, l11, l22 = np.exp(out[:, 3]), np.exp(out[:, 4]), np.exp(out[:, 5])
    z0 = l00 * d[:, 0] + out[:, 6] * d[:, 1] + out[:, 7] * d[:, 2]
    z1 = l11 * d[:, 1] + out[:, 8] * d[:, 2]
    z2 = l22 * d[:, 2]
    return float(np.mean(0.5 * (z0 ** 2 + z1 ** 2 + z2 ** 2) - out[:, 3:6].sum(1)))


def evaluate(cfg, net, ngal, seed=9000, batch=64, log=print):
    """Fresh galaxies through the whole loop with and without the proposal."""
    obs = geometry(cfg)
    cat = synth.make_catalog(ngal, seed=seed, g=(0.0, 0.0))
    out = {}
    for name, prop, mode in (("plain", None, 0), ("learned_start", net.h, 0),
                             ("learned_speculation", net.h, 1)):
        dev = simulate(obs, cat, noise_seed=seed + 1)
        t0 = time.perf_counter()
        res, st = dev.process(cat, batch=batch, proposal=prop, proposal_mode=mode)
        dt = time.perf_counter() - t0
        dev.close()
        r = P.results_array(res)
        out[name] = {"galaxies_per_s": ngal / dt, "evals_per_galaxy": float(r["n_evals"].mean()),
                     "rounds_per_galaxy": float(r["rounds"].mean()),
                     "proposal_start_fraction": float(r["start"].mean()),
                     "bad_final": st.n_bad_final, "e1": r["e1"], "e2": r["e2"],
                     "error": r["error"]}
        log(f"[proposal] {name}: {out[name]['galaxies_per_s']:.1f} gal/s, "
            f"{out[name]['evals_per_galaxy']:.1f} evaluations and "
            f"{out[name]['rounds_per_galaxy']:.1f} launches per galaxy, proposal start "
            f"{out[name]['proposal_start_fraction']:.3f}")
    a = out["plain"]
    summary = {k: {kk: vv for kk, vv in v.items() if kk not in ("e1", "e2", "error")}
               for k, v in out.items()}
    for k in ("learned_start", "learned_speculation"):
        b = out[k]
        de = np.maximum(np.abs(a["e1"] - b["e1"]), np.abs(a["e2"] - b["e2"]))
        ok = (de <= 1e-4) & (a["error"] == b["error"])
        summary[k]["within_1e-4_of_plain"] = float(ok.mean())
        summary[k]["median_abs_de_vs_plain"] = float(np.median(de))
        summary[k]["within_1e-2_of_plain"] = float(((de <= 1e-2) & (a["error"] == b["error"])).mean())
        # galaxies/s counting only galaxies inside the |de| <= 1e-4 gate
        summary[k]["galaxies_per_s_within_gate"] = summary[k]["galaxies_per_s"] * float(ok.mean())
    # galaxies whose fit kept the reference's start (mode 1 never uses the proposal start)
    summary["learned_start"]["fallback_fraction"] = \
        1.0 - summary["learned_start"]["proposal_start_fraction"]
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser(description="learned simplex proposal")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--obs", type=int, default=24)
    t.add_argument("--ngal", type=int, default=1000)
    t.add_argument("--epochs", type=int, default=40)
    t.add_argument("--out", required=True)
    t.add_argument("--eval-ngal", type=int, default=1000)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", required=True)
    e.add_argument("--ngal", type=int, default=1000)
    args = ap.parse_args(argv)
    cfg = synth.ARRAY_64

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.cmd == "train":
        X, Y = collect(cfg, args.obs, args.ngal, log=log)
        log(f"[proposal] {len(X)} training galaxies")
        net, hist = train(X, Y, epochs=args.epochs, log=log)
        net.save(args.out)
        summary = evaluate(cfg, net, args.eval_ngal, log=log)
        summary["training_galaxies"] = int(len(X))
        summary["history"] = hist
        print(json.dumps(summary), flush=True)
    else:
        net = Proposal(0)
        net.load(args.weights)
        print(json.dumps(evaluate(cfg, net, args.ngal, log=log)), flush=True)


if __name__ == "__main__":
    main()
