#!/usr/bin/env python3
"""Galaxies fitted per second on one MI355X.

Workload (default): 197 dishes (dense 1 km core, log-radial to 150 km), 19,306 baselines x
480 times x 16 channels of 15 MHz from 950 MHz = 148.3M visibilities with float64 uvw;
1,000 exponential galaxies (seeded catalog, shear g = (0.03, 0)), simulated on the device
(sheared models + Gaussian noise), sky model of round models subtracted.  The per-galaxy
path is RadioLensfit2's: extract -> fit (simplex + sampling) -> subtract, galaxies in
decreasing flux order, failed fits re-fitted after the others.

A step is one batch of --batch galaxies through that path.  --warmup steps run first
(untimed), then exactly --steps are timed (their bad-list re-fit included), inputs
resident in HBM.  One JSON line is printed.  A CPU baseline (the FP64 oracle, OpenMP over
channels / scalelengths, on the host cores) processes the first --cpu-sample galaxies of the
timed set from the same residual; their ellipticities also give a full-size parity figure.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from radiolensfit2_amd import _native as N  # noqa: E402
from radiolensfit2_amd import pipeline as P  # noqa: E402
from radiolensfit2_amd import synth  # noqa: E402

METRIC = ("galaxies/s @|Δe|≤1e-4 vs FP64: learned+batched vs plain GPU; HBM GB/s")
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def subset(cat, lo, hi):
    return synth.Catalog(*(np.ascontiguousarray(getattr(cat, f)[lo:hi]) for f in
                           ("flux", "scale", "e1", "e2", "l", "m", "snr")))


def make_workload(cfg, ngal, precision, seed=1, shear=(0.03, 0.0)):
    t0 = time.time()
    u, v, w, wn, spec, sigma2, flag, len_lambda = synth.make_geometry(cfg)
    obs = synth.Observation(cfg, u, v, w, wn, spec, sigma2, flag, len_lambda,
                            np.zeros(0, np.complex128))
    cat = synth.make_catalog(ngal, seed=seed, g=shear)
    log(f"[bench] geometry + catalog: {time.time() - t0:.1f}s nrow={obs.nrow} nvis={obs.nvis}")
    t0 = time.time()
    H = N.hip()
    N.check(H.rl2_init(0), "rl2_init")
    h = C.c_void_p()
    N.check(H.rl2_obs_create(C.byref(h), obs.nch, obs.nrow, N.dptr(u), N.dptr(v), N.dptr(w),
                             N.dptr(wn), N.dptr(spec), N.fptr(sigma2), None, len_lambda,
                             precision), "rl2_obs_create")
    dev = P.DeviceObservation.__new__(P.DeviceObservation)
    dev.obs, dev.precision, dev.h = obs, precision, h
    # data = sum of sheared models + noise; residual = data - sum of round models
    dev.add_sources(cat, 1.0)
    dev.add_noise(float(sigma2[0]), 12345)
    dev.add_sources(cat, -1.0, round_models=True)
    N.check(H.rl2_synchronize(), "sync")
    log(f"[bench] device simulation + sky model: {time.time() - t0:.1f}s")
    return obs, cat, dev


def cpu_baseline(obs, cat, resid, lo, n):
    """FP64 oracle on galaxies [lo, lo+n) from the residual the GPU started from."""
    from oracle import oracle as O
    L = O.lib()
    sub = subset(cat, lo, lo + n)
    vd = resid
    vs = np.zeros_like(vd)
    res = (O.GalResult * n)()
    st = O.RunStats()
    t0 = time.time()
    L.orl2_data_processing(n, 1, obs.len, obs.nch, O.ptr(obs.wn), O.ptr(obs.spec), obs.nrow,
                           O.ptr(obs.u), O.ptr(obs.v), O.ptr(obs.w),
                           O.ptr(obs.sigma2, O.c_float_p), None, O.cptr(vd), O.cptr(vs),
                           O.ptr(sub.l), O.ptr(sub.m), O.ptr(sub.flux), O.ptr(sub.scale), res, 1,
                           C.byref(st))
    dt = time.time() - t0
    return dt, [(r.e1, r.e2, r.error) for r in res]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--config", choices=["197dish", "64dish"], default="197dish")
    ap.add_argument("--ngal", type=int, default=1000)
    ap.add_argument("--cpu-sample", type=int, default=1,
                    help="galaxies of the timed set re-run on the CPU oracle (0: skip)")
    ap.add_argument("--no-speculate", action="store_true")
    ap.add_argument("--line-chunk", type=int, default=4)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}); the multi-GPU sharding is a design, not a build.", file=sys.stderr)
        sys.exit(2)

    cfg = synth.ARRAY_197 if args.config == "197dish" else synth.ARRAY_64
    prec = N.RL2_PREC_FP64 if args.precision == "fp64" else N.RL2_PREC_MIXED
    B = args.batch
    need = (args.warmup + args.steps) * B
    if need > args.ngal:
        sys.exit(f"warmup+steps need {need} galaxies, catalog has {args.ngal}")
    obs, cat, dev = make_workload(cfg, args.ngal, prec)
    opts = dict(batch=B, do_refit=1, nm_speculate=0 if args.no_speculate else 1,
                line_chunk=args.line_chunk)

    if args.warmup:
        t0 = time.time()
        dev.process(subset(cat, 0, args.warmup * B), **opts)
        log(f"[bench] warmup {args.warmup} steps: {time.time() - t0:.2f}s")
    lo, hi = args.warmup * B, need
    resid = dev.residual() if args.cpu_sample > 0 else None
    timed = subset(cat, lo, hi)
    P.stats_enable(True)
    P.stats_reset()
    N.check(N.hip().rl2_synchronize(), "sync")
    t0 = time.perf_counter()
    res, st = dev.process(timed, **opts)
    N.check(N.hip().rl2_synchronize(), "sync")
    elapsed = time.perf_counter() - t0
    kstats = P.stats()
    P.stats_enable(False)
    n_gal = hi - lo
    value = n_gal / elapsed
    log(f"[bench] timed {args.steps} steps ({n_gal} galaxies): {elapsed:.3f}s -> {value:.2f} gal/s;"
        f" bad first/final {st.n_bad_first}/{st.n_bad_final}")
    for k, s in kstats.items():
        log(f"[bench]   {k:10s} {s['ms']:10.2f} ms  {s['launches']:7d} launches")

    # rooflines from HIP-event kernel times measured on the library's stream
    lik = kstats["likelihood"]
    ext = kstats["extract"]
    peak_valu = PEAK_FP64_TFLOPS if prec == N.RL2_PREC_FP64 else PEAK_FP32_TFLOPS
    lik_tf = lik["flops"] / (lik["ms"] * 1e-3) / 1e12 if lik["ms"] > 0 else 0.0
    ext_gbs = ext["bytes"] / (ext["ms"] * 1e-3) / 1e9 if ext["ms"] > 0 else 0.0
    roof_lik = {"kernel": "likelihood_kernel", "bound": "valu", "achieved": round(lik_tf, 4),
                "peak": peak_valu, "unit": "TFLOP/s", "frac": round(lik_tf / peak_valu, 5),
                "traffic": None, "launches": lik["launches"],
                "avg_launch_us": round(1e3 * lik["ms"] / max(1, lik["launches"]), 3)}
    roof_ext = {"kernel": "grid_kernel (extraction)", "bound": "hbm", "achieved": round(ext_gbs, 2),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(ext_gbs / PEAK_HBM_GBS, 5),
                "traffic": None, "launches": ext["launches"],
                "avg_launch_us": round(1e3 * ext["ms"] / max(1, ext["launches"]), 3)}
    dominant = roof_lik if lik["ms"] >= ext["ms"] else roof_ext

    cpu = None
    parity = None
    if args.cpu_sample > 0:
        ncpu = min(args.cpu_sample, n_gal)
        dt, ores = cpu_baseline(obs, cat, resid, lo, ncpu)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(ncpu / dt, 5), "unit": "galaxies/s", "cores": cores, "kind": "port",
               "sample": f"{ncpu} galaxies (the first of the timed set: extraction, fit, "
                         f"subtraction, bad-list re-fit) on the FP64 oracle, {dt:.1f}s"}
        de = [max(abs(res[k].e1 - ores[k][0]), abs(res[k].e2 - ores[k][1])) for k in range(ncpu)
              if res[k].error == ores[k][2] == 0]
        parity = {"galaxies": ncpu, "max_abs_de_vs_fp64_oracle": (max(de) if de else None),
                  "error_flags_equal": all(res[k].error == ores[k][2] for k in range(ncpu))}
        log(f"[bench] cpu baseline {cpu['value']} gal/s; parity {parity}")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "galaxies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == N.RL2_PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{cfg.n_dish} dishes, {obs.nrow // cfg.n_times} baselines x "
                         f"{cfg.n_times} times x {cfg.n_chan} ch ({obs.nvis / 1e6:.1f}M vis), "
                         f"{args.ngal}-galaxy catalog; timed galaxies {lo}..{hi - 1}"),
            "batch": B,
            "path": ("plain" if B == 1 else "batched") + "-" + args.precision,
            "parallelism": "1 GPU",
        },
        "roofline": dominant,
        "roofline_likelihood": roof_lik,
        "roofline_extraction": roof_ext,
        "cpu_baseline": cpu,
        "parity": parity,
        "bad_fits": {"first_pass": st.n_bad_first, "after_refit": st.n_bad_final},
    }
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
