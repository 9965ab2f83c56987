"""Python view of the device-resident observation (include/rl2.h).

``DeviceObservation`` owns an ``rl2_obs`` handle: the observation in HBM plus facet slots.
Every method is a thin call into librl2hip.so; nothing is computed in Python.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

NUMR = 24


def r_grid():
    ro = np.zeros(40)
    n = N.hip().rl2_r_grid(N.dptr(ro))
    return ro[:n].copy()


def rprior_for(flux, ro):
    """distributions.c:59-87 prior on the scalelength grid."""
    H = N.hip()
    mu = H.scale_mean(flux)
    rp = np.zeros(len(ro))
    for k in range(1, len(ro)):
        rp[k] = H.rfunc(mu, 0.3136, ro[k])
    return rp


def fit_options(batch=1, do_refit=1, nm_speculate=1, line_chunk=4, lik_block=0,
                lik_precision=0, proposal=None, features_out=None, labels_out=None,
                proposal_mode=0, verbose=0):
    """rl2_fit_options; lik_block 0: tiled likelihood, else threads per (e, r) work-group;
    lik_precision 0: FP64, 1: mixed throughout, 2: mixed simplex + FP64 sampling, 3: FP64
    simplex + mixed sampling, 4: as 3 with FP64 verification; proposal: an rl2_proposal
    handle (proposal_mode 0: simplex start and speculation depth, 1: speculation depth
    only, results identical to the reference start); features_out / labels_out: float32
    (ngal, 770) / float64 (ngal, 3) arrays; verbose: a progress line on stderr every
    `verbose` batches."""
    return N.FitOptions(batch, do_refit, nm_speculate, line_chunk, lik_block, verbose, lik_precision,
                        proposal, None if features_out is None else N.fptr(features_out),
                        None if labels_out is None else N.dptr(labels_out), proposal_mode)


class DeviceObservation:
    def __init__(self, obs, precision=N.RL2_PREC_FP64, vis_sky=None, device=0):
        H = N.hip()
        N.check(H.rl2_init(device), "rl2_init")
        self.obs = obs
        self.precision = precision
        h = C.c_void_p()
        flag = np.ascontiguousarray(obs.flag, dtype=np.uint8)
        N.check(H.rl2_obs_create(C.byref(h), obs.nch, obs.nrow, N.dptr(obs.u), N.dptr(obs.v),
                                 N.dptr(obs.w), N.dptr(obs.wn), N.dptr(obs.spec),
                                 N.fptr(obs.sigma2), flag.ctypes.data_as(N.c_ubyte_p), obs.len,
                                 precision), "rl2_obs_create")
        self.h = h
        N.check(H.rl2_obs_set_vis(self.h, N.cplx(obs.vis), N.cplx(vis_sky)), "rl2_obs_set_vis")

    def close(self):
        if self.h:
            N.hip().rl2_obs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- residual
    def residual(self):
        out = np.zeros(self.obs.nvis, np.complex128)
        N.check(N.hip().rl2_obs_get_residual(self.h, N.cplx(out)), "rl2_obs_get_residual")
        return out

    def sky_model(self, cat):
        """visSkyMod += sum of round models (data_simulation.c:131-165) -> R -= ..."""
        N.check(N.hip().rl2_obs_add_sources(self.h, len(cat), N.dptr(cat.flux),
                                             N.dptr(cat.scale), N.dptr(cat.l), N.dptr(cat.m),
                                             None, None, -1.0), "rl2_obs_add_sources")

    def add_sources(self, cat, sign=1.0, round_models=False):
        N.check(N.hip().rl2_obs_add_sources(
            self.h, len(cat), N.dptr(cat.flux), N.dptr(cat.scale), N.dptr(cat.l), N.dptr(cat.m),
            None if round_models else N.dptr(cat.e1), None if round_models else N.dptr(cat.e2),
            sign), "rl2_obs_add_sources")

    def checkpoint(self):
        N.check(N.hip().rl2_obs_checkpoint(self.h), "rl2_obs_checkpoint")

    def restore(self):
        N.check(N.hip().rl2_obs_restore(self.h), "rl2_obs_restore")

    def loglik_precision(self, prec):
        N.check(N.hip().rl2_obs_loglik_precision(self.h, prec), "rl2_obs_loglik_precision")

    def add_noise(self, sigma2, seed):
        N.check(N.hip().rl2_obs_add_noise(self.h, float(sigma2), int(seed)), "rl2_obs_add_noise")

    # ---- one galaxy
    def extract(self, l0, m0, flux, R_mu, facet, slot=0, want_sum_w=True):
        H = N.hip()
        fmax = H.facet_size(3.5, self.obs.len)
        cap = fmax * fmax
        fv = np.zeros(cap, np.complex128)
        fs = np.zeros(cap)
        uu = np.zeros(cap)
        vv = np.zeros(cap)
        sw = np.zeros(facet * facet) if want_sum_w else None
        n = C.c_uint(0)
        N.check(H.rl2_obs_extract(self.h, slot, l0, m0, flux, R_mu, facet, N.cplx(fv), N.dptr(fs),
                                  N.dptr(uu), N.dptr(vv), N.dptr(sw), C.byref(n)),
                "rl2_obs_extract")
        n = n.value
        return fv[:n].copy(), fs[:n].copy(), uu[:n].copy(), vv[:n].copy(), sw

    def extract_only(self, l0, m0, flux, R_mu, facet, slot=0):
        """Extraction into a device facet slot without copying the facet back; returns n."""
        n = C.c_uint(0)
        N.check(N.hip().rl2_obs_extract(self.h, slot, l0, m0, flux, R_mu, facet, None, None,
                                        None, None, None, C.byref(n)), "rl2_obs_extract")
        return n.value

    def set_prior(self, rprior, slot=0):
        rp = np.ascontiguousarray(rprior, dtype=np.float64)
        N.check(N.hip().rl2_obs_set_prior(self.h, slot, N.dptr(rp)), "rl2_obs_set_prior")

    def load_facet(self, facet_vis, facet_sigma2, uu, vv, slot=0):
        N.check(N.hip().rl2_obs_load_facet(self.h, slot, len(uu), N.cplx(facet_vis),
                                           N.dptr(facet_sigma2), N.dptr(uu), N.dptr(vv)),
                "rl2_obs_load_facet")

    def loglik(self, e1, e2, slot=0, detail=False):
        e1 = np.ascontiguousarray(e1, dtype=np.float64)
        e2 = np.ascontiguousarray(e2, dtype=np.float64)
        n = len(e1)
        Le = np.zeros(n)
        err = np.zeros(n, np.int32)
        Lr = np.zeros(n * NUMR) if detail else None
        nit = np.zeros(n * NUMR, np.int32) if detail else None
        N.check(N.hip().rl2_obs_loglik(self.h, slot, n, N.dptr(e1), N.dptr(e2), N.dptr(Le),
                                       err.ctypes.data_as(N.c_int_p), N.dptr(Lr),
                                       None if nit is None else nit.ctypes.data_as(N.c_int_p)),
                "rl2_obs_loglik")
        if detail:
            return Le, err, Lr.reshape(n, NUMR), nit.reshape(n, NUMR)
        return Le, err

    def fit(self, slot=0, **opts):
        res = N.GalResult()
        o = fit_options(**opts)
        N.check(N.hip().rl2_obs_fit(self.h, slot, C.byref(o), C.byref(res)), "rl2_obs_fit")
        return res

    def subtract(self, l0, m0, flux, R_mu, e1, e2):
        N.check(N.hip().rl2_obs_subtract(self.h, l0, m0, flux, R_mu, e1, e2), "rl2_obs_subtract")

    # ---- the whole loop
    def process(self, cat, **opts):
        n = len(cat)
        res = (N.GalResult * n)()
        st = N.RunStats()
        o = fit_options(**opts)
        N.check(N.hip().rl2_process(self.h, n, N.dptr(cat.l), N.dptr(cat.m), N.dptr(cat.flux),
                                    N.dptr(cat.scale), C.byref(o), res, C.byref(st)),
                "rl2_process")
        return res, st


KNOB_NAMES = {N.RL2_TUNE_GRID_VARIANT: "grid_variant", N.RL2_TUNE_TILE_CELLS: "tile_cells",
              N.RL2_TUNE_TILE_STREAMS: "tile_streams", N.RL2_TUNE_BATCH_GROUPS: "batch_groups",
              N.RL2_TUNE_ITEM_GROUPS: "item_groups",
              -1: "lik_block"}


def autotune(dev, cat, **opts):
    """rl2_autotune over the galaxies of `cat` (first pass, the given fit options): sets the
    fastest launch knobs and returns them with every timed trial."""
    n = len(cat)
    trials = (N.TuneTrial * 64)()
    nt = C.c_int(0)
    lb = C.c_int(0)
    o = fit_options(**opts)
    N.check(N.hip().rl2_autotune(dev.h, n, N.dptr(cat.l), N.dptr(cat.m), N.dptr(cat.flux),
                                 N.dptr(cat.scale), C.byref(o), trials, 64, C.byref(nt),
                                 C.byref(lb)), "rl2_autotune")
    chosen = {"lik_block": lb.value}
    for k, name in KNOB_NAMES.items():
        if k >= 0:
            v = C.c_int()
            N.check(N.hip().rl2_tune_get(k, C.byref(v)), "rl2_tune_get")
            chosen[name] = v.value
    return {"chosen": chosen,
            "trials": [{"knob": KNOB_NAMES.get(t.knob, t.knob), "value": t.value,
                        "seconds": round(t.seconds, 5)} for t in trials[:min(nt.value, 64)]]}


def results_array(res):
    """GalResult ctypes array -> dict of numpy arrays."""
    names = [f[0] for f in N.GalResult._fields_]
    return {k: np.array([getattr(r, k) for r in res]) for k in names}


def stats_enable(on=True):
    N.check(N.hip().rl2_stats_enable(int(on)), "rl2_stats_enable")


def stats_reset():
    N.check(N.hip().rl2_stats_reset(), "rl2_stats_reset")


def stats():
    out = {}
    for k, name in enumerate(N.KERNEL_NAMES):
        s = N.KernelStat()
        N.check(N.hip().rl2_stats_get(k, C.byref(s)), "rl2_stats_get")
        out[name] = {"ms": s.ms, "launches": s.launches, "flops": s.flops, "bytes": s.bytes}
    return out
