"""ctypes bindings for the FP64 CPU oracle (liborl2.so) and the reference gridding build.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, which compare the HIP path against it.  The product (radiolensfit2_amd)
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborl2.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_uvgrid.so")

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_bool_p = C.POINTER(C.c_bool)
c_int_p = C.POINTER(C.c_int)


class Complexd(C.Structure):
    _fields_ = [("real", C.c_double), ("imag", C.c_double)]


class Params(C.Structure):
    """likelihood_params, FACET layout (datatype.h:49-74)."""
    _fields_ = [
        ("numr", C.c_int), ("ro", c_double_p), ("rprior", c_double_p),
        ("nchannels", C.c_uint), ("band_factor", C.c_double), ("acc_time", C.c_double),
        ("spec", c_double_p), ("wavenumbers", c_double_p), ("uu", c_double_p),
        ("vv", c_double_p), ("l0", C.c_double), ("m0", C.c_double),
        ("data", C.POINTER(Complexd)), ("sigma2", c_double_p), ("mod", c_double_p),
        ("facet", C.c_uint), ("ncoords", C.c_uint),
    ]


class GalResult(C.Structure):
    _fields_ = [
        ("e1", C.c_double), ("e2", C.c_double), ("var_e1", C.c_double), ("var_e2", C.c_double),
        ("oneDimvar", C.c_double), ("maxL", C.c_double), ("error", C.c_int), ("pass_", C.c_int),
        ("facet", C.c_uint), ("ncoords", C.c_uint), ("n_evals", C.c_int), ("subtracted", C.c_int),
    ]


class RunStats(C.Structure):
    _fields_ = [("extraction_s", C.c_double), ("fitting_s", C.c_double),
                ("subtraction_s", C.c_double), ("n_bad_first", C.c_int),
                ("n_bad_final", C.c_int)]


class FitTrace(C.Structure):
    _fields_ = [("nm_iter", C.c_int), ("n_evals", C.c_int), ("nm_evals", C.c_int),
                ("nm_e1", C.c_double), ("nm_e2", C.c_double), ("nm_size", C.c_double)]


LOGLIK_CB = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, c_int_p, C.c_void_p)
F2_CB = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_void_p)
F1_CB = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


def ptr(a, t=c_double_p):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def cptr(a):
    """complex128 numpy array -> Complexd*"""
    if a is None:
        return None
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(Complexd))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        sig = {
            "orl2_channel_constants": (None, [C.c_uint, C.c_double, C.c_double, c_double_p, c_double_p]),
            "orl2_default_sigma2": (C.c_float, [C.c_int, C.c_double]),
            "orl2_len_lambda": (C.c_double, [c_double_p, c_double_p, C.c_ulong, C.c_double]),
            "orl2_r_grid": (C.c_int, [c_double_p]),
            "orl2_scale_mean": (C.c_double, [C.c_double]),
            "orl2_rfunc": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "orl2_sersic_shape": (C.c_double, [C.c_double] * 4),
            "orl2_model_at_zero": (None, [C.c_double, C.c_double, C.c_double, C.c_uint, c_double_p,
                                          c_double_p, c_double_p, c_double_p]),
            "orl2_data_galaxy_vis": (None, [C.c_double] * 8 + [C.c_ulong, c_double_p, c_double_p,
                                                                c_double_p, C.POINTER(Complexd)]),
            "orl2_phase_shift": (None, [C.c_double, C.c_double, C.c_double, C.c_ulong, c_double_p,
                                        c_double_p, c_double_p, C.POINTER(Complexd)]),
            "orl2_facet_size": (C.c_uint, [C.c_double, C.c_double]),
            "orl2_uv_grid_size": (C.c_uint, [C.c_double, c_double_p, C.c_uint, C.c_ulong, c_double_p,
                                             c_double_p, C.c_uint, c_bool_p]),
            "orl2_facet_coords": (C.c_uint, [c_double_p, c_double_p, C.c_double, C.c_uint, c_double_p]),
            "orl2_gridding": (None, [c_double_p, C.c_uint, C.c_ulong, c_double_p, c_double_p,
                                     C.POINTER(Complexd), c_float_p, C.c_double, C.c_uint,
                                     C.POINTER(Complexd), c_double_p, c_bool_p, c_double_p]),
            "orl2_average_facets": (None, [C.c_ulong, C.POINTER(Complexd), c_double_p, c_double_p]),
            "orl2_source_extraction": (None, [C.c_uint, C.c_uint, c_double_p, c_double_p,
                                              C.POINTER(Complexd), c_double_p, c_double_p,
                                              C.c_double, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.POINTER(Complexd),
                                              C.POINTER(Complexd), C.POINTER(Complexd), c_float_p,
                                              c_bool_p, C.c_ulong, c_double_p, c_double_p,
                                              c_double_p, C.c_double]),
            "orl2_expint_ei": (C.c_double, [C.c_double]),
            "orl2_marg_position": (C.c_double, [C.c_double]),
            "orl2_cross_correlation": (C.c_int, [C.c_ulong, c_double_p, c_double_p,
                                                 C.POINTER(Complexd), c_double_p, c_double_p,
                                                 c_double_p, c_int_p]),
            "orl2_loglikelihood_r": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_ulong,
                                                  c_double_p, c_double_p, c_double_p,
                                                  C.POINTER(Complexd), c_double_p, c_int_p]),
            "orl2_loglikelihood": (C.c_double, [C.POINTER(Params), C.c_double, C.c_double, c_int_p]),
            "orl2_loglikelihood_detail": (C.c_double, [C.POINTER(Params), C.c_double, C.c_double,
                                                       c_int_p, c_double_p, c_int_p]),
            "orl2_likelihood_sampling": (None, [c_double_p, c_double_p, C.c_double, C.POINTER(Params),
                                                C.c_int, c_double_p, c_double_p, c_double_p]),
            "orl2_set_posterior_values": (None, [C.c_int, c_double_p, c_double_p, c_double_p,
                                                 c_double_p, c_double_p, c_int_p]),
            "orl2_marginalise_posterior_r": (C.c_double, [C.c_int, c_double_p, c_double_p]),
            "orl2_cspline_eval_many": (C.c_int, [C.c_int, c_double_p, c_double_p, C.c_int,
                                                 c_double_p, c_double_p, C.c_int]),
            "orl2_qk41": (None, [F1_CB, C.c_void_p, C.c_double, C.c_double, c_double_p, c_double_p,
                                 c_double_p, c_double_p]),
            "orl2_qag41": (C.c_int, [F1_CB, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_ulong, c_double_p, c_double_p,
                                     C.POINTER(C.c_ulong)]),
            "orl2_source_fitting": (C.c_int, [C.POINTER(Params)] + [c_double_p] * 6),
            "orl2_source_fitting_cb": (C.c_int, [LOGLIK_CB, C.c_void_p, C.c_double, C.c_double,
                                                 C.c_double] + [c_double_p] * 6
                                       + [C.POINTER(FitTrace)]),
            "orl2_likelihood_sampling_cb": (None, [LOGLIK_CB, C.c_void_p, c_double_p, c_double_p,
                                                   C.c_double, C.c_int, c_double_p, c_double_p,
                                                   c_double_p, c_int_p]),
            "orl2_nmsimplex2_minimize": (C.c_int, [F2_CB, C.c_void_p, c_double_p, c_double_p,
                                                   C.c_double, C.c_int, c_double_p, c_double_p,
                                                   c_double_p, c_double_p]),
            "orl2_sky_model": (None, [c_double_p, c_double_p, C.c_uint, C.c_uint, c_double_p,
                                      c_double_p, c_double_p, c_double_p, C.c_ulong, c_double_p,
                                      c_double_p, c_double_p, C.POINTER(Complexd),
                                      C.POINTER(Complexd)]),
            "orl2_data_processing": (C.c_int, [C.c_uint, C.c_int, C.c_double, C.c_uint, c_double_p,
                                               c_double_p, C.c_ulong, c_double_p, c_double_p,
                                               c_double_p, c_float_p, c_bool_p,
                                               C.POINTER(Complexd), C.POINTER(Complexd),
                                               c_double_p, c_double_p, c_double_p, c_double_p,
                                               C.POINTER(GalResult), C.c_int,
                                               C.POINTER(RunStats)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref():
    """The reference's own evaluate_uv_grid.c, compiled in place (None if unavailable)."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.facet_size.restype = C.c_uint
        R.facet_size.argtypes = [C.c_double, C.c_double]
        R.evaluate_uv_grid_size.restype = C.c_uint
        R.evaluate_uv_grid_size.argtypes = [C.c_int, C.c_int, C.c_double, c_double_p, C.c_uint,
                                            C.c_uint, c_double_p, c_double_p, C.c_uint, c_bool_p]
        R.evaluate_facet_coords.restype = C.c_uint
        R.evaluate_facet_coords.argtypes = [c_double_p, c_double_p, C.c_double, C.c_uint, c_double_p]
        R.gridding_visibilities.restype = None
        R.gridding_visibilities.argtypes = [c_double_p, C.c_uint, C.c_uint, c_double_p, c_double_p,
                                            C.POINTER(Complexd), c_float_p, C.c_double, C.c_uint,
                                            C.POINTER(Complexd), c_double_p, c_bool_p, c_double_p]
        R.average_facets.restype = None
        R.average_facets.argtypes = [C.c_ulong, C.POINTER(Complexd), c_double_p, c_double_p]
        _ref = R
    return _ref


# ---------------------------------------------------------------- convenience wrappers

def r_grid():
    ro = np.zeros(40)
    n = lib().orl2_r_grid(ptr(ro))
    return ro[:n].copy(), n


def channel_constants(nch, freq_start, bw):
    wn = np.zeros(nch)
    sp = np.zeros(nch)
    lib().orl2_channel_constants(nch, freq_start, bw, ptr(wn), ptr(sp))
    return wn, sp


def make_params(facet_vis, facet_sigma2, uu, vv, ro, rprior, numr, facet=0):
    """Build a Params struct over numpy buffers; returns (params, keepalive)."""
    n = len(uu)
    mod = np.zeros(max(1, (numr - 1) * n))
    p = Params()
    p.numr = numr
    p.ro = ptr(ro)
    p.rprior = ptr(rprior)
    p.nchannels = 0
    p.uu = ptr(uu)
    p.vv = ptr(vv)
    p.data = cptr(facet_vis)
    p.sigma2 = ptr(facet_sigma2)
    p.mod = ptr(mod)
    p.facet = facet
    p.ncoords = n
    return p, (facet_vis, facet_sigma2, uu, vv, ro, rprior, mod)


def rprior_for(flux, ro, numr):
    L = lib()
    mu = L.orl2_scale_mean(flux)
    rp = np.zeros(40)
    for k in range(1, numr):
        rp[k] = L.orl2_rfunc(mu, 0.3136, ro[k])
    return rp


def extract(obs, gal_l, gal_m, flux, scale, facet, vis_data, vis_sky):
    """Run the oracle's source_extraction + evaluate_facet_coords for one galaxy."""
    L = lib()
    ncells = facet * facet
    fv = np.zeros(ncells, np.complex128)
    fs = np.zeros(ncells)
    sw = np.zeros(ncells)
    gal = np.zeros(obs.nvis, np.complex128)
    L.orl2_source_extraction(facet, obs.nch, ptr(obs.wn), ptr(obs.spec), cptr(fv), ptr(fs), ptr(sw),
                             gal_l, gal_m, flux, scale, 0.0, 0.0, cptr(vis_sky), cptr(vis_data),
                             cptr(gal), ptr(obs.sigma2, c_float_p), ptr(obs.flag, c_bool_p),
                             obs.nrow, ptr(obs.u), ptr(obs.v), ptr(obs.w), obs.len)
    uu = np.zeros(ncells)
    vv = np.zeros(ncells)
    n = L.orl2_facet_coords(ptr(uu), ptr(vv), obs.len, facet, ptr(sw))
    return fv[:n].copy(), fs[:n].copy(), uu[:n].copy(), vv[:n].copy(), sw
