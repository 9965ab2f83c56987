"""ctypes bindings of librl2hip.so (include/rl2.h, include/rl2_reference.h) and librl2sim.so
(include/rl2_sim.h).

There is no Python or CPU fallback: if the native library is missing, importing the
bindings raises.  Build with ``python -m radiolensfit2_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(PKG, "librl2hip.so")
SIM_LIB = os.path.join(PKG, "librl2sim.so")

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_int_p = C.POINTER(C.c_int)
c_uint_p = C.POINTER(C.c_uint)
c_ubyte_p = C.POINTER(C.c_ubyte)
c_bool_p = C.POINTER(C.c_bool)

RL2_OK = 0
(RL2_ERR_NO_DEVICE, RL2_ERR_ARG, RL2_ERR_HIP, RL2_ERR_RANGE, RL2_ERR_STATE) = (-1, -2, -3, -4, -5)
RL2_PREC_FP64 = 0
RL2_PREC_MIXED = 1
RL2_TUNE_GRID_VARIANT = 0
RL2_TUNE_TILE_CELLS = 1
RL2_TUNE_PHASE_TABLES = 2
RL2_TUNE_FUSE_SUBTRACT = 3
RL2_TUNE_TILE_BLOCKS = 4
RL2_TUNE_TILE_STREAMS = 5
RL2_TUNE_BATCH_EXTRACT = 6
RL2_TUNE_BATCH_GROUPS = 7
RL2_TUNE_BATCH_GRAPH = 8
RL2_TUNE_ITEM_GROUPS = 9
RL2_TUNE_FUSE_STEP = 10
RL2_TUNE_TILE_PASSES = 11
RL2_TUNE_P0_TASKS = 12
(RL2_K_EXTRACT, RL2_K_COMPACT, RL2_K_LIKELIHOOD, RL2_K_RMARG, RL2_K_SUBTRACT,
 RL2_K_SOURCES, RL2_K_STEP) = range(7)
KERNEL_NAMES = ["extract", "compact", "likelihood", "rmarg", "subtract", "sources", "step"]


class GalResult(C.Structure):
    _fields_ = [("e1", C.c_double), ("e2", C.c_double), ("var_e1", C.c_double),
                ("var_e2", C.c_double), ("oneDimvar", C.c_double), ("maxL", C.c_double),
                ("error", C.c_int), ("pass_", C.c_int), ("facet", C.c_uint),
                ("ncoords", C.c_uint), ("n_evals", C.c_int), ("subtracted", C.c_int),
                ("start", C.c_int), ("rounds", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("extraction_s", C.c_double), ("fitting_s", C.c_double),
                ("subtraction_s", C.c_double), ("n_bad_first", C.c_int),
                ("n_bad_final", C.c_int)]


class FitOptions(C.Structure):
    _fields_ = [("batch", C.c_int), ("do_refit", C.c_int), ("nm_speculate", C.c_int),
                ("line_chunk", C.c_int), ("lik_block", C.c_int), ("verbose", C.c_int),
                ("lik_precision", C.c_int), ("proposal", C.c_void_p),
                ("features_out", C.POINTER(C.c_float)), ("labels_out", C.POINTER(C.c_double)),
                ("proposal_mode", C.c_int)]


class TuneTrial(C.Structure):
    _fields_ = [("knob", C.c_int), ("value", C.c_int), ("seconds", C.c_double)]


class KernelStat(C.Structure):
    _fields_ = [("ms", C.c_double), ("launches", C.c_longlong), ("flops", C.c_double),
                ("bytes", C.c_double)]


class Complexd(C.Structure):
    _fields_ = [("real", C.c_double), ("imag", C.c_double)]


class LikelihoodParams(C.Structure):
    """likelihood_params, FACET layout (datatype.h:49-74)."""
    _fields_ = [
        ("numr", C.c_int), ("ro", c_double_p), ("rprior", c_double_p),
        ("nchannels", C.c_uint), ("band_factor", C.c_double), ("acc_time", C.c_double),
        ("spec", c_double_p), ("wavenumbers", c_double_p), ("uu", c_double_p),
        ("vv", c_double_p), ("l0", C.c_double), ("m0", C.c_double),
        ("data", C.POINTER(Complexd)), ("sigma2", c_double_p), ("mod", c_double_p),
        ("facet", C.c_uint), ("ncoords", C.c_uint),
    ]


class GslVector(C.Structure):
    _fields_ = [("size", C.c_size_t), ("stride", C.c_size_t), ("data", c_double_p),
                ("block", C.c_void_p), ("owner", C.c_int)]


LOGLIK_CB = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, c_int_p, C.c_void_p)

_HIP_SIGS = {
    "rl2_last_error": (C.c_char_p, []),
    "rl2_init": (C.c_int, [C.c_int]),
    "rl2_device_count": (C.c_int, [c_int_p]),
    "rl2_obs_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint, C.c_ulong, c_double_p,
                                 c_double_p, c_double_p, c_double_p, c_double_p, c_float_p,
                                 c_ubyte_p, C.c_double, C.c_int]),
    "rl2_obs_destroy": (C.c_int, [C.c_void_p]),
    "rl2_obs_set_vis": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "rl2_obs_get_residual": (C.c_int, [C.c_void_p, c_double_p]),
    "rl2_obs_sync": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "rl2_obs_add_sources": (C.c_int, [C.c_void_p, C.c_uint, c_double_p, c_double_p, c_double_p,
                                      c_double_p, c_double_p, c_double_p, C.c_double]),
    "rl2_obs_add_noise": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64]),
    "rl2_obs_checkpoint": (C.c_int, [C.c_void_p]),
    "rl2_obs_restore": (C.c_int, [C.c_void_p]),
    "rl2_obs_loglik_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "rl2_obs_extract": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_uint, c_double_p, c_double_p, c_double_p,
                                  c_double_p, c_double_p, c_uint_p]),
    "rl2_obs_load_facet": (C.c_int, [C.c_void_p, C.c_int, C.c_uint, c_double_p, c_double_p,
                                     c_double_p, c_double_p]),
    "rl2_obs_set_prior": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "rl2_obs_loglik": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p, c_double_p,
                                 c_double_p, c_int_p, c_double_p, c_int_p]),
    "rl2_obs_fit": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(FitOptions), C.POINTER(GalResult)]),
    "rl2_obs_subtract": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double]),
    "rl2_process": (C.c_int, [C.c_void_p, C.c_uint, c_double_p, c_double_p, c_double_p,
                              c_double_p, C.POINTER(FitOptions), C.POINTER(GalResult),
                              C.POINTER(RunStats)]),
    "rl2_r_grid": (C.c_int, [c_double_p]),
    "rl2_marginalise_posterior_r_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                                      C.c_void_p]),
    "rl2_rmarg_eval": (C.c_int, [C.c_int, C.c_int, c_double_p, c_double_p, c_double_p,
                                 c_double_p, C.POINTER(C.c_int)]),
    "rl2_marg_position_eval": (C.c_int, [C.c_int, c_double_p, c_double_p]),
    "rl2_stats_enable": (C.c_int, [C.c_int]),
    "rl2_stats_reset": (C.c_int, []),
    "rl2_stats_get": (C.c_int, [C.c_int, C.POINTER(KernelStat)]),
    "rl2_synchronize": (C.c_int, []),
    "rl2_tune_set": (C.c_int, [C.c_int, C.c_int]),
    "rl2_test_sincos_rev64": (None, [C.c_int, c_double_p, c_double_p, c_double_p]),
    "rl2_test_philox4x32_10": (None, [C.c_int, c_uint_p, c_uint_p, c_uint_p]),
    "rl2_test_cell_index": (C.c_int, [C.c_int, c_double_p, c_double_p, C.c_double, C.c_double,
                                      C.c_double, C.c_uint, c_uint_p, c_uint_p]),
    "rl2_tune_get": (C.c_int, [C.c_int, c_int_p]),
    "rl2_autotune": (C.c_int, [C.c_void_p, C.c_uint, c_double_p, c_double_p, c_double_p,
                               c_double_p, C.POINTER(FitOptions), C.POINTER(TuneTrial), C.c_int,
                               c_int_p, c_int_p]),
    "rl2_fit_with_callback": (C.c_int, [LOGLIK_CB, C.c_void_p, C.c_int, C.c_int,
                                        C.POINTER(GalResult)]),
    # learned proposal
    "rl2_proposal_num_params": (C.c_int, []),
    "rl2_proposal_num_inputs": (C.c_int, []),
    "rl2_proposal_num_outputs": (C.c_int, []),
    "rl2_proposal_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint64]),
    "rl2_proposal_destroy": (C.c_int, [C.c_void_p]),
    "rl2_proposal_get_params": (C.c_int, [C.c_void_p, c_float_p]),
    "rl2_proposal_set_params": (C.c_int, [C.c_void_p, c_float_p]),
    "rl2_proposal_get_grads": (C.c_int, [C.c_void_p, c_float_p]),
    "rl2_proposal_forward": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p]),
    "rl2_proposal_gradients": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p, c_float_p]),
    "rl2_proposal_train_step": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p, C.c_float,
                                          c_float_p]),
    "rl2_obs_features": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, c_float_p]),
    # reference-signature drop-ins (include/rl2_reference.h)
    "sersic_shape": (C.c_double, [C.c_double] * 4),
    "gaussian_shape": (C.c_double, [C.c_double] * 4),
    "facet_size": (C.c_uint, [C.c_double, C.c_double]),
    "scale_mean": (C.c_double, [C.c_double]),
    "rfunc": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "marginalise_over_position_shift": (C.c_double, [C.c_double]),
    "set_posterior_values": (None, [C.c_int, c_double_p, c_double_p, c_double_p, c_double_p,
                                    c_double_p, c_int_p]),
    "marginalise_posterior_r": (C.c_double, [C.c_int, c_double_p, c_double_p]),
    "marf": (C.c_double, [C.c_double, C.c_void_p]),
    "model_galaxy_visibilities_at_zero": (None, [C.c_double, C.c_double, C.c_double, C.c_uint,
                                                 c_double_p, c_double_p, c_double_p, c_double_p]),
    "data_galaxy_visibilities": (None, [C.c_double] * 10 + [C.c_uint, c_double_p, c_double_p,
                                                            c_double_p, C.POINTER(Complexd)]),
    "data_visibilities_phase_shift": (None, [C.c_double, C.c_double, C.c_double, C.c_uint,
                                             c_double_p, c_double_p, c_double_p,
                                             C.POINTER(Complexd)]),
    "evaluate_uv_grid_size": (C.c_uint, [C.c_int, C.c_int, C.c_double, c_double_p, C.c_uint,
                                         C.c_uint, c_double_p, c_double_p, C.c_uint, c_bool_p]),
    "evaluate_facet_coords": (C.c_uint, [c_double_p, c_double_p, C.c_double, C.c_uint,
                                         c_double_p]),
    "gridding_visibilities": (None, [c_double_p, C.c_uint, C.c_uint, c_double_p, c_double_p,
                                     C.POINTER(Complexd), c_float_p, C.c_double, C.c_uint,
                                     C.POINTER(Complexd), c_double_p, c_bool_p, c_double_p]),
    "average_facets": (None, [C.c_ulong, C.POINTER(Complexd), c_double_p, c_double_p]),
    "source_extraction": (None, [C.c_int, C.c_uint, C.POINTER(LikelihoodParams),
                                 C.POINTER(Complexd), c_double_p, c_double_p, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.POINTER(Complexd), C.POINTER(Complexd), C.POINTER(Complexd),
                                 c_float_p, c_bool_p, C.c_ulong, c_double_p, c_double_p,
                                 c_double_p, C.c_double]),
    "f_likelihood": (C.c_double, [C.POINTER(GslVector), C.c_void_p]),
    "loglikelihood": (C.c_double, [C.c_void_p, C.c_double, C.c_double, c_int_p]),
    "loglikelihood_r": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_ulong, c_double_p,
                                     c_double_p, c_double_p, C.POINTER(Complexd), c_double_p]),
    "cross_correlation": (C.c_int, [C.c_ulong, c_double_p, c_double_p, C.POINTER(Complexd),
                                    c_double_p, c_double_p, c_double_p]),
    "likelihood_sampling": (None, [c_double_p, c_double_p, C.c_double, C.c_void_p, C.c_int,
                                   c_double_p, c_double_p, c_double_p]),
    "source_fitting": (C.c_int, [C.c_int, C.POINTER(LikelihoodParams)] + [c_double_p] * 6),
    "sky_model": (None, [c_double_p, c_double_p, C.c_double, C.c_int, C.c_uint, C.c_uint,
                         c_double_p, c_double_p, c_double_p, c_double_p, C.c_ulong, c_double_p,
                         c_double_p, c_double_p, C.POINTER(Complexd), C.POINTER(Complexd)]),
    "data_processing": (None, [C.c_bool, c_uint_p, C.c_int, C.c_double, C.c_uint, C.c_void_p,
                               C.POINTER(LikelihoodParams), c_double_p, c_double_p, c_double_p,
                               c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                               C.POINTER(Complexd), C.POINTER(Complexd), C.POINTER(Complexd),
                               c_float_p, c_bool_p, c_double_p, c_double_p, c_double_p,
                               c_double_p, c_int_p]),
}

_SIM_SIGS = {
    "rl2sim_mt_new": (C.c_void_p, [C.c_ulong]),
    "rl2sim_mt_free": (None, [C.c_void_p]),
    "rl2sim_mt_get": (C.c_uint32, [C.c_void_p]),
    "rl2sim_mt_uniform": (C.c_double, [C.c_void_p]),
    "rl2sim_mt_uniform_pos": (C.c_double, [C.c_void_p]),
    "rl2sim_mt_gaussian": (C.c_double, [C.c_void_p, c_double_p]),
    "rl2sim_layout": (C.c_int, [C.c_int, C.c_uint, C.c_ulong, C.c_double, C.c_double, C.c_double,
                                C.c_double, c_double_p, c_double_p]),
    "rl2sim_uvw": (C.c_ulong, [C.c_uint, c_double_p, c_double_p, C.c_double, C.c_double, C.c_uint,
                               C.c_double, C.c_double, c_double_p, c_double_p, c_double_p]),
    "rl2sim_catalog": (C.c_uint, [C.c_uint, C.c_ulong, C.c_double, C.c_double, C.c_double,
                                  c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                  c_double_p]),
    "rl2sim_shear": (None, [C.c_uint, C.c_double, C.c_double, c_double_p, c_double_p]),
    "rl2sim_simulate": (C.c_int, [C.c_uint, c_double_p, c_double_p, C.c_ulong, c_double_p,
                                  c_double_p, c_double_p, C.c_uint, c_double_p, c_double_p,
                                  c_double_p, c_double_p, c_double_p, c_double_p, C.c_float,
                                  C.c_int, C.c_ulong, C.c_int, c_double_p, c_double_p]),
    "rl2ms_write": (C.c_int, [C.c_char_p, C.c_uint, C.c_uint, C.c_ulong, C.c_double, C.c_double,
                              C.c_double, c_double_p, c_double_p, c_double_p, c_ubyte_p,
                              c_float_p, c_float_p]),
    "rl2ms_info": (C.c_int, [C.c_char_p, c_uint_p, c_uint_p, C.POINTER(C.c_ulong), c_double_p,
                             c_double_p, c_double_p]),
    "rl2ms_read": (C.c_double, [C.c_char_p, c_double_p, c_double_p, c_double_p, c_ubyte_p,
                                c_double_p, c_float_p, c_int_p]),
}

_hip = None
_sim = None


def _load(path, sigs, what):
    if not os.path.exists(path):
        raise RuntimeError(
            f"{what} is not built ({path} missing): run `python -m radiolensfit2_amd.build`. "
            "There is no CPU fallback.")
    lib = C.CDLL(path)
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def hip():
    """librl2hip.so (HIP kernels for gfx950 + C-ABI)."""
    global _hip
    if _hip is None:
        _hip = _load(HIP_LIB, _HIP_SIGS, "librl2hip.so")
    return _hip


def sim():
    """librl2sim.so (host generators)."""
    global _sim
    if _sim is None:
        _sim = _load(SIM_LIB, _SIM_SIGS, "librl2sim.so")
    return _sim


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(c_double_p)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_float_p)


def cplx(a):
    """complex128 array -> double* (re, im interleaved)"""
    if a is None:
        return None
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def cplx_s(a):
    if a is None:
        return None
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(Complexd))


def check(status, what="rl2"):
    if status != RL2_OK:
        msg = hip().rl2_last_error()
        raise RuntimeError(f"{what} failed ({status}): {msg.decode() if msg else ''}")
