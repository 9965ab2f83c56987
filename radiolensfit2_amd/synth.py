"""Synthetic observations and catalogs for the shape-measurement path (librl2sim.so).

An ``Observation`` holds exactly what RadioLensfit2 reads from its Measurement Set and
derives at start-up (RadioLensfit2.cpp:157-351, :408): uvw per row in metres, channel
wavenumbers and spectral factors, sigma^2 per visibility (theoretical value, SIGMA = 0),
flags, the channel-major visibilities in micro-Jy and ``len`` in wavelengths.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _native as N

PI = 3.14159265358979323846
C0 = 299792458.0
EFFICIENCY = 0.9
SEFD_SKA = 400e6
REF_FREQ = 1.4e9


@dataclasses.dataclass
class ArrayConfig:
    """Antenna layout + observation set-up."""
    n_dish: int
    seed: int = 1
    kind: int = 0              # 0: log-uniform radius; 1: dense core + log-radial
    r_min: float = 20.0
    r_max: float = 10e3
    core_frac: float = 0.5
    r_core: float = 1e3
    lat_deg: float = -30.0
    dec_deg: float = -30.0
    n_times: int = 480
    dt_s: float = 60.0
    h0_hours: float = -4.0
    n_chan: int = 4
    freq_start_hz: float = 950e6
    bw_hz: float = 60e6


# Configurations named by the workload (sizes follow the project's benchmark configs).
ARRAY_64 = ArrayConfig(n_dish=64)                                  # 64 dishes, 4 x 60 MHz
ARRAY_197 = ArrayConfig(n_dish=197, kind=1, r_max=150e3, n_chan=16, bw_hz=15e6)


def small_array(n_dish=12, n_times=40, n_chan=2, seed=3, r_max=3e3, bw_hz=60e6, dt_s=360.0):
    """A reduced array the oracle handles in seconds."""
    return ArrayConfig(n_dish=n_dish, seed=seed, r_max=r_max, n_times=n_times, dt_s=dt_s,
                       n_chan=n_chan, bw_hz=bw_hz)


def channel_constants(n_chan, freq_start_hz, bw_hz):
    """RadioLensfit2.cpp:342-351"""
    wn = np.zeros(n_chan)
    spec = np.zeros(n_chan)
    ch_freq = freq_start_hz + 0.5 * bw_hz
    for i in range(n_chan):
        wn[i] = 2.0 * PI * ch_freq / C0
        spec[i] = math.pow(ch_freq / REF_FREQ, -0.7)
        ch_freq += bw_hz
    return wn, spec


def theoretical_sigma2(time_acc, bw_hz):
    """RadioLensfit2.cpp:255, stored as float."""
    return np.float32((SEFD_SKA * SEFD_SKA) / (2. * int(time_acc) * bw_hz * EFFICIENCY * EFFICIENCY))


@dataclasses.dataclass
class Observation:
    cfg: ArrayConfig
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    wn: np.ndarray
    spec: np.ndarray
    sigma2: np.ndarray      # float32, nch*nrow
    flag: np.ndarray        # bool, nch*nrow
    len: float              # wavelengths
    vis: np.ndarray         # complex128 nch*nrow (micro-Jy), channel-major

    @property
    def nrow(self):
        return len(self.u)

    @property
    def nch(self):
        return len(self.wn)

    @property
    def nvis(self):
        return self.nrow * self.nch


def make_geometry(cfg: ArrayConfig):
    S = N.sim()
    east = np.zeros(cfg.n_dish)
    north = np.zeros(cfg.n_dish)
    S.rl2sim_layout(cfg.kind, cfg.n_dish, cfg.seed, cfg.r_min, cfg.r_max, cfg.core_frac,
                    cfg.r_core, N.dptr(east), N.dptr(north))
    nbl = cfg.n_dish * (cfg.n_dish - 1) // 2
    nrow = nbl * cfg.n_times
    u = np.zeros(nrow)
    v = np.zeros(nrow)
    w = np.zeros(nrow)
    S.rl2sim_uvw(cfg.n_dish, N.dptr(east), N.dptr(north), cfg.lat_deg, cfg.dec_deg, cfg.n_times,
                 cfg.dt_s, cfg.h0_hours, N.dptr(u), N.dptr(v), N.dptr(w))
    wn, spec = channel_constants(cfg.n_chan, cfg.freq_start_hz, cfg.bw_hz)
    # measurement_set.cpp:223-240 then RadioLensfit2.cpp:408
    umax = math.ceil(max(np.abs(u).max(), np.abs(v).max()))
    len_lambda = umax * wn[-1] / (2 * PI)
    s2 = theoretical_sigma2(cfg.dt_s, cfg.bw_hz)
    sigma2 = np.full(nrow * cfg.n_chan, s2, np.float32)
    flag = np.zeros(nrow * cfg.n_chan, bool)
    return u, v, w, wn, spec, sigma2, flag, len_lambda


@dataclasses.dataclass
class Catalog:
    flux: np.ndarray
    scale: np.ndarray
    e1: np.ndarray       # sheared (what the simulation used and the catalog file holds)
    e2: np.ndarray
    l: np.ndarray
    m: np.ndarray
    snr: np.ndarray

    def __len__(self):
        return len(self.flux)


def make_catalog(n_gal, seed=1, fov_arcmin=20.0, fmin=10.0, fmax=200.0, g=(0.0, 0.0)):
    S = N.sim()
    a = [np.zeros(n_gal) for _ in range(6)]
    n = S.rl2sim_catalog(n_gal, seed, fov_arcmin, fmin, fmax, *[N.dptr(x) for x in a])
    flux, scale, e1, e2, l, m = [x[:n].copy() for x in a]
    S.rl2sim_shear(n, g[0], g[1], N.dptr(e1), N.dptr(e2))
    return Catalog(flux, scale, e1, e2, l, m, np.zeros(n))


def make_observation(cfg: ArrayConfig, cat: Catalog | None = None, noise=True, noise_seed=7,
                     ms_roundtrip=True) -> Observation:
    """Simulate on the host (data_simulation.c:41-127); fine for small arrays."""
    S = N.sim()
    u, v, w, wn, spec, sigma2, flag, len_lambda = make_geometry(cfg)
    nrow = len(u)
    vis = np.zeros(nrow * cfg.n_chan, np.complex128)
    if cat is not None and len(cat):
        S.rl2sim_simulate(cfg.n_chan, N.dptr(wn), N.dptr(spec), nrow, N.dptr(u), N.dptr(v),
                          N.dptr(w), len(cat), N.dptr(cat.flux), N.dptr(cat.scale),
                          N.dptr(cat.e1), N.dptr(cat.e2), N.dptr(cat.l), N.dptr(cat.m),
                          float(sigma2[0]), int(noise), noise_seed, int(ms_roundtrip),
                          N.cplx(vis), N.dptr(cat.snr))
    elif noise:
        S.rl2sim_simulate(cfg.n_chan, N.dptr(wn), N.dptr(spec), nrow, N.dptr(u), N.dptr(v),
                          N.dptr(w), 0, None, None, None, None, None, None, float(sigma2[0]), 1,
                          noise_seed, int(ms_roundtrip), N.cplx(vis), None)
    return Observation(cfg, u, v, w, wn, spec, sigma2, flag, len_lambda, vis)
