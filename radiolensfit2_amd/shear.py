"""Shear estimate from measured ellipticities (the reference's shear.py:1-148, numpy only).

The measurement file is the reference's output table (header line, then one line per
galaxy written by data_processing.c:102,136):
    flux | e1 | m_e1 | err1 | e2 | m_e2 | err2 | 1D var | SNR | l | m
Catalogs come in orientation pairs (consecutive lines, generate_catalog.c): a bad
measurement removes its partner too (shear.py:58-100).

Behaviour kept from the reference:
- acceptance: 1D var > 1e-5, err1 > 1e-3, err2 > 1e-3, SNR >= 10 (shear.py:64);
- pair removal, including how a bad odd line overwrites its even partner, and how a bad
  last odd line removes the entry before its partner (shear.py:65-100);
- the shape-noise moments are unweighted: the reference's prior_ellipticity has no return
  statement, so numpy.average receives weights=None (shear.py:32-34,109-117);
- weights 1 / (var e_max^2 / (e_max^2 - 2 var) + sigma_shape) (shear.py:77,126);
- bootstrap over galaxy pairs (rows of the (ngal, 2, 2, norient) array), 1000 resamples,
  the mean and standard deviation of the weighted-mean shear (shear.py:135-145).
"""
from __future__ import annotations

import numpy as np

E_MAX = 0.804
HEADER = "flux | e1 | m_e1 | err1 | e2 | m_e2 | err2 | 1D var | SNR |   l  |  m  | \n"


def write_ellipticities(path, flux, e1, e2, results, snr, l, m):
    """Reference output table for per-galaxy results (rl2_gal_result records)."""
    arcs2rad = np.pi / 180.0 / 3600.0
    with open(path, "w") as f:
        f.write(HEADER)
        for k, r in enumerate(results):
            me1, me2 = (0.0, 0.0) if r.error else (r.e1, r.e2)
            f.write("%f | %f | %f | %f | %f | %f | %f | %f | %f | %f | %f \n" % (
                flux[k], e1[k], me1, np.sqrt(max(r.var_e1, 0.0)), e2[k], me2,
                np.sqrt(max(r.var_e2, 0.0)), r.oneDimvar, snr[k], l[k] / arcs2rad,
                m[k] / arcs2rad))


def read_ellipticities(path):
    """Columns (m_e1, err1, m_e2, err2, 1D var, SNR) of a measurement file."""
    return np.loadtxt(path, skiprows=1, delimiter="|", usecols=(2, 3, 5, 6, 7, 8), ndmin=2)


def select(data):
    """Accepted (m_e1, m_e2, w, SNR) after pair removal, in the reference's order.

    `data` rows: (m_e1, err1, m_e2, err2, 1D var, SNR)."""
    me1, me2, w, snr = [], [], [], []
    k, remove, bad, i = 0, 0, 0, 0
    while i < len(data):
        error1, error2, var, snrv = data[i, 1], data[i, 3], data[i, 4], data[i, 5]
        if var > 1e-5 and error1 > 1e-3 and error2 > 1e-3 and snrv >= 10:
            wk = var * E_MAX * E_MAX / (E_MAX * E_MAX - 2 * var)
            if remove == 1:
                me1[k], me2[k], w[k], snr[k] = data[i, 0], data[i, 2], wk, snrv
            else:
                me1.append(data[i, 0])
                me2.append(data[i, 2])
                w.append(wk)
                snr.append(snrv)
            k += 1
            remove = 0
        else:
            if i % 2 != 0:
                remove = 1
                k -= 1
            else:
                i += 1
            bad += 1
        i += 1
    if remove == 1:
        k -= 1
        del me1[k], me2[k], w[k], snr[k]
    return np.array(me1), np.array(me2), np.array(w), np.array(snr), bad


def shape_noise(me1, me2):
    m1, m2 = np.mean(me1), np.mean(me2)
    var12 = np.mean(me1 * me2) - m1 * m2
    var1 = np.mean(me1 * me1) - m1 * m1
    var2 = np.mean(me2 * me2) - m2 * m2
    return float(np.sqrt(var1 * var2 - var12 * var12))


def shear(me1, me2, w, bootnum=1000, seed=None):
    """Bootstrap weighted-mean shear: ((mean, std) of g1, (mean, std) of g2, sigma_shape)."""
    sigma = shape_noise(me1, me2)
    ngal = len(me1) // 2
    n = 2 * ngal
    wt = 1.0 / (w[:n] + sigma)
    d1, d2, ww = me1[:n].reshape(ngal, 2), me2[:n].reshape(ngal, 2), wt.reshape(ngal, 2)
    rng = np.random.default_rng(seed)
    g1, g2 = np.zeros(bootnum), np.zeros(bootnum)
    for b in range(bootnum):
        idx = rng.integers(0, ngal, ngal)
        g1[b] = np.average(d1[idx], weights=ww[idx])
        g2[b] = np.average(d2[idx], weights=ww[idx])
    return (float(g1.mean()), float(g1.std())), (float(g2.mean()), float(g2.std())), sigma


def shear_from_file(path, bootnum=1000, seed=None):
    me1, me2, w, snr, bad = select(read_ellipticities(path))
    return shear(me1, me2, w, bootnum, seed), bad
