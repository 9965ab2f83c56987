"""The likelihood engine's r-marginalisation kernel (rmarg_kernel, through rl2_rmarg_eval)
against the oracle's set_posterior_values + marginalise_posterior_r (likelihood.c:126-146,
marginalise_r.c:29-223) on caller-given L_r rows, including the cases the fits reach only
rarely: points at kBad or NaN, fewer than 5 points, all values below the -10 floor of ymax,
a flat posterior, a rough one, one whose cubic spline overflows exp (the retry on the linear
interpolant), one whose single sharp peak the first 41-point rule misses (L_e = -inf, as
QUADPACK's acceptance of a zero first estimate gives in the reference).  Tolerance: L_e within 1e-8 absolute (the wave's Gauss-Kronrod
sums differ from QUADPACK's sequential order only by rounding), error flags equal."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from radiolensfit2_amd import _native as N

pytestmark = pytest.mark.gpu

KMAXR = 32
KBAD = -1.0e10


def _oracle_le(numr, ro, rprior, row):
    x = np.zeros(40)
    y = np.zeros(40)
    num = C.c_int()
    lr = np.zeros(40)
    lr[:numr] = row[:numr]
    rp = np.zeros(40)
    rp[:numr] = rprior[:numr]
    r40 = np.zeros(40)
    r40[:numr] = ro[:numr]
    L = O.lib()
    L.orl2_set_posterior_values(numr, O.ptr(lr), O.ptr(rp), O.ptr(r40), O.ptr(x), O.ptr(y),
                                C.byref(num))
    if num.value > 4:
        le = L.orl2_marginalise_posterior_r(num.value, O.ptr(x), O.ptr(y))
        err = 1 if le == KBAD else 0
    else:
        le, err = KBAD, 1
    return le, err


def _rows(numr, ro, rng):
    """L_r rows (KMAXR each) covering the cases named in the module docstring."""
    lro = np.log(ro[1:numr])
    rows = []
    for k in range(24):  # smooth peaks of varied width, height and position, with noise
        c = rng.uniform(lro.min(), lro.max())
        w = rng.uniform(0.2, 2.0)
        h = rng.uniform(-50.0, 5e4)
        rows.append(h - 0.5 * ((lro - c) / w) ** 2 * rng.uniform(1.0, 400.0)
                    + rng.normal(0.0, 0.02, numr - 1))
    # alternating 0 / -40000: the natural cubic overshoots beyond exp's range, QAG on it fails
    # and the integral is retried on the linear interpolant (marginalise_r.c:171-210)
    rows.append(np.where(np.arange(numr - 1) % 2 == 0, 0.0, -40000.0))
    # one point at 0, the rest at -5000: every node of the first 41-point rule sees exp(y) = 0,
    # so QAG accepts an integral of 0 and L_e is -inf with no error flag, as in the reference
    r = np.full(numr - 1, -5000.0); r[10] = 0.0; rows.append(r)
    base = 200.0 - 30.0 * (lro - lro.mean()) ** 2
    r = base.copy(); r[[0, 3, 7]] = KBAD; rows.append(r)               # scattered kBad
    r = base.copy(); r[:4] = KBAD; r[-3:] = KBAD; rows.append(r)        # ends at kBad
    r = base.copy(); r[[2, 5]] = np.nan; rows.append(r)                 # NaN points
    r = np.full(numr - 1, KBAD); r[[1, 4, 9, 12]] = base[[1, 4, 9, 12]]; rows.append(r)  # 4 pts
    r = np.full(numr - 1, KBAD); r[[1, 4, 9, 12, 15]] = base[[1, 4, 9, 12, 15]]; rows.append(r)
    rows.append(np.full(numr - 1, KBAD))                                # none
    rows.append(-100.0 - 5.0 * (lro - lro.mean()) ** 2)                 # below the -10 floor
    rows.append(np.full(numr - 1, 12.5))                                # flat
    rows.append(np.where(np.arange(numr - 1) % 2 == 0, 0.0, -8.0))      # rough
    out = np.zeros((len(rows), KMAXR))
    for i, r in enumerate(rows):
        out[i, 1:numr] = r
    return out


def test_rmarg_kernel_matches_the_oracle(gpu):
    ro, numr = O.r_grid()
    rng = np.random.default_rng(21)
    rows = _rows(numr, ro, rng)
    rprior = np.zeros(KMAXR)
    rprior[:numr] = O.rprior_for(40.0, np.pad(ro, (0, 40 - numr)), numr)[:numr]
    ro32 = np.zeros(KMAXR)
    ro32[:numr] = ro[:numr]
    n = rows.shape[0]
    le = np.zeros(n)
    err = (C.c_int * n)()
    rc = N.hip().rl2_rmarg_eval(n, numr, O.ptr(ro32), O.ptr(rprior), O.ptr(rows), O.ptr(le),
                                err)
    assert rc == 0, N.hip().rl2_last_error()
    for i in range(n):
        ole, oerr = _oracle_le(numr, ro32, rprior, rows[i])
        assert err[i] == oerr, (i, err[i], oerr)
        if np.isnan(ole):
            assert np.isnan(le[i]), (i, le[i])
        elif np.isinf(ole):
            assert le[i] == ole, (i, le[i], ole)
        else:
            assert abs(le[i] - ole) <= 1e-8, (i, le[i], ole)
    assert np.isneginf(le[25]) and err[25] == 0  # the peak the first rule misses
    # the cases whose outcome the reference fixes without quadrature
    assert err[-6] == 1 and le[-6] == KBAD      # 4 points: no marginalisation
    assert err[-4] == 1 and le[-4] == KBAD      # no point at all
    assert err[-5] == 0 and np.isfinite(le[-5])  # 5 points suffice


def test_rmarg_eval_rejects_bad_arguments(gpu):
    H = N.hip()
    ro, numr = O.r_grid()
    z = np.zeros(KMAXR)
    le = np.zeros(1)
    err = (C.c_int * 1)()
    for items, nr in ((0, numr), (1, 1), (1, 33)):
        assert H.rl2_rmarg_eval(items, nr, O.ptr(z), O.ptr(z), O.ptr(z), O.ptr(le), err) != 0
        assert b"rl2_rmarg_eval" in H.rl2_last_error()
