"""Host-side logic of librl2hip.so vs the oracle, without a GPU.

The product's simplex + likelihood-sampling driver (rl2_fit.cpp) is run through its
callback test hook on analytic log-likelihoods; with and without speculative batching it
must reproduce the oracle's sequential trajectory exactly (same ellipticities, variances,
maximum and number of loglikelihood calls).
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from radiolensfit2_amd import _native as N


def _surfaces():
    def gauss(mu, s, rho=0.0, Lmax=40.0):
        def f(e1, e2):
            d1, d2 = (e1 - mu[0]) / s[0], (e2 - mu[1]) / s[1]
            return Lmax - 0.5 * (d1 * d1 - 2 * rho * d1 * d2 + d2 * d2) / (1 - rho * rho), 0
        return f

    def banana(e1, e2):
        return 30.0 - 8 * (e2 - 0.3 * e1 * e1 - 0.05) ** 2 * 40 - 3 * (e1 - 0.2) ** 2, 0

    def holed(e1, e2):  # loglikelihood errors (numvals <= 4) outside |e| < 0.45
        if e1 * e1 + e2 * e2 > 0.2025:
            return -1e10, 1
        return 20.0 - 40 * ((e1 + 0.1) ** 2 + (e2 - 0.15) ** 2), 0

    def edge(e1, e2):  # maximum near the |e|^2 = 0.65 boundary of f_likelihood
        return 25.0 - 30 * ((e1 - 0.55) ** 2 + (e2 + 0.5) ** 2), 0

    def flat(e1, e2):  # no information: fval >= 0 stops the simplex at once
        return -3.0, 0

    def bumpy(e1, e2):
        return (15.0 - 20 * ((e1 - 0.3) ** 2 + (e2 - 0.1) ** 2)
                + 0.4 * math.sin(40 * e1) * math.cos(33 * e2)), 0

    return {
        "gauss": gauss((0.12, -0.07), (0.05, 0.08)),
        "gauss_corr": gauss((-0.2, 0.3), (0.06, 0.04), rho=0.6),
        "gauss_tight": gauss((0.1, 0.05), (0.004, 0.004), Lmax=500.0),
        "banana": banana,
        "holed": holed,
        "edge": edge,
        "flat": flat,
        "bumpy": bumpy,
    }


SURFACES = _surfaces()


def _cb(f, counter):
    def inner(e1, e2, err, ctx):
        counter[0] += 1
        L, e = f(e1, e2)
        err[0] = e
        return L
    return inner


def _oracle(f):
    cnt = [0]
    cb = O.LOGLIK_CB(_cb(f, cnt))
    out = [C.c_double() for _ in range(6)]
    tr = O.FitTrace()
    err = O.lib().orl2_source_fitting_cb(cb, None, 0.0, 0.0, 0.1, *[C.byref(x) for x in out],
                                         C.byref(tr))
    return err, [x.value for x in out], tr.n_evals


def _product(f, speculate, chunk):
    cnt = [0]
    cb = N.LOGLIK_CB(_cb(f, cnt))
    res = N.GalResult()
    launches = N.hip().rl2_fit_with_callback(cb, None, speculate, chunk, C.byref(res))
    assert launches > 0
    return res, launches, cnt[0]


@pytest.mark.parametrize("name", sorted(SURFACES))
@pytest.mark.parametrize("speculate,chunk", [(0, 1), (1, 1), (1, 4), (1, 16)])
def test_fit_control_flow_matches_oracle(name, speculate, chunk):
    f = SURFACES[name]
    oerr, (e1, e2, v1, v2, od, maxL), onev = _oracle(f)
    res, launches, calls = _product(f, speculate, chunk)
    assert res.error == oerr
    # identical trajectory: bit-identical results
    got = [res.e1, res.e2, res.var_e1, res.var_e2, res.oneDimvar, res.maxL]
    want = [e1, e2, v1, v2, od, maxL]
    for g, w in zip(got, want):
        assert g == w or (math.isnan(g) and math.isnan(w)), (name, got, want)
    assert res.n_evals == onev
    if speculate or chunk > 1:
        assert calls >= onev  # speculation may evaluate points the replay never uses


def test_speculation_reduces_launches():
    f = SURFACES["gauss_corr"]
    _, seq, _ = _product(f, 0, 1)
    _, spec, _ = _product(f, 1, 4)
    assert spec < seq


def test_hip_library_has_no_cpu_path(monkeypatch, tmp_path):
    """Without the native library the product refuses to run (no silent fallback)."""
    monkeypatch.setattr(N, "_hip", None)
    monkeypatch.setattr(N, "HIP_LIB", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        N.hip()
