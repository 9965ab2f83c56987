"""Pins of the oracle's model, gridding, cross-correlation and fit control flow (no GPU).

- gridding / facet coordinates / uv-grid size / facet averaging: bit-exact against the
  reference's own evaluate_uv_grid.c compiled in place (oracle/_ref), and brute force;
- visibility model: closed-form invariants (sersic(0) = 1, unit normalisation, a round
  source phase-shifted to its own position is real, l = m = 0 is the identity);
- cross-correlation: noiseless and shifted point-source data with known peak, and a
  brute-force local maximisation for random data;
- simplex: hand-derived first moves and convergence on a quadratic;
- likelihood sampling: the symmetry of a Gaussian likelihood fixes the mean exactly;
- a bright noiseless galaxy is recovered.
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.optimize as so

from oracle import oracle as O
from radiolensfit2_amd import synth
from tests import _common as T

L = O.lib()
R = O.ref()


def _random_uv(n, seed, nch=2, flagged=0.1, len_m=3000.0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-len_m, len_m, n)
    v = rng.uniform(-len_m, len_m, n)
    wn, spec = synth.channel_constants(nch, 950e6, 60e6)
    vis = (rng.normal(size=n * nch) + 1j * rng.normal(size=n * nch)) * 100
    s2 = rng.uniform(0.5e6, 2e6, n * nch).astype(np.float32)
    flag = rng.uniform(size=n * nch) < flagged
    umax = math.ceil(max(np.abs(u).max(), np.abs(v).max()))
    len_l = umax * wn[-1] / (2 * math.pi)
    return u, v, wn, vis, s2, flag, len_l


def _oracle_grid(u, v, wn, vis, s2, flag, len_l, facet):
    ncell = facet * facet
    nv, ns, sw = np.zeros(ncell, np.complex128), np.zeros(ncell), np.zeros(ncell)
    L.orl2_gridding(O.ptr(wn), len(wn), len(u), O.ptr(u), O.ptr(v), O.cptr(vis),
                    O.ptr(s2, O.c_float_p), len_l, facet, O.cptr(nv), O.ptr(ns),
                    O.ptr(flag, O.c_bool_p), O.ptr(sw))
    return nv, ns, sw


need_ref = pytest.mark.skipif(R is None, reason="oracle/_ref not built (reference tree absent)")


@need_ref
@pytest.mark.parametrize("seed,facet", [(0, 16), (1, 30), (2, 58), (3, 7)])
def test_gridding_bit_exact_vs_reference_build(seed, facet):
    u, v, wn, vis, s2, flag, len_l = _random_uv(3000, seed)
    nv, ns, sw = _oracle_grid(u, v, wn, vis, s2, flag, len_l, facet)
    ncell = facet * facet
    rv, rs, rw = np.zeros(ncell, np.complex128), np.zeros(ncell), np.zeros(ncell)
    R.gridding_visibilities(O.ptr(wn), len(wn), len(u), O.ptr(u), O.ptr(v), O.cptr(vis),
                            O.ptr(s2, O.c_float_p), len_l, facet, O.cptr(rv), O.ptr(rs),
                            O.ptr(flag, O.c_bool_p), O.ptr(rw))
    assert np.array_equal(nv, rv) and np.array_equal(ns, rs) and np.array_equal(sw, rw)
    gu, gv, ru, rvv = (np.zeros(ncell) for _ in range(4))
    n1 = L.orl2_facet_coords(O.ptr(gu), O.ptr(gv), len_l, facet, O.ptr(sw))
    n2 = R.evaluate_facet_coords(O.ptr(ru), O.ptr(rvv), len_l, facet, O.ptr(rw))
    assert n1 == n2 and np.array_equal(gu, ru) and np.array_equal(gv, rvv)
    a = L.orl2_uv_grid_size(len_l, O.ptr(wn), len(wn), len(u), O.ptr(u), O.ptr(v), facet,
                            O.ptr(flag, O.c_bool_p))
    b = R.evaluate_uv_grid_size(0, 1, len_l, O.ptr(wn), len(wn), len(u), O.ptr(u), O.ptr(v),
                                facet, O.ptr(flag, O.c_bool_p))
    assert a == b == n1
    for theta in (0.3, 1.0, 3.5):
        assert L.orl2_facet_size(theta, len_l) == R.facet_size(theta, len_l)


@need_ref
def test_average_facets_bit_exact_vs_reference_build():
    rng = np.random.default_rng(9)
    size = 40 * 40
    w = np.where(rng.uniform(size=size) < 0.4, rng.integers(1, 9, size), 0).astype(float)
    s2 = w * rng.uniform(1e6, 2e6, size)
    vis = (rng.normal(size=size) + 1j * rng.normal(size=size)) * w
    a_v, a_s = vis.copy(), s2.copy()
    b_v, b_s = vis.copy(), s2.copy()
    L.orl2_average_facets(size, O.cptr(a_v), O.ptr(a_s), O.ptr(w))
    R.average_facets(size, O.cptr(b_v), O.ptr(b_s), O.ptr(w))
    assert np.array_equal(a_v, b_v) and np.array_equal(a_s, b_s)


def test_gridding_brute_force_and_invariants():
    u, v, wn, vis, s2, flag, len_l = _random_uv(2000, 11)
    facet = 24
    nv, ns, sw = _oracle_grid(u, v, wn, vis, s2, flag, len_l, facet)
    inc = 2 * len_l / facet
    acc = {}
    for c in range(len(wn)):
        il = wn[c] / (2 * math.pi)
        for k in range(len(u)):
            n = c * len(u) + k
            if flag[n]:
                continue
            pu = int((u[k] * il + len_l) / inc)
            pv = int((v[k] * il + len_l) / inc)
            a = acc.setdefault(pv * facet + pu, [0j, 0.0, 0])
            a[0] += vis[n]
            a[1] += float(s2[n])
            a[2] += 1
    cells = sorted(acc)
    assert int(sw.sum()) == int((~flag).sum())  # every unflagged visibility lands once
    assert len(cells) == int((sw != 0).sum())
    for j, p in enumerate(cells):  # compacted in row-major cell order
        S, V, W = acc[p]
        assert sw[p] == W
        assert nv[j] == pytest.approx(S * W / V, rel=1e-12)
        assert ns[j] == pytest.approx(V / (W * W), rel=1e-12)


def test_visibility_model_invariants():
    assert L.orl2_sersic_shape(0.0, 0.0, 1.7, 3.1) == 1.0
    rng = np.random.default_rng(2)
    n = 500
    gu, gv = rng.uniform(-3e4, 3e4, n), rng.uniform(-3e4, 3e4, n)
    s2 = rng.uniform(0.1, 2.0, n)
    mod = np.zeros(n)
    L.orl2_model_at_zero(0.3, -0.2, 1.1, n, O.ptr(gu), O.ptr(gv), O.ptr(s2), O.ptr(mod))
    assert (mod ** 2 / s2).sum() == pytest.approx(1.0, rel=1e-14)
    # e -> -e with (u, v) -> (v, u) swaps k1 <-> k2: the shape is symmetric
    m2 = np.zeros(n)
    L.orl2_model_at_zero(-0.3, -0.2, 1.1, n, O.ptr(gv), O.ptr(gu), O.ptr(s2), O.ptr(m2))
    np.testing.assert_allclose(m2, mod, rtol=1e-13)
    # a round source shifted to its own position is real, with the analytic amplitude
    u, v, w = (rng.uniform(-5e3, 5e3, n) for _ in range(3))
    vis = np.zeros(n, np.complex128)
    l0, m0, k = 3e-4, -2e-4, 21.0
    L.orl2_data_galaxy_vis(0.9, k, 0.0, 0.0, 1.5, 40.0, l0, m0, n, O.ptr(u), O.ptr(v), O.ptr(w),
                           O.cptr(vis))
    amp = np.abs(vis).copy()
    L.orl2_phase_shift(k, l0, m0, n, O.ptr(u), O.ptr(v), O.ptr(w), O.cptr(vis))
    assert np.abs(vis.imag).max() <= 1e-13 * amp.max()
    sf = (1.5 * math.pi / 648000.0) ** 2
    np.testing.assert_allclose(vis.real, 36.0 * (1 + sf * k * k * (u * u + v * v)) ** -1.5,
                               rtol=1e-13)
    # l = m = 0: the phase shift is the identity
    z = (rng.normal(size=n) + 1j * rng.normal(size=n))
    z0 = z.copy()
    L.orl2_phase_shift(k, 0.0, 0.0, n, O.ptr(u), O.ptr(v), O.ptr(w), O.cptr(z))
    assert np.array_equal(z, z0)


def _xcorr(gu, gv, data, mod):
    ho, det, it = C.c_double(), C.c_double(), C.c_int()
    err = L.orl2_cross_correlation(len(gu), O.ptr(gu), O.ptr(gv), O.cptr(data), O.ptr(mod),
                                   C.byref(ho), C.byref(det), C.byref(it))
    return err, ho.value, det.value, it.value


def _facet(n=400, seed=0):
    rng = np.random.default_rng(seed)
    gu, gv = rng.uniform(-2e4, 2e4, n), rng.uniform(-2e4, 2e4, n)
    s2 = rng.uniform(0.5, 1.5, n)
    mod = np.zeros(n)
    L.orl2_model_at_zero(0.2, 0.1, 0.8, n, O.ptr(gu), O.ptr(gv), O.ptr(s2), O.ptr(mod))
    return gu, gv, s2, mod


def test_cross_correlation_noiseless_at_origin():
    gu, gv, s2, mod = _facet()
    A = 37.0
    data = (A * mod / s2).astype(np.complex128)
    err, ho, det, it = _xcorr(gu, gv, data, mod)
    assert err == 0 and it == 1  # zero gradient: one Newton step of length 0
    assert ho == pytest.approx(A, rel=1e-13)
    hr = A * mod * mod / s2
    U, V = 2 * math.pi * gu, 2 * math.pi * gv
    want = (hr * U * U).sum() * (hr * V * V).sum() - (hr * U * V).sum() ** 2
    assert det == pytest.approx(want, rel=1e-12)


def test_cross_correlation_finds_shifted_peak():
    gu, gv, s2, mod = _facet(seed=1)
    A, x0, y0 = 25.0, 2.1e-7, -1.3e-7  # radians; well inside the Newton basin
    data = A * mod / s2 * np.exp(1j * 2 * math.pi * (gu * x0 + gv * y0))
    err, ho, det, it = _xcorr(gu, gv, data, mod)
    assert err == 0 and it > 1
    assert ho == pytest.approx(A, rel=1e-10)


def test_cross_correlation_brute_force_local_max():
    gu, gv, s2, mod = _facet(n=60, seed=2)
    rng = np.random.default_rng(4)
    data = (5 * mod / s2 * np.exp(1j * 2 * math.pi * (gu * 1e-7 - gv * 0.5e-7))
            + (rng.normal(size=60) + 1j * rng.normal(size=60)) * 0.3)
    err, ho, det, it = _xcorr(gu, gv, data, mod)
    hF = np.conj(data) * mod
    U, V = 2 * math.pi * gu, 2 * math.pi * gv

    def h(X):
        return -(hF * np.exp(1j * (X[0] * U + X[1] * V))).real.sum()

    res = so.minimize(h, [0.0, 0.0], method="Nelder-Mead",
                      options={"xatol": 1e-14, "fatol": 1e-14, "maxiter": 20000})
    assert err == 0
    assert ho == pytest.approx(-res.fun, rel=1e-9)
    X = res.x
    ph = np.exp(1j * (X[0] * U + X[1] * V))
    re = (hF * ph).real
    want = (re * U * U).sum() * (re * V * V).sum() - (re * U * V).sum() ** 2
    assert det == pytest.approx(want, rel=1e-5)


def _nm(f, x0, step, tol=1e-3, max_iter=50):
    cb = O.F2_CB(lambda a, b, ctx: f(a, b))
    xo, fv, sz = np.zeros(2), C.c_double(), C.c_double()
    trace = np.zeros(2 * max_iter)
    it = L.orl2_nmsimplex2_minimize(cb, None, O.ptr(np.array(x0, float)),
                                    O.ptr(np.array(step, float)), tol, max_iter, O.ptr(xo),
                                    C.byref(fv), C.byref(sz), O.ptr(trace))
    return it, xo, fv.value, sz.value, trace[:2 * it].reshape(-1, 2)


def test_simplex_first_moves_by_hand():
    # f = x + 2y from (0,0), step 0.1: worst corner (0,0.1) reflects to (0.1,-0.1), which
    # beats the best, so the expansion 4.5*c - 3.5*x_hi = (0.15,-0.2) is taken.
    it, xo, fv, sz, tr = _nm(lambda x, y: x + 2 * y, [0, 0], [0.1, 0.1], max_iter=1)
    np.testing.assert_allclose(tr[0], [0.15, -0.2], atol=1e-15)
    # f = x^2 + y^2 from (1,1), step 1: worst (2,1) reflects to (0,2) (f=4): not better than
    # the best (2) nor worse than the second worst (5) -> plain reflection, best unchanged.
    it, xo, fv, sz, tr = _nm(lambda x, y: x * x + y * y, [1, 1], [1, 1], max_iter=1)
    np.testing.assert_allclose(tr[0], [1, 1], atol=1e-15)
    assert fv == 2.0
    # size is the RMS distance of the corners to their centroid
    it, xo, fv, sz, tr = _nm(lambda x, y: x * x + y * y, [1, 1], [1, 1], max_iter=1)
    P = np.array([[1, 1], [0, 2], [1, 2]], float)
    assert sz == pytest.approx(math.sqrt(((P - P.mean(0)) ** 2).sum() / 3), rel=1e-14)


def test_simplex_converges_on_quadratic():
    a, b = 0.21, -0.13
    f = lambda x, y: (x - a) ** 2 + 2 * (y - b) ** 2 + 0.5 * (x - a) * (y - b)  # noqa: E731
    it, xo, fv, sz, tr = _nm(f, [0, 0], [0.1, 0.1], tol=1e-6, max_iter=200)
    assert sz < 1e-6 and it < 200
    assert abs(xo[0] - a) < 1e-5 and abs(xo[1] - b) < 1e-5
    vals = [f(*p) for p in tr]
    assert all(v2 <= v1 for v1, v2 in zip(vals, vals[1:]))  # best value never increases


def _gauss_cb(mu, sig, rho=0.0, Lmax=50.0):
    s1, s2 = sig
    det = 1 - rho * rho

    def f(e1, e2, err, ctx):
        err[0] = 0
        d1, d2 = (e1 - mu[0]) / s1, (e2 - mu[1]) / s2
        return Lmax - 0.5 * (d1 * d1 - 2 * rho * d1 * d2 + d2 * d2) / det
    return O.LOGLIK_CB(f)


def test_sampling_gaussian_symmetry_fixes_mean():
    mu = (0.12, -0.07)
    cb = _gauss_cb(mu, (0.04, 0.06))
    e1, e2 = C.c_double(mu[0]), C.c_double(mu[1])
    v1, v2, od = C.c_double(), C.c_double(), C.c_double()
    nev = C.c_int()
    L.orl2_likelihood_sampling_cb(cb, None, C.byref(e1), C.byref(e2), 50.0, 30, C.byref(v1),
                                  C.byref(v2), C.byref(od), C.byref(nev))
    assert e1.value == pytest.approx(mu[0], abs=1e-14)
    assert e2.value == pytest.approx(mu[1], abs=1e-14)
    # diagonal covariance on a symmetric grid: zero covariance, 1-D variance = sqrt(v1 v2)
    assert od.value == pytest.approx(math.sqrt(v1.value * v2.value), rel=1e-10)
    # truncated at L >= 0.05 Lmax: variances below the Gaussian's, same order of magnitude
    assert 0.3 * 0.04 ** 2 < v1.value < 0.04 ** 2 and 0.3 * 0.06 ** 2 < v2.value < 0.06 ** 2
    assert nev.value > 30


def test_bright_noiseless_galaxy_is_recovered():
    cfg = synth.small_array(n_dish=32, n_times=60, n_chan=2, r_max=8e3, dt_s=480.0)
    cat = synth.make_catalog(10, seed=4)
    one = synth.Catalog(*(np.ascontiguousarray(getattr(cat, f)[:1]) for f in
                          ("flux", "scale", "e1", "e2", "l", "m", "snr")))
    one.flux[:] = 1000.0  # SNR ~150
    one.scale[:] = 1.2
    one.e1[:] = 0.3
    one.e2[:] = -0.15
    obs = synth.make_observation(cfg, one, noise=False, ms_roundtrip=False)
    vs = T.oracle_sky_model(obs, one)
    facet, fv, fs, uu, vv, sw = T.oracle_extract(obs, one, 0, obs.vis, vs)
    par, keep = T.oracle_params(fv, fs, uu, vv, one.flux[0])
    err, (e1, e2, v1, v2, od, maxL) = T.oracle_fit(par)
    assert err == 0
    assert abs(e1 - 0.3) < 1.5 * math.sqrt(v1) and abs(e2 + 0.15) < 1.5 * math.sqrt(v2)
    assert abs(e1 - 0.3) < 0.03 and abs(e2 + 0.15) < 0.03


def _fit_cb(cb):
    out = [C.c_double() for _ in range(6)]
    tr = O.FitTrace()
    err = L.orl2_source_fitting_cb(cb, None, 0.0, 0.0, 0.1, *[C.byref(x) for x in out],
                                   C.byref(tr))
    return err, [x.value for x in out], tr


def test_fit_flags_tiny_posterior_variance_as_bad():
    # galaxy_fitting.c:100: var < VAR = 1e-4 marks the fit bad, even for very bright sources
    err, r, tr = _fit_cb(_gauss_cb((0.1, 0.05), (0.004, 0.004), Lmax=500.0))
    assert err == 1 and r[2] < 1e-4 and abs(r[0] - 0.1) < 1e-3
    err, r, tr = _fit_cb(_gauss_cb((0.1, 0.05), (0.05, 0.05), Lmax=500.0))
    assert err == 0 and r[2] > 1e-4


def test_fit_flags_unbounded_or_failed_likelihood():
    def bad(e1, e2, err, ctx):
        err[0] = 1
        return -1e10
    err, r, tr = _fit_cb(O.LOGLIK_CB(bad))
    assert err == 1 and r[5] == -1e10 and tr.nm_iter == 1
