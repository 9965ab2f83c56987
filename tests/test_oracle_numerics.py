"""Pins of the oracle's numerical building blocks (no GPU).

GSL 2.5 is absent, so each restated routine is checked against an independent source:
scipy / mpmath for Ei and splines, polynomial exactness for the Gauss-Kronrod rule,
adaptive quadrature of the same spline for QAG, closed forms for the scalelength grid,
facet size and the position marginalisation branches.
"""
import ctypes as C
import math

import mpmath as mp
import numpy as np
import pytest
import scipy.integrate as si
import scipy.interpolate as sint
import scipy.special as sp

from oracle import oracle as O

L = O.lib()
DELTA = 0.5 * 5.991


def test_expint_ei_matches_scipy():
    xs = np.concatenate([np.geomspace(1e-4, 0.3, 60), np.linspace(0.45, 100.0, 800)])
    ours = np.array([L.orl2_expint_ei(x) for x in xs])
    # scipy's expi itself is only good to ~3e-14 near x = 40; mpmath pins the last digits
    np.testing.assert_allclose(ours, sp.expi(xs), rtol=5e-14, atol=0)
    for x in np.linspace(0.45, 100.0, 120):
        assert L.orl2_expint_ei(x) == pytest.approx(float(mp.ei(x)), rel=3e-15)


@pytest.mark.parametrize("x", [1e-4, 0.5, 2.9955, 3.0, 10.0, 39.9, 40.1, 77.7, 99.99])
def test_expint_ei_matches_mpmath(x):
    assert L.orl2_expint_ei(x) == pytest.approx(float(mp.ei(x)), rel=1e-14)


def test_marg_position_branches():
    f = L.orl2_marg_position
    # unconstrained position: no shape information (likelihood.c:392-396)
    assert f(DELTA) == -1e10 and f(1.0) == -1e10
    # x too close to Delta is clamped to Delta + 1e-4 (:405)
    assert f(DELTA + 1e-6) == f(DELTA + 1e-4)

    def exact(x):
        x = mp.mpf(x)
        d = mp.mpf(DELTA)
        return float(mp.log((mp.ei(x) - mp.ei(x - d)) / mp.log(x / (x - d))))

    for x in (3.5, 10.0, 50.0, 100.0):
        assert f(x) == pytest.approx(exact(x), rel=1e-13, abs=1e-13)
    assert f(100.0) == pytest.approx(98.844926, abs=1e-6)
    # asymptotic branch above 100 (:399-403): x - ln(Delta) - exp(-Delta)
    for x in (100.0000001, 150.0, 1e4):
        assert f(x) == pytest.approx(x - math.log(DELTA) - math.exp(-DELTA), rel=1e-15)
    assert f(100.0000001) == pytest.approx(98.852877, abs=1e-6)


def _qk41(fun, a, b):
    cb = O.F1_CB(lambda x, p: fun(x))
    r, e, ra, rs = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    L.orl2_qk41(cb, None, a, b, C.byref(r), C.byref(e), C.byref(ra), C.byref(rs))
    return r.value, e.value, ra.value, rs.value


def test_gauss_kronrod_41_exact_to_degree_61():
    for k in range(62):
        val, _, _, _ = _qk41(lambda x, k=k: x ** k, -1.0, 1.0)
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        assert abs(val - exact) <= 4e-16, k


def test_gauss_kronrod_error_estimate_is_embedded_gauss20():
    """abserr = rescale(|K - G|) with G the 20-point Gauss-Legendre rule (QUADPACK qk41)."""
    xg, wg = np.polynomial.legendre.leggauss(20)  # independent Gauss-Legendre nodes
    for k in (40, 46, 50, 60):
        _, err, resabs, resasc = _qk41(lambda x, k=k: x ** k, -1.0, 1.0)
        eg = abs((wg * xg ** k).sum() - 2.0 / (k + 1))  # Kronrod is exact here
        scale = (200 * eg / resasc) ** 1.5
        want = max(resasc * min(1.0, scale), 50 * 2.220446049250313e-16 * resabs)
        assert err == pytest.approx(want, rel=2e-3), k
    # below degree 40 both rules are exact and only the roundoff floor remains
    _, err, resabs, _ = _qk41(lambda x: x ** 30, -1.0, 1.0)
    assert err == pytest.approx(50 * 2.220446049250313e-16 * resabs, rel=1e-12)


def test_qag_matches_mpmath_quad():
    cases = [(lambda x: math.exp(-x * x), -2.0, 3.0, lambda x: mp.exp(-x * x)),
             (lambda x: 1.0 / (1e-3 + x * x), -1.0, 1.0, lambda x: 1 / (mp.mpf('1e-3') + x * x)),
             (lambda x: math.sqrt(abs(x)), -1.0, 2.0, lambda x: mp.sqrt(abs(x)))]
    for f, a, b, fm in cases:
        cb = O.F1_CB(lambda x, p, f=f: f(x))
        r, e = C.c_double(), C.c_double()
        nint = C.c_ulong()
        st = L.orl2_qag41(cb, None, a, b, 0.0, 1e-10, 1000, C.byref(r), C.byref(e), C.byref(nint))
        exact = float(mp.quad(fm, [a, 0, b] if a < 0 < b else [a, b]))
        assert st == 0
        assert r.value == pytest.approx(exact, rel=1e-9)
        assert e.value <= 1e-10 * abs(r.value) * 1.0001


def _spline_eval(x, y, xq, linear=False):
    out = np.zeros(len(xq))
    st = L.orl2_cspline_eval_many(len(x), O.ptr(x), O.ptr(y), len(xq), O.ptr(xq), O.ptr(out),
                                  int(linear))
    assert st == 0
    return out


def test_natural_cubic_spline_matches_scipy():
    rng = np.random.default_rng(0)
    for n in (3, 5, 11, 23):
        x = np.sort(rng.uniform(0.3, 3.8, n))
        y = rng.normal(size=n) * 5
        xq = np.concatenate([x, rng.uniform(x[0], x[-1], 200)])
        ref = sint.CubicSpline(x, y, bc_type="natural")(xq)
        np.testing.assert_allclose(_spline_eval(x, y, xq), ref, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(_spline_eval(x, y, xq, linear=True), np.interp(xq, x, y),
                                   rtol=1e-13, atol=1e-13)


def _oracle_rmarg(x, y):
    xx, yy = np.array(x, float), np.array(y, float)
    return L.orl2_marginalise_posterior_r(len(xx), O.ptr(xx), O.ptr(yy)), yy


def test_marginalise_posterior_r_matches_quadrature():
    ro, numr = O.r_grid()
    x = ro[1:numr]
    rng = np.random.default_rng(5)
    for trial in range(6):
        mu, s = rng.uniform(0.5, 2.5), rng.uniform(0.2, 1.0)
        y = -0.5 * ((x - mu) / s) ** 2 + rng.uniform(-30, 40)
        got, yshift = _oracle_rmarg(x, y)
        ymax = max(-10.0, y.max())
        spl = sint.CubicSpline(x, y - ymax, bc_type="natural")
        ref = math.log(si.quad(lambda t: math.exp(spl(t)), x[0], x[-1], epsabs=0,
                               epsrel=1e-12, limit=500)[0]) + ymax
        assert got == pytest.approx(ref, abs=2e-3)  # QAG epsrel 1e-3 -> |d log I| <= 1e-3
        np.testing.assert_allclose(yshift, y - ymax)  # y is shifted in place (:117)


def test_marginalise_posterior_r_ymax_floor():
    # ymax starts at -10 (marginalise_r.c:94): for log-posteriors below -10 nothing is shifted
    ro, numr = O.r_grid()
    x = ro[1:numr]
    y = np.full(len(x), -50.0)
    got, yshift = _oracle_rmarg(x, y)
    np.testing.assert_array_equal(yshift, y + 10.0)
    assert got == pytest.approx(-50.0 + math.log(x[-1] - x[0]), rel=1e-12)


def _qag(f, a, b):
    cb = O.F1_CB(lambda t, p: f(t))
    r, e, n = C.c_double(), C.c_double(), C.c_ulong()
    st = L.orl2_qag41(cb, None, a, b, 0.0, 1e-3, 1000, C.byref(r), C.byref(e), C.byref(n))
    return st, r.value


def test_marginalise_posterior_r_retries_with_linear_interpolation():
    # marginalise_r.c:171-210: when QAG on the cubic spline fails, the integral is taken again
    # on the linear interpolant.  Log-posteriors alternating 0 / -40000 make the natural cubic
    # overshoot beyond exp's range (scipy's spline, not the oracle's, shows it), so QAG on it
    # fails; QAG on numpy's linear interpolant succeeds, and the oracle returns that integral.
    ro, numr = O.r_grid()
    x = ro[1:numr]
    y = np.where(np.arange(len(x)) % 2 == 0, 0.0, -40000.0)
    spl = sint.CubicSpline(x, y, bc_type="natural")
    assert spl(np.linspace(x[0], x[-1], 20001)).max() > math.log(np.finfo(float).max)
    with np.errstate(over="ignore"):
        st_cubic, _ = _qag(lambda t: float(np.exp(spl(t))), x[0], x[-1])
    assert st_cubic != 0
    st_lin, lin = _qag(lambda t: math.exp(float(np.interp(t, x, y))), x[0], x[-1])
    assert st_lin == 0 and lin > 0
    got, _ = _oracle_rmarg(x, y)
    assert got == pytest.approx(math.log(lin), abs=1e-9)  # ymax = 0


def test_set_posterior_values_drops_invalid():
    ro, numr = O.r_grid()
    Lr = np.linspace(-5, 5, 40)
    Lr[[3, 7, 8]] = -1e10
    rp = np.arange(40) * 0.1
    xm, ym = np.zeros(40), np.zeros(40)
    n = C.c_int()
    L.orl2_set_posterior_values(numr, O.ptr(Lr), O.ptr(rp), O.ptr(np.ascontiguousarray(
        np.pad(ro, (0, 40 - numr)))), O.ptr(xm), O.ptr(ym), C.byref(n))
    keep = [k for k in range(1, numr) if k not in (3, 7, 8)]
    assert n.value == len(keep)
    np.testing.assert_array_equal(xm[:n.value], ro[keep])
    np.testing.assert_array_equal(ym[:n.value], Lr[keep] + rp[keep])


def test_scalelength_grid_closed_form():
    ro, numr = O.r_grid()
    assert numr == 24  # 23 models (RadioLensfit2.cpp:370-392 stops once Ro > RMAX)
    r = [0.0, 0.3]
    while len(r) < 29 and r[-1] < 3.5:
        r.append(r[-1] + 0.08 + 0.4 * ((r[-1] - 0.3) / 3.2) ** 2)
    np.testing.assert_allclose(ro, r, rtol=1e-15)
    np.testing.assert_allclose(ro[[1, 2, 3, 4, 21, 22, 23]],
                               [0.3, 0.38, 0.4602, 0.5413, 2.9641, 3.3213, 3.7579], atol=5e-5)


def test_facet_size_closed_form():
    f = L.orl2_facet_size
    assert [f(r, 77387.0) for r in (0.5, 1.0, 2.0, 3.5)] == [16, 30, 58, 100]
    for len_, r in [(1.18e6, 1.0), (9.7e5, 3.5), (12345.6, 0.3)]:
        assert f(r, len_) == 2 * math.ceil(len_ * 38 * r * math.pi / 648000.0)


def test_priors():
    mu = L.orl2_scale_mean(50.0)
    assert mu == pytest.approx(-0.93 + 0.33 * math.log(50.0), rel=1e-15)
    # rfunc is the log of a log-normal density in r up to a constant
    r = np.array([0.3, 0.9, 2.7])
    lp = np.array([L.orl2_rfunc(mu, 0.3136, x) for x in r])
    import scipy.stats as ss
    ref = np.log(ss.lognorm.pdf(r, s=0.3136, scale=math.exp(mu))) + 0.5 * math.log(2 * math.pi)
    np.testing.assert_allclose(lp, ref, rtol=1e-13)
    assert math.isnan(L.orl2_rfunc(mu, 0.3136, 0.0))
