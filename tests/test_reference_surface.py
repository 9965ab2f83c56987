"""librl2hip's reference-signature functions (include/rl2_reference.h), called through ctypes
with the reference's own argument lists, against the FP64 oracle on the same inputs.

Each call goes through the binding a RadioLensfit2 build would use (a likelihood_params
struct, caller-owned host arrays, in-place outputs).  Tolerances: facet cell counts, cell
order and coordinates bit-exact; gridded sums 1e-11 relative; log-likelihoods 1e-8 absolute;
fitted ellipticities 1e-8; the whole loop 1e-6 in e and 1e-10 relative in the residual; the
output file of data_processing line by line (the reference's "%f" columns, so 6 decimals).
"""
import ctypes as C
import math
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from radiolensfit2_amd import _native as N
from radiolensfit2_amd import pipeline as P
from tests import _common as T

pytestmark = pytest.mark.gpu

ARCS2RAD = math.pi / 648000.0


@pytest.fixture(scope="module")
def case(gpu):
    return T.small_case()


class Par:
    """One likelihood_params (datatype.h:49-74) over numpy buffers sized for facet_size(RMAX),
    viewed both as the library's struct and as the oracle's."""

    def __init__(self, obs, fmax):
        ro, numr = O.r_grid()
        self.ro = np.zeros(40)
        self.ro[:numr] = ro
        self.numr = numr
        self.rprior = np.zeros(40)
        n = fmax * fmax
        self.data = np.zeros(n, np.complex128)
        self.sigma2 = np.zeros(n)
        self.uu = np.zeros(n)
        self.vv = np.zeros(n)
        self.mod = np.zeros(max(1, (numr - 1) * n))
        self.sum_w = np.zeros(n)
        self.wn, self.spec = obs.wn, obs.spec
        self.nch = obs.nch

    def set_prior(self, flux):
        L = O.lib()
        mu = L.orl2_scale_mean(flux)
        for k in range(1, self.numr):
            self.rprior[k] = L.orl2_rfunc(mu, 0.3136, self.ro[k])

    def _fill(self, p, cptr, facet, n):
        p.numr = self.numr
        p.ro = O.ptr(self.ro)
        p.rprior = O.ptr(self.rprior)
        p.nchannels = self.nch
        p.spec = O.ptr(self.spec)
        p.wavenumbers = O.ptr(self.wn)
        p.uu = O.ptr(self.uu)
        p.vv = O.ptr(self.vv)
        p.data = cptr(self.data)
        p.sigma2 = O.ptr(self.sigma2)
        p.mod = O.ptr(self.mod)
        p.facet = facet
        p.ncoords = n
        return p

    def lib(self, facet=0, n=0):
        return self._fill(N.LikelihoodParams(), N.cplx_s, facet, n)

    def oracle(self, facet=0, n=0):
        return self._fill(O.Params(), O.cptr, facet, n)


def _extract_dropin(obs, cat, g, par, vd, vs):
    """data_processing.c:68-84: prior, facet size, source_extraction, evaluate_facet_coords."""
    H = N.hip()
    par.set_prior(cat.flux[g])
    facet = H.facet_size(cat.scale[g], obs.len)
    lp = par.lib(facet)
    gal = np.zeros(obs.nvis, np.complex128)
    H.source_extraction(0, facet, C.byref(lp), lp.data, O.ptr(par.sigma2), O.ptr(par.sum_w),
                        cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], 0.0, 0.0, N.cplx_s(vs),
                        N.cplx_s(vd), N.cplx_s(gal), O.ptr(obs.sigma2, O.c_float_p),
                        O.ptr(obs.flag, O.c_bool_p), obs.nrow, O.ptr(obs.u), O.ptr(obs.v),
                        O.ptr(obs.w), obs.len)
    n = H.evaluate_facet_coords(O.ptr(par.uu), O.ptr(par.vv), obs.len, facet, O.ptr(par.sum_w))
    lp.ncoords = n
    return facet, n, lp


@pytest.mark.parametrize("g", [0, 11])
def test_source_extraction_and_facet_coords(case, g):
    obs, cat, vs = case
    fmax = N.hip().facet_size(3.5, obs.len)
    par = Par(obs, fmax)
    facet, n, lp = _extract_dropin(obs, cat, g, par, obs.vis, vs)
    ofacet, fv, fs, uu, vv, sw = T.oracle_extract(obs, cat, g, obs.vis, vs)
    assert (facet, n) == (ofacet, len(fv))
    np.testing.assert_array_equal(par.sum_w[:facet * facet], sw)
    np.testing.assert_array_equal(par.uu[:n], uu)
    np.testing.assert_array_equal(par.vv[:n], vv)
    np.testing.assert_allclose(par.sigma2[:n], fs, rtol=1e-13, atol=0)
    assert np.abs(par.data[:n] - fv).max() <= 1e-11 * np.abs(fv).max()


@pytest.mark.parametrize("resident", [True, False])
def test_loglikelihood_fit_and_sampling(case, resident):
    """loglikelihood / f_likelihood / source_fitting / likelihood_sampling on the facet the
    last source_extraction left on the device (resident) and on a host-only copy of it."""
    obs, cat, vs = case
    H = N.hip()
    g = 3
    fmax = H.facet_size(3.5, obs.len)
    par = Par(obs, fmax)
    facet, n, lp = _extract_dropin(obs, cat, g, par, obs.vis, vs)
    if not resident:
        other = Par(obs, fmax)
        other.rprior[:] = par.rprior
        for k in ("data", "sigma2", "uu", "vv"):
            getattr(other, k)[:] = getattr(par, k)
        par = other
        lp = par.lib(facet, n)
    op = par.oracle(facet, n)
    rng = np.random.default_rng(11)
    pts = rng.uniform(-0.55, 0.55, (10, 2))
    for e1, e2 in pts:
        err, oerr = C.c_int(-1), C.c_int(-1)
        L = H.loglikelihood(C.byref(lp), e1, e2, C.byref(err))
        oL = O.lib().orl2_loglikelihood(C.byref(op), e1, e2, C.byref(oerr))
        assert err.value == oerr.value
        assert abs(L - oL) <= 1e-8, (e1, e2, L, oL)
        x = np.array([e1, e2])
        gv = N.GslVector(2, 1, O.ptr(x), None, 0)
        f = H.f_likelihood(C.byref(gv), C.byref(lp))
        assert f == (-L if e1 * e1 + e2 * e2 < 0.65 else 1e10)
    out = [C.c_double() for _ in range(6)]
    oout = [C.c_double() for _ in range(6)]
    err = H.source_fitting(0, C.byref(lp), *[C.byref(v) for v in out])
    oerr = O.lib().orl2_source_fitting(C.byref(op), *[C.byref(v) for v in oout])
    assert err == oerr
    a, b = [v.value for v in out], [v.value for v in oout]
    assert abs(a[0] - b[0]) <= 1e-8 and abs(a[1] - b[1]) <= 1e-8, (a, b)
    np.testing.assert_allclose(a[2:5], b[2:5], rtol=1e-6, atol=1e-12)
    assert abs(a[5] - b[5]) <= 1e-8
    # likelihood_sampling alone, from the simplex's end point
    m1, m2 = C.c_double(b[0] + 0.01), C.c_double(b[1] - 0.02)
    om1, om2 = C.c_double(m1.value), C.c_double(m2.value)
    err0 = C.c_int()
    maxL = O.lib().orl2_loglikelihood(C.byref(op), m1.value, m2.value, C.byref(err0))
    v = [C.c_double() for _ in range(3)]
    ov = [C.c_double() for _ in range(3)]
    H.likelihood_sampling(C.byref(m1), C.byref(m2), maxL, C.byref(lp), 30, *[C.byref(x) for x in v])
    O.lib().orl2_likelihood_sampling(C.byref(om1), C.byref(om2), maxL, C.byref(op), 30,
                                     *[C.byref(x) for x in ov])
    assert abs(m1.value - om1.value) <= 1e-8 and abs(m2.value - om2.value) <= 1e-8
    np.testing.assert_allclose([x.value for x in v], [x.value for x in ov], rtol=1e-6, atol=1e-12)


def test_loglikelihood_r_cross_correlation_and_model(case):
    obs, cat, vs = case
    H = N.hip()
    g = 1
    facet, fv, fs, uu, vv, sw = T.oracle_extract(obs, cat, g, obs.vis, vs)
    n = len(fv)
    for e1, e2, r in ((0.1, -0.2, 0.54), (-0.3, 0.05, 1.7), (0.0, 0.0, 3.3)):
        mod, omod = np.zeros(n), np.zeros(n)
        L = H.loglikelihood_r(e1, e2, r, n, O.ptr(fs), O.ptr(uu), O.ptr(vv), N.cplx_s(fv),
                              O.ptr(mod))
        nit = C.c_int()
        oL = O.lib().orl2_loglikelihood_r(e1, e2, r, n, O.ptr(fs), O.ptr(uu), O.ptr(vv), O.cptr(fv),
                                          O.ptr(omod), C.byref(nit))
        assert abs(L - oL) <= 1e-8, (e1, e2, r, L, oL)
        np.testing.assert_allclose(mod, omod, rtol=1e-13, atol=0)
        # model_galaxy_visibilities_at_zero alone (galaxy_visibilities.c:59-91)
        m2 = np.zeros(n)
        H.model_galaxy_visibilities_at_zero(e1, e2, r, n, O.ptr(uu), O.ptr(vv), O.ptr(fs), O.ptr(m2))
        np.testing.assert_allclose(m2, omod, rtol=1e-13, atol=0)
        # cross_correlation on the data / variance and that model (likelihood.c:195-358)
        dn = np.ascontiguousarray(fv)
        ho, det = C.c_double(), C.c_double()
        oho, odet = C.c_double(), C.c_double()
        st = H.cross_correlation(n, O.ptr(uu), O.ptr(vv), N.cplx_s(dn), O.ptr(omod), C.byref(ho),
                                 C.byref(det))
        ost = O.lib().orl2_cross_correlation(n, O.ptr(uu), O.ptr(vv), O.cptr(dn), O.ptr(omod),
                                             C.byref(oho), C.byref(odet), C.byref(nit))
        assert st == ost
        assert abs(ho.value - oho.value) <= 1e-10 * max(1.0, abs(oho.value))
        assert abs(det.value - odet.value) <= 1e-9 * abs(odet.value)


def test_scalelength_marginalisation(case):
    H = N.hip()
    ro, numr = O.r_grid()
    rp = O.rprior_for(60.0, np.pad(ro, (0, 40 - numr)), numr)
    rng = np.random.default_rng(2)
    Lr = np.full(40, -1e10)
    Lr[1:numr] = 30.0 - 8.0 * (np.log(ro[1:numr]) - 0.1) ** 2 + rng.normal(0, 0.05, numr - 1)
    Lr[3] = -1e10
    ro40 = np.pad(ro, (0, 40 - numr))
    x, y, ox, oy = np.zeros(40), np.zeros(40), np.zeros(40), np.zeros(40)
    num, onum = C.c_int(), C.c_int()
    H.set_posterior_values(numr, O.ptr(Lr), O.ptr(rp), O.ptr(ro40), O.ptr(x), O.ptr(y), C.byref(num))
    O.lib().orl2_set_posterior_values(numr, O.ptr(Lr), O.ptr(rp), O.ptr(ro40), O.ptr(ox), O.ptr(oy),
                                      C.byref(onum))
    assert num.value == onum.value == numr - 2
    np.testing.assert_array_equal(x, ox)
    np.testing.assert_array_equal(y, oy)
    k = num.value
    val = H.marginalise_posterior_r(k, O.ptr(x), O.ptr(y))
    oval = O.lib().orl2_marginalise_posterior_r(k, O.ptr(ox), O.ptr(oy))
    assert abs(val - oval) <= 1e-9 * max(1.0, abs(oval))
    np.testing.assert_array_equal(y[:k], oy[:k])            # y shifted by ymax in place
    # marf: exp of the natural cubic spline through the shifted values, 0 outside the range
    xq = np.linspace(x[0] - 0.1, x[k - 1] + 0.1, 57)
    sq = np.zeros_like(xq)
    O.lib().orl2_cspline_eval_many(k, O.ptr(ox), O.ptr(oy), len(xq), O.ptr(xq), O.ptr(sq), 0)
    for xi, si in zip(xq, sq):
        want = math.exp(si) if x[0] <= xi <= x[k - 1] else 0.0
        assert abs(H.marf(xi, None) - want) <= 1e-12 * max(1.0, want)


def test_visibility_model_and_phase_shift(case):
    obs, cat, vs = case
    H = N.hip()
    u, v, w = obs.u, obs.v, obs.w
    n = obs.nrow
    for c in range(obs.nch):
        for (e1, e2, s, f, l, m) in ((0.0, 0.0, 1.2, 150.0, 1e-3, -2e-3),
                                     (0.3, -0.25, 0.6, 40.0, -2.5e-3, 1.5e-3)):
            vis, ovis = np.zeros(n, np.complex128), np.zeros(n, np.complex128)
            H.data_galaxy_visibilities(obs.spec[c], obs.wn[c], 0.0, 0.0, e1, e2, s, f, l, m, n,
                                       O.ptr(u), O.ptr(v), O.ptr(w), N.cplx_s(vis))
            O.lib().orl2_data_galaxy_vis(obs.spec[c], obs.wn[c], e1, e2, s, f, l, m, n, O.ptr(u),
                                         O.ptr(v), O.ptr(w), O.cptr(ovis))
            assert np.abs(vis - ovis).max() <= 1e-12 * f
            H.data_visibilities_phase_shift(obs.wn[c], l, m, n, O.ptr(u), O.ptr(v), O.ptr(w),
                                            N.cplx_s(vis))
            O.lib().orl2_phase_shift(obs.wn[c], l, m, n, O.ptr(u), O.ptr(v), O.ptr(w), O.cptr(ovis))
            assert np.abs(vis - ovis).max() <= 1e-12 * f


@pytest.mark.parametrize("init", ["zero", "accumulate"])
def test_sky_model(case, init):
    obs, cat, vs = case
    H = N.hip()
    rng = np.random.default_rng(5)
    vm = np.zeros(obs.nvis, np.complex128)
    if init == "accumulate":
        vm = (rng.normal(size=obs.nvis) + 1j * rng.normal(size=obs.nvis)) * 100.0
    want = vm.copy()
    gal = np.zeros(obs.nvis, np.complex128)
    O.lib().orl2_sky_model(O.ptr(obs.wn), O.ptr(obs.spec), obs.nch, len(cat), O.ptr(cat.flux),
                           O.ptr(cat.scale), O.ptr(cat.l), O.ptr(cat.m), obs.nrow, O.ptr(obs.u),
                           O.ptr(obs.v), O.ptr(obs.w), O.cptr(gal), O.cptr(want))
    H.sky_model(O.ptr(obs.wn), O.ptr(obs.spec), 60e6, 480, obs.nch, len(cat), O.ptr(cat.flux),
                O.ptr(cat.scale), O.ptr(cat.l), O.ptr(cat.m), obs.nrow, O.ptr(obs.u), O.ptr(obs.v),
                O.ptr(obs.w), N.cplx_s(gal), N.cplx_s(vm))
    assert np.abs(vm - want).max() <= 1e-12 * np.abs(want).max()


def _libc():
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    return libc


def _expected_lines(cat, snr, ores):
    """data_processing.c:102,136: the first pass writes every good fit in catalog order; the
    re-fit pass writes every galaxy of the bad list, with e = 0 when it failed again."""
    def line(g):
        r = ores[g]
        return [cat.flux[g], cat.e1[g], r.e1, math.sqrt(r.var_e1) if r.var_e1 >= 0 else math.nan,
                cat.e2[g], r.e2, math.sqrt(r.var_e2) if r.var_e2 >= 0 else math.nan, r.oneDimvar,
                snr[g], cat.l[g] / ARCS2RAD, cat.m[g] / ARCS2RAD]
    first = [line(g) for g in range(len(cat)) if ores[g].pass_ == 0]
    second = [line(g) for g in range(len(cat)) if ores[g].pass_ == 1]
    return first, second


def _read_lines(path):
    rows = []
    for ln in open(path).read().splitlines():
        f = [x.strip() for x in ln.split("|") if x.strip()]
        rows.append([float(x) for x in f])
    return rows


def test_data_processing_writes_the_reference_lines(case):
    """RadioLensfit2.cpp:597-606 calls data_processing twice (the list, then the bad list);
    on return visData holds the residual visData - visSkyMod and visSkyMod is zero."""
    obs, cat, vs = case
    H = N.hip()
    ores, ost, ores_vis = T.oracle_process(obs, cat, vs)
    fmax = H.facet_size(3.5, obs.len)
    par = Par(obs, fmax)
    lp = par.lib()
    vd = obs.vis.copy()
    vsk = vs.copy()
    gal = np.zeros(obs.nvis, np.complex128)
    snr = np.linspace(5.0, 50.0, len(cat))
    bad_list = np.zeros(len(cat), np.uint32)
    bad = C.c_int()
    fit_t = C.c_double(0.0)
    libc = _libc()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "ellipticities.txt")
        fp = libc.fopen(path.encode(), b"w")
        assert fp
        args = lambda re, lst, n: (re, lst.ctypes.data_as(N.c_uint_p), n, obs.len, obs.nrow, fp,  # noqa: E731
                                   C.byref(lp), O.ptr(cat.l), O.ptr(cat.m), O.ptr(cat.flux),
                                   O.ptr(cat.scale), O.ptr(cat.e1), O.ptr(cat.e2), O.ptr(snr),
                                   O.ptr(par.sum_w), N.cplx_s(gal), N.cplx_s(vsk), N.cplx_s(vd),
                                   O.ptr(obs.sigma2, O.c_float_p), O.ptr(obs.flag, O.c_bool_p),
                                   O.ptr(obs.u), O.ptr(obs.v), O.ptr(obs.w), C.byref(fit_t),
                                   C.byref(bad))
        H.data_processing(*args(False, bad_list, len(cat)))
        nbad = bad.value
        assert nbad == ost.n_bad_first
        assert sorted(bad_list[:nbad]) == [g for g in range(len(cat)) if ores[g].pass_ == 1]
        H.data_processing(*args(True, bad_list, nbad))
        assert bad.value == ost.n_bad_final
        libc.fclose(fp)
        rows = _read_lines(path)
    first, second = _expected_lines(cat, snr, ores)
    assert len(rows) == len(first) + len(second)
    for got, want in zip(rows, first + second):
        for k, (a, b) in enumerate(zip(got, want)):
            if k in (3, 6, 7) and (math.isnan(a) or math.isnan(b)):
                # a square root of a variance that cancels to zero: its rounding residue may
                # be +-1e-17 on either side, printed as 0.000000 or nan
                other = b if math.isnan(a) else a
                assert math.isnan(other) or abs(other) <= 1.5e-6, (k, got, want)
            else:
                assert abs(a - b) <= 1.5e-6 + 1e-9 * abs(b), (k, got, want)
    assert np.all(vsk == 0)
    assert np.abs(vd - ores_vis).max() <= 1e-10 * np.abs(ores_vis).max()
    assert fit_t.value > 0


def test_loglikelihood_refuses_a_scalelength_grid_beyond_the_device_buffers(case):
    """numr outside [2, 32] (the device holds 32 scalelength points per request) is refused
    with an error instead of being indexed past the buffers; a valid call still works after."""
    obs, cat, vs = case
    H = N.hip()
    fmax = H.facet_size(3.5, obs.len)
    par = Par(obs, fmax)
    facet, n, lp = _extract_dropin(obs, cat, 3, par, obs.vis, vs)
    for bad in (40, 33, 1):
        lp.numr = bad
        err = C.c_int(-1)
        L = H.loglikelihood(C.byref(lp), 0.1, -0.05, C.byref(err))
        assert L == -1.0e10 and err.value == 1, (bad, L, err.value)
        assert b"numr" in H.rl2_last_error()
    lp.numr = par.numr
    err = C.c_int(-1)
    L = H.loglikelihood(C.byref(lp), 0.1, -0.05, C.byref(err))
    op = par.oracle(facet, n)
    oerr = C.c_int(-1)
    oL = O.lib().orl2_loglikelihood(C.byref(op), 0.1, -0.05, C.byref(oerr))
    assert err.value == oerr.value and abs(L - oL) <= 1e-8
