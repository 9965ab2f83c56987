"""HIP path vs the FP64 oracle on seeded synthetic observations (FP64 precision mode).

Tolerances: cell indices, cell order, weights and facet coordinates are bit-exact; gridded
sums agree to ~1e-12 relative (the oracle's sequential rounding and last-ulp sin/cos
differences; the device's integer accumulation is exact and run-to-run identical);
log-likelihoods to 1e-8 absolute; fitted ellipticities to 1e-8 (identical simplex
trajectory); the whole per-galaxy loop to 1e-6 in e (residual differences accumulate
through the subtractions).
"""
import numpy as np
import pytest

from radiolensfit2_amd import _native as N
from radiolensfit2_amd import pipeline as P
from tests import _common as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    return T.small_case()


@pytest.fixture(scope="module")
def dev(gpu, case):
    obs, cat, vs = case
    d = P.DeviceObservation(obs, vis_sky=vs)
    yield d
    d.close()


def test_residual_roundtrip(dev, case):
    obs, cat, vs = case
    r = dev.residual()
    np.testing.assert_array_equal(r, obs.vis - vs)


@pytest.mark.parametrize("g", [0, 5, 10, 15])
def test_extraction_matches_oracle(dev, case, g):
    obs, cat, vs = case
    facet, fv, fs, uu, vv, sw = T.oracle_extract(obs, cat, g, obs.vis, vs)
    gv, gs, guu, gvv, gsw = dev.extract(cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], facet)
    assert len(gv) == len(fv)
    np.testing.assert_array_equal(gsw, sw)       # per-cell counts
    np.testing.assert_array_equal(guu, uu)       # cell centres and compacted order
    np.testing.assert_array_equal(gvv, vv)
    np.testing.assert_allclose(gs, fs, rtol=1e-13, atol=0)
    scale = np.abs(fv).max()
    assert np.abs(gv - fv).max() <= 1e-11 * scale


def test_likelihood_matches_oracle(dev, case):
    obs, cat, vs = case
    g = 0
    facet, fv, fs, uu, vv, sw = T.oracle_extract(obs, cat, g, obs.vis, vs)
    params, keep = T.oracle_params(fv, fs, uu, vv, cat.flux[g])
    ro = P.r_grid()
    dev.load_facet(fv, fs, uu, vv, slot=1)
    dev.set_prior(P.rprior_for(cat.flux[g], ro), slot=1)
    rng = np.random.default_rng(3)
    e = rng.uniform(-0.6, 0.6, size=(24, 2))
    e = e[(e ** 2).sum(1) < 0.65]
    Le, err, Lr, nit = dev.loglik(e[:, 0], e[:, 1], slot=1, detail=True)
    from oracle import oracle as O
    import ctypes as C
    for k in range(len(e)):
        oLr = np.zeros(40)
        onit = np.zeros(40, np.int32)
        oerr = C.c_int()
        oLe = O.lib().orl2_loglikelihood_detail(C.byref(params), e[k, 0], e[k, 1], C.byref(oerr),
                                                O.ptr(oLr), onit.ctypes.data_as(O.c_int_p))
        assert err[k] == oerr.value
        np.testing.assert_array_equal(nit[k, 1:24], onit[1:24])
        np.testing.assert_allclose(Lr[k, 1:24], oLr[1:24], rtol=0, atol=1e-8)
        assert abs(Le[k] - oLe) <= 1e-8, (k, Le[k], oLe)


@pytest.mark.parametrize("g", [2, 9])
def test_likelihood_on_extracted_facet_matches_oracle(dev, case, g):
    """A facet extracted on the device keeps its grid structure, so the FP64 Newton passes
    take their phases from per-pass column x row tables.  On the device's own facet (copied
    back and given to the oracle) L_r, the Newton step counts and L_e match the oracle, and
    the per-cell sin/cos path (RL2_TUNE_PHASE_TABLES 0) gives the same values and fit."""
    obs, cat, vs = case
    facet = N.hip().facet_size(cat.scale[g], obs.len)
    gv, gs, guu, gvv, gsw = dev.extract(cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], facet,
                                        slot=3)
    params, keep = T.oracle_params(gv, gs, guu, gvv, cat.flux[g])
    dev.set_prior(P.rprior_for(cat.flux[g], P.r_grid()), slot=3)
    rng = np.random.default_rng(30 + g)
    e = rng.uniform(-0.6, 0.6, size=(24, 2))
    e = e[(e ** 2).sum(1) < 0.65]
    print(f"extracted facet {facet} of galaxy {g}: {len(gv)} cells, {len(e)} points", flush=True)
    Le, err, Lr, nit = dev.loglik(e[:, 0], e[:, 1], slot=3, detail=True)
    from oracle import oracle as O
    import ctypes as C
    for k in range(len(e)):
        oLr = np.zeros(40)
        onit = np.zeros(40, np.int32)
        oerr = C.c_int()
        oLe = O.lib().orl2_loglikelihood_detail(C.byref(params), e[k, 0], e[k, 1], C.byref(oerr),
                                                O.ptr(oLr), onit.ctypes.data_as(O.c_int_p))
        assert err[k] == oerr.value
        np.testing.assert_array_equal(nit[k, 1:24], onit[1:24])
        np.testing.assert_allclose(Lr[k, 1:24], oLr[1:24], rtol=0, atol=1e-8)
        assert abs(Le[k] - oLe) <= 1e-8, (k, Le[k], oLe)
    fit_t = dev.fit(slot=3)
    H = N.hip()
    mode = C.c_int()
    N.check(H.rl2_tune_get(N.RL2_TUNE_PHASE_TABLES, C.byref(mode)), "rl2_tune_get")
    # the row tables read from LDS (2) or from global memory (1): the same table entries,
    # so the same values
    N.check(H.rl2_tune_set(N.RL2_TUNE_PHASE_TABLES, 3 - mode.value), "rl2_tune_set")
    try:
        Le2, err2, Lr2, nit2 = dev.loglik(e[:, 0], e[:, 1], slot=3, detail=True)
    finally:
        N.check(H.rl2_tune_set(N.RL2_TUNE_PHASE_TABLES, mode.value), "rl2_tune_set")
    np.testing.assert_array_equal(err2, err)
    np.testing.assert_array_equal(nit2, nit)
    np.testing.assert_array_equal(Lr2[:, 1:24], Lr[:, 1:24])
    np.testing.assert_array_equal(Le2, Le)
    N.check(H.rl2_tune_set(N.RL2_TUNE_PHASE_TABLES, 0), "rl2_tune_set")
    try:
        Le0, err0, Lr0, nit0 = dev.loglik(e[:, 0], e[:, 1], slot=3, detail=True)
        fit_0 = dev.fit(slot=3)
    finally:
        N.check(H.rl2_tune_set(N.RL2_TUNE_PHASE_TABLES, mode.value), "rl2_tune_set")
    np.testing.assert_array_equal(err0, err)
    np.testing.assert_array_equal(nit0, nit)
    np.testing.assert_allclose(Lr0[:, 1:24], Lr[:, 1:24], rtol=0, atol=1e-9)
    assert fit_0.error == fit_t.error
    assert abs(fit_0.e1 - fit_t.e1) <= 1e-8 and abs(fit_0.e2 - fit_t.e2) <= 1e-8
    # the scalelength tasks in two groups on two streams: every task's chunk sums are added
    # in the same order, so the values are identical
    N.check(H.rl2_tune_set(N.RL2_TUNE_TILE_STREAMS, 2), "rl2_tune_set")
    try:
        Le1, err1, Lr1, nit1 = dev.loglik(e[:, 0], e[:, 1], slot=3, detail=True)
    finally:
        N.check(H.rl2_tune_set(N.RL2_TUNE_TILE_STREAMS, 1), "rl2_tune_set")
    np.testing.assert_array_equal(err1, err)
    np.testing.assert_array_equal(nit1, nit)
    np.testing.assert_array_equal(Lr1[:, 1:24], Lr[:, 1:24])
    np.testing.assert_array_equal(Le1, Le)
    # pass 0 with one or two tasks per work-group instead of three (RL2_TUNE_P0_TASKS): each
    # task's chunk sums are the same expressions in the same order, so the same values bit
    # for bit
    for tpw in (1, 2):
        N.check(H.rl2_tune_set(N.RL2_TUNE_P0_TASKS, tpw), "rl2_tune_set")
        try:
            Le2, err2, Lr2, nit2 = dev.loglik(e[:, 0], e[:, 1], slot=3, detail=True)
        finally:
            N.check(H.rl2_tune_set(N.RL2_TUNE_P0_TASKS, 3), "rl2_tune_set")
        np.testing.assert_array_equal(err2, err)
        np.testing.assert_array_equal(nit2, nit)
        np.testing.assert_array_equal(Lr2[:, 1:24], Lr[:, 1:24])
        np.testing.assert_array_equal(Le2, Le)
    # Newton steps in a step kernel after every pass (RL2_TUNE_FUSE_STEP 0), in the next
    # pass's work-groups (1), or every Newton pass of an evaluation in one launch (2); the
    # fused passes also queued one at a time (every evaluation then resumes from the step
    # kernel's state after each status check): the same sums, steps and tables, so the same
    # values bit for bit
    for fuse, passes in ((0, 0), (1, 1), (0, 1), (2, 0), (1, 0)):
        N.check(H.rl2_tune_set(N.RL2_TUNE_FUSE_STEP, fuse), "rl2_tune_set")
        N.check(H.rl2_tune_set(N.RL2_TUNE_TILE_PASSES, passes), "rl2_tune_set")
        try:
            Le3, err3, Lr3, nit3 = dev.loglik(e[:, 0], e[:, 1], slot=3, detail=True)
            fit_3 = dev.fit(slot=3)
        finally:
            N.check(H.rl2_tune_set(N.RL2_TUNE_FUSE_STEP, 1), "rl2_tune_set")
            N.check(H.rl2_tune_set(N.RL2_TUNE_TILE_PASSES, 0), "rl2_tune_set")
        np.testing.assert_array_equal(err3, err)
        np.testing.assert_array_equal(nit3, nit)
        np.testing.assert_array_equal(Lr3[:, 1:24], Lr[:, 1:24])
        np.testing.assert_array_equal(Le3, Le)
        assert (fit_3.error, fit_3.e1, fit_3.e2) == (fit_t.error, fit_t.e1, fit_t.e2), (fuse, passes)


@pytest.mark.parametrize("g", [0, 3, 7])
def test_fit_matches_oracle(dev, case, g):
    obs, cat, vs = case
    facet, fv, fs, uu, vv, sw = T.oracle_extract(obs, cat, g, obs.vis, vs)
    params, keep = T.oracle_params(fv, fs, uu, vv, cat.flux[g])
    oerr, (e1, e2, v1, v2, od, maxL) = T.oracle_fit(params)
    ro = P.r_grid()
    dev.load_facet(fv, fs, uu, vv, slot=1)
    dev.set_prior(P.rprior_for(cat.flux[g], ro), slot=1)
    # tiled likelihood without / with speculation, then one work-group per (e, r) task
    for spec, blk in ((0, 0), (1, 0), (1, 256)):
        r = dev.fit(slot=1, nm_speculate=spec, line_chunk=1 + 3 * spec, lik_block=blk)
        assert r.error == oerr
        assert abs(r.e1 - e1) <= 1e-8 and abs(r.e2 - e2) <= 1e-8, (r.e1, e1, r.e2, e2)
        assert abs(r.maxL - maxL) <= 1e-8
        np.testing.assert_allclose([r.var_e1, r.var_e2, r.oneDimvar], [v1, v2, od], rtol=1e-6,
                                   atol=1e-12)


@pytest.mark.parametrize("batch", [1, 4])
def test_process_matches_oracle(gpu, case, batch):
    obs, cat, vs = case
    ores, ost, ores_vis = T.oracle_process(obs, cat, vs, batch=batch)
    d = P.DeviceObservation(obs, vis_sky=vs)
    res, st = d.process(cat, batch=batch)
    resid = d.residual()
    d.close()
    assert st.n_bad_first == ost.n_bad_first and st.n_bad_final == ost.n_bad_final
    for g in range(len(cat)):
        a, b = res[g], ores[g]
        assert (a.error, a.pass_, a.subtracted) == (b.error, b.pass_, b.subtracted), g
        assert abs(a.e1 - b.e1) <= 1e-6 and abs(a.e2 - b.e2) <= 1e-6, (g, a.e1, b.e1, a.e2, b.e2)
    scale = np.abs(ores_vis).max()
    assert np.abs(resid - ores_vis).max() <= 1e-10 * scale


def test_wide_facet_fused_passes_match_step_kernels(gpu):
    """A facet wider than 1024 columns (a 3.4 arcsec galaxy on 500 km baselines: 1150) takes the
    fused Newton passes with column records and a cache of its chunk's row entries (or, for
    a chunk spanning more rows than the cache, row entries evaluated per cell); the values
    equal the step-kernel passes bit for bit and the oracle to 1e-8 with equal Newton step
    counts, at the default chunks and at 256-cell chunks."""
    from radiolensfit2_amd import synth
    cfg = synth.ArrayConfig(n_dish=24, seed=3, r_max=500e3, n_times=40, n_chan=2, bw_hz=15e6)
    cat = synth.Catalog(*(np.array(x, dtype=np.float64) for x in
                          ([150.0], [3.4], [0.2], [-0.1], [2e-5], [-1e-5], [0.0])))
    obs = synth.make_observation(cfg, cat, noise=True)
    vs = T.oracle_sky_model(obs, cat)
    H = N.hip()
    facet = H.facet_size(cat.scale[0], obs.len)
    assert facet > 1024
    d = P.DeviceObservation(obs, vis_sky=vs)
    try:
        gv, gs, guu, gvv, gsw = d.extract(cat.l[0], cat.m[0], cat.flux[0], cat.scale[0], facet,
                                          slot=0)
        d.set_prior(P.rprior_for(cat.flux[0], P.r_grid()), slot=0)
        params, keep = T.oracle_params(gv, gs, guu, gvv, cat.flux[0])
        e = np.array([[0.0, 0.0], [0.2, -0.1], [-0.35, 0.25], [0.5, 0.1]])
        print(f"wide facet {facet}: {len(gv)} cells", flush=True)
        from oracle import oracle as O
        import ctypes as C
        for cells in (0, 256):
            out = {}
            try:
                N.check(H.rl2_tune_set(N.RL2_TUNE_TILE_CELLS, cells), "rl2_tune_set")
                for fuse in (1, 0):
                    N.check(H.rl2_tune_set(N.RL2_TUNE_FUSE_STEP, fuse), "rl2_tune_set")
                    out[fuse] = d.loglik(e[:, 0], e[:, 1], slot=0, detail=True)
            finally:
                N.check(H.rl2_tune_set(N.RL2_TUNE_FUSE_STEP, 1), "rl2_tune_set")
                N.check(H.rl2_tune_set(N.RL2_TUNE_TILE_CELLS, 0), "rl2_tune_set")
            for a, b in zip(out[1], out[0]):
                np.testing.assert_array_equal(a, b)
            Le, err, Lr, nit = out[1]
            for k in range(len(e)):
                oLr = np.zeros(40)
                onit = np.zeros(40, np.int32)
                oerr = C.c_int()
                oLe = O.lib().orl2_loglikelihood_detail(C.byref(params), e[k, 0], e[k, 1],
                                                        C.byref(oerr), O.ptr(oLr),
                                                        onit.ctypes.data_as(O.c_int_p))
                assert err[k] == oerr.value
                np.testing.assert_array_equal(nit[k, 1:24], onit[1:24])
                np.testing.assert_allclose(Lr[k, 1:24], oLr[1:24], rtol=0, atol=1e-8)
                assert abs(Le[k] - oLe) <= 1e-8, (cells, k, Le[k], oLe)
    finally:
        d.close()


@pytest.mark.parametrize("batch", [1, 4])
def test_fused_subtraction_matches_separate_passes(gpu, case, batch):
    """A subtraction made in the same pass as the next extraction (RL2_TUNE_FUSE_SUBTRACT)
    gives the results and residual of the separate subtract and extract passes; only the
    gridding's atomic summation order differs between runs."""
    obs, cat, vs = case
    H = N.hip()
    out = {}
    try:
        for fuse in (1, 0):
            N.check(H.rl2_tune_set(N.RL2_TUNE_FUSE_SUBTRACT, fuse), "rl2_tune_set")
            print(f"process batch {batch} fused {fuse}", flush=True)
            d = P.DeviceObservation(obs, vis_sky=vs)
            res, st = d.process(cat, batch=batch)
            out[fuse] = (res, st, d.residual())
            d.close()
    finally:
        N.check(H.rl2_tune_set(N.RL2_TUNE_FUSE_SUBTRACT, 1), "rl2_tune_set")
    (ra, sa, va), (rb, sb, vb) = out[1], out[0]
    assert (sa.n_bad_first, sa.n_bad_final) == (sb.n_bad_first, sb.n_bad_final)
    for g in range(len(cat)):
        a, b = ra[g], rb[g]
        assert (a.error, a.pass_, a.subtracted) == (b.error, b.pass_, b.subtracted), g
        assert abs(a.e1 - b.e1) <= 1e-9 and abs(a.e2 - b.e2) <= 1e-9, (g, a.e1, b.e1)
    assert np.abs(va - vb).max() <= 1e-11 * np.abs(vb).max()


def test_mixed_likelihood_tracks_fp64(dev, case):
    """FP32 model/cross-correlation terms with FP64 phase reduction and sums: the
    log-likelihood stays within a few parts in 1e7 of the FP64 value (relative to its size)."""
    obs, cat, vs = case
    g = 0
    facet, fv, fs, uu, vv, sw = T.oracle_extract(obs, cat, g, obs.vis, vs)
    dev.load_facet(fv, fs, uu, vv, slot=2)
    dev.set_prior(P.rprior_for(cat.flux[g], P.r_grid()), slot=2)
    rng = np.random.default_rng(7)
    e = rng.uniform(-0.6, 0.6, size=(40, 2))
    e = e[(e ** 2).sum(1) < 0.65]
    L64, err64, Lr64, nit64 = dev.loglik(e[:, 0], e[:, 1], slot=2, detail=True)
    dev.loglik_precision(1)
    try:
        L32, err32, Lr32, nit32 = dev.loglik(e[:, 0], e[:, 1], slot=2, detail=True)
    finally:
        dev.loglik_precision(0)
    assert (err32 == err64).all()
    ok = err64 == 0
    scale = np.maximum(1.0, np.abs(L64[ok]))
    assert (np.abs(L32[ok] - L64[ok]) / scale).max() < 5e-6
    valid = Lr64[:, 1:] > -1e9
    assert (np.abs(Lr32[:, 1:][valid] - Lr64[:, 1:][valid])
            / np.maximum(1, np.abs(Lr64[:, 1:][valid]))).max() < 5e-6


def test_gridding_is_deterministic(dev, case):
    """Integer accumulation with one fixed-point scale for every gridding path: the same
    extraction twice, and with other variants (5: runs of 2 rows kept across channels, the
    default; 3 and 4: runs of 4 rows flushed per channel, 1024 / 2048 slots), gives
    bit-identical facets; the whole loop run twice gives bit-identical results and
    residual."""
    obs, cat, vs = case
    H = N.hip()
    g = 4
    facet = H.facet_size(cat.scale[g], obs.len)
    d0 = dev.extract(cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], facet, slot=2)
    d1 = dev.extract(cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], facet, slot=2)
    try:
        N.check(H.rl2_tune_set(N.RL2_TUNE_GRID_VARIANT, 3), "rl2_tune_set")
        a = dev.extract(cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], facet, slot=2)
        b = dev.extract(cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], facet, slot=2)
        N.check(H.rl2_tune_set(N.RL2_TUNE_GRID_VARIANT, 4), "rl2_tune_set")
        c = dev.extract(cat.l[g], cat.m[g], cat.flux[g], cat.scale[g], facet, slot=2)
    finally:
        N.check(H.rl2_tune_set(N.RL2_TUNE_GRID_VARIANT, 5), "rl2_tune_set")
    for x, y, z, p, q in zip(d0, d1, a, b, c):
        np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(x, z)
        np.testing.assert_array_equal(x, p)
        np.testing.assert_array_equal(x, q)
    runs = []
    for k in range(2):
        d = P.DeviceObservation(obs, vis_sky=vs)
        res, st = d.process(cat, batch=1)
        runs.append((P.results_array(res), d.residual()))
        d.close()
    (r0, v0), (r1, v1) = runs
    for k in ("e1", "e2", "var_e1", "var_e2", "oneDimvar", "maxL", "error", "pass_"):
        np.testing.assert_array_equal(r0[k], r1[k])
    np.testing.assert_array_equal(v0, v1)


@pytest.mark.parametrize("prec,batch", [(0, 4), (0, 16), (1, 4)])
def test_batched_extraction_matches_per_galaxy_passes(gpu, case, prec, batch):
    """A batch extracted in one persistent pass over the residual, which also applies the
    previous batch's subtractions (RL2_TUNE_BATCH_EXTRACT 1), gives the results and residual
    of one extraction and one subtraction pass per galaxy (0), bit for bit, for complex128
    and complex64 residuals; so does replaying the batch's launches as a hipGraph
    (RL2_TUNE_BATCH_GRAPH 1, one replay per batch, the fits' own launches in between)."""
    obs, cat, vs = case
    H = N.hip()
    out = {}
    try:
        for be, graph in ((1, 1), (1, 0), (0, 0)):
            N.check(H.rl2_tune_set(N.RL2_TUNE_BATCH_EXTRACT, be), "rl2_tune_set")
            N.check(H.rl2_tune_set(N.RL2_TUNE_BATCH_GRAPH, graph), "rl2_tune_set")
            print(f"process prec {prec} batch {batch} batched extraction {be} graph {graph}",
                  flush=True)
            d = P.DeviceObservation(obs, precision=prec, vis_sky=vs)
            res, st = d.process(cat, batch=batch)
            out[(be, graph)] = (P.results_array(res), d.residual())
            d.close()
    finally:
        N.check(H.rl2_tune_set(N.RL2_TUNE_BATCH_EXTRACT, 1), "rl2_tune_set")
        N.check(H.rl2_tune_set(N.RL2_TUNE_BATCH_GRAPH, 1), "rl2_tune_set")
    ref_r, ref_v = out[(0, 0)]
    for key in ((1, 1), (1, 0)):
        r, v = out[key]
        for k in ("e1", "e2", "var_e1", "var_e2", "oneDimvar", "maxL", "error", "pass_",
                  "subtracted", "ncoords"):
            np.testing.assert_array_equal(r[k], ref_r[k])
        np.testing.assert_array_equal(v, ref_v)


@pytest.mark.parametrize("batch", [1, 8])
def test_autotune_keeps_residual_and_results(gpu, case, batch):
    """rl2_autotune times every candidate of each knob, sets the fastest and leaves the
    residual as it was; the loop with the tuned knobs gives the default knobs' results up to
    rounding."""
    obs, cat, vs = case
    H = N.hip()
    d = P.DeviceObservation(obs, vis_sky=vs)
    try:
        r0, st0 = d.process(cat, batch=batch)
        ref = P.results_array(r0)
        d.close()
        d = P.DeviceObservation(obs, vis_sky=vs)
        before = d.residual()
        out = P.autotune(d, T.subset(cat, 0, 8), batch=batch)
        print(f"autotune batch {batch}: {out['chosen']}, {len(out['trials'])} trials", flush=True)
        assert len(out["trials"]) >= 12
        assert all(t["seconds"] > 0 for t in out["trials"])
        np.testing.assert_array_equal(d.residual(), before)
        r1, st1 = d.process(cat, batch=batch, lik_block=out["chosen"]["lik_block"])
        got = P.results_array(r1)
        np.testing.assert_array_equal(got["error"], ref["error"])
        assert np.abs(got["e1"] - ref["e1"]).max() <= 1e-6
        assert np.abs(got["e2"] - ref["e2"]).max() <= 1e-6
    finally:
        d.close()
        for k, v in ((N.RL2_TUNE_GRID_VARIANT, 5), (N.RL2_TUNE_TILE_CELLS, 0),
                     (N.RL2_TUNE_TILE_STREAMS, 1), (N.RL2_TUNE_BATCH_GROUPS, 0),
                     (N.RL2_TUNE_ITEM_GROUPS, 0)):
            N.check(H.rl2_tune_set(k, v), "rl2_tune_set")
