"""The C-ABI libraries load and export every declared symbol; generators and the RLMS
container (no GPU compute)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from radiolensfit2_amd import _native as N
from radiolensfit2_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = "\n".join(ln for ln in text.splitlines() if not ln.lstrip().startswith("#"))
    names = set()
    for stmt in text.split(";"):
        s = " ".join(stmt.split())
        if not s or s.startswith(("typedef", "#")) or "{" in s or "}" in s:
            s = s.split("}")[-1].strip()
        m = re.match(r"^(?:extern\s+\"C\"\s*)?(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", s)
        if m and not s.startswith("typedef"):
            names.add(m.group(1))
    return names


@pytest.mark.parametrize("header,lib", [("rl2.h", "hip"), ("rl2_reference.h", "hip"),
                                        ("rl2_sim.h", "sim")])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) >= 10
    L = N.hip() if lib == "hip" else N.sim()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_reference_surface_is_complete():
    """Every function of the reference's per-galaxy headers (FACET build) is declared."""
    want = {  # galaxy_visibilities.h:31-55, evaluate_uv_grid.h:29-33, source_extraction.h:31,
        # likelihood.h:32-54, marginalise_r.h:24-27, galaxy_fitting.h:30, data_processing.h:32
        "gaussian_shape", "sersic_shape", "model_galaxy_visibilities_at_zero",
        "data_galaxy_visibilities", "data_visibilities_phase_shift", "facet_size",
        "evaluate_uv_grid_size", "evaluate_facet_coords", "gridding_visibilities",
        "average_facets", "source_extraction", "f_likelihood", "loglikelihood",
        "loglikelihood_r", "cross_correlation", "marginalise_over_position_shift",
        "likelihood_sampling", "set_posterior_values", "marginalise_posterior_r", "marf",
        "source_fitting", "data_processing", "sky_model", "scale_mean", "rfunc"}
    assert want <= _declared("rl2_reference.h")


def test_no_device_calls_fail_cleanly_without_gpu():
    n = C.c_int(-1)
    assert N.hip().rl2_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        z = np.zeros(4)
        st = N.hip().rl2_obs_create(C.byref(h), 1, 4, N.dptr(z), N.dptr(z), N.dptr(z),
                                    N.dptr(np.ones(1)), N.dptr(np.ones(1)),
                                    N.fptr(np.ones(4, np.float32)), None, 1.0, 0)
        assert st == -1 and b"no HIP device" in N.hip().rl2_last_error()


def test_mt19937_known_answers():
    S = N.sim()
    g = S.rl2sim_mt_new(5489)
    for _ in range(9999):
        S.rl2sim_mt_get(g)
    assert S.rl2sim_mt_get(g) == 4123659995  # C++11 [rand.predef] 10000th output
    S.rl2sim_mt_free(g)
    g = S.rl2sim_mt_new(1)
    bg = np.random.MT19937()
    bg._legacy_seeding(1)
    assert [S.rl2sim_mt_get(g) for _ in range(1000)] == list(bg.random_raw(1000))
    assert S.rl2sim_mt_uniform(g) == bg.random_raw(1)[0] / 2 ** 32  # gsl_rng_uniform
    S.rl2sim_mt_free(g)
    g0, g1 = S.rl2sim_mt_new(0), S.rl2sim_mt_new(4357)  # GSL maps seed 0 to 4357
    assert [S.rl2sim_mt_get(g0) for _ in range(5)] == [S.rl2sim_mt_get(g1) for _ in range(5)]


def test_catalog_properties():
    cat = synth.make_catalog(200, seed=3, fov_arcmin=20.0, fmin=10.0, fmax=200.0)
    assert len(cat) == 200
    f = cat.flux.reshape(20, 10)
    assert (f == f[:, :1]).all() and (np.diff(f[:, 0]) <= 0).all()  # groups of 10, descending
    assert (cat.flux >= 10).all() and (cat.flux <= 200).all()
    s = cat.scale.reshape(20, 10)
    assert (s == s[:, :1]).all() and (cat.scale >= 0.3).all() and (cat.scale <= 3.5).all()
    e1, e2 = cat.e1.reshape(100, 2), cat.e2.reshape(100, 2)  # +/- pairs (no shear applied)
    np.testing.assert_allclose(e1[:, 0], -e1[:, 1], atol=1.5e-6)
    np.testing.assert_allclose(e2[:, 0], -e2[:, 1], atol=1.5e-6)
    mod = np.hypot(cat.e1, cat.e2).reshape(20, 10)
    np.testing.assert_allclose(mod, np.broadcast_to(mod[:, :1], mod.shape), atol=3e-6)  # one modulus per group
    assert mod.max() < 0.804
    r = np.hypot(cat.l, cat.m)
    assert r.max() <= 10 * 60 * math.pi / 648000 + 1e-6  # inside the 20' disk
    assert np.allclose(cat.l * 1e6, np.round(cat.l * 1e6), rtol=0, atol=1e-6)  # %f rounding


def test_shear_closed_form():
    e1 = np.array([0.0, 0.3, -0.2])
    e2 = np.array([0.0, 0.1, 0.4])
    a, b = e1.copy(), e2.copy()
    N.sim().rl2sim_shear(3, 0.02, -0.01, N.dptr(a), N.dptr(b))
    z = (e1 + 1j * e2 + (0.02 - 0.01j)) / (1 + np.conj(0.02 - 0.01j) * (e1 + 1j * e2))
    np.testing.assert_allclose(a, z.real, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(b, z.imag, rtol=1e-14, atol=1e-16)


def test_uvw_geometry():
    cfg = synth.small_array(n_dish=10, n_times=5)
    u, v, w, wn, spec, s2, flag, L = synth.make_geometry(cfg)
    nbl = 45
    assert len(u) == nbl * 5
    # baseline length is invariant under earth rotation
    b = np.sqrt(u * u + v * v + w * w).reshape(5, nbl)
    np.testing.assert_allclose(b, np.broadcast_to(b[:1], b.shape), rtol=1e-12)
    assert L == math.ceil(max(abs(u).max(), abs(v).max())) * wn[-1] / (2 * math.pi)
    # sigma^2 = SEFD^2 / (2 t dnu eta^2) as float (RadioLensfit2.cpp:255)
    assert s2[0] == np.float32(400e6 ** 2 / (2 * 360 * 60e6 * 0.81))


def test_simulation_is_linear_and_noise_has_sigma():
    cfg = synth.small_array(n_dish=8, n_times=10)
    cat = synth.make_catalog(20, seed=2)
    a = synth.make_observation(cfg, synth.Catalog(*(getattr(cat, f)[:10].copy() for f in (
        "flux", "scale", "e1", "e2", "l", "m", "snr"))), noise=False, ms_roundtrip=False)
    b = synth.make_observation(cfg, synth.Catalog(*(getattr(cat, f)[10:].copy() for f in (
        "flux", "scale", "e1", "e2", "l", "m", "snr"))), noise=False, ms_roundtrip=False)
    ab = synth.make_observation(cfg, cat, noise=False, ms_roundtrip=False)
    np.testing.assert_allclose(ab.vis, a.vis + b.vis, rtol=1e-12, atol=1e-9)
    cfg2 = synth.small_array(n_dish=30, n_times=40)
    n = synth.make_observation(cfg2, None, noise=True, ms_roundtrip=False)
    sd = math.sqrt(n.sigma2[0])
    assert np.std(n.vis.real) == pytest.approx(sd, rel=0.03)
    assert np.std(n.vis.imag) == pytest.approx(sd, rel=0.03)


def test_rlms_roundtrip(tmp_path):
    cfg = synth.small_array(n_dish=6, n_times=4, n_chan=3)
    cat = synth.make_catalog(10, seed=5)
    obs = synth.make_observation(cfg, cat, ms_roundtrip=False)
    nrow, nch = obs.nrow, obs.nch
    vis_rc = obs.vis.reshape(nch, nrow).T  # row-major (row, channel) like the MS DATA column
    data = np.empty((nrow, nch, 2), np.float32)
    data[..., 0] = (vis_rc.real * 1e-6).astype(np.float32)
    data[..., 1] = (vis_rc.imag * 1e-6).astype(np.float32)
    flag_rc = np.zeros((nrow, nch), np.uint8)
    flag_rc[::7, 1] = 1
    sig = np.full(nrow, 2.5, np.float32)
    p = str(tmp_path / "obs.rlms").encode()
    S = N.sim()
    assert S.rl2ms_write(p, 6, nch, nrow, 950e6, 60e6, 360.0, N.dptr(obs.u), N.dptr(obs.v),
                         N.dptr(obs.w), flag_rc.ctypes.data_as(N.c_ubyte_p),
                         data.ctypes.data_as(N.c_float_p), N.fptr(sig)) == 0
    u, v, w = np.zeros(nrow), np.zeros(nrow), np.zeros(nrow)
    fl = np.zeros(nrow * nch, np.uint8)
    vis = np.zeros(nrow * nch, np.complex128)
    s2 = np.zeros(nrow * nch, np.float32)
    st = C.c_int()
    umax = S.rl2ms_read(p, N.dptr(u), N.dptr(v), N.dptr(w), fl.ctypes.data_as(N.c_ubyte_p),
                        N.cplx(vis), N.fptr(s2), C.byref(st))
    assert st.value == 0 and umax == math.ceil(max(abs(obs.u).max(), abs(obs.v).max()))
    np.testing.assert_array_equal(u, obs.u)
    np.testing.assert_array_equal(fl.reshape(nch, nrow).T, flag_rc)
    want = (data[..., 0].astype(np.float64) * 1e6 + 1j * data[..., 1].astype(np.float64) * 1e6)
    np.testing.assert_array_equal(vis.reshape(nch, nrow).T, want)
    assert (s2[:nrow] == np.float32(6.25)).all() and (s2[nrow:] == 0).all()


def test_kernel_sincos_in_revolutions_matches_mpmath():
    """sin/cos(2 pi t) as the FP64 kernels evaluate them, against 50-digit mpmath, for
    phases up to 1e4 revolutions; the error budget is a few ulp of 1."""
    import mpmath
    H = N.hip()
    rng = np.random.default_rng(21)
    t = np.concatenate([rng.uniform(-1e4, 1e4, 300), rng.uniform(-1, 1, 300),
                        np.array([0.0, 0.125, 0.25, 0.375, 0.5, -0.125, 1e-9, 123456.875])])
    s, c = np.zeros_like(t), np.zeros_like(t)
    H.rl2_test_sincos_rev64(len(t), N.dptr(t), N.dptr(s), N.dptr(c))
    mpmath.mp.dps = 50
    es = max(abs(float(mpmath.sin(2 * mpmath.pi * mpmath.mpf(float(x)))) - y) for x, y in zip(t, s))
    ec = max(abs(float(mpmath.cos(2 * mpmath.pi * mpmath.mpf(float(x)))) - y) for x, y in zip(t, c))
    assert es < 1e-15 and ec < 1e-15, (es, ec)


def test_kernel_cell_index_equals_reference_division():
    """The reciprocal-product cell index equals the reference's division everywhere,
    including points exactly on cell edges and one ulp either side of them."""
    H = N.hip()
    facet, len_l, inv_lambda = 874, 9.7e5, 3.33
    inc = 2 * len_l / facet
    rng = np.random.default_rng(22)
    u = rng.uniform(-len_l, len_l, 20000) / inv_lambda
    v = rng.uniform(-len_l, len_l, 20000) / inv_lambda
    # edges: (u inv_lambda + len) = k inc, and its neighbours
    k = rng.integers(0, facet + 1, 3000)
    edge = (k * inc - len_l) / inv_lambda
    ue = np.concatenate([edge, np.nextafter(edge, np.inf), np.nextafter(edge, -np.inf)])
    u = np.concatenate([u, ue])
    v = np.concatenate([v, ue[::-1]])
    fast = np.zeros(len(u), np.uint32)
    exact = np.zeros(len(u), np.uint32)
    bad = H.rl2_test_cell_index(len(u), N.dptr(u), N.dptr(v), inv_lambda, len_l, inc, facet,
                                fast.ctypes.data_as(N.c_uint_p), exact.ctypes.data_as(N.c_uint_p))
    assert bad == 0
    assert (fast == exact).all()


def test_tune_knobs_validate_and_round_trip():
    """rl2_tune_set accepts exactly the documented values and rl2_tune_get reads them back;
    host-only, so it runs without a device."""
    H = N.hip()
    v = C.c_int()
    cases = [(N.RL2_TUNE_GRID_VARIANT, [0, 1, 2, 3, 4, 5, 6], [-1, 7]),
             (N.RL2_TUNE_TILE_CELLS, [0, 256, 2048, 4096, 65536], [128, 3000, 131072, -2048]),
             (N.RL2_TUNE_PHASE_TABLES, [0, 1, 2], [3, -1]),
             (N.RL2_TUNE_FUSE_SUBTRACT, [0, 1], [2, -1]),
             (N.RL2_TUNE_TILE_BLOCKS, [8, 1024, 0], [4, -8, 1001]),
             (N.RL2_TUNE_TILE_STREAMS, [2, 1], [0, 3]),
             (N.RL2_TUNE_BATCH_EXTRACT, [0, 1], [2, -1]),
             (N.RL2_TUNE_BATCH_GROUPS, [512, 0], [-1, 70000]),
             (N.RL2_TUNE_BATCH_GRAPH, [0, 1], [2, -1]),
             (N.RL2_TUNE_ITEM_GROUPS, [256, 0], [-1, 70000]),
             (N.RL2_TUNE_FUSE_STEP, [0, 2, 1], [3, -1]),
             (N.RL2_TUNE_TILE_PASSES, [1, 50, 0], [-1, 51]),
             (N.RL2_TUNE_P0_TASKS, [1, 2, 3], [0, 4])]
    try:
        for knob, good, bad in cases:
            for x in good:
                assert H.rl2_tune_set(knob, x) == 0
                assert H.rl2_tune_get(knob, C.byref(v)) == 0 and v.value == x
            for x in bad:
                assert H.rl2_tune_set(knob, x) == N.RL2_ERR_ARG
                assert v.value == good[-1]
        assert H.rl2_tune_set(99, 0) == N.RL2_ERR_ARG
    finally:  # defaults
        H.rl2_tune_set(N.RL2_TUNE_GRID_VARIANT, 5)
        H.rl2_tune_set(N.RL2_TUNE_TILE_CELLS, 0)
        H.rl2_tune_set(N.RL2_TUNE_PHASE_TABLES, 2)
        H.rl2_tune_set(N.RL2_TUNE_FUSE_SUBTRACT, 1)
        H.rl2_tune_set(N.RL2_TUNE_TILE_BLOCKS, 0)
        H.rl2_tune_set(N.RL2_TUNE_TILE_STREAMS, 1)
        H.rl2_tune_set(N.RL2_TUNE_BATCH_EXTRACT, 1)
        H.rl2_tune_set(N.RL2_TUNE_BATCH_GROUPS, 0)
        H.rl2_tune_set(N.RL2_TUNE_BATCH_GRAPH, 1)
        H.rl2_tune_set(N.RL2_TUNE_ITEM_GROUPS, 0)
        H.rl2_tune_set(N.RL2_TUNE_FUSE_STEP, 1)
        H.rl2_tune_set(N.RL2_TUNE_TILE_PASSES, 0)
        H.rl2_tune_set(N.RL2_TUNE_P0_TASKS, 3)


def _philox_py(ctr, key, rounds=10):
    """Philox4x32 from its published definition (Salmon et al., SC'11, Table 2 constants):
    each round maps (c0, c1, c2, c3) to (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1,
    lo(M0 c0)); the key advances by the Weyl constants after every round."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k0, k1 = key
    mask = 0xFFFFFFFF
    for _ in range(rounds):
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & mask, p1 & mask, ((p0 >> 32) ^ c[3] ^ k1) & mask,
             p0 & mask]
        k0, k1 = (k0 + W0) & mask, (k1 + W1) & mask
    return c


def test_noise_generator_philox_known_answers():
    """The simulated noise's Philox4x32-10 (host evaluation of the kernel's function) against
    the Random123 known-answer vectors and a restatement of the published algorithm."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    H = N.hip()
    for ctr, key, want in kat:
        c = np.array(ctr, np.uint32)
        k = np.array(key, np.uint32)
        out = np.zeros(4, np.uint32)
        H.rl2_test_philox4x32_10(1, c.ctypes.data_as(N.c_uint_p), k.ctypes.data_as(N.c_uint_p),
                                 out.ctypes.data_as(N.c_uint_p))
        assert tuple(int(x) for x in out) == want
        assert tuple(_philox_py(ctr, key)) == want
    rng = np.random.default_rng(8)
    ctr = rng.integers(0, 2 ** 32, (64, 4), dtype=np.uint64).astype(np.uint32)
    key = np.array([12345, 0], np.uint32)
    out = np.zeros((64, 4), np.uint32)
    H.rl2_test_philox4x32_10(64, ctr.ctypes.data_as(N.c_uint_p), key.ctypes.data_as(N.c_uint_p),
                             out.ctypes.data_as(N.c_uint_p))
    for i in range(64):
        assert [int(x) for x in out[i]] == _philox_py([int(x) for x in ctr[i]], (12345, 0))


def test_catalog_draws_follow_the_priors():
    """Kolmogorov-Smirnov distance of the catalog's draws (one per group of ten) from the
    priors they sample (distributions.c:35-101): the flux power law S^-1.34 on [10, 200] uJy,
    the truncated log-normal scalelength around scale_mean(flux), the ellipticity modulus
    density e_pdf on [0, 0.804]; each below the 0.1% critical value."""
    from scipy import integrate, special
    cat = synth.make_catalog(30000, seed=21, fmin=10.0, fmax=200.0)
    flux, scale = cat.flux[::10], cat.scale[::10]
    mod = np.hypot(cat.e1[::10], cat.e2[::10])
    n = len(flux)
    crit = 1.95 / math.sqrt(n)

    def ks(x, cdf):
        x = np.sort(x)
        f = cdf(x)
        i = np.arange(1, len(x) + 1)
        return max((i / len(x) - f).max(), (f - (i - 1) / len(x)).max())

    a = -0.34
    assert ks(flux, lambda s: (s ** a - 10 ** a) / (200 ** a - 10 ** a)) < crit
    # scalelength: transform each draw to its own truncated-normal probability (the mean
    # depends on the galaxy's flux), which must be uniform
    mu = -0.93 + 0.33 * np.log(flux)
    phi = lambda x: 0.5 * (1 + special.erf((x - mu) / (math.sqrt(2) * 0.3136)))  # noqa: E731
    u = (phi(np.log(scale)) - phi(np.log(0.3))) / (phi(np.log(3.5)) - phi(np.log(0.3)))
    assert ks(u, lambda x: x) < crit + 1e-6 / 0.05           # %f rounding of the catalog
    emax, e0, disp = 0.804, 0.0732, 0.2298
    pdf = lambda e: e * (1 - np.exp((e - emax) / disp)) / ((1 + e) * np.sqrt(e * e + e0 * e0))  # noqa: E731
    grid = np.linspace(0, emax, 4001)
    F = np.concatenate([[0], np.cumsum([integrate.quad(pdf, grid[k], grid[k + 1])[0]
                                        for k in range(len(grid) - 1)])])
    F /= F[-1]
    assert ks(mod, lambda x: np.interp(x, grid, F)) < crit


def test_header_constants_equal_the_bindings():
    """Every enumerator of include/rl2.h has the value radiolensfit2_amd/_native.py binds."""
    text = open(os.path.join(ROOT, "include", "rl2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    consts = {}
    for body in re.findall(r"enum\s*\{([^}]*)\}", text):
        for name, val in re.findall(r"(RL2_\w+)\s*=\s*(-?\d+)", body):
            consts[name] = int(val)
    assert len(consts) >= 29  # status codes, precisions, kernel timers, launch knobs
    for name, val in consts.items():
        if name == "RL2_K_COUNT":
            assert val == len(N.KERNEL_NAMES)
            continue
        assert getattr(N, name) == val, name
