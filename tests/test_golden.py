"""Committed fixtures of the reference's own evaluate_uv_grid.c (tests/golden/uvgrid.npz,
written by tests/golden/make_golden.py from the reference sources compiled in place).

CPU: the oracle reproduces them bit for bit.  GPU: librl2hip's drop-in gridding,
facet coordinates, uv-grid size and facet averaging reproduce them (cell counts, cell
order and coordinates bit-exact; gridded sums to 1e-12 relative, since device atomics
add in another order).
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from radiolensfit2_amd import _native as N

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uvgrid.npz"))
CASES = [int(c) for c in GOLD["cases"]]


def _case(c):
    p = f"c{c}_"
    return {k[len(p):]: GOLD[k] for k in GOLD.files if k.startswith(p)}


def _grid(lib, g, fn="gridding_visibilities"):
    """Grid a fixture case with the oracle (fn orl2_gridding) or librl2hip's drop-in."""
    facet = int(g["facet"])
    ncell = facet * facet
    nv, ns, sw = np.zeros(ncell, np.complex128), np.zeros(ncell), np.zeros(ncell)
    vis = np.ascontiguousarray(g["vis"])
    s2 = np.ascontiguousarray(g["s2"])
    flag = np.ascontiguousarray(g["flag"])
    cp = O.cptr if fn.startswith("orl2") else N.cplx_s
    getattr(lib, fn)(O.ptr(g["wn"]), len(g["wn"]), len(g["u"]), O.ptr(g["u"]), O.ptr(g["v"]),
                     cp(vis), O.ptr(s2, O.c_float_p), float(g["len"]), facet, cp(nv),
                     O.ptr(ns), O.ptr(flag, O.c_bool_p), O.ptr(sw))
    return nv, ns, sw


@pytest.mark.parametrize("c", CASES)
def test_oracle_reproduces_reference_gridding(c):
    g = _case(c)
    L = O.lib()
    nv, ns, sw = _grid(L, g, "orl2_gridding")
    n = int(g["n"])
    assert np.array_equal(sw, g["sum_w"])
    assert np.array_equal(nv[:n], g["grid_vis"]) and np.array_equal(ns[:n], g["grid_s2"])
    facet = int(g["facet"])
    gu, gv = np.zeros(facet * facet), np.zeros(facet * facet)
    assert L.orl2_facet_coords(O.ptr(gu), O.ptr(gv), float(g["len"]), facet, O.ptr(sw)) == n
    assert np.array_equal(gu[:n], g["uu"]) and np.array_equal(gv[:n], g["vv"])
    size = L.orl2_uv_grid_size(float(g["len"]), O.ptr(g["wn"]), len(g["wn"]), len(g["u"]),
                               O.ptr(g["u"]), O.ptr(g["v"]), facet,
                               O.ptr(np.ascontiguousarray(g["flag"]), O.c_bool_p))
    assert size == int(g["grid_size"])
    fs = [L.orl2_facet_size(t, float(g["len"])) for t in (0.3, 1.0, 2.0, 3.5)]
    assert fs == list(g["facet_sizes"])


def test_oracle_reproduces_reference_average_facets():
    v, s = GOLD["avg_vis_in"].copy(), GOLD["avg_s2_in"].copy()
    O.lib().orl2_average_facets(len(v), O.cptr(v), O.ptr(s), O.ptr(GOLD["avg_w"]))
    assert np.array_equal(v, GOLD["avg_vis"]) and np.array_equal(s, GOLD["avg_s2"])


@pytest.mark.gpu
@pytest.mark.parametrize("c", CASES)
def test_gpu_gridding_matches_reference_fixtures(gpu, c):
    H = N.hip()
    g = _case(c)
    print(f"golden case {c}: facet {int(g['facet'])}, rows {len(g['u'])}", flush=True)
    nv, ns, sw = _grid(H, g)
    n = int(g["n"])
    np.testing.assert_array_equal(sw, g["sum_w"])
    np.testing.assert_allclose(nv[:n], g["grid_vis"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(ns[:n], g["grid_s2"], rtol=1e-12, atol=0)
    facet = int(g["facet"])
    gu, gv = np.zeros(facet * facet), np.zeros(facet * facet)
    assert H.evaluate_facet_coords(O.ptr(gu), O.ptr(gv), float(g["len"]), facet, O.ptr(sw)) == n
    np.testing.assert_array_equal(gu[:n], g["uu"])
    np.testing.assert_array_equal(gv[:n], g["vv"])
    size = H.evaluate_uv_grid_size(0, 1, float(g["len"]), O.ptr(g["wn"]), len(g["wn"]),
                                   len(g["u"]), O.ptr(g["u"]), O.ptr(g["v"]), facet,
                                   O.ptr(np.ascontiguousarray(g["flag"]), O.c_bool_p))
    assert size == int(g["grid_size"])
    assert [H.facet_size(t, float(g["len"])) for t in (0.3, 1.0, 2.0, 3.5)] == \
        list(g["facet_sizes"])


@pytest.mark.gpu
def test_gpu_average_facets_matches_reference_fixture(gpu):
    H = N.hip()
    v, s = GOLD["avg_vis_in"].copy(), GOLD["avg_s2_in"].copy()
    H.average_facets(len(v), N.cplx_s(v), O.ptr(s), O.ptr(GOLD["avg_w"]))
    np.testing.assert_array_equal(v, GOLD["avg_vis"])
    np.testing.assert_array_equal(s, GOLD["avg_s2"])
