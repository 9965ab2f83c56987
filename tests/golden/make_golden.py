#!/usr/bin/env python3
"""Writes tests/golden/uvgrid.npz: inputs and outputs of the reference's own
evaluate_uv_grid.c (facet_size, evaluate_uv_grid_size, gridding_visibilities,
evaluate_facet_coords, average_facets), compiled from the reference sources by
oracle/Makefile into oracle/_ref/libref_uvgrid.so.

Run from the repository root in a tree where oracle/_ref has been built:
    python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from radiolensfit2_amd import synth  # noqa: E402

CASES = [(0, 16, 600, 2), (1, 30, 600, 2), (3, 7, 400, 3), (4, 58, 500, 1)]


def case_inputs(seed, nrow, nch):
    rng = np.random.default_rng(1000 + seed)
    len_m = 3000.0
    u = rng.uniform(-len_m, len_m, nrow)
    v = rng.uniform(-len_m, len_m, nrow)
    u[0] = len_m   # integer |u| in metres: the pu == facet edge at the top channel
    v[1] = -len_m
    wn, spec = synth.channel_constants(nch, 950e6, 60e6)
    vis = (rng.normal(size=nrow * nch) + 1j * rng.normal(size=nrow * nch)) * 100
    s2 = rng.uniform(0.5e6, 2e6, nrow * nch).astype(np.float32)
    flag = rng.uniform(size=nrow * nch) < 0.1
    umax = math.ceil(max(np.abs(u).max(), np.abs(v).max()))
    len_l = umax * wn[-1] / (2 * math.pi)
    return u, v, wn, vis, s2, flag, len_l


def main():
    R = O.ref()
    if R is None:
        sys.exit("oracle/_ref/libref_uvgrid.so is not built (needs the reference tree)")
    out = {}
    for seed, facet, nrow, nch in CASES:
        u, v, wn, vis, s2, flag, len_l = case_inputs(seed, nrow, nch)
        ncell = facet * facet
        gv = np.zeros(ncell, np.complex128)
        gs, sw = np.zeros(ncell), np.zeros(ncell)
        R.gridding_visibilities(O.ptr(wn), nch, nrow, O.ptr(u), O.ptr(v), O.cptr(vis),
                                O.ptr(s2, O.c_float_p), len_l, facet, O.cptr(gv), O.ptr(gs),
                                O.ptr(flag, O.c_bool_p), O.ptr(sw))
        gu, gvv = np.zeros(ncell), np.zeros(ncell)
        n = R.evaluate_facet_coords(O.ptr(gu), O.ptr(gvv), len_l, facet, O.ptr(sw))
        size = R.evaluate_uv_grid_size(0, 1, len_l, O.ptr(wn), nch, nrow, O.ptr(u), O.ptr(v),
                                       facet, O.ptr(flag, O.c_bool_p))
        fs = np.array([R.facet_size(t, len_l) for t in (0.3, 1.0, 2.0, 3.5)], np.int64)
        p = f"c{seed}_"
        out.update({p + "u": u, p + "v": v, p + "wn": wn, p + "vis": vis, p + "s2": s2,
                    p + "flag": flag, p + "len": np.float64(len_l), p + "facet": np.int64(facet),
                    p + "grid_vis": gv[:n], p + "grid_s2": gs[:n], p + "sum_w": sw,
                    p + "uu": gu[:n], p + "vv": gvv[:n], p + "n": np.int64(n),
                    p + "grid_size": np.int64(size), p + "facet_sizes": fs})
    # average_facets on a dense accumulator
    rng = np.random.default_rng(9)
    size = 40 * 40
    w = np.where(rng.uniform(size=size) < 0.4, rng.integers(1, 9, size), 0).astype(float)
    s2 = w * rng.uniform(1e6, 2e6, size)
    vis = (rng.normal(size=size) + 1j * rng.normal(size=size)) * w
    av, as2 = vis.copy(), s2.copy()
    R.average_facets(size, O.cptr(av), O.ptr(as2), O.ptr(w))
    out.update({"avg_w": w, "avg_s2_in": s2, "avg_vis_in": vis, "avg_vis": av, "avg_s2": as2})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "uvgrid.npz")
    np.savez_compressed(dst, cases=np.array([c[0] for c in CASES]), **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
