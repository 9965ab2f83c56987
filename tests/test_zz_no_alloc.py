"""rl2_marginalise_posterior_r_noalloc (no allocation; fixed per-device scratch) against the
allocating marginalise_posterior_r on the same inputs, two calls in a row, the second with
a smaller input, plus its argument checks.  The test's device buffers come from the HIP
runtime librl2hip.so is linked against (torch carries a runtime of its own, which cannot share
the device with it in one process)."""
import ctypes as C
import math

import numpy as np
import pytest

from radiolensfit2_amd import _native as N

pytestmark = pytest.mark.gpu


def _points(num, seed):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0.3, 3.7, num))
    y = 25.0 - 6.0 * (np.log(x) - rng.uniform(-0.3, 0.3)) ** 2 + rng.normal(0, 0.05, num)
    return x, y


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime librl2hip.so uses."""

    def __init__(self):
        N.hip()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        return p

    def get(self, p, n):
        a = np.empty(n, np.float64)
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0
        return a

    def free(self, *ps):
        for p in ps:
            self.L.hipFree(p)


def _reference(x, y):
    H = N.hip()
    yy = y.copy()
    val = H.marginalise_posterior_r(len(x), N.dptr(np.ascontiguousarray(x)), N.dptr(yy))
    return val, yy


def test_noalloc_matches_allocating_entry_point(gpu):
    H = N.hip()
    N.check(H.rl2_init(0), "rl2_init")
    D = _Hip()
    for num, seed in ((23, 1), (9, 2), (23, 3)):
        x, y = _points(num, seed)
        want, want_y = _reference(x, y)
        dx, dy = D.put(x), D.put(y)
        dout = D.put(np.full(num + 1, math.nan))
        st = H.rl2_marginalise_posterior_r_noalloc(num, dx, dy, dout, None)
        assert st == N.RL2_OK, H.rl2_last_error()
        N.check(H.rl2_synchronize(), "rl2_synchronize")
        out = D.get(dout, num + 1)
        got_y = D.get(dy, num)
        D.free(dx, dy, dout)
        print(f"num {num}: {out[0]!r} vs {want!r}", flush=True)
        assert out[0] == want
        np.testing.assert_array_equal(out[1:], want_y)
        np.testing.assert_array_equal(got_y, y)      # inputs untouched


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    H = N.hip()
    N.check(H.rl2_init(0), "rl2_init")
    D = _Hip()
    p = D.put(np.zeros(64))
    try:
        assert H.rl2_marginalise_posterior_r_noalloc(33, p, p, p, None) == N.RL2_ERR_RANGE
        assert b"fit its fixed scratch" in H.rl2_last_error()
        assert H.rl2_marginalise_posterior_r_noalloc(2, p, p, p, None) == N.RL2_ERR_RANGE
        assert H.rl2_marginalise_posterior_r_noalloc(5, None, p, p, None) == N.RL2_ERR_ARG
    finally:
        D.free(p)
