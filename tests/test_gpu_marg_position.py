"""The position marginalisation on the device (marg_position, the function the likelihood
kernels apply to B = ho^2 / 2; through rl2_marg_position_eval) against the oracle's
orl2_marg_position (likelihood.c:376-421) on every branch: x <= Delta (-1e10), the clamp of
x - Delta < 1e-4 to Delta + 1e-4 (likelihood.c:405), Ei by its series (arguments up to 40),
by its asymptotic form (above 40, also for x only), x = 100 and just above it (the closed
form x - ln Delta - exp(-Delta), likelihood.c:399-403), and arguments spread over
(Delta, 1e4).  Tolerance: 1e-13 relative (the device's exp and log may differ from the
host's by an ulp; the Ei sums are the same operations in the same order)."""
import numpy as np
import pytest

from oracle import oracle as O
from radiolensfit2_amd import _native as N

pytestmark = pytest.mark.gpu

DELTA = 0.5 * 5.991
KBAD = -1.0e10


def test_position_marginalisation_matches_the_oracle_on_every_branch(gpu):
    rng = np.random.default_rng(8)
    x = np.concatenate([
        [-5.0, 0.0, 1.0, DELTA],                                   # no shape information
        [DELTA + 1e-9, DELTA + 1e-6, DELTA + 5e-5, DELTA + 1e-4],  # clamped
        [DELTA + 2e-4, 3.5, 10.0, 39.999, 40.0],                   # Ei series for x and x - Delta
        [40.001, 41.0, 42.9, 43.0, 43.5, 50.0, 99.9],              # asymptotic Ei
        [100.0, 100.0000001, 150.0, 1e4],                          # the closed form above 100
        np.exp(rng.uniform(np.log(DELTA + 1e-3), np.log(1e4), 200)),
    ])
    out = np.zeros_like(x)
    H = N.hip()
    assert H.rl2_marg_position_eval(len(x), N.dptr(x), N.dptr(out)) == 0, H.rl2_last_error()
    ref = np.array([O.lib().orl2_marg_position(float(v)) for v in x])
    bad = x <= DELTA
    assert np.all(out[bad] == KBAD) and np.all(ref[bad] == KBAD)
    np.testing.assert_allclose(out[~bad], ref[~bad], rtol=1e-13, atol=1e-13)
    # the clamp: every argument within 1e-4 of Delta gives the value at Delta + 1e-4
    clamp = (x > DELTA) & (x - DELTA < 1e-4)
    assert clamp.sum() == 3 and np.all(out[clamp] == out[7])
    # above 100 the closed form
    hi = x > 100.0
    np.testing.assert_allclose(out[hi], x[hi] - np.log(DELTA) - np.exp(-DELTA), rtol=1e-15)


def test_marg_position_eval_rejects_bad_arguments(gpu):
    H = N.hip()
    z = np.zeros(1)
    assert H.rl2_marg_position_eval(0, N.dptr(z), N.dptr(z)) != 0
    assert b"rl2_marg_position_eval" in H.rl2_last_error()
    assert H.rl2_marg_position_eval(1, None, N.dptr(z)) != 0
