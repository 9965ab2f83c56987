"""Shear estimator (radiolensfit2_amd/shear.py, the reference's shear.py) on the CPU."""
import numpy as np

from radiolensfit2_amd import shear as S


def _row(e1, e2, var=1e-3, err=0.01, snr=20.0):
    return [e1, err, e2, err, var, snr]


def test_orientation_pairs_give_the_shear_exactly():
    rng = np.random.default_rng(0)
    g = (0.03, -0.01)
    rows = []
    for _ in range(50):
        e = rng.uniform(-0.4, 0.4, 2)
        var = rng.uniform(1e-4, 1e-2)
        rows.append(_row(e[0] + g[0], e[1] + g[1], var))
        rows.append(_row(-e[0] + g[0], -e[1] + g[1], var))
    me1, me2, w, snr, bad = S.select(np.array(rows))
    assert bad == 0 and len(me1) == 100
    (g1, s1), (g2, s2), sigma = S.shear(me1, me2, w, bootnum=50, seed=1)
    assert abs(g1 - g[0]) < 1e-12 and abs(g2 - g[1]) < 1e-12
    assert s1 < 1e-12 and s2 < 1e-12 and sigma > 0


def test_bad_measurements_remove_their_partner():
    good = [_row(0.1 * k, -0.1 * k) for k in range(8)]
    rows = [list(r) for r in good]
    rows[2][5] = 5.0          # even line bad (SNR < 10): its partner (line 3) is skipped
    rows[5][4] = 1e-6         # odd line bad (1D var): its partner (line 4) is overwritten
    me1, me2, w, snr, bad = S.select(np.array(rows))
    assert bad == 2
    np.testing.assert_allclose(me1, [0.0, 0.1, 0.6, 0.7])


def test_bad_last_odd_line_drops_the_entry_before_its_partner():
    rows = [_row(0.1 * k, 0.0) for k in range(6)]
    rows[5][1] = 0.0          # err1 <= 1e-3 on the last (odd) line
    me1, *_ = S.select(np.array(rows))
    # the reference deletes index k-1 after its own k -= 1, i.e. line 3, keeping line 4
    np.testing.assert_allclose(me1, [0.0, 0.1, 0.2, 0.4])


def test_shape_noise_closed_form():
    me1 = np.array([0.1, -0.1, 0.3, -0.3])
    me2 = np.zeros(4)
    assert S.shape_noise(me1, me2) == 0.0
    me2 = np.array([0.2, -0.2, 0.0, 0.0])
    v1, v2, v12 = np.mean(me1 ** 2), np.mean(me2 ** 2), np.mean(me1 * me2)
    assert abs(S.shape_noise(me1, me2) - np.sqrt(v1 * v2 - v12 ** 2)) < 1e-15


def test_file_roundtrip(tmp_path):
    class R:
        def __init__(self, e1, e2, err):
            self.e1, self.e2, self.error = e1, e2, err
            self.var_e1, self.var_e2, self.oneDimvar = 1e-4, 4e-4, 2e-4

    res = [R(0.11, -0.2, 0), R(0.3, 0.1, 1)]
    p = tmp_path / "ellipticities0.txt"
    S.write_ellipticities(p, [50.0, 60.0], [0.1, 0.3], [-0.2, 0.1], res, [20.0, 30.0],
                          [1e-4, 2e-4], [3e-4, -1e-4])
    d = S.read_ellipticities(p)
    np.testing.assert_allclose(d[0], [0.11, 0.01, -0.2, 0.02, 2e-4, 20.0], atol=1e-6)
    np.testing.assert_allclose(d[1, [0, 2]], [0.0, 0.0])  # bad fits are written as e = 0
