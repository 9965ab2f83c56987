"""Learned proposal network: the CPU oracle's backward pass against finite differences, the
device forward / backward kernels against the oracle, and what the trained network gains.

CPU: the oracle (oracle/mlp_oracle.py, float64, no bf16 rounding) is differentiated
numerically at random parameter coordinates; the analytic gradient must agree to 1e-6
relative.  GPU: the device network (BF16 MFMA operands, FP32 accumulation) loaded with the
same parameters gives the oracle's outputs and gradients with bf16-rounded operands to a
relative Frobenius error below 1e-2 per layer (bf16 has an 8-bit mantissa); Adam steps on a
learnable target lower the loss; the committed trained weights cut the launches per fit on
fresh simulated galaxies.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as M


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, M.N_IN)).astype(np.float32)
    y = np.column_stack([rng.uniform(-0.5, 0.5, B), rng.uniform(-0.5, 0.5, B),
                         rng.uniform(-1, 1, B)]).astype(np.float32)
    return x, y


def test_oracle_layout():
    assert M.n_params() == 256 * 770 + 256 + 2 * (256 * 256 + 256) + 9 * 256 + 9


def test_oracle_gradient_matches_finite_differences():
    p = M.init_params(3)
    # larger output layer so every term of the loss is exercised
    W4, b4 = M.unpack(p)[3]
    W4 *= 10
    b4[:] = np.random.default_rng(1).normal(scale=0.3, size=b4.shape)
    x, y = _batch(6, 4)
    x, y = x.astype(np.float64), y.astype(np.float64)
    loss, g = M.gradients(p, x, y, bf16_ops=False)
    rng = np.random.default_rng(5)
    # coordinates from every layer's weights and biases
    idx = []
    k = 0
    for i, o in M.DIMS:
        idx += list(k + rng.integers(0, o * i, 6))
        idx += list(k + o * i + rng.integers(0, o, 2))
        k += o * i + o
    for j in idx:
        h = 1e-6 * max(1.0, abs(p[j]))
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        lp = M.nll(M.forward(pp, x, False)[0], y)[0].mean()
        lm = M.nll(M.forward(pm, x, False)[0], y)[0].mean()
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (j, fd, g[j])


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7         # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) < 3.0e-3 * 2 ** -8


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.fixture(scope="module")
def net(gpu):
    from radiolensfit2_amd import _native as N
    H = N.hip()
    h = C.c_void_p()
    N.check(H.rl2_proposal_create(C.byref(h), 7), "rl2_proposal_create")
    yield H, h
    H.rl2_proposal_destroy(h)


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(net):
    H, h = net
    p = M.init_params(11).astype(np.float32)
    W4, b4 = M.unpack(p)[3]
    W4 *= 10
    H.rl2_proposal_set_params(h, p.ctypes.data_as(C.POINTER(C.c_float)))
    x, _ = _batch(40, 12)
    out = np.zeros((40, M.N_OUT), np.float32)
    print("proposal forward, batch 40", flush=True)
    assert H.rl2_proposal_forward(h, 40, x.ctypes.data_as(C.POINTER(C.c_float)),
                                  out.ctypes.data_as(C.POINTER(C.c_float))) == 0
    ref, _ = M.forward(p.astype(np.float64), x.astype(np.float64), bf16_ops=True)
    assert _rel(out, ref) < 1e-2


@pytest.mark.gpu
def test_gpu_gradients_match_oracle(net):
    H, h = net
    p = M.init_params(13).astype(np.float32)
    W4, b4 = M.unpack(p)[3]
    W4 *= 10
    H.rl2_proposal_set_params(h, p.ctypes.data_as(C.POINTER(C.c_float)))
    x, y = _batch(64, 14)
    loss = C.c_float()
    print("proposal gradients, batch 64", flush=True)
    assert H.rl2_proposal_gradients(h, 64, x.ctypes.data_as(C.POINTER(C.c_float)),
                                    y.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)) == 0
    g = np.zeros(M.n_params(), np.float32)
    H.rl2_proposal_get_grads(h, g.ctypes.data_as(C.POINTER(C.c_float)))
    ref_loss, ref_g = M.gradients(p.astype(np.float64), x.astype(np.float64),
                                  y.astype(np.float64), bf16_ops=True)
    assert abs(loss.value - ref_loss) <= 1e-2 * max(1.0, abs(ref_loss))
    k = 0
    for l, (i, o) in enumerate(M.DIMS):  # per layer: weights, then biases
        assert _rel(g[k:k + o * i], ref_g[k:k + o * i]) < 2e-2, ("W", l)
        k += o * i
        assert _rel(g[k:k + o], ref_g[k:k + o]) < 2e-2, ("b", l)
        k += o


@pytest.mark.gpu
def test_gpu_training_reduces_loss(net):
    """A linear-in-features target is learnable: Adam steps lower the loss."""
    H, h = net
    p = M.init_params(17).astype(np.float32)
    H.rl2_proposal_set_params(h, p.ctypes.data_as(C.POINTER(C.c_float)))
    rng = np.random.default_rng(18)
    x = rng.normal(size=(256, M.N_IN)).astype(np.float32)
    A = rng.normal(scale=0.05, size=(M.N_IN, 3))
    y = np.tanh(x @ A).astype(np.float32)
    fp = C.POINTER(C.c_float)
    losses = []
    for it in range(60):
        loss = C.c_float()
        assert H.rl2_proposal_train_step(h, 256, x.ctypes.data_as(fp), y.ctypes.data_as(fp),
                                         C.c_float(2e-3), C.byref(loss)) == 0
        losses.append(loss.value)
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < losses[0] - 1.0


@pytest.mark.gpu
def test_gpu_trained_proposal_gains_over_reference_start(gpu):
    """The committed network (trained on the GPU on 161k simulated 64-dish galaxies labelled
    by exact-likelihood fits) on fresh galaxies: starting the simplex at its peak cuts the
    batched launches a fit waits for, and the proposal is taken for most galaxies (the rest
    fall back to the reference start); in speculation-only mode the results are the
    reference start's exactly."""
    import os
    from radiolensfit2_amd import proposal as PR
    from radiolensfit2_amd import synth
    w = os.path.join(os.path.dirname(PR.__file__), "proposal_c1.npy")
    net = PR.Proposal(0)
    net.load(w)
    print("proposal evaluation on 300 fresh 64-dish galaxies", flush=True)
    s = PR.evaluate(synth.ARRAY_64, net, 300, seed=4242, batch=64, log=lambda *a: None)
    net.close()
    plain, start, spec = s["plain"], s["learned_start"], s["learned_speculation"]
    print(s, flush=True)
    assert start["rounds_per_galaxy"] < 0.95 * plain["rounds_per_galaxy"]
    assert start["proposal_start_fraction"] > 0.5
    assert spec["rounds_per_galaxy"] < plain["rounds_per_galaxy"]
    # speculation only: every likelihood value is a function of its point and facet alone, so
    # the fits, and through them the shared residual, are the plain path's bit for bit
    assert spec["median_abs_de_vs_plain"] == 0.0
    assert spec["within_1e-4_of_plain"] == 1.0
    # the learned start moves the simplex end point within the simplex tolerance, which the
    # sampling grid anchored there turns into a different posterior mean: statistically the
    # same measurement, not the same numbers
    assert start["within_1e-2_of_plain"] >= 0.5
