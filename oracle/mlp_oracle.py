"""CPU oracle of the learned proposal network (test infrastructure only).

Restates radiolensfit2_amd/csrc/rl2_mlp.hip in float64 numpy: the same parameter layout
(W row-major [out][in], then b, per layer), GELU (erf form), the Gaussian negative
log-likelihood with a Cholesky factor of the precision, and its backward pass.  With
`bf16=True` every matrix-product operand is rounded to bfloat16 (round to nearest even) as
the device kernel rounds them; accumulation stays float64 here (float32 on the device).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT, N_LAYERS = 770, 256, 9, 4
DIMS = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]


def n_params():
    return sum(o * i + o for i, o in DIMS)


def unpack(p):
    """[(W[out][in], b[out]) per layer] views of a flat parameter vector."""
    out, k = [], 0
    for i, o in DIMS:
        W = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def _mm(a, b, rnd):
    return (bf16(a) @ bf16(b)) if rnd else (a @ b)


def forward(p, x, bf16_ops=True):
    """Returns (output[B][9], cache)."""
    layers = unpack(p)
    acts, pres = [x], []
    h = x
    for l, (W, b) in enumerate(layers):
        z = _mm(h, W.T, bf16_ops) + b
        pres.append(z)
        h = gelu(z) if l < N_LAYERS - 1 else z
        acts.append(h)
    return h, (acts, pres)


def nll(out, y):
    """Per-sample loss and d loss / d out (mean over the batch)."""
    B = out.shape[0]
    d = y - out[:, :3]
    l00, l11, l22 = np.exp(out[:, 3]), np.exp(out[:, 4]), np.exp(out[:, 5])
    l10, l20, l21 = out[:, 6], out[:, 7], out[:, 8]
    z0 = l00 * d[:, 0] + l10 * d[:, 1] + l20 * d[:, 2]
    z1 = l11 * d[:, 1] + l21 * d[:, 2]
    z2 = l22 * d[:, 2]
    loss = 0.5 * (z0 ** 2 + z1 ** 2 + z2 ** 2) - (out[:, 3] + out[:, 4] + out[:, 5])
    g = np.zeros_like(out)
    g[:, 0] = -(l00 * z0)
    g[:, 1] = -(l10 * z0 + l11 * z1)
    g[:, 2] = -(l20 * z0 + l21 * z1 + l22 * z2)
    g[:, 3] = d[:, 0] * z0 * l00 - 1
    g[:, 4] = d[:, 1] * z1 * l11 - 1
    g[:, 5] = d[:, 2] * z2 * l22 - 1
    g[:, 6] = d[:, 1] * z0
    g[:, 7] = d[:, 2] * z0
    g[:, 8] = d[:, 2] * z1
    return loss, g / B


def gradients(p, x, y, bf16_ops=True):
    """(mean loss, flat gradient) of the mean negative log-likelihood."""
    out, (acts, pres) = forward(p, x, bf16_ops)
    loss, dpre = nll(out, y)
    layers = unpack(p)
    grads = []
    for l in range(N_LAYERS - 1, -1, -1):
        W, b = layers[l]
        dW = _mm(dpre.T, acts[l], bf16_ops)
        db = dpre.sum(0)
        grads.append((dW, db))
        if l > 0:
            dact = _mm(dpre, W, bf16_ops)
            dpre = dact * gelu_grad(pres[l - 1])
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in reversed(grads)])
    return float(loss.mean()), flat


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for l, (i, o) in enumerate(DIMS):
        sd = (0.1 if l == N_LAYERS - 1 else 1.0) / np.sqrt(i)
        parts += [rng.normal(scale=sd, size=o * i), np.zeros(o)]
    return np.concatenate(parts)
