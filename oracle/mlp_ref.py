"""Test infrastructure: numpy reference of the learned-guess MLP (forward, loss, backward,
Adam), used only by the tests to check the HIP kernels.

The network has no counterpart in the reference code.  This restatement follows the
definition the library documents: 60 -> 256 -> 256 -> 256 -> 55, GELU in the tanh form
0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) (Hendrycks & Gimpel 2016), mean-squared
error of the normalised outputs, Adam (Kingma & Ba 2015, beta 0.9/0.999, eps 1e-8).

With ``bf16=True`` the values the device holds in BF16 (normalised inputs, weights,
hidden activations, output and hidden gradients) are rounded to BF16 (round to nearest
even) at the same points, and the rest is computed in float64; with ``bf16=False``
everything is float64 (for finite-difference checks of the backward pass).
"""
from __future__ import annotations

import numpy as np

NIN, NOUT, PAD, WIDTH = 60, 55, 64, 256
SHAPES = [("W1", (PAD, WIDTH)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
          ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH, PAD)), ("b4", (PAD,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
C_GELU = np.sqrt(2.0 / np.pi)


def bf16(x) -> np.ndarray:
    """Round float32 values to BF16 (nearest, ties to even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def split(flat):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def loss_and_grads(flat, X, T, norm, use_bf16=True):
    """Loss and gradient (flat, float64) for raw features X (n, 60) and raw targets T (n, 55).
    norm = (in_mu, in_sd, out_mu, out_sd)."""
    r = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P = split(flat)
    in_mu, in_sd, out_mu, out_sd = [np.asarray(v, dtype=np.float64) for v in norm]
    n = X.shape[0]
    x = np.zeros((n, PAD))
    if use_bf16:
        xs = (X.astype(np.float32) - in_mu.astype(np.float32)) * (np.float32(1.0) / in_sd.astype(np.float32))
        x[:, :NIN] = bf16(xs)
    else:
        x[:, :NIN] = (X - in_mu) / in_sd
    W = [r(P["W1"]), r(P["W2"]), r(P["W3"]), r(P["W4"])]
    b = [P["b1"], P["b2"], P["b3"], P["b4"]]
    acts, zs = [x], []
    for l in range(3):
        z = acts[-1] @ W[l] + b[l]
        zs.append(z)
        acts.append(r(gelu(z)))
    y = acts[-1] @ W[3] + b[3]
    y[:, NOUT:] = 0.0
    t = np.zeros((n, PAD))
    t[:, :NOUT] = (T - out_mu) / out_sd
    d = y - t
    d[:, NOUT:] = 0.0
    loss = float(np.sum(d * d) / (n * NOUT))
    dyf = 2.0 * d / (n * NOUT)
    grads = {}
    dzf, dzb = dyf, r(dyf)
    for l in (3, 2, 1, 0):
        grads[f"W{l + 1}"] = acts[l].T @ dzb
        grads[f"b{l + 1}"] = dzf.sum(axis=0)
        if l == 0:
            break
        da = dzb @ W[l].T
        dzf = da * gelu_grad(zs[l - 1])
        dzb = r(dzf)
    flat_g = np.concatenate([grads[name].ravel() for name, _ in SHAPES])
    return loss, flat_g


def forward(flat, X, norm, use_bf16=True):
    """Raw outputs (n, 55)."""
    r = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P = split(flat)
    in_mu, in_sd, out_mu, out_sd = [np.asarray(v, dtype=np.float64) for v in norm]
    n = X.shape[0]
    x = np.zeros((n, PAD))
    if use_bf16:
        xs = (X.astype(np.float32) - in_mu.astype(np.float32)) * (np.float32(1.0) / in_sd.astype(np.float32))
        x[:, :NIN] = bf16(xs)
    else:
        x[:, :NIN] = (X - in_mu) / in_sd
    a = x
    for l, (w, bb) in enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3"))):
        a = r(gelu(a @ r(P[w]) + P[bb]))
    y = a @ r(P["W4"]) + P["b4"]
    return y[:, :NOUT] * out_sd + out_mu


def adam_step(flat, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v
