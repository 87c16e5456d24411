"""CPU restatement of the learned surrogate's MLP (test infrastructure): forward pass, MSE loss and
back-propagation of a 22-128-128-128-16 ReLU network in float64 numpy, with an optional rounding
function applied where the device kernels round to bf16 (the inputs, the weights, each hidden
activation after ReLU, the output gradient and each back-propagated hidden gradient). With
`rnd=None` it is the exact float64 network that the finite-difference test pins; with
`rnd=bf16` it states what the matrix-core kernels compute (FP32 accumulation aside).

Parameter order (the library's): W1 [128 x 22], b1, W2 [128 x 128], b2, W3, b3, W4 [16 x 128], b4.
"""
from __future__ import annotations

import numpy as np

SHAPES = [("W1", (128, 22)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("W3", (128, 128)), ("b3", (128,)),
          ("W4", (16, 128)), ("b4", (16,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def _r(rnd, x):
    return x if rnd is None else rnd(x)


def forward(p, X, rnd=None):
    """Returns (Y, cache)."""
    w = unpack(p)
    W = {k: _r(rnd, w[k]) for k in ("W1", "W2", "W3", "W4")}
    X0 = _r(rnd, np.asarray(X, np.float64))
    H1 = _r(rnd, np.maximum(X0 @ W["W1"].T + w["b1"], 0))
    H2 = _r(rnd, np.maximum(H1 @ W["W2"].T + w["b2"], 0))
    H3 = _r(rnd, np.maximum(H2 @ W["W3"].T + w["b3"], 0))
    Y = H3 @ W["W4"].T + w["b4"]
    return Y, dict(W=W, X0=X0, H1=H1, H2=H2, H3=H3)


def loss_and_grad(p, X, T, rnd=None):
    """Mean-squared error over all outputs and its gradient (packed in the parameter order)."""
    Y, c = forward(p, X, rnd)
    n = Y.shape[0]
    D = Y - np.asarray(T, np.float64)
    loss = float((D * D).mean())
    dY = _r(rnd, 2.0 * D / (n * Y.shape[1]))
    W = c["W"]
    G3 = _r(rnd, (dY @ W["W4"]) * (c["H3"] > 0))
    G2 = _r(rnd, (G3 @ W["W3"]) * (c["H2"] > 0))
    G1 = _r(rnd, (G2 @ W["W2"]) * (c["H1"] > 0))
    g = dict(W4=dY.T @ c["H3"], b4=dY.sum(0), W3=G3.T @ c["H2"], b3=G3.sum(0), W2=G2.T @ c["H1"], b2=G2.sum(0),
             W1=G1.T @ c["X0"], b1=G1.sum(0))
    return loss, pack(g)


def he_init(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            b = np.sqrt(6.0 / shape[1])
            d[name] = rng.uniform(-b, b, shape)
        else:
            d[name] = rng.uniform(-0.1, 0.1, shape)
    return pack(d)
