"""FP64 reference of the learned relaxation network (test infrastructure only).

The network maps 16 residual statistics of a SIRT iteration to (o0, o1):
lambda_k * rho = LAMBDA_MAX sigmoid(o0) (LAMBDA_MAX = 12) and a stop logit o1.  Layers 16 -> 64 -> 64 -> 2
with ReLU; parameters are 5378 values laid out as W1[64][16], b1[64],
W2[64][64], b2[64], W3[2][64], b3[2] (row = output unit).

The loss of n samples with targets (t0, t1) is
    mean[(LAMBDA_MAX sigmoid(o0) - t0)^2 * [t0 >= 0] + stop_w * BCE(o1, t1)].

`emulate_bf16=True` rounds to BF16 at the points where the device kernels feed
their BF16 matrix units (inputs, weights, activations and back-propagated
gradients), so device results can be compared with a tight tolerance; with
`False` this is the exact FP64 function and gradient, which the tests pin with
finite differences.  There is no reference counterpart: AIR Tools II has no
learned component.
"""
from __future__ import annotations

import numpy as np

NI, NH, NO = 16, 64, 2
NPARAM = NH * NI + NH + NH * NH + NH + NO * NH + NO
LAMBDA_MAX = 12.0


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + NH * NI].reshape(NH, NI); o += NH * NI
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NO * NH].reshape(NO, NH); o += NO * NH
    b3 = p[o:o + NO]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


def init_params(seed=0):
    """He-style initialisation (uniform), zero biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (NH, NI)) * np.sqrt(6.0 / NI)
    W2 = rng.uniform(-1, 1, (NH, NH)) * np.sqrt(6.0 / NH)
    W3 = rng.uniform(-1, 1, (NO, NH)) * np.sqrt(3.0 / NH)
    return pack(W1, np.zeros(NH), W2, np.zeros(NH), W3, np.zeros(NO)).astype(np.float32)


def forward(p, X, emulate_bf16=False):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = r(X)
    z1 = x @ r(W1).T + b1
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ r(W2).T + b2
    h2 = r(np.maximum(z2, 0))
    o = h2 @ r(W3).T + b3
    return o, (x, h1, h2)


def _sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss_grad(p, X, T, stop_w=1.0, emulate_bf16=False):
    """(mean loss, gradient of the mean loss w.r.t. the 5378 parameters)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    T = np.asarray(T, np.float64)
    n = X.shape[0]
    o, (x, h1, h2) = forward(p, X, emulate_bf16)
    s0, s1 = _sigmoid(o[:, 0]), _sigmoid(o[:, 1])
    keep = T[:, 0] >= 0
    e = LAMBDA_MAX * s0 - T[:, 0]
    z = o[:, 1]
    bce = np.maximum(z, 0) - z * T[:, 1] + np.log1p(np.exp(-np.abs(z)))
    loss = (np.sum(np.where(keep, e * e, 0.0)) + stop_w * np.sum(bce)) / n
    dO = np.zeros((n, NO))
    dO[:, 0] = np.where(keep, 2 * e * LAMBDA_MAX * s0 * (1 - s0), 0.0) / n
    dO[:, 1] = stop_w * (s1 - T[:, 1]) / n
    db3 = dO.sum(0)
    dOr = r(dO)
    dz2 = (dOr @ r(W3)) * (h2 > 0)
    db2 = dz2.sum(0)
    d2 = r(dz2)
    dW3 = dOr.T @ h2
    dz1 = (d2 @ r(W2)) * (h1 > 0)
    db1 = dz1.sum(0)
    d1 = r(dz1)
    dW2 = d2.T @ h1
    dW1 = d1.T @ x
    return loss, pack(dW1, db1, dW2, db2, dW3, db3)
