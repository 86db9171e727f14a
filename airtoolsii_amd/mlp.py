"""The learned relaxation network on the engine (hand-written HIP kernels).

A 16-64-64-2 ReLU perceptron maps residual statistics of a SIRT iteration to
(o0, o1): lambda_k * rho = 12 sigmoid(o0) and a stop logit o1.  Parameters are
AIR_MLP_PARAMS (5378) float32 values laid out as W1[64][16], b1[64],
W2[64][64], b2[64], W3[2][64], b3[2].  Training runs on the GPU: BF16 MFMA
forward/backward with FP32 accumulation and Adam on FP32 master weights.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Engine, check, load, ptr

NIN, NHID, NOUT, NPARAM = 16, 64, 2, 5378
LAMBDA_MAX = 12.0         # lambda_k rho = LAMBDA_MAX sigmoid(o0) (AIR_MLP_LAMBDA_MAX)


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def init_params(seed=0):
    """He-style uniform initialisation (float32); zero biases except the lambda
    output's, which starts the network at lambda rho = 1.9 (the default of
    calc_relaxpar.m:110)."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (NHID, NIN)) * np.sqrt(6.0 / NIN)
    W2 = rng.uniform(-1, 1, (NHID, NHID)) * np.sqrt(6.0 / NHID)
    W3 = rng.uniform(-1, 1, (NOUT, NHID)) * np.sqrt(3.0 / NHID)
    b3 = np.array([np.log(1.9 / (LAMBDA_MAX - 1.9)), 0.0])
    return np.concatenate([W1.ravel(), np.zeros(NHID), W2.ravel(), np.zeros(NHID), W3.ravel(),
                           b3]).astype(np.float32)


class MLP:
    """Device entry points of the network, bound to one engine handle."""

    def __init__(self, engine: Engine | None = None, device: int = 0):
        self.engine = engine or Engine(device)

    def forward(self, params, X):
        X = _f32(X)
        n = X.shape[0]
        if X.ndim != 2 or X.shape[1] != NIN:
            raise ValueError("X must be n x 16")
        out = np.empty((n, NOUT), np.float32)
        check(load().air_mlp_forward(self.engine.h, ptr(_f32(params, (NPARAM,))), ptr(X), n, ptr(out)))
        return out

    def grad(self, params, X, T, stop_weight=1.0):
        X, T = _f32(X), _f32(T)
        n = X.shape[0]
        g = np.empty(NPARAM, np.float32)
        loss = C.c_double()
        check(load().air_mlp_grad(self.engine.h, ptr(_f32(params, (NPARAM,))), ptr(X), ptr(T), n,
                                  float(stop_weight), ptr(g), C.byref(loss)))
        return loss.value, g

    def train(self, params, X, T, epochs=10, batch=4096, lr=1e-3, stop_weight=1.0, seed=0):
        """Adam training; returns (trained params, per-epoch mean loss)."""
        X, T = _f32(X), _f32(T)
        n = X.shape[0]
        p = _f32(params, (NPARAM,)).copy()
        hist = np.zeros(max(epochs, 1))
        check(load().air_mlp_train(self.engine.h, ptr(p), ptr(X), ptr(T), n, int(epochs), int(batch), float(lr),
                                   float(stop_weight), int(seed), ptr(hist)))
        return p, hist[:epochs]
